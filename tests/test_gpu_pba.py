"""Discrete Voronoi texture (PBA on gfx950) against brute-force nearest sites: the distance
field must be exact at every pixel (integer squared distances); labels may differ only on ties."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,T,seed", [(1000, 512, 1), (300, 256, 5), (5000, 512, 6)])
def test_distance_field_exact(oracle, n, T, seed):
    import galpha2d_amd as G
    print("case: voronoi", n, T, seed, flush=True)
    ctx = G.Context(0)
    xy = oracle.gen_uniform(seed, n)
    labels, pix, nf = ctx.voronoi(xy, T)
    px = (pix >> 16).astype(np.int64)
    py = (pix & 0xffff).astype(np.int64)
    sites = np.unique(labels)
    X, Y = np.meshgrid(np.arange(T), np.arange(T), indexing="ij")
    best = np.full((T, T), 1 << 60, dtype=np.int64)
    for s in sites:
        np.minimum(best, (X - px[s]) ** 2 + (Y - py[s]) ** 2, out=best)
    lab_d = (X - px[labels]) ** 2 + (Y - py[labels]) ** 2
    assert int((lab_d != best).sum()) == 0
    # every site in the map labels its own pixel; frame sites own the texture border
    assert (labels[px[sites], py[sites]] == sites).all()
    border = np.concatenate([labels[0, :], labels[-1, :], labels[:, 0], labels[:, -1]])
    assert (border >= n).all()
    ctx.close()
