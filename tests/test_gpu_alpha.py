"""Alpha-shape stage on the GPU against the oracle: spectrum values bit-identical (0 ulp),
ranks, intervals, master list and spec pointers identical."""
import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu

FIELDS = ("spectrum", "tri_ro_rank", "edge_t", "edge_vopp", "edge_ro_rank", "edge_nu_min", "edge_nu_max",
          "vert_nu_min", "vert_nu_max", "master_simplex", "master_st", "master_at", "spec_pointers")


def _alpha_gpu(xy, T):
    import torch
    import galpha2d_amd as G
    ctx = G.Context(0)
    d = torch.from_numpy(np.ascontiguousarray(xy)).to("cuda:0")
    torch.cuda.synchronize()
    ctx.compute_device(d.data_ptr(), xy.shape[0], T, with_alpha=True,
                       stream=torch.cuda.current_stream().cuda_stream)
    tris, a, st = ctx.triangles(), ctx.alpha(), ctx.stats()
    ctx.close()
    return tris, a, st


def _check(oracle, xy, T):
    tris, a, st = _alpha_gpu(xy, T)
    ref, _ = oracle.delaunay(xy)
    assert np.array_equal(tris, ref)
    ra = oracle.alpha(xy, ref)
    for k in FIELDS:
        g, r = np.asarray(a[k]), np.asarray(ra[k])
        assert g.shape == r.shape, k
        if k == "spectrum":
            assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), "spectrum differs in ulps"
        else:
            assert np.array_equal(g, r), k
    return st


def test_alpha_c1(oracle):
    print("case: alpha C1", flush=True)
    _check(oracle, oracle.gen_uniform(1, 1000), 512)


@pytest.mark.parametrize("name,T", [("bird.txt", 256), ("alpha.txt", 256), ("10000_1.txt", 1024)])
def test_alpha_fixtures(oracle, name, T):
    print("case: alpha fixture", name, flush=True)
    _check(oracle, load_fixture(name), T)


def test_alpha_exact_ties(oracle):
    """A lattice: many exactly equal rho values, resolved by the device exact path."""
    print("case: alpha lattice ties", flush=True)
    g = np.array([(i / 30.0, j / 30.0) for i in range(30) for j in range(30)])
    st = _check(oracle, g, 512)
    assert st["alphaExactCompares"] > 0


def test_alpha_100k(oracle):
    print("case: alpha 100k", flush=True)
    _check(oracle, oracle.gen_uniform(4, 100000), 2048)


def test_alpha_lattice_clusters(oracle):
    """Lattice half + Gaussian clusters: large clusters of near-equal rho are ordered on the
    device by exact 256-bit quotient keys, key-equal neighbours by the exact comparator."""
    from conftest import lattice_clusters
    print("case: alpha lattice+clusters 60k", flush=True)
    xy = lattice_clusters(30000, 30000, 200, 8, 0.005, 3)
    st = _check(oracle, xy, 2048)
    assert st["alphaExactCompares"] > 0
