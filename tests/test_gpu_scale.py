"""Full-size workloads on the GPU against the oracle's committed digests (canonical triangles
and every alpha array, bit-exact): 1M uniform points, and the lattice + Gaussian-cluster
stress sets at 400K and 2M points (many co-circular quadruples and near-equal rho values)."""
import numpy as np
import pytest

from conftest import digest, golden_digests

pytestmark = pytest.mark.gpu


def _run(xy, T):
    import galpha2d_amd as G
    ctx = G.Context(0)
    d = G.DeviceBuffer(np.ascontiguousarray(xy, dtype=np.float64))
    ctx.compute_device(d.ptr, xy.shape[0], T, with_alpha=True)
    tris, a, st = ctx.triangles(), ctx.alpha(), ctx.stats()
    ctx.close()
    return tris, a, st


def _check(name, xy, T):
    print("case:", name, xy.shape[0], "points, texture", T, flush=True)
    want = golden_digests()[name]
    assert want["n"] == xy.shape[0]
    tris, a, st = _run(xy, T)
    got = digest(tris, a)
    bad = [k for k in want if k != "n" and got.get(k) != want[k]]
    print("stats", {k: st[k] for k in ("missing", "flips", "incircleExact", "alphaExactCompares", "frameBulk", "earRounds", "ms_boundary",
                                       "alphaMaxCluster", "ms_total")}, flush=True)
    assert not bad, "differs from the oracle in: %s" % bad
    return st


def test_uniform_1m():
    from galpha2d_amd import datagen
    _check("uniform_1m", datagen.uniform(2, 1048576), 4096)


def test_lattice_400k():
    from galpha2d_amd import datagen
    st = _check("lattice_400k", datagen.lattice_clusters(400000, side=632, nclusters=16, sigma=0.005, seed=3),
                4096)
    assert st["alphaExactCompares"] > 0


def test_lattice_2m():
    from galpha2d_amd import datagen
    st = _check("lattice_2m", datagen.lattice_clusters(2000000, side=1414, nclusters=64, sigma=0.005, seed=3),
                8192)
    assert st["incircleExact"] > 0
    # the lattice's straight hull sides leave long thin pockets under hull edges to cluster points;
    # fan ears fill each in a few rounds instead of one round per lattice point
    assert st["frameBulk"] == 1 and 0 < st["earRounds"] <= 64
