"""Delaunay without device allocation (galphaComputeDeviceNoAlloc): its triangles equal those of
the allocating entry point galphaComputeDevice on the same points, over calls in a row that
shrink (the scratch keeps the previous call's contents), and sizes beyond the scratch are refused
with a message."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _allocating(G, xy, T):
    ctx = G.Context(0)
    d = G.DeviceBuffer(xy)
    try:
        ctx.compute_device(d.ptr, xy.shape[0], T)
        return ctx.triangles()
    finally:
        d.free()
        ctx.close()


def _no_alloc(G, xy, T, capacity=None):
    n = xy.shape[0]
    cap = 2 * n if capacity is None else capacity
    d = G.DeviceBuffer(xy)
    out = G.DeviceBuffer(nbytes=36 * max(cap, 1))
    try:
        F = G.compute_device_noalloc(d.ptr, n, T, out.ptr, cap)
        return out.download(np.int32, (max(cap, 1), 9))[:F]
    finally:
        d.free()
        out.free()


def test_no_alloc_matches_allocating_entry_point(oracle):
    import galpha2d_amd as G
    lattice = np.array([(i / 60.0, j / 60.0) for i in range(60) for j in range(60)])
    cases = [("uniform", oracle.gen_uniform(31, 65536), 1024),   # the largest size the header promises
             ("uniform", oracle.gen_uniform(32, 1000), 512),     # a smaller second call
             ("lattice", lattice, 512)]                          # collinear hull: round-by-round removal
    for name, xy, T in cases:
        print(f"case: no-allocation DT, {name} {xy.shape[0]} points, texture {T}", flush=True)
        got = _no_alloc(G, xy, T)
        want = _allocating(G, xy, T)
        assert got.shape[0] > 0 and np.array_equal(got, want), name
    ref, _ = oracle.delaunay(cases[1][1])
    assert np.array_equal(_no_alloc(G, cases[1][1], 512), ref)


def test_no_alloc_refuses_what_does_not_fit():
    import galpha2d_amd as G
    xy = np.random.default_rng(5).random((200000, 2))
    print("case: no-allocation DT beyond the scratch (200000 points, texture 4096)", flush=True)
    with pytest.raises(G.GalphaError, match="scratch"):
        _no_alloc(G, xy, 4096)
    print("case: no-allocation DT into an output array that is too small", flush=True)
    with pytest.raises(G.GalphaError, match="output holds"):
        _no_alloc(G, xy[:1000], 512, capacity=100)
    # the scratch still serves a run after a refused one
    small = xy[:500]
    assert np.array_equal(_no_alloc(G, small, 256), _allocating(G, small, 256))
