"""GPU Delaunay path (HIP, gfx950) against the oracle: bit-exact canonical triangles and
adjacency, plus size-independent validity checks at full scale."""
import numpy as np
import pytest

from conftest import ALPHA_FIELDS, load_fixture, lattice_clusters

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import galpha2d_amd as G
    c = G.Context(0)
    yield c
    c.close()


def _check_same(ctx, oracle, xy, T):
    tris = ctx.delaunay(xy, T)
    ref, dup = oracle.delaunay(xy)
    assert tris.shape == ref.shape, (tris.shape, ref.shape)
    assert np.array_equal(tris, ref)
    return tris


def test_c1_uniform_1000(ctx, oracle):
    print("case: C1 1000 uniform seed 1 T=512", flush=True)
    _check_same(ctx, oracle, oracle.gen_uniform(1, 1000), 512)


@pytest.mark.parametrize("name,T", [("bird.txt", 256), ("alpha.txt", 256), ("1000_1.txt", 512),
                                    ("10000_1.txt", 1024)])
def test_reference_fixtures(ctx, oracle, name, T):
    print("case: fixture", name, T, flush=True)
    _check_same(ctx, oracle, load_fixture(name), T)


@pytest.mark.parametrize("n,T", [(3, 64), (4, 64), (50, 64), (5000, 256), (5000, 2048), (100000, 2048)])
def test_uniform_sizes_and_textures(ctx, oracle, n, T):
    print("case: uniform", n, T, flush=True)
    _check_same(ctx, oracle, oracle.gen_uniform(7 + n, n), T)


def test_duplicates_and_lattice(ctx, oracle):
    print("case: lattice with duplicates", flush=True)
    g = np.array([(i / 60.0, j / 60.0) for i in range(60) for j in range(60)])
    xy = np.concatenate([g, g[::7], g[:3]])
    tris = _check_same(ctx, oracle, xy, 512)
    assert tris[:, 3:6].max() < len(g)      # duplicates keep their smallest index


def test_collinear_and_tiny(ctx, oracle):
    print("case: degenerate small inputs", flush=True)
    col = np.array([[i * 0.25, 0.5 + i * 0.125] for i in range(30)])
    assert ctx.delaunay(col, 256).shape[0] == 0
    assert ctx.delaunay(np.zeros((0, 2)), 256).shape[0] == 0
    assert ctx.delaunay(np.array([[0.0, 0.0], [1.0, 1.0]]), 256).shape[0] == 0
    tri = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    _check_same(ctx, oracle, tri, 64)


def test_lattice_clusters_small(ctx, oracle):
    print("case: lattice+clusters 200k T=4096", flush=True)
    _check_same(ctx, oracle, lattice_clusters(100000, 100000, 500, 16, 0.005, 3), 4096)


def test_c2_full_scale(ctx, oracle):
    """1,048,576 uniform points, texture 4096: bit-exact vs the oracle and valid."""
    print("case: C2 1M uniform seed 2 T=4096", flush=True)
    xy = oracle.gen_uniform(2, 1048576)
    tris = _check_same(ctx, oracle, xy, 4096)
    rc, msg = oracle.validate(xy, tris, None)
    assert rc == 0, msg


def test_deterministic_repeat(ctx, oracle):
    print("case: repeat determinism 200k", flush=True)
    xy = oracle.gen_uniform(9, 200000)
    a = ctx.delaunay(xy, 2048)
    b = ctx.delaunay(xy, 2048)
    assert np.array_equal(a, b)


def test_reference_api_entry_point(oracle):
    """gpudtComputeDT / gpudtReleaseDTOutput through the C ABI with the reference's structs."""
    print("case: gpudtComputeDT", flush=True)
    import ctypes as C
    import galpha2d_amd as G
    L = G.lib()

    class Vertex(C.Structure):
        _fields_ = [("x", C.c_double), ("y", C.c_double)]

    class Tri(C.Structure):
        _fields_ = [("tri", C.c_int * 3), ("vtx", C.c_int * 3), ("tmp", C.c_int * 3)]

    class Params(C.Structure):
        _fields_ = [("minX", C.c_double), ("minY", C.c_double), ("maxX", C.c_double), ("maxY", C.c_double),
                    ("nPoints", C.c_int), ("points", C.POINTER(Vertex)), ("nConstraints", C.c_int),
                    ("constraints", C.POINTER(C.c_int)), ("fboSize", C.c_int)]

    class Output(C.Structure):
        _fields_ = [("triangles", C.POINTER(Tri)), ("nTris", C.c_int)]

    xy = load_fixture("1000_1.txt")
    pts = (Vertex * len(xy))(*[Vertex(x, y) for x, y in xy])
    p = Params(xy[:, 0].min(), xy[:, 1].min(), xy[:, 0].max(), xy[:, 1].max(), len(xy), pts, 0, None, 512)
    L.gpudtComputeDT.argtypes = [C.POINTER(Params)]
    L.gpudtComputeDT.restype = C.POINTER(Output)
    L.gpudtReleaseDTOutput.argtypes = [C.POINTER(Output)]
    out = L.gpudtComputeDT(C.byref(p))
    assert out
    F = out.contents.nTris
    got = np.ctypeslib.as_array(C.cast(out.contents.triangles, C.POINTER(C.c_int)), shape=(F * 9,)).reshape(F, 9).copy()
    L.gpudtReleaseDTOutput(out)
    ref, _ = oracle.delaunay(xy)
    assert np.array_equal(got, ref)
    # a constraint count without a list is an error with a message, not a crash
    p.nConstraints = 1
    assert not L.gpudtComputeDT(C.byref(p))
    assert b"constraint list" in L.gpudtGetLastError()
    # with a list: the constrained DT (a chord across the set, between its two extreme points)
    c = np.array([int(np.argmin(xy[:, 0])), int(np.argmax(xy[:, 0]))], dtype=np.int32)
    p.constraints = c.ctypes.data_as(C.POINTER(C.c_int))
    out = L.gpudtComputeDT(C.byref(p))
    assert out, L.gpudtGetLastError()
    F = out.contents.nTris
    got = np.ctypeslib.as_array(C.cast(out.contents.triangles, C.POINTER(C.c_int)), shape=(F * 9,)).reshape(F, 9).copy()
    L.gpudtReleaseDTOutput(out)
    assert np.array_equal(got, oracle.cdt(xy, c.reshape(1, 2))[0])
    p.nConstraints = 0
    p.constraints = None
    p.fboSize = 1000
    assert not L.gpudtComputeDT(C.byref(p))


def _degenerate_cases(oracle):
    """Exactly co-circular integer points (x^2 + y^2 = 25^2, 65^2), a fan over collinear points,
    a large offset, a tiny scale, a full integer grid and the grid with duplicates."""
    out = {}
    r = 25
    pts = sorted({(x, y) for x in range(-r, r + 1) for y in range(-r, r + 1) if x * x + y * y == r * r})
    out["circle25"] = np.array(pts, dtype=np.float64)
    out["circle25_center"] = np.array(pts + [(0, 0), (1, 2)], dtype=np.float64)
    r = 65
    pts65 = sorted({(x, y) for x in range(-r, r + 1) for y in range(-r, r + 1) if x * x + y * y == r * r})
    out["circle65_two"] = np.array(pts65 + [(x // 5, y // 5) for x, y in pts if (x // 5, y // 5) not in pts65],
                                   dtype=np.float64)
    out["fan"] = np.array([(i, 0.0) for i in range(40)] + [(19.5, 7.0)], dtype=np.float64)
    # vertices of degree 2000: an apex over 2000 collinear points (open fan, on the hull) and
    # the centre of a ring of 2000 (closed fan)
    out["fan_2000"] = np.array([(i, 0.0) for i in range(2000)] + [(999.5, 7.0)], dtype=np.float64)
    th = 2 * np.pi * np.arange(2000) / 2000
    out["wheel_2000"] = np.concatenate([np.stack([np.cos(th), np.sin(th)], 1), [[0.001, 0.002]]])
    # the same with the high-degree vertex first: it is the smallest vertex of all its triangles,
    # which overflows the per-vertex buckets of the canonical order (the keys are sorted instead)
    out["wheel_2000_centre_first"] = np.concatenate([[[0.001, 0.002]], np.stack([np.cos(th), np.sin(th)], 1)])
    out["fan_2000_apex_first"] = np.array([(999.5, 7.0)] + [(i, 0.0) for i in range(2000)], dtype=np.float64)
    u = oracle.gen_uniform(91, 3000)
    out["offset"] = u + 1.0e6
    out["scaled_tiny"] = u * 1e-9
    g = np.array([(i, j) for i in range(30) for j in range(30)], dtype=np.float64)
    out["grid_full"] = g
    out["grid_dups"] = np.concatenate([g, g[::7]])
    return out


def _same_alpha(a, ra, name):
    for k in ALPHA_FIELDS:
        g, r = np.asarray(a[k]), np.asarray(ra[k])
        assert g.shape == r.shape, (name, k)
        if k == "spectrum":
            assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), (name, "spectrum differs in ulps")
        else:
            assert np.array_equal(g, r), (name, k)


def test_degenerate_geometry(oracle):
    """Every case through the single-set path and all of them through one batch: triangles and
    every alpha array bit-exact against the oracle."""
    import galpha2d_amd as G
    cases = _degenerate_cases(oracle)
    ctx = G.Context(0)
    refs = {}
    try:
        for name, xy in cases.items():
            print("case: degenerate", name, xy.shape[0], flush=True)
            ref, _ = oracle.delaunay(xy)
            ra = oracle.alpha(xy, ref)
            refs[name] = (ref, ra)
            d = G.DeviceBuffer(np.ascontiguousarray(xy))
            ctx.compute_device(d.ptr, xy.shape[0], G.auto_texture(xy.shape[0]), with_alpha=True)
            tris, a = ctx.triangles(), ctx.alpha()
            d.free()
            assert np.array_equal(tris, ref), name
            _same_alpha(a, ra, name)
        print("case: degenerate batch", len(cases), flush=True)
        got = ctx.batch(list(cases.values()), with_alpha=True)
        for k, name in enumerate(cases):
            ref, ra = refs[name]
            assert np.array_equal(got[k], ref), name
            _same_alpha(ctx.batch_alpha(k), ra, name)
    finally:
        ctx.close()


def test_slot_capacity_guard():
    """Links name triangle slots as (t << 2) | j in int32, so one pass holds at most 3 V + 8192 <
    2^29 slots.  Sizes past that bound are refused from the sizes alone, before any kernel reads
    the (here tiny) point buffer."""
    import galpha2d_amd as G
    print("case: slot capacity guard", flush=True)
    ctx = G.Context(0)
    d = G.DeviceBuffer(np.zeros((16, 2)))
    try:
        with pytest.raises(G.GalphaError, match="2\\^29"):
            ctx.compute_device(d.ptr, 179_000_000, 16384)
        with pytest.raises(G.GalphaError, match="2\\^29"):
            ctx.batch_device(d.ptr, np.array([0, 90_000_000, 180_000_000]), fbo_sizes=[16384, 16384])
        # the context still works afterwards
        xy = np.random.default_rng(3).random((500, 2))
        d2 = G.DeviceBuffer(xy)
        assert ctx.compute_device(d2.ptr, 500, 256) > 0
        d2.free()
    finally:
        d.free()
        ctx.close()
