"""Alpha-shape stage on the GPU against the oracle: spectrum values bit-identical (0 ulp),
ranks, intervals, master list and spec pointers identical."""
import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu

FIELDS = ("spectrum", "tri_ro_rank", "edge_t", "edge_vopp", "edge_ro_rank", "edge_nu_min", "edge_nu_max",
          "vert_nu_min", "vert_nu_max", "master_simplex", "master_st", "master_at", "spec_pointers")


def _alpha_gpu(xy, T):
    import galpha2d_amd as G
    ctx = G.Context(0)
    d = G.DeviceBuffer(np.ascontiguousarray(xy, dtype=np.float64))
    ctx.compute_device(d.ptr, xy.shape[0], T, with_alpha=True)
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


def _classes_by_definition(tris9, a, r):
    """Alpha-complex classes at rank r from the definition: a triangle is in when its rank <= r;
    an edge is in when an incident triangle is, or when it is unattached with rank <= r; an edge
    in the complex is singular with no incident triangle in, interior with both, else regular;
    a vertex of the triangulation is singular with no incident edge in, interior when every
    incident triangle is in and it is not on the hull, else regular; other points exterior."""
    tri_in = a["tri_ro_rank"] <= r
    et, ev, ero = a["edge_t"], a["edge_vopp"], a["edge_ro_rank"]
    nb = tris9[et, ev]
    has2 = nb >= 0
    k = tri_in[et].astype(np.int64) + np.where(has2, tri_in[np.maximum(nb, 0) >> 2], False)
    ein = (k >= 1) | ((ero != 0) & (ero <= r))
    ecls = np.where(~ein, 0, np.where(k == 0, 1, np.where(has2 & (k == 2), 3, 2)))
    n = len(a["vert_nu_min"])
    u = tris9[et, 3 + (ev + 1) % 3]
    w = tris9[et, 3 + (ev + 2) % 3]
    in_dt = np.zeros(n, bool)
    in_dt[tris9[:, 3:6].ravel()] = True
    any_edge = np.zeros(n, bool)
    any_edge[u[ein]] = True
    any_edge[w[ein]] = True
    hull = np.zeros(n, bool)
    hull[u[~has2]] = True
    hull[w[~has2]] = True
    all_tri = np.ones(n, bool)
    all_tri[tris9[~tri_in, 3:6].ravel()] = False
    vcls = np.where(~in_dt, 0, np.where(~any_edge, 1, np.where(all_tri & ~hull, 3, 2)))
    tcls = np.where(tri_in, 3, 0)
    return vcls.astype(np.uint8), ecls.astype(np.uint8), tcls.astype(np.uint8)


@pytest.mark.parametrize("case", ["uniform", "lattice", "bird"])
def test_alpha_classify(oracle, case):
    """Alpha-complex classes on the device (from the interval ranks) equal the classes derived
    from the definition (triangle ranks, edge ranks, adjacency), at ranks across the spectrum;
    the rank lookup returns the last rank of each spectrum value."""
    import galpha2d_amd as G
    print("case: alpha classification", case, flush=True)
    if case == "uniform":
        xy, T = oracle.gen_uniform(9, 1500), 512
    elif case == "lattice":
        xy, T = np.array([(i / 20.0, j / 20.0) for i in range(20) for j in range(20)]), 512
    else:
        xy, T = load_fixture("bird.txt"), 256
    ctx = G.Context(0)
    d = G.DeviceBuffer(np.ascontiguousarray(xy, dtype=np.float64))
    ctx.compute_device(d.ptr, xy.shape[0], T, with_alpha=True)
    tris, a = ctx.triangles(), ctx.alpha()
    ref, _ = oracle.delaunay(xy)
    assert np.array_equal(tris, ref)
    spec = np.asarray(a["spectrum"])
    R = len(spec) - 1
    for r in sorted({0, 1, R // 7, R // 3, R // 2, R - 1, R}):
        got = ctx.classify(r, a)
        want = _classes_by_definition(tris, a, r)
        for g, w_, name in zip(got, want, ("vertex", "edge", "triangle")):
            assert np.array_equal(g, w_), (case, r, name)
    for r in sorted({1, R // 2, R}):
        # the FP64 spectrum need not be monotone (near-equal exact values can round either way)
        last = int(np.nonzero(spec <= spec[r])[0].max())
        assert ctx.alpha_rank(spec[r]) == last
    assert ctx.alpha_rank(-1.0) == -1
    ctx.close()


@pytest.mark.gpu
def test_compute_on_caller_stream(oracle):
    """galphaComputeDevice on a stream the caller owns gives the result of the context's own
    stream: a stream of the library's runtime always, and a PyTorch stream with torch's device
    memory when torch's own HIP runtime sees the GPU."""
    import galpha2d_amd as G
    print("case: caller stream", flush=True)
    xy = oracle.gen_uniform(13, 3000)
    want = _alpha_gpu(xy, 512)

    def check(ctx):
        tris, a = ctx.triangles(), ctx.alpha()
        assert np.array_equal(tris, want[0])
        for k in FIELDS:
            assert np.array_equal(np.asarray(a[k]), np.asarray(want[1][k])), k

    ctx = G.Context(0)
    s = G.Stream(0)
    d = G.DeviceBuffer(xy)
    ctx.compute_device(d.ptr, xy.shape[0], 512, with_alpha=True, stream=s.handle)
    s.synchronize()
    check(ctx)
    s.close()
    import torch
    if torch.cuda.is_available():
        d2 = torch.from_numpy(xy).to("cuda:0")
        ts = torch.cuda.Stream()
        torch.cuda.synchronize()
        with torch.cuda.stream(ts):
            ctx.compute_device(d2.data_ptr(), xy.shape[0], 512, with_alpha=True, stream=ts.cuda_stream)
        ts.synchronize()
        check(ctx)
    else:
        print("torch's HIP runtime sees no GPU here: PyTorch stream part not run", flush=True)
    ctx.close()


def test_alpha_of_caller_triangles(oracle):
    """gpudtComputeAlpha over triangles the caller passes in: the oracle's DT gives the oracle's
    alpha result; triangles that are not a mesh of the points are an error with a message."""
    print("case: gpudtComputeAlpha with caller triangles", flush=True)
    import ctypes as C
    import galpha2d_amd as G
    L = G.lib()
    L.gpudtComputeAlpha.argtypes = [C.POINTER(G._Params), C.POINTER(G._Output), C.POINTER(C.POINTER(G._Alpha))]
    L.gpudtComputeAlpha.restype = C.c_int
    xy = np.ascontiguousarray(load_fixture("1000_1.txt"))
    n = xy.shape[0]
    prm = G._Params(0, 0, 1, 1, n, xy.ctypes.data_as(C.POINTER(C.c_double)), 0, None, 512)
    ref, _ = oracle.delaunay(xy)

    def run(tris):
        tris = np.ascontiguousarray(tris, dtype=np.int32)
        out = G._Output(tris.ctypes.data_as(C.POINTER(C.c_int)), tris.shape[0])
        A = C.POINTER(G._Alpha)()
        if L.gpudtComputeAlpha(C.byref(prm), C.byref(out), C.byref(A)) < 0:
            return L.gpudtGetLastError().decode()
        try:
            return G._alpha_to_numpy(A)
        finally:
            L.gpudtReleaseAlpha(A)

    got, ra = run(ref), oracle.alpha(xy, ref)
    for k in FIELDS:
        g, r = np.asarray(got[k]), np.asarray(ra[k])
        if k == "spectrum":
            assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), k
        else:
            assert np.array_equal(g, r), k
    t = 10 if ref[10, 0] >= 0 else int(np.argmax(ref[:, 0] >= 0))
    bad = {}
    b = ref.copy(); b[5, 3] = n; bad["vertex out of range"] = b
    b = ref.copy(); b[5, 4] = b[5, 3]; bad["repeated vertex"] = b
    b = ref.copy(); b[t, 0] = 4 * ref.shape[0]; bad["neighbour out of range"] = b
    b = ref.copy(); b[t, 0] = b[t, 0] ^ 1 if (b[t, 0] & 3) != 2 else b[t, 0] - 2; bad["link that does not link back"] = b
    for name, b in bad.items():
        msg = run(b)
        assert isinstance(msg, str) and "not a mesh" in msg, (name, msg)
    # the context still works after a rejected call
    got = run(ref)
    assert np.array_equal(np.asarray(got["vert_nu_min"]), np.asarray(ra["vert_nu_min"]))


def test_alpha_of_own_output_and_of_an_edited_copy(oracle):
    """gpudtComputeAlpha on gpudtComputeDT's output uses the points and triangles still on the
    device; after the caller edits that output in place (here: the vertex order of one triangle
    rotated, with its links) the edited mesh is the one analysed."""
    print("case: gpudtComputeDT then gpudtComputeAlpha, unchanged and edited in place", flush=True)
    import ctypes as C
    import galpha2d_amd as G
    L = G.lib()
    L.gpudtComputeDT.argtypes = [C.POINTER(G._Params)]
    L.gpudtComputeDT.restype = C.POINTER(G._Output)
    L.gpudtReleaseDTOutput.argtypes = [C.POINTER(G._Output)]
    L.gpudtComputeAlpha.argtypes = [C.POINTER(G._Params), C.POINTER(G._Output), C.POINTER(C.POINTER(G._Alpha))]
    L.gpudtComputeAlpha.restype = C.c_int
    xy = np.ascontiguousarray(oracle.gen_uniform(41, 3000))
    n = xy.shape[0]
    prm = G._Params(0, 0, 1, 1, n, xy.ctypes.data_as(C.POINTER(C.c_double)), 0, None, 512)
    out = L.gpudtComputeDT(C.byref(prm))
    assert out
    F = out.contents.nTris
    view = np.ctypeslib.as_array(C.cast(out.contents.triangles, C.POINTER(C.c_int)), shape=(F * 9,)).reshape(F, 9)
    ref, _ = oracle.delaunay(xy)
    assert np.array_equal(view, ref)

    def alpha():
        A = C.POINTER(G._Alpha)()
        assert L.gpudtComputeAlpha(C.byref(prm), out, C.byref(A)) == 0, L.gpudtGetLastError().decode()
        try:
            return G._alpha_to_numpy(A)
        finally:
            L.gpudtReleaseAlpha(A)

    def check(got, tris):
        want = oracle.alpha(xy, tris)
        for k in FIELDS:
            g, r = np.asarray(got[k]), np.asarray(want[k])
            assert g.shape == r.shape, k
            if k == "spectrum":
                assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), k
            else:
                assert np.array_equal(g, r), k

    check(alpha(), ref)
    # rotate triangle t's corners by one (vtx and tri together) and fix its neighbours' back-links
    t = F // 2
    old = view[t].copy()
    for i in range(3):
        view[t, i] = old[(i + 1) % 3]
        view[t, 3 + i] = old[3 + (i + 1) % 3]
    for i in range(3):
        e = view[t, i]
        if e >= 0:
            u, j = e >> 2, e & 3
            view[u, j] = (t << 2) | i
    edited = view.copy()
    assert not np.array_equal(edited, ref)
    check(alpha(), edited)
    L.gpudtReleaseDTOutput(out)
