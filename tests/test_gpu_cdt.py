"""Constrained DT on the GPU (cavity insertion + constrained flips in csrc/mesh.hip) against the
oracle (oracle/cdt.c, pinned in test_oracle_cdt.py): bit-exact triangles, order independence, the
reference-named entry point with GPUDTPARAMS::constraints, alpha shapes of a constrained DT, the
definition at a larger size, and the errors for constraints the reference leaves undefined."""
import ctypes as C

import numpy as np
import pytest

import galpha2d_amd as G
from cdt_cases import cell_polygons, lattice_knights, star_polygon
from conftest import ALPHA_FIELDS, load_fixture, lattice_clusters


def _case(oracle, name):
    if name == "uniform_star":
        xy = oracle.gen_uniform(5, 3000)
        return xy, star_polygon(xy, 60, 1)
    if name == "file_cells":
        xy = load_fixture("10000_1.txt")
        return xy, cell_polygons(xy, 8, 7, 2)
    if name == "uniform_cells":
        xy = oracle.gen_uniform(11, 20000)
        return xy, cell_polygons(xy, 12, 8, 6)
    if name == "lattice_knights":
        side = 40
        kept, nodes = lattice_knights(side, 900, 150, 3)
        pos = {v: k for k, v in enumerate(nodes)}
        xy = np.array([divmod(v, side) for v in nodes], dtype=np.float64) / side
        return xy, np.array([(pos[a], pos[b]) for a, b in kept], dtype=np.int32)
    if name == "clusters_long":
        xy = lattice_clusters(0, 6000, 64, 6, 0.05, 12)
        return xy, star_polygon(xy, 40, 3, center=xy.mean(0))
    raise ValueError(name)


CASES = ["uniform_star", "file_cells", "uniform_cells", "lattice_knights", "clusters_long"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_cdt_matches_oracle(oracle, name):
    print("case: cdt", name, flush=True)
    xy, cons = _case(oracle, name)
    want, dup = oracle.cdt(xy, cons)
    ctx = G.Context(0)
    ctx.set_constraints(cons)
    got = ctx.delaunay(xy)
    st = ctx.stats()
    ctx.close()
    assert got.shape == want.shape and np.array_equal(got, want), name
    assert st["constraintRounds"] >= 1
    rc, msg = oracle.validate_constrained(xy, got, dup, cons)
    assert rc == 0, msg


@pytest.mark.gpu
def test_cdt_order_and_direction(oracle):
    print("case: cdt order", flush=True)
    xy, cons = _case(oracle, "file_cells")
    ctx = G.Context(0)
    ctx.set_constraints(cons)
    a = ctx.delaunay(xy)
    rng = np.random.default_rng(1)
    ctx.set_constraints(cons[rng.permutation(len(cons))][:, ::-1].copy())
    b = ctx.delaunay(xy)
    # constraints that are Delaunay edges already leave the DT as it is
    ctx.set_constraints(None)
    dt = ctx.delaunay(xy)
    ctx.set_constraints(dt[::5, 3:5].copy())
    c = ctx.delaunay(xy)
    ctx.close()
    assert np.array_equal(a, b)
    assert np.array_equal(c, dt)


@pytest.mark.gpu
def test_cdt_reference_entry_point(oracle):
    """GPUDTPARAMS::constraints through gpudtComputeDT (the reference's own call)."""
    print("case: cdt gpudtComputeDT", flush=True)
    xy, cons = _case(oracle, "uniform_star")
    L = G.lib()

    class Params(C.Structure):
        _fields_ = [("minX", C.c_double), ("minY", C.c_double), ("maxX", C.c_double), ("maxY", C.c_double),
                    ("nPoints", C.c_int), ("points", C.c_void_p), ("nConstraints", C.c_int),
                    ("constraints", C.POINTER(C.c_int)), ("fboSize", C.c_int)]

    class Output(C.Structure):
        _fields_ = [("triangles", C.POINTER(C.c_int)), ("nTris", C.c_int)]

    L.gpudtComputeDT.argtypes = [C.POINTER(Params)]
    L.gpudtComputeDT.restype = C.POINTER(Output)
    L.gpudtReleaseDTOutput.argtypes = [C.POINTER(Output)]
    p_xy = np.ascontiguousarray(xy)
    c = np.ascontiguousarray(cons, dtype=np.int32)
    prm = Params(0, 0, 1, 1, len(xy), p_xy.ctypes.data, len(c), c.ctypes.data_as(C.POINTER(C.c_int)),
                 G.auto_texture(len(xy)))
    out = L.gpudtComputeDT(C.byref(prm))
    assert out, L.gpudtGetLastError().decode()
    got = np.ctypeslib.as_array(out.contents.triangles, shape=(out.contents.nTris, 9)).copy()
    L.gpudtReleaseDTOutput(out)
    want, _ = oracle.cdt(xy, cons)
    assert np.array_equal(got, want)
    # the default context drops the constraints after the call
    prm.nConstraints = 0
    out = L.gpudtComputeDT(C.byref(prm))
    plain = np.ctypeslib.as_array(out.contents.triangles, shape=(out.contents.nTris, 9)).copy()
    L.gpudtReleaseDTOutput(out)
    assert np.array_equal(plain, oracle.delaunay(xy)[0])


@pytest.mark.gpu
def test_cdt_alpha(oracle):
    """Alpha shapes of a constrained DT: the device computes them from the constrained mesh,
    with the same results as the oracle's alpha stage on the oracle's constrained DT."""
    print("case: cdt alpha", flush=True)
    xy, cons = _case(oracle, "file_cells")
    want, _ = oracle.cdt(xy, cons)
    wa = oracle.alpha(xy, want)
    ctx = G.Context(0)
    ctx.set_constraints(cons)
    d = G.DeviceBuffer(np.ascontiguousarray(xy, dtype=np.float64))
    ctx.compute_device(d.ptr, len(xy), G.auto_texture(len(xy)), with_alpha=True)
    assert np.array_equal(ctx.triangles(), want)
    ga = ctx.alpha()
    ctx.close()
    for k in ALPHA_FIELDS:
        assert np.array_equal(np.asarray(ga[k]), np.asarray(wa[k])), k


@pytest.mark.gpu
def test_cdt_large_by_definition(oracle):
    """200,000 uniform points with 20,000 cell-polygon constraints: the device result satisfies
    the definition (validator), and a second run gives the same triangles."""
    print("case: cdt large", flush=True)
    xy = oracle.gen_uniform(21, 200000)
    cons = cell_polygons(xy, 60, 6, 8)
    ctx = G.Context(0)
    ctx.set_constraints(cons)
    a = ctx.delaunay(xy)
    st = ctx.stats()
    b = ctx.delaunay(xy)
    ctx.close()
    assert np.array_equal(a, b)
    dup = np.full(len(xy), -1, dtype=np.int32)
    rc, msg = oracle.validate_constrained(xy, a, dup, cons)
    assert rc == 0, msg
    print(f"constraints {len(cons)}: rounds {st['constraintRounds']}, {st['ms_constraints']:.2f} ms", flush=True)


_SQ = np.array([[0, 0], [4, 0], [4, 4], [0, 4], [2, 1.5], [1, 3.5], [0, 0], [2, 2]], dtype=np.float64)
INVALID = [
    ("cross", _SQ[:7], [[0, 2], [1, 3]], "cross"),
    ("through", _SQ, [[0, 2]], "inside a constraint"),
    ("dup", _SQ[:7], [[6, 1]], "duplicate"),
    ("range", _SQ[:7], [[0, 9]], "outside"),
    ("same", _SQ[:7], [[1, 1]], "both endpoints"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case,xy,cons,msg", INVALID, ids=[c[0] for c in INVALID])
def test_cdt_invalid(case, xy, cons, msg):
    print("case: cdt invalid", case, flush=True)
    ctx = G.Context(0)
    ctx.set_constraints(np.array(cons))
    with pytest.raises(G.GalphaError, match=msg):
        ctx.delaunay(xy, 256)
    # the context stays usable
    ctx.set_constraints(None)
    assert ctx.delaunay(xy, 256).shape[0] > 0
    ctx.close()


@pytest.mark.gpu
def test_cdt_rejected_by_batches(oracle):
    ctx = G.Context(0)
    ctx.set_constraints(np.array([[0, 1]]))
    xy = oracle.gen_uniform(1, 300)
    with pytest.raises(G.GalphaError, match="single-set"):
        ctx.batch([xy, xy])
    ctx.close()


@pytest.mark.gpu
def test_cdt_long_rays(oracle):
    """16 rays from the point nearest the centre of 200,000 uniform points to points near the
    border: each crosses hundreds of triangles (one large cavity per ray, rays share an endpoint).
    The device result equals the oracle's and satisfies the definition."""
    print("case: cdt long rays", flush=True)
    xy = oracle.gen_uniform(23, 200000)
    c = int(np.argmin(((xy - 0.5) ** 2).sum(1)))
    ang = np.arctan2(xy[:, 1] - xy[c, 1], xy[:, 0] - xy[c, 0])
    far = np.nonzero(((xy - 0.5) ** 2).sum(1) > 0.2)[0]
    picks = []
    for k in range(16):   # one far point per angular sector, so rays never overlap
        lo, hi = -np.pi + k * np.pi / 8, -np.pi + (k + 1) * np.pi / 8
        sel = far[(ang[far] > lo + 0.05) & (ang[far] < hi - 0.05)]
        picks.append(int(sel[0]))
    cons = np.array([(c, p) for p in picks], dtype=np.int32)
    ctx = G.Context(0)
    ctx.set_constraints(cons)
    a = ctx.delaunay(xy)
    st = ctx.stats()
    ctx.close()
    want, dup = oracle.cdt(xy, cons)
    assert np.array_equal(a, want)
    rc, msg = oracle.validate_constrained(xy, a, dup, cons)
    assert rc == 0, msg
    print(f"rays: rounds {st['constraintRounds']}, {st['ms_constraints']:.2f} ms", flush=True)
