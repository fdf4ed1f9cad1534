"""ctypes bindings to liboracle.so -- the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  The product package never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.ora_delaunay.argtypes = [dp, C.c_int, C.POINTER(C.c_void_p), ip]
        L.ora_delaunay.restype = C.c_int
        L.ora_validate.argtypes = [dp, C.c_int, ip, C.c_int, ip, C.c_char_p, C.c_int]
        L.ora_validate.restype = C.c_int
        L.ora_canonicalize.argtypes = [ip, C.c_int]
        L.ora_alpha_compute.argtypes = [dp, C.c_int, ip, C.c_int, C.c_void_p]
        L.ora_alpha_compute.restype = C.c_int
        L.ora_alpha_free.argtypes = [C.c_void_p]
        L.ora_gen_uniform.argtypes = [C.c_uint64, C.c_int, dp]
        for f in ("ora_orient2d", "ora_incircle"):
            getattr(L, f).restype = C.c_double
        L.ora_orient2d.argtypes = [dp, dp, dp]
        L.ora_incircle.argtypes = [dp, dp, dp, dp]
        L.ora_orient_sign.argtypes = [dp, dp, dp]
        L.ora_orient_sign.restype = C.c_int
        L.ora_incircle_sign.argtypes = [dp, dp, dp, dp]
        L.ora_incircle_sign.restype = C.c_int
        L.ora_incircle_sos.argtypes = [dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int]
        L.ora_incircle_sos.restype = C.c_int
        L.ora_rho_compare.argtypes = [C.c_int, dp, dp, dp, C.c_int, dp, dp, dp]
        L.ora_rho_compare.restype = C.c_int
        L.ora_rho_tri.argtypes = [dp, dp, dp]
        L.ora_rho_tri.restype = C.c_double
        L.ora_rho_edge.argtypes = [dp, dp]
        L.ora_rho_edge.restype = C.c_double
        L.ora_attached.argtypes = [dp, dp, dp]
        L.ora_attached.restype = C.c_int
        L.ora_pred_counters.argtypes = [C.POINTER(C.c_longlong)] * 5
        L.ora_cdt.argtypes = [dp, C.c_int, ip, C.c_int, C.POINTER(C.c_void_p), ip, C.c_char_p, C.c_int]
        L.ora_cdt.restype = C.c_int
        L.ora_validate_constrained.argtypes = [dp, C.c_int, ip, C.c_int, ip, ip, C.c_int, C.c_char_p, C.c_int]
        L.ora_validate_constrained.restype = C.c_int
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def gen_uniform(seed, n):
    xy = np.empty((n, 2), dtype=np.float64)
    lib().ora_gen_uniform(C.c_uint64(seed), n, _dp(xy))
    return xy


def delaunay(xy):
    """Canonical DT: returns (tris int32 [F, 9] in gpudtTriangle layout tri[3],vtx[3],tmp[3], dup_of)."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = xy.shape[0]
    out = C.c_void_p()
    dup = np.empty(max(n, 1), dtype=np.int32)
    nt = lib().ora_delaunay(_dp(xy), n, C.byref(out), _ip(dup))
    if nt < 0:
        raise RuntimeError("oracle DT failed")
    tris = np.zeros((nt, 9), dtype=np.int32)
    if nt > 0:
        C.memmove(tris.ctypes.data, out.value, nt * 36)
    if out.value:
        C.CDLL(None).free(out)
    return tris, dup[:n].copy()


def canonicalize(tris):
    t = np.ascontiguousarray(tris, dtype=np.int32).copy()
    if t.shape[0]:
        lib().ora_canonicalize(_ip(t), t.shape[0])
    return t


def dt_alpha_points(xy):
    """DT + alpha of one set, for a pool of worker processes; returns the number of points."""
    tris, _ = delaunay(xy)
    alpha(xy, tris)
    return int(xy.shape[0])


def validate(xy, tris, dup_of=None):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    d = None if dup_of is None else np.ascontiguousarray(dup_of, dtype=np.int32)
    msg = C.create_string_buffer(512)
    rc = lib().ora_validate(_dp(xy), xy.shape[0], _ip(t), t.shape[0],
                            _ip(d) if d is not None else None, msg, 512)
    return rc, msg.value.decode()


def cdt(xy, cons):
    """Constrained DT (cdt.c): (tris int32 [F, 9] canonical, dup_of); raises ValueError with the
    oracle's message for invalid constraints."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    c = np.ascontiguousarray(cons, dtype=np.int32).reshape(-1, 2)
    n = xy.shape[0]
    out = C.c_void_p()
    dup = np.empty(max(n, 1), dtype=np.int32)
    msg = C.create_string_buffer(512)
    nt = lib().ora_cdt(_dp(xy), n, _ip(c), c.shape[0], C.byref(out), _ip(dup), msg, 512)
    if nt < 0:
        raise ValueError(msg.value.decode())
    tris = np.zeros((nt, 9), dtype=np.int32)
    if nt > 0:
        C.memmove(tris.ctypes.data, out.value, nt * 36)
    if out.value:
        C.CDLL(None).free(out)
    return tris, dup[:n].copy()


def validate_constrained(xy, tris, dup_of, cons):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    d = None if dup_of is None else np.ascontiguousarray(dup_of, dtype=np.int32)
    c = np.ascontiguousarray(cons, dtype=np.int32).reshape(-1, 2)
    msg = C.create_string_buffer(512)
    rc = lib().ora_validate_constrained(_dp(xy), xy.shape[0], _ip(t), t.shape[0],
                                        _ip(d) if d is not None else None, _ip(c), c.shape[0], msg, 512)
    return rc, msg.value.decode()


class _Alpha(C.Structure):
    _fields_ = [("nEdges", C.c_int), ("nTris", C.c_int), ("nPoints", C.c_int),
                ("specSize", C.c_int), ("spectrum", C.POINTER(C.c_double)),
                ("tri_ro_rank", C.POINTER(C.c_int)),
                ("edge_t", C.POINTER(C.c_int)), ("edge_vopp", C.POINTER(C.c_int)),
                ("edge_ro_rank", C.POINTER(C.c_int)), ("edge_nu_min", C.POINTER(C.c_int)),
                ("edge_nu_max", C.POINTER(C.c_int)), ("edge_attached", C.POINTER(C.c_int)),
                ("vert_nu_min", C.POINTER(C.c_int)), ("vert_nu_max", C.POINTER(C.c_int)),
                ("masterSize", C.c_int), ("master_simplex", C.POINTER(C.c_int)),
                ("master_st", C.POINTER(C.c_int)), ("master_at", C.POINTER(C.c_int)),
                ("nSpecPointers", C.c_int), ("spec_pointers", C.POINTER(C.c_int)),
                ("exact_compares", C.c_longlong), ("exact_attach", C.c_longlong)]


def _arr(ptr, n, dt):
    if n <= 0:
        return np.zeros(0, dtype=dt)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)


def alpha(xy, tris):
    """Alpha spectrum + intervals + master list of a canonical DT (dict of numpy arrays)."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    A = _Alpha()
    lib().ora_alpha_compute(_dp(xy), xy.shape[0], _ip(t), t.shape[0], C.byref(A))
    r = dict(
        spectrum=_arr(A.spectrum, A.specSize, np.float64),
        tri_ro_rank=_arr(A.tri_ro_rank, A.nTris, np.int32),
        edge_t=_arr(A.edge_t, A.nEdges, np.int32),
        edge_vopp=_arr(A.edge_vopp, A.nEdges, np.int32),
        edge_ro_rank=_arr(A.edge_ro_rank, A.nEdges, np.int32),
        edge_nu_min=_arr(A.edge_nu_min, A.nEdges, np.int32),
        edge_nu_max=_arr(A.edge_nu_max, A.nEdges, np.int32),
        edge_attached=_arr(A.edge_attached, A.nEdges, np.int32),
        vert_nu_min=_arr(A.vert_nu_min, A.nPoints, np.int32),
        vert_nu_max=_arr(A.vert_nu_max, A.nPoints, np.int32),
        master_simplex=_arr(A.master_simplex, A.masterSize, np.int32),
        master_st=_arr(A.master_st, A.masterSize, np.int32),
        master_at=_arr(A.master_at, A.masterSize, np.int32),
        spec_pointers=_arr(A.spec_pointers, A.nSpecPointers, np.int32),
        exact_compares=int(A.exact_compares), exact_attach=int(A.exact_attach))
    lib().ora_alpha_free(C.byref(A))
    return r


def pt(p):
    a = np.ascontiguousarray(p, dtype=np.float64)
    return a


def orient_sign(a, b, c):
    a, b, c = (pt(v) for v in (a, b, c))
    return lib().ora_orient_sign(_dp(a), _dp(b), _dp(c))


def incircle_sign(a, b, c, d):
    a, b, c, d = (pt(v) for v in (a, b, c, d))
    return lib().ora_incircle_sign(_dp(a), _dp(b), _dp(c), _dp(d))


def incircle_sos(a, ia, b, ib, c, ic, d, id_):
    a, b, c, d = (pt(v) for v in (a, b, c, d))
    return lib().ora_incircle_sos(_dp(a), ia, _dp(b), ib, _dp(c), ic, _dp(d), id_)


def rho_compare(sta, A, stb, B):
    def three(P):
        P = [pt(p) for p in P] + [np.zeros(2)] * (3 - len(P))
        return P
    A3, B3 = three(A), three(B)
    return lib().ora_rho_compare(sta, _dp(A3[0]), _dp(A3[1]), _dp(A3[2]),
                                 stb, _dp(B3[0]), _dp(B3[1]), _dp(B3[2]))


def rho_tri(a, b, c):
    return lib().ora_rho_tri(_dp(pt(a)), _dp(pt(b)), _dp(pt(c)))


def rho_edge(a, b):
    return lib().ora_rho_edge(_dp(pt(a)), _dp(pt(b)))


def attached(i, j, k):
    return lib().ora_attached(_dp(pt(i)), _dp(pt(j)), _dp(pt(k)))


def pred_counters():
    v = [C.c_longlong() for _ in range(5)]
    lib().ora_pred_counters(*[C.byref(x) for x in v])
    return [x.value for x in v]
