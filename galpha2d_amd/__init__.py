"""galpha2d_amd -- MI355X-native 2D Delaunay triangulation + alpha shapes.

Python mirror of the reference host interface (gpudt/gpudt.h: gpudtComputeDT /
gpudtReleaseDTOutput, and AlphaShape/Alpha_Shapes_2D.h: Alpha_Shapes_2D::computeAlphaShapes)
over the C ABI of libgalpha.so (include/gpudt.h).  All compute runs in hand-written HIP
kernels for gfx950; this module only marshals arrays.  There is no CPU fallback: when the
shared library or a GPU is missing, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgalpha.so")
_lib = None


class GalphaError(RuntimeError):
    pass


class _Stats(C.Structure):
    _fields_ = [("nTris", C.c_int), ("nEdges", C.c_int), ("specSize", C.c_int),
                ("islandSites", C.c_int), ("pbaIterations", C.c_int),
                ("reconTris", C.c_int), ("shiftBad", C.c_int), ("deleted", C.c_int),
                ("missing", C.c_int), ("inserted", C.c_int), ("duplicates", C.c_int),
                ("insertRounds", C.c_int), ("flipPasses", C.c_int),
                ("flips", C.c_longlong), ("incircleTests", C.c_longlong),
                ("incircleExact", C.c_longlong), ("orientExact", C.c_longlong),
                ("alphaExactCompares", C.c_longlong), ("alphaExactAttach", C.c_longlong),
                ("alphaClusters", C.c_longlong), ("alphaMaxCluster", C.c_longlong),
                ("ms_pba", C.c_double), ("ms_recon", C.c_double), ("ms_shift", C.c_double),
                ("ms_missing", C.c_double), ("ms_boundary", C.c_double), ("ms_flip", C.c_double),
                ("ms_canon", C.c_double), ("ms_alpha", C.c_double), ("ms_total", C.c_double),
                ("frameSeen", C.c_longlong), ("frameChain", C.c_longlong), ("flipListed", C.c_longlong),
                ("frameBulk", C.c_int), ("earRounds", C.c_int),
                ("constraintRounds", C.c_int), ("ms_constraints", C.c_double),
                ("ms_alpha_rho", C.c_double), ("ms_alpha_order", C.c_double), ("ms_alpha_lists", C.c_double),
                ("stageListed", C.c_longlong * 3), ("stageIncircleTests", C.c_longlong * 3),
                ("stageIncircleExact", C.c_longlong * 3), ("stageOrientExact", C.c_longlong * 3),
                ("flipTailPasses", C.c_int), ("flipTailListed", C.c_longlong)]


class _Alpha(C.Structure):
    _fields_ = [("nPoints", C.c_int), ("nEdges", C.c_int), ("nTris", C.c_int),
                ("specSize", C.c_int), ("spectrum", C.POINTER(C.c_double)),
                ("tri_ro_rank", C.POINTER(C.c_int)), ("edge_t", C.POINTER(C.c_int)),
                ("edge_vopp", C.POINTER(C.c_int)), ("edge_ro_rank", C.POINTER(C.c_int)),
                ("edge_nu_min_rank", C.POINTER(C.c_int)), ("edge_nu_max_rank", C.POINTER(C.c_int)),
                ("vert_nu_min_rank", C.POINTER(C.c_int)), ("vert_nu_max_rank", C.POINTER(C.c_int)),
                ("masterSize", C.c_int), ("master_simplex", C.POINTER(C.c_int)),
                ("master_st", C.POINTER(C.c_int)), ("master_at", C.POINTER(C.c_int)),
                ("nSpecPointers", C.c_int), ("spec_pointers", C.POINTER(C.c_int))]


# every symbol include/gpudt.h declares
EXPORTS = ["gpudtComputeDT", "gpudtReleaseDTOutput", "gpudtGetLastError", "gpudtComputeAlpha",
           "gpudtReleaseAlpha", "galphaCreate", "galphaDestroy", "galphaLastError",
           "galphaComputeHost", "galphaComputeDevice", "galphaDeviceTriangles",
           "galphaCopyTriangles", "galphaGetStats", "galphaVoronoi", "galphaCopyAlpha",
           "galphaTextureFor", "galphaAutotune", "galphaBatch", "galphaBatchDevice", "galphaBatchTriangles", "galphaBatchCopyAlpha", "galphaFree", "galphaAutoTexture", "galphaKernelProbe",
           "galphaKernelProbeResult", "galphaServerCreate", "galphaServerDestroy", "galphaServerSubmit",
           "galphaServerWait", "galphaServerTriangles", "galphaServerAlpha", "galphaServerRelease",
           "galphaServerStats", "galphaAlphaRank", "galphaClassify", "galphaClassifyDevice",
           "galphaReadPoints", "galphaReleaseParams", "galphaWritePoints", "galphaDumpValues",
           "galphaPackResult", "galphaUnpackResult", "galphaSetConstraints", "galphaDeviceCount",
           "galphaDeviceMalloc", "galphaDeviceFree", "galphaDeviceCopy", "galphaDeviceSynchronize",
           "galphaStreamCreate", "galphaStreamDestroy", "galphaStreamSynchronize", "galphaServerSetGroupPoints",
           "galphaServerFootprint", "galphaComputeDeviceNoAlloc"]


class _Params(C.Structure):
    """GPUDTPARAMS (gpudt/gpudt.h:106-119)."""
    _fields_ = [("minX", C.c_double), ("minY", C.c_double), ("maxX", C.c_double), ("maxY", C.c_double),
                ("nPoints", C.c_int), ("points", C.POINTER(C.c_double)), ("nConstraints", C.c_int),
                ("constraints", C.POINTER(C.c_int)), ("fboSize", C.c_int)]


class _Output(C.Structure):
    """GPUDTOUTPUT (gpudt/gpudt.h:125-129)."""
    _fields_ = [("triangles", C.POINTER(C.c_int)), ("nTris", C.c_int)]


def lib():
    """Load libgalpha.so (raises GalphaError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GalphaError("libgalpha.so not built: run __graft_entry__.build() or make -C galpha2d_amd/csrc")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    L.galphaCreate.argtypes = [C.c_int]
    L.galphaCreate.restype = vp
    L.galphaDestroy.argtypes = [vp]
    L.galphaLastError.argtypes = [vp]
    L.galphaLastError.restype = C.c_char_p
    L.gpudtGetLastError.restype = C.c_char_p
    L.galphaComputeHost.argtypes = [vp, dp, C.c_int, C.c_int, ip, ip]
    L.galphaComputeHost.restype = C.c_int
    L.galphaComputeDevice.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, vp]
    L.galphaComputeDevice.restype = C.c_int
    L.galphaComputeDeviceNoAlloc.argtypes = [C.c_int, vp, C.c_int, C.c_int, vp, C.c_int, vp]
    L.galphaComputeDeviceNoAlloc.restype = C.c_int
    L.galphaDeviceTriangles.argtypes = [vp, C.POINTER(vp), ip]
    L.galphaCopyTriangles.argtypes = [vp, ip, C.c_int]
    L.galphaCopyTriangles.restype = C.c_int
    L.galphaGetStats.argtypes = [vp, C.POINTER(_Stats)]
    L.galphaVoronoi.argtypes = [vp, dp, C.c_int, C.c_int, ip, ip, C.c_int]
    L.galphaVoronoi.restype = C.c_int
    L.galphaCopyAlpha.argtypes = [vp, C.POINTER(C.POINTER(_Alpha))]
    L.galphaCopyAlpha.restype = C.c_int
    L.galphaAlphaRank.argtypes = [vp, C.c_double, ip]
    L.galphaAlphaRank.restype = C.c_int
    L.galphaClassify.argtypes = [vp, C.c_int, vp, vp, vp]
    L.galphaClassify.restype = C.c_int
    L.galphaClassifyDevice.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    L.galphaClassifyDevice.restype = C.c_int
    L.gpudtReleaseAlpha.argtypes = [C.POINTER(_Alpha)]
    L.galphaKernelProbe.argtypes = [C.c_char_p]
    L.galphaKernelProbeResult.argtypes = [ip, C.POINTER(C.c_double)]
    L.galphaAutoTexture.argtypes = [C.c_int]
    L.galphaAutoTexture.restype = C.c_int
    L.galphaTextureFor.argtypes = [vp, C.c_int]
    L.galphaTextureFor.restype = C.c_int
    L.galphaAutotune.argtypes = [vp, dp, ip, C.c_int, C.c_int, dp]
    L.galphaAutotune.restype = C.c_int
    L.galphaBatch.argtypes = [vp, dp, ip, C.c_int, ip, C.c_int, C.POINTER(ip), C.POINTER(ip)]
    L.galphaBatch.restype = C.c_int
    L.galphaBatchDevice.argtypes = [vp, vp, ip, C.c_int, ip, C.c_int, vp]
    L.galphaBatchDevice.restype = C.c_int
    L.galphaBatchTriangles.argtypes = [vp, C.c_int, ip, C.c_int]
    L.galphaBatchTriangles.restype = C.c_int
    L.galphaBatchCopyAlpha.argtypes = [vp, C.c_int, C.POINTER(C.POINTER(_Alpha))]
    L.galphaBatchCopyAlpha.restype = C.c_int
    L.galphaFree.argtypes = [vp]
    L.galphaServerCreate.argtypes = [C.c_int, C.c_int, C.c_size_t, C.c_int]
    L.galphaServerCreate.restype = vp
    L.galphaServerDestroy.argtypes = [vp]
    L.galphaServerSubmit.argtypes = [vp, dp, C.c_int, C.c_int, C.c_int]
    L.galphaServerSubmit.restype = C.c_longlong
    L.galphaServerWait.argtypes = [vp, C.c_longlong, dp, dp, ip]
    L.galphaServerWait.restype = C.c_int
    L.galphaServerTriangles.argtypes = [vp, C.c_longlong, ip, C.c_int]
    L.galphaServerTriangles.restype = C.c_int
    L.galphaServerAlpha.argtypes = [vp, C.c_longlong, C.POINTER(C.POINTER(_Alpha))]
    L.galphaServerAlpha.restype = C.c_int
    L.galphaServerRelease.argtypes = [vp, C.c_longlong]
    L.galphaServerRelease.restype = C.c_int
    L.galphaServerStats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_size_t)]
    L.galphaServerStats.restype = C.c_int
    L.galphaServerSetGroupPoints.argtypes = [vp, C.c_longlong]
    L.galphaServerSetGroupPoints.restype = C.c_int
    L.galphaServerFootprint.argtypes = [C.c_int, C.c_int]
    L.galphaServerFootprint.restype = C.c_size_t
    L.galphaDeviceCount.restype = C.c_int
    L.galphaDeviceMalloc.argtypes = [C.c_int, C.c_size_t]
    L.galphaDeviceMalloc.restype = vp
    L.galphaDeviceFree.argtypes = [vp]
    L.galphaDeviceCopy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    L.galphaDeviceCopy.restype = C.c_int
    L.galphaDeviceSynchronize.restype = C.c_int
    L.galphaStreamCreate.argtypes = [C.c_int]
    L.galphaStreamCreate.restype = vp
    L.galphaStreamDestroy.argtypes = [vp]
    L.galphaStreamDestroy.restype = C.c_int
    L.galphaStreamSynchronize.argtypes = [vp]
    L.galphaStreamSynchronize.restype = C.c_int
    L.galphaSetConstraints.argtypes = [vp, ip, C.c_int]
    L.galphaSetConstraints.restype = C.c_int
    L.galphaReadPoints.argtypes = [C.c_char_p, C.POINTER(C.POINTER(_Params))]
    L.galphaReadPoints.restype = C.c_int
    L.galphaReleaseParams.argtypes = [C.POINTER(_Params)]
    L.galphaWritePoints.argtypes = [C.c_char_p, dp, C.c_int]
    L.galphaWritePoints.restype = C.c_int
    L.galphaDumpValues.argtypes = [C.c_char_p, C.POINTER(_Params), C.POINTER(_Output), C.POINTER(_Alpha), C.c_int]
    L.galphaDumpValues.restype = C.c_int
    L.galphaPackResult.argtypes = [C.c_int, ip, C.c_int, C.POINTER(_Alpha), vp, C.c_longlong]
    L.galphaPackResult.restype = C.c_longlong
    L.galphaUnpackResult.argtypes = [vp, C.c_longlong, ip, C.POINTER(ip), ip, C.POINTER(C.POINTER(_Alpha))]
    L.galphaUnpackResult.restype = C.c_int
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class DeviceBuffer:
    """Device memory from the library's own HIP runtime, filled from a numpy array (or `nbytes`
    uninitialised bytes): `.ptr` is the device address for the device-resident entry points."""

    def __init__(self, array=None, nbytes=None, device=0):
        self._L = lib()
        self.array = None if array is None else np.ascontiguousarray(array)
        self.nbytes = int(self.array.nbytes if array is not None else nbytes)
        self.ptr = self._L.galphaDeviceMalloc(int(device), self.nbytes)
        if not self.ptr:
            raise GalphaError("galphaDeviceMalloc: " + self._L.gpudtGetLastError().decode())
        if self.array is not None and self.nbytes:
            self.upload(self.array)

    def upload(self, array):
        a = np.ascontiguousarray(array)
        if a.nbytes > self.nbytes:
            raise GalphaError("upload larger than the buffer")
        if self._L.galphaDeviceCopy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, 0) < 0:
            raise GalphaError("galphaDeviceCopy: " + self._L.gpudtGetLastError().decode())

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if self._L.galphaDeviceCopy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes, 1) < 0:
            raise GalphaError("galphaDeviceCopy: " + self._L.gpudtGetLastError().decode())
        return out

    def copy_from(self, other, nbytes=None):
        """Device-to-device copy of `nbytes` (default: all of this buffer) from another buffer."""
        nb = self.nbytes if nbytes is None else int(nbytes)
        if self._L.galphaDeviceCopy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), nb, 2) < 0:
            raise GalphaError("galphaDeviceCopy: " + self._L.gpudtGetLastError().decode())

    def free(self):
        if self.ptr:
            self._L.galphaDeviceFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream of the library's own runtime (`.handle` for the `stream` arguments)."""

    def __init__(self, device=0):
        self._L = lib()
        self.handle = self._L.galphaStreamCreate(int(device))
        if not self.handle:
            raise GalphaError("galphaStreamCreate: " + self._L.gpudtGetLastError().decode())

    def synchronize(self):
        if self._L.galphaStreamSynchronize(C.c_void_p(self.handle)) < 0:
            raise GalphaError(self._L.gpudtGetLastError().decode())

    def close(self):
        if self.handle:
            self._L.galphaStreamDestroy(C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count():
    return lib().galphaDeviceCount()


def synchronize():
    if lib().galphaDeviceSynchronize() < 0:
        raise GalphaError(lib().gpudtGetLastError().decode())


def compute_device_noalloc(d_xy_ptr, n, fbo_size, d_tris_ptr, capacity, device=0, stream=None):
    """Delaunay triangulation of n device points (x, y doubles at d_xy_ptr) that allocates no
    device memory: the run's buffers come from the library's static per-device scratch and the
    canonical triangles (9 ints each) go to d_tris_ptr (room for `capacity` of them).  Returns the
    triangle count."""
    L = lib()
    F = L.galphaComputeDeviceNoAlloc(int(device), C.c_void_p(d_xy_ptr), int(n), int(fbo_size),
                                     C.c_void_p(d_tris_ptr), int(capacity), C.c_void_p(stream or 0))
    if F < 0:
        raise GalphaError("galphaComputeDeviceNoAlloc: " + L.gpudtGetLastError().decode())
    return F


def auto_texture(n):
    return lib().galphaAutoTexture(int(n))


def _alpha_to_numpy(A):
    a = A.contents

    def arr(p, n, dt):
        if n <= 0:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)
    return dict(
        spectrum=arr(a.spectrum, a.specSize, np.float64),
        tri_ro_rank=arr(a.tri_ro_rank, a.nTris, np.int32),
        edge_t=arr(a.edge_t, a.nEdges, np.int32), edge_vopp=arr(a.edge_vopp, a.nEdges, np.int32),
        edge_ro_rank=arr(a.edge_ro_rank, a.nEdges, np.int32),
        edge_nu_min=arr(a.edge_nu_min_rank, a.nEdges, np.int32),
        edge_nu_max=arr(a.edge_nu_max_rank, a.nEdges, np.int32),
        vert_nu_min=arr(a.vert_nu_min_rank, a.nPoints, np.int32),
        vert_nu_max=arr(a.vert_nu_max_rank, a.nPoints, np.int32),
        master_simplex=arr(a.master_simplex, a.masterSize, np.int32),
        master_st=arr(a.master_st, a.masterSize, np.int32),
        master_at=arr(a.master_at, a.masterSize, np.int32),
        spec_pointers=arr(a.spec_pointers, a.nSpecPointers, np.int32))


_ALPHA_ARRAYS = (("spectrum", "spectrum", C.c_double), ("tri_ro_rank", "tri_ro_rank", C.c_int),
                 ("edge_t", "edge_t", C.c_int), ("edge_vopp", "edge_vopp", C.c_int),
                 ("edge_ro_rank", "edge_ro_rank", C.c_int), ("edge_nu_min_rank", "edge_nu_min", C.c_int),
                 ("edge_nu_max_rank", "edge_nu_max", C.c_int), ("vert_nu_min_rank", "vert_nu_min", C.c_int),
                 ("vert_nu_max_rank", "vert_nu_max", C.c_int), ("master_simplex", "master_simplex", C.c_int),
                 ("master_st", "master_st", C.c_int), ("master_at", "master_at", C.c_int),
                 ("spec_pointers", "spec_pointers", C.c_int))


def _alpha_struct(alpha, n_points, n_tris):
    """GPUDTALPHA viewing the arrays of an alpha dict; returns (struct, arrays kept alive)."""
    A = _Alpha()
    keep = []
    for field, key, ct in _ALPHA_ARRAYS:
        a = np.ascontiguousarray(alpha[key], dtype=np.float64 if ct is C.c_double else np.int32)
        keep.append(a)
        setattr(A, field, a.ctypes.data_as(C.POINTER(ct)))
    A.nPoints, A.nTris = int(n_points), int(n_tris)
    A.nEdges, A.specSize = len(alpha["edge_t"]), len(alpha["spectrum"])
    A.masterSize, A.nSpecPointers = len(alpha["master_simplex"]), len(alpha["spec_pointers"])
    return A, keep


def read_points(path):
    """Reference point file (main.cu:117-161) -> (points [n, 2] float64, params dict with the
    bounds, nConstraints and the texture side the library picks)."""
    L = lib()
    P = C.POINTER(_Params)()
    if L.galphaReadPoints(os.fsencode(path), C.byref(P)) < 0:
        raise GalphaError(L.gpudtGetLastError().decode())
    try:
        p = P.contents
        n = p.nPoints
        xy = np.ctypeslib.as_array(p.points, shape=(n, 2)).copy() if n else np.zeros((0, 2))
        info = dict(minX=p.minX, minY=p.minY, maxX=p.maxX, maxY=p.maxY, nConstraints=p.nConstraints,
                    fboSize=p.fboSize)
    finally:
        L.galphaReleaseParams(P)
    return xy, info


def write_points(path, points):
    """Write points [n, 2] in the reference point-file format (17 significant digits)."""
    L = lib()
    xy = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    if L.galphaWritePoints(os.fsencode(path), _dp(xy), xy.shape[0]) < 0:
        raise GalphaError(L.gpudtGetLastError().decode())


def dump_values(path, points, tris, alpha, full=False):
    """Alpha_Shapes_2D::dumpValues (AlphaShape/Alpha_Shapes_2D.cu:749-793) of a result."""
    L = lib()
    xy = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    t9 = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 9)
    A, keep = _alpha_struct(alpha, xy.shape[0], t9.shape[0])
    prm = _Params(0, 0, 0, 0, xy.shape[0], _dp(xy), 0, None, 0)
    out = _Output(_ip(t9), t9.shape[0])
    if L.galphaDumpValues(os.fsencode(path), C.byref(prm), C.byref(out), C.byref(A), int(bool(full))) < 0:
        raise GalphaError(L.gpudtGetLastError().decode())
    del keep


def pack_result(n_points, tris, alpha=None):
    """Binary result payload (bytes) of a triangulation and, optionally, its alpha result."""
    L = lib()
    t9 = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 9)
    A, keep = _alpha_struct(alpha, n_points, t9.shape[0]) if alpha is not None else (None, None)
    pa = C.byref(A) if A is not None else None
    size = L.galphaPackResult(int(n_points), _ip(t9), t9.shape[0], pa, None, 0)
    if size < 0:
        raise GalphaError(L.gpudtGetLastError().decode())
    buf = C.create_string_buffer(size)
    if L.galphaPackResult(int(n_points), _ip(t9), t9.shape[0], pa, buf, size) != size:
        raise GalphaError(L.gpudtGetLastError().decode())
    del keep
    return buf.raw


def unpack_result(payload):
    """Inverse of pack_result: (n_points, triangles [F, 9], alpha dict or None)."""
    L = lib()
    b = bytes(payload)
    n, F = C.c_int(), C.c_int()
    t = C.POINTER(C.c_int)()
    A = C.POINTER(_Alpha)()
    if L.galphaUnpackResult(b, len(b), C.byref(n), C.byref(t), C.byref(F), C.byref(A)) < 0:
        raise GalphaError(L.gpudtGetLastError().decode())
    try:
        tris = np.ctypeslib.as_array(t, shape=(F.value, 9)).copy() if F.value else np.zeros((0, 9), np.int32)
        alpha = _alpha_to_numpy(A) if A else None
    finally:
        L.galphaFree(C.cast(t, C.c_void_p))
        if A:
            L.gpudtReleaseAlpha(A)
    return n.value, tris, alpha


class Context:
    """One device context (buffers + HIP stream) -- the analogue of the reference's global
    GPU-DT state (gpudt.cpp:64-77, Cuda/cudaMain.cu:69-92), made reentrant."""

    def __init__(self, device=0):
        self._L = lib()
        self._h = self._L.galphaCreate(int(device))
        if not self._h:
            raise GalphaError("galphaCreate failed: " + self._L.gpudtGetLastError().decode())

    def close(self):
        if self._h:
            self._L.galphaDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, what):
        raise GalphaError(f"{what}: {self._L.galphaLastError(self._h).decode()}")

    def set_constraints(self, pairs):
        """Constraint edges (pairs of point indices, [k, 2]) for the following single-set runs of
        this context; None or an empty array clears them.  The runs then give the constrained DT."""
        c = np.ascontiguousarray(np.zeros((0, 2)) if pairs is None else pairs, dtype=np.int32).reshape(-1, 2)
        self._cons = c   # the library copies the pairs; kept for reference
        if self._L.galphaSetConstraints(self._h, _ip(c) if c.size else None, c.shape[0]) < 0:
            self._err("galphaSetConstraints")

    def delaunay(self, xy, fbo_size=None):
        """Canonical DT of xy ([n, 2] float64) as int32 [F, 9] (tri[3], vtx[3], tmp[3])."""
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        n = xy.shape[0]
        T = int(fbo_size or auto_texture(n))
        cap = C.c_int(0)
        F = self._L.galphaComputeHost(self._h, _dp(xy), n, T, None, C.byref(cap))
        if F < 0:
            self._err("galphaComputeHost")
        out = np.zeros((max(F, 0), 9), dtype=np.int32)
        if F > 0:
            got = self._L.galphaCopyTriangles(self._h, _ip(out), F)
            if got < 0:
                self._err("galphaCopyTriangles")
        return out

    def batch(self, sets, fbo_sizes=None, with_alpha=False):
        """All point sets ([n_i, 2] float64 each) through one batched pass of the pipeline.
        Returns the canonical triangles of every set (int32 [F_i, 9], set-local indices)."""
        sets = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 2) for x in sets]
        ns = len(sets)
        offs = np.zeros(ns + 1, dtype=np.int32)
        offs[1:] = np.cumsum([x.shape[0] for x in sets])
        xy = np.ascontiguousarray(np.concatenate(sets) if ns else np.zeros((0, 2)))
        Ts = None
        if fbo_sizes is not None:
            Ts = np.ascontiguousarray(fbo_sizes, dtype=np.int32)
        tp, op = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        r = self._L.galphaBatch(self._h, _dp(xy), _ip(offs), ns, _ip(Ts) if Ts is not None else None,
                                1 if with_alpha else 0, C.byref(tp), C.byref(op))
        if r < 0:
            self._err("galphaBatch")
        try:
            toff = np.ctypeslib.as_array(op, shape=(ns + 1,)).copy()
            F = int(toff[-1])
            tris = np.ctypeslib.as_array(tp, shape=(max(F, 1) * 9,))[: F * 9].copy().reshape(F, 9)
        finally:
            self._L.galphaFree(C.cast(tp, C.c_void_p))
            self._L.galphaFree(C.cast(op, C.c_void_p))
        return [tris[toff[k]:toff[k + 1]] for k in range(ns)]

    def texture_for(self, n):
        return self._L.galphaTextureFor(self._h, int(n))

    def autotune(self, sets, reps=2):
        """Tunes the pixels-per-point density of each size class on a sample of point sets;
        returns the 4 chosen densities."""
        sets = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 2) for x in sets]
        offs = np.zeros(len(sets) + 1, dtype=np.int32)
        offs[1:] = np.cumsum([x.shape[0] for x in sets])
        xy = np.ascontiguousarray(np.concatenate(sets))
        out = np.zeros(4)
        if self._L.galphaAutotune(self._h, _dp(xy), _ip(offs), len(sets), int(reps), _dp(out)) < 0:
            self._err("galphaAutotune")
        return out

    def batch_device(self, d_xy_ptr, offsets, fbo_sizes=None, with_alpha=False, stream=None):
        """Device-resident batch: d_xy_ptr is the device address of all sets' (x, y) doubles,
        offsets[nsets + 1] in points.  Results stay on the device; returns the triangle total."""
        offs = np.ascontiguousarray(offsets, dtype=np.int32)
        Ts = None if fbo_sizes is None else np.ascontiguousarray(fbo_sizes, dtype=np.int32)
        F = self._L.galphaBatchDevice(self._h, C.c_void_p(d_xy_ptr), _ip(offs), len(offs) - 1,
                                      _ip(Ts) if Ts is not None else None, 1 if with_alpha else 0,
                                      C.c_void_p(stream or 0))
        if F < 0:
            self._err("galphaBatchDevice")
        return F

    def batch_triangles(self, k):
        """Triangles of set k of the last batch (set-local indices)."""
        F = self._L.galphaBatchTriangles(self._h, int(k), None, 0)
        if F < 0:
            self._err("galphaBatchTriangles")
        out = np.zeros((F, 9), dtype=np.int32)
        if F > 0 and self._L.galphaBatchTriangles(self._h, int(k), _ip(out), F) < 0:
            self._err("galphaBatchTriangles")
        return out

    def batch_alpha(self, k):
        """Alpha result of set k of the last batch(with_alpha=True)."""
        A = C.POINTER(_Alpha)()
        if self._L.galphaBatchCopyAlpha(self._h, int(k), C.byref(A)) < 0:
            self._err("galphaBatchCopyAlpha")
        try:
            return _alpha_to_numpy(A)
        finally:
            self._L.gpudtReleaseAlpha(A)

    def compute_device(self, d_xy_ptr, n, fbo_size, with_alpha=False, stream=None):
        """Device-resident run; d_xy_ptr is a device address of n (x, y) doubles."""
        F = self._L.galphaComputeDevice(self._h, C.c_void_p(d_xy_ptr), int(n), int(fbo_size),
                                        1 if with_alpha else 0, C.c_void_p(stream or 0))
        if F < 0:
            self._err("galphaComputeDevice")
        return F

    def triangles(self):
        p, nt = C.c_void_p(), C.c_int()
        self._L.galphaDeviceTriangles(self._h, C.byref(p), C.byref(nt))
        out = np.zeros((nt.value, 9), dtype=np.int32)
        if nt.value:
            if self._L.galphaCopyTriangles(self._h, _ip(out), nt.value) < 0:
                self._err("galphaCopyTriangles")
        return out

    def alpha(self):
        """Alpha result of the last compute_device(with_alpha=True)."""
        A = C.POINTER(_Alpha)()
        if self._L.galphaCopyAlpha(self._h, C.byref(A)) < 0:
            self._err("galphaCopyAlpha")
        try:
            return _alpha_to_numpy(A)
        finally:
            self._L.gpudtReleaseAlpha(A)

    # alpha-complex classes (galphaClassify): 0 exterior, 1 singular, 2 regular, 3 interior
    EXTERIOR, SINGULAR, REGULAR, INTERIOR = 0, 1, 2, 3

    def alpha_rank(self, alpha):
        """Largest rank r with spectrum[r] <= alpha in the last single-set alpha result."""
        r = C.c_int()
        if self._L.galphaAlphaRank(self._h, float(alpha), C.byref(r)) < 0:
            self._err("galphaAlphaRank")
        return r.value

    def classify(self, rank, alpha):
        """Classes of every vertex, edge (alpha-result edge order) and triangle at `rank`:
        (vert, edge, tri) uint8 arrays; `alpha` is this context's alpha() result (array sizes)."""
        import numpy as np
        v = np.zeros(len(alpha["vert_nu_min"]), np.uint8)
        e = np.zeros(len(alpha["edge_t"]), np.uint8)
        t = np.zeros(len(alpha["tri_ro_rank"]), np.uint8)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a.size else None
        if self._L.galphaClassify(self._h, int(rank), ptr(v), ptr(e), ptr(t)) < 0:
            self._err("galphaClassify")
        return v, e, t

    def kernel_probe(self, kernel_name, fn):
        """Run fn() with HIP events around every launch of `kernel_name` (on its stream).
        Returns {"launches": k, "total_ms": t}."""
        L = self._L
        L.galphaKernelProbe(kernel_name.encode())
        try:
            fn()
            k, t = C.c_int(), C.c_double()
            if L.galphaKernelProbeResult(C.byref(k), C.byref(t)) < 0:
                raise GalphaError("kernel probe failed")
            return {"launches": k.value, "total_ms": t.value}
        finally:
            L.galphaKernelProbe(None)

    def stats(self):
        s = _Stats()
        self._L.galphaGetStats(self._h, C.byref(s))
        out = {}
        for k, _ in _Stats._fields_:
            v = getattr(s, k)
            out[k] = list(v) if isinstance(v, C.Array) else v
        return out

    def voronoi(self, xy, fbo_size):
        """Discrete Voronoi texture: (labels [T, T] x-major, pix [n + frame], n_frame)."""
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        n, T = xy.shape[0], int(fbo_size)
        labels = np.zeros(T * T, dtype=np.int32)
        cap = n + 4 * T
        pix = np.zeros(cap, dtype=np.int32)
        nf = self._L.galphaVoronoi(self._h, _dp(xy), n, T, _ip(labels), _ip(pix), cap)
        if nf < 0:
            self._err("galphaVoronoi")
        return labels.reshape(T, T), pix[:n + nf].copy(), nf


_default = None


def server_footprint(n, fbo_size=0):
    """Device bytes the serving admission rule counts for one request of n points."""
    return int(lib().galphaServerFootprint(int(n), int(fbo_size)))


class Server:
    """Request loop on one device: `slots` contexts with their own streams, admission by HBM
    footprint, continuous batching of queued requests (see include/gpudt.h, galphaServer*)."""

    def __init__(self, device=0, slots=4, hbm_budget_bytes=64 << 30, max_group=64, max_group_points=0):
        self._L = lib()
        self._h = self._L.galphaServerCreate(int(device), int(slots), int(hbm_budget_bytes), int(max_group))
        if not self._h:
            raise GalphaError("galphaServerCreate failed: " + self._L.gpudtGetLastError().decode())
        if max_group_points and self._L.galphaServerSetGroupPoints(self._h, int(max_group_points)) < 0:
            raise GalphaError("galphaServerSetGroupPoints failed")

    def submit(self, xy, fbo_size=0, with_alpha=True):
        """Queues one request; returns its ticket."""
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        t = self._L.galphaServerSubmit(self._h, _dp(xy), xy.shape[0], int(fbo_size), 1 if with_alpha else 0)
        if t == -2:
            raise GalphaError("request exceeds the server's HBM budget")
        if t < 0:
            raise GalphaError("galphaServerSubmit failed")
        return t

    def wait(self, ticket):
        """Blocks until the ticket is done: (triangles, latency_ms, queue_ms, group)."""
        lat, q, g = C.c_double(), C.c_double(), C.c_int()
        F = self._L.galphaServerWait(self._h, C.c_longlong(ticket), C.byref(lat), C.byref(q), C.byref(g))
        if F < 0:
            raise GalphaError(f"request {ticket} failed")
        return F, lat.value, q.value, g.value

    def result(self, ticket, alpha=True):
        """Triangles (int32 [F, 9]) and, if requested at submit, the alpha dict; forgets the ticket."""
        F = self._L.galphaServerTriangles(self._h, C.c_longlong(ticket), None, 0)
        if F < 0:
            raise GalphaError(f"request {ticket} has no result")
        tris = np.zeros((F, 9), dtype=np.int32)
        if F > 0:
            self._L.galphaServerTriangles(self._h, C.c_longlong(ticket), _ip(tris), F)
        a = None
        if alpha:
            A = C.POINTER(_Alpha)()
            if self._L.galphaServerAlpha(self._h, C.c_longlong(ticket), C.byref(A)) == 0:
                try:
                    a = _alpha_to_numpy(A)
                finally:
                    self._L.gpudtReleaseAlpha(A)
        self._L.galphaServerRelease(self._h, C.c_longlong(ticket))
        return tris, a

    def release(self, ticket):
        self._L.galphaServerRelease(self._h, C.c_longlong(ticket))

    def stats(self):
        p, s, h = C.c_longlong(), C.c_longlong(), C.c_size_t()
        self._L.galphaServerStats(self._h, C.byref(p), C.byref(s), C.byref(h))
        return {"passes": p.value, "served": s.value, "held_bytes": h.value}

    def close(self):
        if self._h:
            self._L.galphaServerDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def gpudt_compute_dt(points, fbo_size, constraints=None):
    """Mirror of gpudtComputeDT (gpudt/gpudt.h:131): points [n, 2] -> triangles [F, 9]; with
    constraints (pairs of point indices, GPUDTPARAMS::constraints) the constrained DT."""
    ctx = default_context()
    ctx.set_constraints(constraints)
    try:
        return ctx.delaunay(points, fbo_size)
    finally:
        ctx.set_constraints(None)


def compute_alpha_shapes(points, fbo_size):
    """Mirror of Alpha_Shapes_2D(pInput, pDelaunay).computeAlphaShapes()
    (AlphaShape/Alpha_Shapes_2D.cu:206-233): DT + alpha spectrum, intervals and master list."""
    ctx = default_context()
    xy = np.ascontiguousarray(points, dtype=np.float64)
    d = DeviceBuffer(xy)
    ctx.compute_device(d.ptr, xy.shape[0], fbo_size, with_alpha=True)
    tris = ctx.triangles()
    res = ctx.alpha()
    d.free()
    return tris, res
