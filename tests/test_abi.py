"""The C-ABI library: it loads without a GPU, exports every entry point include/gpudt.h
declares, and its structs are byte-compatible with the reference's gpudt/gpudt.h."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gpudt.h")
LIB = os.path.join(ROOT, "galpha2d_amd", "libgalpha.so")
REF_HEADER_DIR = "/root/reference/gpudt"

PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "gpudt.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(gpudtVertex), sizeof(gpudtTriangle), sizeof(GPUDTPARAMS), sizeof(GPUDTOUTPUT));
  printf("%zu %zu %zu\n", offsetof(gpudtTriangle, tri), offsetof(gpudtTriangle, vtx), offsetof(gpudtTriangle, tmp));
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", offsetof(GPUDTPARAMS, minX), offsetof(GPUDTPARAMS, minY),
         offsetof(GPUDTPARAMS, maxX), offsetof(GPUDTPARAMS, maxY), offsetof(GPUDTPARAMS, nPoints),
         offsetof(GPUDTPARAMS, points), offsetof(GPUDTPARAMS, nConstraints), offsetof(GPUDTPARAMS, constraints),
         offsetof(GPUDTPARAMS, fboSize));
  printf("%zu %zu\n", offsetof(GPUDTOUTPUT, triangles), offsetof(GPUDTOUTPUT, nTris));
  return 0;
}
"""


def _probe(include_dir, lang="c"):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe." + ("c" if lang == "c" else "cpp"))
        with open(src, "w") as f:
            f.write(PROBE)
        exe = os.path.join(d, "probe")
        cc = "gcc" if lang == "c" else "g++"
        subprocess.run([cc, "-I", include_dir, src, "-o", exe], check=True, capture_output=True)
        return subprocess.run([exe], check=True, capture_output=True, text=True).stdout


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gpudt|galpha)[A-Za-z0-9_]*)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libgalpha.so not built")
    L = ctypes.CDLL(LIB)
    names = declared_functions()
    assert "gpudtComputeDT" in names and "gpudtReleaseDTOutput" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_mirror_lists_every_export():
    import galpha2d_amd
    assert sorted(galpha2d_amd.EXPORTS) == declared_functions()


def test_struct_layout():
    out = _probe(os.path.join(ROOT, "include")).split("\n")
    assert out[0] == "16 36 72 16"
    assert out[1] == "0 12 24"
    assert out[2] == "0 8 16 24 32 40 48 56 64"
    assert out[3] == "0 8"


def test_struct_layout_matches_reference_header():
    if not os.path.exists(os.path.join(REF_HEADER_DIR, "gpudt.h")):
        pytest.skip("reference tree not present")
    ours = _probe(os.path.join(ROOT, "include"), "cpp")
    ref = _probe(REF_HEADER_DIR, "cpp")
    assert ours == ref


def test_auto_texture_rule():
    if not os.path.exists(LIB):
        pytest.skip("libgalpha.so not built")
    import galpha2d_amd as G
    assert G.auto_texture(1000) == 256
    assert G.auto_texture(1048576) == 4096
    assert G.auto_texture(2000000) == 8192


def test_no_process_global_work_state():
    """Contexts run concurrently (one per serving slot), so the library may keep no work buffers or
    pointers in process globals: the only writable globals are the mutex-protected default context
    and kernel-probe state, the thread-local error text, the no-allocation entry point's per-device
    scratch table (one lock per device), the host-side stand-in of its device array, the
    thread-local pointer to the region a no-allocation run carves from, and C++ runtime
    bookkeeping."""
    import re
    import shutil
    import subprocess
    import galpha2d_amd as G
    if shutil.which("nm") is None:
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-C", "--defined-only", G.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    allowed = {"g_default", "g_default_mu", "g_last_error[abi:cxx11]", "(anonymous namespace)::g_ev",
               "(anonymous namespace)::g_mu", "_ZN12_GLOBAL__N_14g_onE.0", "(anonymous namespace)::g_name[abi:cxx11]",
               "(anonymous namespace)::g_used", "std::__ioinit", "__tls_guard", "__fini", "__init",
               "(anonymous namespace)::g_dev", "scratch", "t_arena"}
    runtime = re.compile(r"rocprim|__hip|^_ZZ|guard variable|typeinfo|vtable|__dso_handle|__TMC_END__|completed\.|"
                         r"__init_array|__fini_array|__frame_dummy|__do_global_dtors|DW\.ref|^_DYNAMIC|"
                         r"^_GLOBAL_OFFSET_TABLE_|^__do_(init|fini)\.")
    bad = []
    for line in out.splitlines():
        parts = line.split(None, 2)
        if len(parts) < 3 or parts[1] not in "bBdD":
            continue
        name = parts[2]
        if re.match(r"k_[a-z0-9_]+\(", name) or runtime.search(name) or name in allowed:
            continue
        bad.append(name)
    assert not bad, f"process-global writable state: {bad}"


def test_python_mirrors_match_c_layouts():
    """The ctypes mirrors of galpha_stats and GPUDTALPHA have the C sizes and field offsets."""
    import ctypes as C
    import subprocess
    import tempfile
    import galpha2d_amd as G
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "gpudt.h"
int main(void) {
    printf("%zu %zu %zu %zu\n", sizeof(galpha_stats), offsetof(galpha_stats, flips),
           offsetof(galpha_stats, ms_pba), offsetof(galpha_stats, frameSeen));
    printf("%zu %zu %zu\n", sizeof(GPUDTALPHA), offsetof(GPUDTALPHA, spectrum),
           offsetof(GPUDTALPHA, spec_pointers));
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        exe = os.path.join(d, "p")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    S, A = G._Stats, G._Alpha
    assert out[0] == f"{C.sizeof(S)} {S.flips.offset} {S.ms_pba.offset} {S.frameSeen.offset}"
    assert out[1] == f"{C.sizeof(A)} {A.spectrum.offset} {A.spec_pointers.offset}"


# ---- mesh navigation (include/gpudt.h, the reference's gpudt/gpudt.h:137-244) ------------------
NAV_PROBE = r"""
#include <stdio.h>
#include <string.h>
#include "gpudt.h"
/* unit square 0 (0,0), 1 (1,0), 2 (1,1), 3 (0,1) split along 0-2: t0 = (0, 1, 2), t1 = (0, 2, 3);
   t0's edge opposite vtx[1] is t1's edge opposite vtx[2] */
static gpudtTriangle T[2] = {{{-1, (1 << 2) | 2, -1}, {0, 1, 2}, {0, 0, 0}},
                             {{-1, -1, (0 << 2) | 1}, {0, 2, 3}, {0, 0, 0}}};
static void show(const char *what, gpudtOrientedTriangle o) {
    printf(" %s(%d,%d)", what, o.tri, o.orient);
    if (o.tri >= 0) { int a, b, c; org(T, o, a); dest(T, o, b); apex(T, o, c); printf("[%d>%d^%d]", a, b, c); }
}
int main(void) {
    gpudtOrientedTriangle o, p, q;
    int ptr = 0;
    for (o.tri = 0; o.tri < 2; o.tri++)
        for (o.orient = 0; o.orient < 3; o.orient++) {
            printf("%d", Encode(o));
            show("o", o);
            lnext(o, p); show("lnext", p);
            lprev(o, p); show("lprev", p);
            sym(T, o, p); show("sym", p); printf(" ptr=%d", ptr);
            int inner = p.tri >= 0;
            onext(T, o, p); show("onext", p);
            if (inner) { oprev(T, o, p); show("oprev", p); }   /* not from a hull edge */
            otricopy(o, q); lnextself(q); lnextself(q); lnextself(q); printf(" eq=%d", otriequal(o, q));
            lprevself(q); show("lprevself", q);
            otricopy(o, q); symself(T, q); show("symself", q);
            otricopy(o, q); onextself(T, q); show("onextself", q);
            if (inner) { otricopy(o, q); oprevself(T, q); show("oprevself", q); }
            printf("\n");
        }
    decode(-1, p); show("decode(-1)", p);
    decode(6, p); show("decode(6)", p);
    printf(" plus=%d%d%d minus=%d%d%d\n", gpudt_plus1mod3[0], gpudt_plus1mod3[1], gpudt_plus1mod3[2],
           gpudt_minus1mod3[0], gpudt_minus1mod3[1], gpudt_minus1mod3[2]);
    o.tri = 0; o.orient = 0; setorg(T, o, 7); setdest(T, o, 8); setapex(T, o, 9);
    p.tri = 1; p.orient = 1; bond(T, o, p);
    q.tri = -1; q.orient = 3; o.orient = 2; bond(T, o, q);
    for (int t = 0; t < 2; t++)
        printf("T%d tri=%d,%d,%d vtx=%d,%d,%d\n", t, T[t].tri[0], T[t].tri[1], T[t].tri[2], T[t].vtx[0], T[t].vtx[1],
               T[t].vtx[2]);
    return 0;
}
"""


def _nav_probe(include_dir, lang):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "nav." + lang)
        open(src, "w").write(NAV_PROBE)
        exe = os.path.join(d, "nav")
        cc = "gcc" if lang == "c" else "g++"
        subprocess.run([cc, "-Wall", "-Wno-unused-variable", "-I", include_dir, src, "-o", exe], check=True,
                       capture_output=True)
        return subprocess.run([exe], check=True, capture_output=True, text=True).stdout


def _nav_macros(path):
    text = open(path).read()
    return set(re.findall(r"#define\s+([A-Za-z_][A-Za-z0-9_]*)\(", text))


def test_navigation_macros_hand_mesh():
    """The macros on a square split into two triangles, against facts worked out by hand."""
    out = _nav_probe(os.path.join(ROOT, "include"), "c").splitlines()
    assert out == _nav_probe(os.path.join(ROOT, "include"), "cpp").splitlines()
    row = {int(line.split()[0]): line for line in out[:6]}
    # (t0, 1): the diagonal seen from t0, directed 2 -> 0 with apex 1; its sym is t1's edge 0 -> 2
    assert " o(0,1)[2>0^1]" in row[1] and " sym(1,2)[0>2^3] ptr=6" in row[1]
    # spinning around 0: counter-clockwise from 0 -> 1 (t0) comes 0 -> 2 (t1), then the hull
    assert " onext(1,2)[0>2^3]" in row[2] and " onext(-1,3)" in row[(1 << 2) | 2]
    assert " oprev(" not in row[0] and " oprevself(" not in row[0]
    # clockwise from 0 -> 2 (t1) comes 0 -> 1 (t0)
    assert " oprev(0,2)[0>1^2]" in row[(1 << 2) | 2]
    # lnext / lprev stay in the triangle; three lnexts are the identity
    assert " lnext(0,2)[0>1^2] lprev(0,0)[1>2^0]" in row[1] and all(" eq=1" in r for r in out[:6])
    assert row[0].count("sym(-1,3)") == 1   # hull edges decode to tri -1
    assert "decode(-1)(-1,3) decode(6)(1,2)[0>2^3]" in out[6] and "plus=120 minus=201" in out[6]
    # set* write the corners of (t0, 0); bond links both sides and skips a hull side
    assert out[7] == "T0 tri=5,6,-1 vtx=9,7,8" and out[8] == "T1 tri=-1,0,1 vtx=0,2,3"


def test_navigation_macros_match_reference_header():
    """Every navigation macro of the reference header exists here, and the same caller code
    gives the same output compiled against either header."""
    ref = os.path.join(REF_HEADER_DIR, "gpudt.h")
    if not os.path.exists(ref):
        pytest.skip("reference tree not present")
    ours = _nav_macros(HEADER)
    theirs = _nav_macros(ref)
    assert theirs and theirs <= ours, sorted(theirs - ours)
    assert _nav_probe(os.path.join(ROOT, "include"), "cpp") == _nav_probe(REF_HEADER_DIR, "cpp")


NAV_DRIVER = r"""
#include <stdio.h>
#include <vector>
#include "navcheck.h"
int main(int argc, char **argv) {
    FILE *f = fopen(argv[1], "rb");
    int n, F;
    if (fread(&n, 4, 1, f) != 1) return 3;
    std::vector<gpudtVertex> P(n);
    if (fread(P.data(), 16, n, f) != (size_t)n || fread(&F, 4, 1, f) != 1) return 3;
    std::vector<gpudtTriangle> T(F > 0 ? F : 1);
    if (fread(T.data(), 36, F, f) != (size_t)F) return 3;
    fclose(f);
    NavReport r = galpha_nav_check(P.data(), n, T.data(), F);
    printf("errors=%d vertices=%d edges=%d hull=%d euler=%d corners=%lld\n", r.errors, r.vertices, r.edges,
           r.hullEdges, r.euler, r.fanCorners);
    return 0;
}
"""


@pytest.fixture(scope="module")
def nav_driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("nav")
    src = d / "drv.cpp"
    src.write_text(NAV_DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "galpha2d_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True, capture_output=True)
    return str(exe)


def _nav_run(exe, tmp_path, xy, tris):
    import numpy as np
    p = tmp_path / "mesh.bin"
    with open(p, "wb") as f:
        f.write(np.int32(len(xy)).tobytes())
        f.write(np.ascontiguousarray(xy, dtype="<f8").tobytes())
        f.write(np.int32(len(tris)).tobytes())
        f.write(np.ascontiguousarray(tris, dtype="<i4").tobytes())
    out = subprocess.run([exe, str(p)], check=True, capture_output=True, text=True).stdout
    return dict(kv.split("=") for kv in out.split("\n")[-2].split())


@pytest.mark.parametrize("name", ["bird.txt", "1000_1.txt"])
def test_navigation_walk_of_oracle_meshes(oracle, nav_driver, tmp_path, name):
    """The macro-based check (the command-line driver's --check) accepts the oracle's DT and
    rejects planted mistakes."""
    import numpy as np
    from conftest import load_fixture
    xy = load_fixture(name)
    tris, _ = oracle.delaunay(xy)
    F = len(tris)
    r = _nav_run(nav_driver, tmp_path, xy, tris)
    hull = int((tris[:, :3] == -1).sum())
    assert r == {"errors": "0", "vertices": str(len(xy)), "edges": str((3 * F + hull) // 2), "hull": str(hull),
                 "euler": "2", "corners": str(3 * F)}
    assert F == 2 * len(xy) - hull - 2   # Euler's count for a triangulated point set
    # a reversed triangle
    bad = tris.copy()
    bad[5, 3:6] = bad[5, [4, 3, 5]]
    assert _nav_run(nav_driver, tmp_path, xy, bad)["errors"] != "0"
    # a link that does not come back
    bad = tris.copy()
    i = int(np.flatnonzero(bad[:, 0] >= 0)[0])
    bad[i, 0] = -1
    assert _nav_run(nav_driver, tmp_path, xy, bad)["errors"] != "0"
    # a missing triangle leaves a vertex fan open and breaks Euler's formula
    assert _nav_run(nav_driver, tmp_path, xy, tris[1:])["errors"] != "0"
