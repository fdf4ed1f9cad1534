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
    and kernel-probe state, the thread-local error text and C++ runtime bookkeeping."""
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
               "(anonymous namespace)::g_used", "std::__ioinit", "__tls_guard", "__fini", "__init"}
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
