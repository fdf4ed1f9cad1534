"""The command-line driver on the GPU (the counterpart of main.cu:364-440): it reads the
reference's point files, runs DT + alpha through the reference-named entry points, checks the
result and writes the payload and the full dump; both equal the oracle's result."""
import os
import subprocess

import numpy as np
import pytest

import galpha2d_amd as G
from conftest import ALPHA_FIELDS, GOLDEN, ROOT, load_fixture

CLI = os.path.join(ROOT, "galpha2d_amd", "galpha_cli")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["bird.txt", "alpha.txt", "1000_1.txt", "10000_1.txt"])
def test_cli_reference_files(oracle, tmp_path, name):
    print("case: galpha_cli", name, flush=True)
    pack, dump = tmp_path / "r.bin", tmp_path / "dump.txt"
    r = subprocess.run([CLI, os.path.join(GOLDEN, name), "--check", "--pack", str(pack), "--dump-full", str(dump)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "GAlpha time:" in r.stdout and ", 0 errors," in r.stdout
    # the macro walk (include/gpudt.h navigation) closed every vertex fan and Euler's formula holds
    assert "Euler characteristic 2," in r.stdout
    xy = load_fixture(name)
    want, _ = oracle.delaunay(xy)
    wa = oracle.alpha(xy, want)
    n, tris, alpha = G.unpack_result(pack.read_bytes())
    assert n == len(xy) and np.array_equal(tris, want)
    for k in ALPHA_FIELDS:
        assert np.array_equal(alpha[k], np.asarray(wa[k])), k
    # the full dump written by the driver equals the dump of the oracle's result
    ref = tmp_path / "ref.txt"
    G.dump_values(str(ref), xy, want, wa, full=True)
    assert dump.read_text() == ref.read_text()


@pytest.mark.gpu
def test_render_device_classes(tmp_path):
    """The picture from the device's classes (galphaClassify) on the reference's bird input."""
    import xml.etree.ElementTree as ET
    from galpha2d_amd.visualize import render_svg
    print("case: render bird", flush=True)
    xy = load_fixture("bird.txt")
    ctx = G.Context(0)
    d = G.DeviceBuffer(np.ascontiguousarray(xy, dtype=np.float64))
    ctx.compute_device(d.ptr, len(xy), 256, with_alpha=True)
    tris, a = ctx.triangles(), ctx.alpha()
    r = len(a["spectrum"]) // 2
    v, e, t = ctx.classify(r, a)
    ctx.close()
    p = tmp_path / "bird.svg"
    render_svg(str(p), xy, tris, a, (v, e, t))
    root = ET.parse(str(p)).getroot()
    n_poly = sum(1 for el in root if el.tag.endswith("polygon"))
    assert n_poly == int((t == 3).sum()) > 0
