"""Formats on the CPU: the reference point file (main.cu:117-161), dumpValues
(AlphaShape/Alpha_Shapes_2D.cu:749-793) and the binary result payload.  The library's format
calls are host code, so these run without a GPU; the triangulations and alpha results come
from the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

import galpha2d_amd as G
from conftest import ALPHA_FIELDS, GOLDEN, ROOT, load_fixture

REF_FILES = ["bird.txt", "alpha.txt", "1000_1.txt", "10000_1.txt"]


@pytest.mark.parametrize("name", REF_FILES)
def test_read_reference_point_files(name):
    """The library reader gives the points of the tests' own reader, the bounds of main.cu:139-159
    and the library's texture choice."""
    xy, info = G.read_points(os.path.join(GOLDEN, name))
    want = load_fixture(name)
    assert xy.dtype == np.float64 and np.array_equal(xy, want)
    assert (info["minX"], info["minY"]) == (want[:, 0].min(), want[:, 1].min())
    assert (info["maxX"], info["maxY"]) == (want[:, 0].max(), want[:, 1].max())
    assert info["nConstraints"] == 0 and info["fboSize"] == G.auto_texture(len(want))


def test_point_file_round_trip(tmp_path):
    """Written with 17 significant digits, every double reads back bit-exact (subnormal, huge,
    negative zero, and values with long expansions)."""
    rng = np.random.default_rng(3)
    xy = np.concatenate([rng.random((500, 2)), rng.normal(size=(500, 2)) * 1e300,
                         np.array([[5e-324, -0.0], [1.0 / 3.0, 2.0 ** 52 + 1], [-1e-310, 0.1]])])
    p = tmp_path / "pts.txt"
    G.write_points(str(p), xy)
    back, _ = G.read_points(str(p))
    assert np.array_equal(back.view(np.uint64), xy.view(np.uint64))
    lines = p.read_text().splitlines()
    assert lines[0] == str(len(xy)) and lines[1].split()[0] == "0"


def test_point_file_out_of_order_indices(tmp_path):
    """Point idx goes to points[idx] whatever the line order (main.cu:131-136)."""
    p = tmp_path / "pts.txt"
    p.write_text("3\n2 5 6\n0 1 2\n1 3 4\n\n")
    xy, _ = G.read_points(str(p))
    assert np.array_equal(xy, [[1, 2], [3, 4], [5, 6]])


@pytest.mark.parametrize("text,msg", [
    ("", "empty"),
    ("x\n", "bad point count"),
    ("3\n0 1 2\n1 3 4\n", "3 points announced, 2 read"),
    ("2\n0 1 2\n2 3 4\n", r"outside \[0, 2\)"),
    ("2\n0 1 2\n0 3 4\n", "no line for point 1"),
    ("2\n0 1 2\n1 3 abc\n", "bad coordinate"),
    ("2\n0 1 2\n1 3 inf\n", "not finite"),
])
def test_point_file_errors(tmp_path, text, msg):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(G.GalphaError, match=msg):
        G.read_points(str(p))


def test_point_file_missing():
    with pytest.raises(G.GalphaError, match="cannot open"):
        G.read_points("/nonexistent/points.txt")


def _fnv1a(data):
    h = 0xcbf29ce484222325
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_fnv1a_published_vectors():
    """The payload hash is 64-bit FNV-1a; this decoder's copy matches its published vectors."""
    assert _fnv1a(b"") == 0xcbf29ce484222325
    assert _fnv1a(b"a") == 0xaf63dc4c8601ec8c
    assert _fnv1a(b"foobar") == 0x85944171f73967e8


def _decode(payload):
    """Independent decoder of the documented payload layout."""
    magic, version, flags, n, F, E, S, M, SP, zero = struct.unpack_from("<10I", payload, 0)
    assert magic == 0x52443247 and version == 1 and zero == 0
    off = 40
    out = {}

    def take(key, dtype, count):
        nonlocal off
        a = np.frombuffer(payload, dtype=dtype, count=count, offset=off)
        off += (a.nbytes + 7) // 8 * 8
        out[key] = a

    if flags & 1:
        take("spectrum", "<f8", S)
    take("tris", "<i4", 9 * F)
    if flags & 1:
        take("tri_ro_rank", "<i4", F)
        for k in ("edge_t", "edge_vopp", "edge_ro_rank", "edge_nu_min", "edge_nu_max"):
            take(k, "<i4", E)
        for k in ("vert_nu_min", "vert_nu_max"):
            take(k, "<i4", n)
        for k in ("master_simplex", "master_st", "master_at"):
            take(k, "<i4", M)
        take("spec_pointers", "<i4", SP)
    assert off + 8 == len(payload)
    assert struct.unpack_from("<Q", payload, off)[0] == _fnv1a(payload[:off])
    return n, out


@pytest.mark.parametrize("name", ["bird.txt", "alpha.txt", "1000_1.txt"])
def test_payload_round_trip(oracle, name):
    xy = load_fixture(name)
    tris, _ = oracle.delaunay(xy)
    alpha = oracle.alpha(xy, tris)
    buf = G.pack_result(len(xy), tris, alpha)
    n, d = _decode(buf)
    assert n == len(xy) and np.array_equal(d["tris"].reshape(-1, 9), tris)
    for k in ALPHA_FIELDS:
        assert np.array_equal(d[k], np.asarray(alpha[k])), k
    n2, t2, a2 = G.unpack_result(buf)
    assert n2 == len(xy) and np.array_equal(t2, tris)
    for k in ALPHA_FIELDS:
        assert np.array_equal(a2[k], np.asarray(alpha[k])), k
    # triangles only
    buf0 = G.pack_result(len(xy), tris)
    n0, d0 = _decode(buf0)
    assert n0 == len(xy) and np.array_equal(d0["tris"].reshape(-1, 9), tris) and "spectrum" not in d0
    n3, t3, a3 = G.unpack_result(buf0)
    assert a3 is None and np.array_equal(t3, tris)


def test_payload_rejects_damage(oracle):
    xy = load_fixture("bird.txt")
    tris, _ = oracle.delaunay(xy)
    buf = bytearray(G.pack_result(len(xy), tris, oracle.alpha(xy, tris)))
    for pos in (0, 4, 60, len(buf) // 2, len(buf) - 1):
        bad = bytearray(buf)
        bad[pos] ^= 0x10
        with pytest.raises(G.GalphaError, match="payload"):
            G.unpack_result(bytes(bad))
    with pytest.raises(G.GalphaError, match="payload"):
        G.unpack_result(bytes(buf[:-8]))
    with pytest.raises(G.GalphaError, match="payload"):
        G.unpack_result(bytes(buf[:20]))


def test_payload_empty():
    buf = G.pack_result(2, np.zeros((0, 9), np.int32))
    assert len(buf) == 48
    n, t, a = G.unpack_result(buf)
    assert n == 2 and t.shape == (0, 9) and a is None


def test_dump_values(oracle, tmp_path):
    """The active output of dumpValues is the spectrum size line; the full dump's sections carry
    spectrum[rank] of every interval, in the point, edge and triangle order of the result."""
    xy = load_fixture("alpha.txt")
    tris, _ = oracle.delaunay(xy)
    a = oracle.alpha(xy, tris)
    spec = np.asarray(a["spectrum"])
    p = tmp_path / "out.txt"
    G.dump_values(str(p), xy, tris, a)
    assert p.read_text() == f"{len(spec)}\n"
    G.dump_values(str(p), xy, tris, a, full=True)
    lines = p.read_text().splitlines()
    nv, ne, nt = len(xy), len(a["edge_t"]), len(tris)
    assert lines[0] == "VERTICES" and lines[1 + nv] == "EDGES" and lines[2 + nv + ne] == "TRIANGLES"
    assert int(lines[3 + nv + ne + nt]) == len(spec)

    def val(r):
        return spec[r] if r >= 0 else np.nan

    def row(s):
        return np.array([float(x) for x in s.split()])

    for v in range(nv):
        want = [xy[v, 0], xy[v, 1], val(a["vert_nu_min"][v]), val(a["vert_nu_max"][v])]
        assert np.array_equal(row(lines[1 + v]), want, equal_nan=True)
    for e in range(ne):
        t, i = a["edge_t"][e], a["edge_vopp"][e]
        p1, p2 = tris[t, 3 + (i + 1) % 3], tris[t, 3 + (i + 2) % 3]
        want = [*xy[p1], *xy[p2], val(a["edge_ro_rank"][e]), val(a["edge_nu_min"][e]), val(a["edge_nu_max"][e])]
        assert np.array_equal(row(lines[2 + nv + e]), want)
    for t in range(nt):
        want = [*xy[tris[t, 3]], *xy[tris[t, 4]], *xy[tris[t, 5]], val(a["tri_ro_rank"][t])]
        assert np.array_equal(row(lines[3 + nv + ne + t]), want)
    assert np.array_equal(np.array([float(x) for x in lines[4 + nv + ne + nt:]]), spec)


def test_cli_usage_and_errors():
    """The command-line driver is built next to the library; argument and file errors exit
    non-zero before any GPU call."""
    cli = os.path.join(ROOT, "galpha2d_amd", "galpha_cli")
    assert os.access(cli, os.X_OK)
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([cli, "/nonexistent.txt"], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_render_svg_counts(oracle, tmp_path):
    """The picture has one filled polygon per interior triangle, one bold line per regular edge,
    one thin red line per singular edge and one dot per singular vertex (classes from the
    definition, on the oracle's result)."""
    import xml.etree.ElementTree as ET
    from galpha2d_amd.visualize import render_svg
    from test_gpu_alpha import _classes_by_definition
    xy = load_fixture("bird.txt")
    tris, _ = oracle.delaunay(xy)
    a = oracle.alpha(xy, tris)
    for r in (1, len(a["spectrum"]) // 3, len(a["spectrum"]) - 1):
        v, e, t = _classes_by_definition(tris, a, r)
        p = tmp_path / f"r{r}.svg"
        render_svg(str(p), xy, tris, a, (v, e, t))
        root = ET.parse(str(p)).getroot()
        kinds = {}
        for el in root:
            key = (el.tag.split("}")[-1], el.get("class"))
            kinds[key] = kinds.get(key, 0) + 1
        assert kinds.get(("polygon", "interior"), 0) == int((t == 3).sum())
        assert kinds.get(("line", "regular"), 0) == int((e == 2).sum())
        assert kinds.get(("line", "singular"), 0) == int((e == 1).sum())
        assert kinds.get(("circle", "singular"), 0) == int((v == 1).sum())
        assert kinds.get(("line", "mesh"), 0) == int((e == 0).sum())
