"""Shared test setup: the `gpu` marker, the oracle library (the checker) and fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")



def load_fixture(name):
    """Reference input format (main.cu:117-161): N, then N lines 'idx x y'."""
    with open(os.path.join(GOLDEN, name)) as f:
        n = int(f.readline())
        xy = np.zeros((n, 2))
        for line in f:
            p = line.split()
            if len(p) == 3:
                xy[int(p[0])] = (float(p[1]), float(p[2]))
    return xy


def lattice_clusters(n_lat, n_cl, side, nclusters, sigma, seed):
    """C3-style input: n_lat nodes drawn without replacement from a side x side lattice
    scaled to [0,1)^2, plus n_cl points from Gaussian clusters (resampled into [0,1)^2)."""
    rng = np.random.default_rng(seed)
    sel = rng.choice(side * side, n_lat, replace=False)
    lat = np.stack([(sel // side) / side, (sel % side) / side], axis=1)
    cent = rng.random((nclusters, 2))
    pts = cent[rng.integers(0, nclusters, n_cl)] + rng.normal(size=(n_cl, 2)) * sigma
    bad = ((pts < 0) | (pts >= 1)).any(axis=1)
    while bad.any():
        k = int(bad.sum())
        pts[bad] = cent[rng.integers(0, nclusters, k)] + rng.normal(size=(k, 2)) * sigma
        bad = ((pts < 0) | (pts >= 1)).any(axis=1)
    return np.concatenate([lat, pts])


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


ALPHA_FIELDS = ("spectrum", "tri_ro_rank", "edge_t", "edge_vopp", "edge_ro_rank", "edge_nu_min",
                "edge_nu_max", "vert_nu_min", "vert_nu_max", "master_simplex", "master_st",
                "master_at", "spec_pointers")


def digest(tris, alpha=None):
    """SHA-256 digests of a canonical triangulation (int32 [F, 9]) and of each alpha array
    (spectrum as little-endian float64, the rest as little-endian int32)."""
    import hashlib
    out = {"nTris": int(tris.shape[0]),
           "tris": hashlib.sha256(np.ascontiguousarray(tris, dtype="<i4").tobytes()).hexdigest()}
    if alpha is not None:
        out["specSize"] = int(len(alpha["spectrum"]))
        for k in ALPHA_FIELDS:
            dt = "<f8" if k == "spectrum" else "<i4"
            out[k] = hashlib.sha256(np.ascontiguousarray(alpha[k], dtype=dt).tobytes()).hexdigest()
    return out


def set_digest(tris, alpha):
    """One SHA-256 over a set's canonical triangles (int32 [F, 9]) and every alpha array in
    ALPHA_FIELDS order (spectrum little-endian float64, the rest little-endian int32)."""
    import hashlib
    h = hashlib.sha256(np.ascontiguousarray(tris, dtype="<i4").tobytes())
    for k in ALPHA_FIELDS:
        h.update(np.ascontiguousarray(alpha[k], dtype="<f8" if k == "spectrum" else "<i4").tobytes())
    return h.hexdigest()


def golden_digests():
    import json
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)
