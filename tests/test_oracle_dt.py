"""Oracle DT pinned by brute force, an independent library DT, invariants and planted
mistakes (no GPU)."""
import itertools
from fractions import Fraction as Fr

import numpy as np
import pytest

from conftest import load_fixture, lattice_clusters


def brute_dt(xy):
    """O(n^4) empty-circle enumeration with exact rationals (general position)."""
    n = len(xy)
    P = [(Fr(x), Fr(y)) for x, y in xy]
    tris = set()
    for i, j, k in itertools.combinations(range(n), 3):
        a, b, c = P[i], P[j], P[k]
        o = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
        if o == 0:
            continue
        t = (i, j, k) if o > 0 else (i, k, j)
        A, B, Cc = (P[v] for v in t)
        ok = True
        for m in range(n):
            if m in (i, j, k):
                continue
            d = P[m]
            rows = [(p[0] - d[0], p[1] - d[1]) for p in (A, B, Cc)]
            rows = [(x, y, x * x + y * y) for x, y in rows]
            (a1, a2, a3), (b1, b2, b3), (c1, c2, c3) = rows
            det = a1 * (b2 * c3 - b3 * c2) - a2 * (b1 * c3 - b3 * c1) + a3 * (b1 * c2 - b2 * c1)
            if det > 0:
                ok = False
                break
        if ok:
            r = t.index(min(t))
            tris.add(t[r:] + t[:r])
    return sorted(tris)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_dt_matches_brute_force(oracle, seed):
    xy = oracle.gen_uniform(seed, 28)
    tris, dup = oracle.delaunay(xy)
    got = sorted(tuple(t) for t in tris[:, 3:6].tolist())
    assert got == brute_dt(xy)


def test_bird_fixture_brute_force(oracle):
    xy = load_fixture("bird.txt")     # 51 points, integer coordinates
    tris, dup = oracle.delaunay(xy)
    assert oracle.validate(xy, tris, dup)[0] == 0
    got = set(tuple(t) for t in tris[:, 3:6].tolist())
    bf = brute_dt(xy)
    # with co-circular quadruples brute force lists both diagonals; the DT picks one of them
    assert got <= set(bf)
    assert len(got) == 2 * len(xy) - int((tris[:, :3] == -1).sum()) - 2


def test_matches_independent_library(oracle):
    scipy_spatial = pytest.importorskip("scipy.spatial")
    xy = oracle.gen_uniform(11, 3000)
    tris, _ = oracle.delaunay(xy)
    q = scipy_spatial.Delaunay(xy)
    ref = set()
    for t in q.simplices.tolist():
        a, b, c = t
        if oracle.orient_sign(xy[a], xy[b], xy[c]) < 0:
            b, c = c, b
        r = [a, b, c]
        i = r.index(min(r))
        ref.add(tuple(r[i:] + r[:i]))
    assert set(tuple(t) for t in tris[:, 3:6].tolist()) == ref


@pytest.mark.parametrize("name", ["bird.txt", "alpha.txt", "1000_1.txt"])
def test_fixtures_valid(oracle, name):
    xy = load_fixture(name)
    tris, dup = oracle.delaunay(xy)
    rc, msg = oracle.validate(xy, tris, dup)
    assert rc == 0, msg
    h = int((tris[:, :3] == -1).sum())
    kept = int((dup < 0).sum())
    assert tris.shape[0] == 2 * kept - h - 2


def test_degenerate_inputs(oracle):
    g = np.array([(i, j) for i in range(15) for j in range(15)], dtype=np.float64)
    tris, dup = oracle.delaunay(g)
    assert oracle.validate(g, tris, dup)[0] == 0
    xy = np.concatenate([g, g[:40], [[3.5, 0.0]]])
    tris, dup = oracle.delaunay(xy)
    assert int((dup >= 0).sum()) == 40 and (dup[225:265] == np.arange(40)).all()
    assert oracle.validate(xy, tris, dup)[0] == 0
    col = np.array([[i, 2.0 * i] for i in range(20)], dtype=np.float64)
    assert oracle.delaunay(col)[0].shape[0] == 0
    assert oracle.delaunay(np.zeros((2, 2)))[0].shape[0] == 0


def test_lattice_clusters_small(oracle):
    xy = lattice_clusters(3000, 3000, 80, 8, 0.01, 5)
    tris, dup = oracle.delaunay(xy)
    rc, msg = oracle.validate(xy, tris, dup)
    assert rc == 0, msg


def test_validator_catches_planted_mistakes(oracle):
    xy = oracle.gen_uniform(3, 300)
    tris, dup = oracle.delaunay(xy)
    assert oracle.validate(xy, tris, dup)[0] == 0
    # 1) flip one interior edge (valid triangulation, no longer Delaunay)
    t = tris.copy()
    k = next(k for k in range(len(t)) if (t[k, :3] >= 0).all())
    e = t[k, 0]
    u, j = e >> 2, e & 3
    c, x, y = t[k, 3], t[k, 4], t[k, 5]
    d = t[u, 3 + j]
    if oracle.orient_sign(xy[c], xy[x], xy[d]) > 0 and oracle.orient_sign(xy[d], xy[y], xy[c]) > 0:
        t[k, 3:6] = [c, x, d]
        t[u, 3:6] = [d, y, c]
        assert oracle.validate(xy, oracle.canonicalize(t), dup)[0] != 0
    # 2) reverse one triangle's orientation
    t = tris.copy()
    t[5, 4], t[5, 5] = t[5, 5], t[5, 4]
    assert oracle.validate(xy, t, dup)[0] != 0
    # 3) drop a triangle
    assert oracle.validate(xy, tris[1:], dup)[0] != 0
    # 4) corrupt a neighbour link
    t = tris.copy()
    t[7, 0] = t[8, 0] if t[8, 0] != t[7, 0] else t[9, 1]
    assert oracle.validate(xy, t, dup)[0] != 0
