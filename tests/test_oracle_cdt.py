"""The oracle's constrained DT (oracle/cdt.c), pinned on the CPU.

- Brute force on tiny inputs: every triangulation that holds the constraints is reached from the
  oracle's by flips of the other edges (the flip graph of constrained triangulations is
  connected), and the constrained DT is the one whose ascending angle vector is lexicographically
  largest (each Lawson flip raises it strictly in general position).  Integer coordinates keep
  these checks exact.
- The definition on larger inputs: ora_validate_constrained (every constraint an edge, every other
  edge locally Delaunay under the tie-break, a valid triangulation of the hull), which is shown to
  reject a planted mistake.
- Uniqueness: the result does not depend on the order or direction of the constraints.
- Errors for the inputs the reference leaves undefined."""
import itertools
import math

import numpy as np
import pytest

from cdt_cases import cell_polygons, lattice_knights, star_polygon
from conftest import load_fixture


def _orient(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def _incircle(a, b, c, d):
    m = [[p[0] - d[0], p[1] - d[1], (p[0] - d[0]) ** 2 + (p[1] - d[1]) ** 2] for p in (a, b, c)]
    return (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
            + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))


def _general_position_ints(n, seed):
    rng = np.random.default_rng(seed)
    while True:
        P = [tuple(int(v) for v in rng.integers(0, 1024, 2)) for _ in range(n)]
        if any(_orient(*t) == 0 for t in itertools.combinations(P, 3)):
            continue
        if any(_incircle(*q) == 0 for q in itertools.combinations(P, 4)):
            continue
        return P


def _proper_cross(P, s, t):
    if len({*s, *t}) < 4:
        return False
    a, b, c, d = (P[v] for v in (*s, *t))
    return _orient(a, b, c) * _orient(a, b, d) < 0 and _orient(c, d, a) * _orient(c, d, b) < 0


def _flips(P, tris, fixed):
    """All triangulations one flip away, flipping any convex interior edge not in `fixed`."""
    edge_tris = {}
    for t in tris:
        for e in itertools.combinations(t, 2):
            edge_tris.setdefault(frozenset(e), []).append(t)
    for e, ts in edge_tris.items():
        if len(ts) != 2 or e in fixed:
            continue
        x, y = tuple(e)
        p = (set(ts[0]) - e).pop()
        q = (set(ts[1]) - e).pop()
        if not _proper_cross(P, (p, q), (x, y)):
            continue
        new = set(tris) - set(ts)
        new |= {tuple(sorted((p, q, x))), tuple(sorted((p, q, y)))}
        yield frozenset(new)


def _angles(P, tris):
    out = []
    for t in tris:
        for i in range(3):
            a, b, c = P[t[i]], P[t[(i + 1) % 3]], P[t[(i + 2) % 3]]
            v1, v2 = (b[0] - a[0], b[1] - a[1]), (c[0] - a[0], c[1] - a[1])
            out.append(math.atan2(abs(v1[0] * v2[1] - v1[1] * v2[0]), v1[0] * v2[0] + v1[1] * v2[1]))
    return sorted(out)


@pytest.mark.parametrize("seed", range(6))
def test_cdt_brute_force_max_angle_vector(oracle, seed):
    n = 8
    P = _general_position_ints(n, 100 + seed)
    xy = np.array(P, dtype=np.float64)
    rng = np.random.default_rng(seed)
    dt, _ = oracle.delaunay(xy)
    dt_edges = {frozenset((int(t[3 + (i + 1) % 3]), int(t[3 + (i + 2) % 3]))) for t in dt for i in range(3)}
    # one or two constraints that are not Delaunay edges and do not cross
    pairs = [p for p in itertools.combinations(range(n), 2) if frozenset(p) not in dt_edges]
    rng.shuffle(pairs)
    cons = [pairs[0]]
    for p in pairs[1:]:
        if len(cons) < 1 + seed % 2 and not _proper_cross(P, p, cons[0]):
            cons.append(p)
    tris, _ = oracle.cdt(xy, np.array(cons))
    got = frozenset(tuple(sorted(int(v) for v in t[3:6])) for t in tris)
    fixed = {frozenset(c) for c in cons}
    seen, todo = {got}, [got]
    while todo:
        cur = todo.pop()
        for nxt in _flips(P, cur, fixed):
            if nxt not in seen:
                seen.add(nxt)
                todo.append(nxt)
    assert len(seen) > 1
    for T in seen:
        edges = {frozenset(e) for t in T for e in itertools.combinations(t, 2)}
        assert fixed <= edges
    best = max(seen, key=lambda T: _angles(P, T))
    assert best == got, (seed, cons)


def _check(oracle, xy, cons):
    tris, dup = oracle.cdt(xy, cons)
    rc, msg = oracle.validate_constrained(xy, tris, dup, cons)
    assert rc == 0, msg
    return tris, dup


def test_cdt_uniform_star_polygon(oracle):
    xy = oracle.gen_uniform(5, 3000)
    cons = star_polygon(xy, 60, 1)
    tris, dup = _check(oracle, xy, cons)
    dt, _ = oracle.delaunay(xy)
    assert not np.array_equal(tris, dt)
    rc, _ = oracle.validate(xy, tris, dup)
    assert rc == 7   # some edge is forced away from local Delaunay


def test_cdt_reference_file_cells(oracle):
    xy = load_fixture("10000_1.txt")
    cons = cell_polygons(xy, 8, 7, 2)
    _check(oracle, xy, cons)


def test_cdt_lattice_knights(oracle):
    side = 40
    kept, nodes = lattice_knights(side, 900, 150, 3)
    pos = {v: k for k, v in enumerate(nodes)}
    xy = np.array([divmod(v, side) for v in nodes], dtype=np.float64) / side
    cons = np.array([(pos[a], pos[b]) for a, b in kept], dtype=np.int32)
    _check(oracle, xy, cons)


def test_cdt_independent_of_constraint_order(oracle):
    xy = oracle.gen_uniform(7, 2000)
    cons = cell_polygons(xy, 5, 6, 4)
    a, _ = oracle.cdt(xy, cons)
    rng = np.random.default_rng(0)
    c2 = cons[rng.permutation(len(cons))][:, ::-1].copy()
    b, _ = oracle.cdt(xy, c2)
    assert np.array_equal(a, b)


def test_cdt_with_delaunay_edges_only_is_the_dt(oracle):
    xy = oracle.gen_uniform(8, 500)
    dt, _ = oracle.delaunay(xy)
    cons = dt[::7, 3:5].copy()
    tris, _ = oracle.cdt(xy, cons)
    assert np.array_equal(tris, dt)


def test_validator_rejects_planted_mistakes(oracle):
    """The validator is what pins the oracle at scale, so it must reject (a) a missing constraint
    and (b) a mesh with one non-constraint edge flipped away from local Delaunay."""
    xy = oracle.gen_uniform(9, 800)
    cons = star_polygon(xy, 30, 5)
    dt, dup = oracle.delaunay(xy)
    rc, msg = oracle.validate_constrained(xy, dt, dup, cons)
    assert rc == 9 and "not an edge" in msg
    tris, dup = oracle.cdt(xy, cons)
    consset = {frozenset(map(int, c)) for c in cons}
    bad = tris.copy()
    for t in range(len(bad)):   # flip the first convex non-constraint interior edge
        for i in range(3):
            e = bad[t, i]
            x, y = bad[t, 3 + (i + 1) % 3], bad[t, 3 + (i + 2) % 3]
            if e < 0 or frozenset((int(x), int(y))) in consset:
                continue
            u, j = e >> 2, e & 3
            p, q = bad[t, 3 + i], bad[u, 3 + j]
            P = {v: xy[v] for v in (x, y, p, q)}
            o = lambda a, b, c: (P[b][0] - P[a][0]) * (P[c][1] - P[a][1]) - (P[b][1] - P[a][1]) * (P[c][0] - P[a][0])
            if o(p, q, x) * o(p, q, y) < 0:
                bad[t, 3:6] = (p, x, q)
                bad[u, 3:6] = (q, y, p)
                break
        else:
            continue
        break
    rc, msg = oracle.validate_constrained(xy, oracle.canonicalize(bad), dup, cons)
    assert rc != 0


# square corners, two interior points off both diagonals, a duplicate of point 0, and the centre
_SQ = np.array([[0, 0], [4, 0], [4, 4], [0, 4], [2, 1.5], [1, 3.5], [0, 0], [2, 2]], dtype=np.float64)

INVALID = [
    ("cross", _SQ[:7], [[0, 2], [1, 3]], "crosses constraint"),
    ("through", _SQ, [[0, 2]], "passes through point 7"),
    ("dup", _SQ[:7], [[6, 1]], "duplicate point"),
    ("range", _SQ[:7], [[0, 9]], "out of range"),
    ("same", _SQ[:7], [[1, 1]], "both endpoints"),
]


@pytest.mark.parametrize("case,xy,cons,msg", INVALID, ids=[c[0] for c in INVALID])
def test_cdt_invalid_constraints(oracle, case, xy, cons, msg):
    with pytest.raises(ValueError, match=msg):
        oracle.cdt(xy, np.array(cons))
