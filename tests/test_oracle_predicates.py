"""Oracle predicates pinned against exact rationals, the reference's own predicates and
algebraic invariants (no GPU)."""
import ctypes as C
import itertools
import json
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def _cases():
    with open(os.path.join(GOLDEN, "predicates.json")) as f:
        return json.load(f)


def test_orient_known_answers(oracle):
    for a, b, c, s in _cases()["orient"]:
        assert oracle.orient_sign(a, b, c) == s


def test_incircle_known_answers(oracle):
    for a, b, c, d, s in _cases()["incircle"]:
        assert oracle.incircle_sign(a, b, c, d) == s


def test_exact_stage_is_exercised(oracle):
    oracle.lib().ora_pred_reset_counters()
    for a, b, c, d, s in _cases()["incircle"]:
        oracle.incircle_sign(a, b, c, d)
    calls, exact = oracle.pred_counters()[2:4]
    assert exact > calls // 2   # the grid is built so the FP64 filter rarely decides


def test_antisymmetry_and_cyclic(oracle):
    rnd = random.Random(7)
    for a, b, c, d, _ in _cases()["incircle"][:150]:
        o = oracle.orient_sign(a, b, c)
        assert oracle.orient_sign(b, c, a) == o and oracle.orient_sign(b, a, c) == -o
        i = oracle.incircle_sign(a, b, c, d)
        assert oracle.incircle_sign(b, c, a, d) == i
        assert oracle.incircle_sign(b, a, c, d) == -i
        assert oracle.incircle_sign(a, b, d, c) == -i   # row swap of the lifted determinant


def _perm_sign(p):
    s, p = 1, list(p)
    for i in range(len(p)):
        while p[i] != i:
            j = p[i]
            p[i], p[j] = p[j], p[i]
            s = -s
    return s


def _det(m):
    """Exact determinant (integers or Fractions) by cofactor expansion."""
    if len(m) == 1:
        return m[0][0]
    return sum((-1) ** c * m[0][c] * _det([row[:c] + row[c + 1:] for row in m[1:]]) for c in range(len(m)))


def _sos_by_definition(P, idx, s0):
    """Sign of det[x, y, x^2 + y^2 + eps^(i + 1), 1] for an infinitesimal eps > 0: the determinant
    is linear in each lifted height, so the coefficient of eps^(i + 1) is the cofactor of row i's
    height, and the first non-zero one in index order decides.  s0 maps the 4x4 determinant's
    sign to incircle's."""
    rows = [[x, y, x * x + y * y, 1] for x, y in P]
    assert _det(rows) == 0
    for r in sorted(range(4), key=lambda k: idx[k]):
        minor = [row[:2] + row[3:] for k, row in enumerate(rows) if k != r]
        cof = (-1) ** (r + 2) * _det(minor)
        if cof:
            return s0 * (1 if cof > 0 else -1)
    return 0


def test_sos_tiebreak_consistent(oracle):
    """Exactly co-circular quadruples: the tie-break never returns 0, is a determinant sign over
    all 24 argument orders (the sign of the permutation times the identity order's), and equals
    the sign of the height-perturbed lifted determinant worked out exactly."""
    circ = [(5.0, 0.0), (4.0, 3.0), (3.0, 4.0), (0.0, 5.0), (-3.0, 4.0), (-4.0, 3.0),
            (-5.0, 0.0), (-4.0, -3.0), (-3.0, -4.0), (0.0, -5.0), (3.0, -4.0), (4.0, -3.0)]
    # the 4x4 lifted determinant's sign against incircle's, on a point strictly inside
    a, b, c, d = (5, 0), (0, 5), (-5, 0), (0, 0)
    s0 = oracle.incircle_sign(a, b, c, d) * (1 if _det([[x, y, x * x + y * y, 1] for x, y in (a, b, c, d)]) > 0 else -1)
    rnd = random.Random(11)
    labels = list(range(40))
    for q in itertools.combinations(range(12), 4):
        P = [circ[i] for i in q]
        assert oracle.incircle_sign(*P) == 0
        idx = rnd.sample(labels, 4)   # point indices in any order relative to the positions
        s = oracle.incircle_sos(P[0], idx[0], P[1], idx[1], P[2], idx[2], P[3], idx[3])
        assert s != 0
        assert s == _sos_by_definition([tuple(map(int, p)) for p in P], idx, s0)
        for perm in itertools.permutations(range(4)):
            args = []
            for k in perm:
                args += [P[k], idx[k]]
            assert oracle.incircle_sos(*args) == _perm_sign(perm) * s, (q, idx, perm)


def _ref_lib():
    path = os.path.join(ROOT, "oracle", "_ref", "libref_predicates.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (the reference tree is not present)")
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    L.ref_orient2d.argtypes = [dp, dp, dp]
    L.ref_orient2d.restype = C.c_double
    L.ref_incircle.argtypes = [dp, dp, dp, dp]
    L.ref_incircle.restype = C.c_double
    return L


def _sgn(v):
    return (v > 0) - (v < 0)


def test_oracle_matches_reference_predicates(oracle):
    """oracle/_ref: the reference's own gpudt/predicates.cpp, compiled from its sources."""
    L = _ref_lib()
    P = lambda p: np.ascontiguousarray(p, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))
    cases = _cases()
    keep = []
    for a, b, c, s in cases["orient"]:
        A, B, Cc = (np.array(v, dtype=np.float64) for v in (a, b, c))
        keep += [A, B, Cc]
        assert _sgn(L.ref_orient2d(P(A), P(B), P(Cc))) == oracle.orient_sign(a, b, c) == s
    for a, b, c, d, s in cases["incircle"]:
        A, B, Cc, D = (np.array(v, dtype=np.float64) for v in (a, b, c, d))
        assert _sgn(L.ref_incircle(P(A), P(B), P(Cc), P(D))) == oracle.incircle_sign(a, b, c, d) == s
