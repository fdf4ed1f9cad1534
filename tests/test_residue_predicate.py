"""The residue zero tests of the device's exact predicates (galpha2d_amd/csrc/dev_common.h),
compiled for the host and checked against exact rational arithmetic:
- `incircle_zero_res`: co-circular quadruples (lattice points on circles with many of them,
  lattice squares at a binary snapping unit) come out zero, perturbed and random ones nonzero;
- `rho_equal_res`: equal alpha values (congruent lattice triangles moved and turned, a right
  triangle against its hypotenuse) come out equal, perturbed and random pairs different;
- `attached_i128` (an exact sign in 128-bit integers): apexes on, inside and outside the circle
  over an edge;
and coordinate spreads beyond their bit limits defer (-1) instead of answering.  A wrong prime, a
dropped term or a bound that lets a nonzero value vanish fails here.  The fixed-width forms
(incircle_fixed, rho_cmp_fixed, rho_key_fixed and its long division mw_div_17_8) and the
big-integer fallbacks (orient_exact, incircle_exact) are checked the same way below."""
import math
import os
import random
import shutil
import subprocess
import tempfile
from fractions import Fraction

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "galpha2d_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

HARNESS = r"""
#include <hip/hip_runtime.h>
#include <cstdio>
#include "dev_common.h"
int main(int argc, char **argv)
{
    FILE *f = fopen(argv[1], "r");
    int n;
    if (!f || fscanf(f, "%d", &n) != 1) return 2;
    for (int i = 0; i < n; i++) {
        int kind;
        double v[12];
        if (fscanf(f, "%d", &kind) != 1) return 2;
        for (int k = 0; k < 12; k++)
            if (fscanf(f, "%lf", &v[k]) != 1) return 2;
        double2 P[6];
        for (int k = 0; k < 6; k++) P[k] = make_double2(v[2 * k], v[2 * k + 1]);
        int r;
        if (kind == 10) {   // fixed-width key of entity st = v[0] with points P[1..3] (v[2..7])
            unsigned long long q[4];
            int E;
            const bool ok = rho_key_fixed((int)v[0], P + 1, q, &E);
            printf("%d %d %llu %llu %llu %llu\n", ok ? 1 : 0, E, q[0], q[1], q[2], q[3]);
            continue;
        }
        if (kind >= 21) r = rho_equal_res((kind - 21) & 1, P, ((kind - 21) >> 1) & 1, P + 3, 1);   // key-equal
        else if (kind == 16) r = incircle_fixed(P[0], P[1], P[2], P[3]);
        else if (kind >= 11) r = rho_cmp_fixed((kind - 11) & 1, P, ((kind - 11) >> 1) & 1, P + 3);   // kind 11 + sta + 2 stb
        else if (kind == 0) r = incircle_zero_res(P[0], P[1], P[2], P[3]);
        else if (kind == 1) r = attached_i128(P[0], P[1], P[2]);
        else r = rho_equal_res((kind - 2) & 1, P, ((kind - 2) >> 1) & 1, P + 3);   // kind 2 + sta + 2 stb
        printf("%d\n", r);
    }
    return 0;
}
"""


def incircle_det(pts):
    (ax, ay), (bx, by), (cx, cy), (dx, dy) = pts
    adx, ady, bdx, bdy, cdx, cdy = ax - dx, ay - dy, bx - dx, by - dy, cx - dx, cy - dy
    la, lb, lc = adx * adx + ady * ady, bdx * bdx + bdy * bdy, cdx * cdx + cdy * cdy
    return la * (bdx * cdy - cdx * bdy) + lb * (cdx * ady - adx * cdy) + lc * (adx * bdy - bdx * ady)


def circle_points(r2):
    r = math.isqrt(r2)
    pts = []
    for x in range(-r, r + 1):
        y2 = r2 - x * x
        y = math.isqrt(y2)
        if y * y == y2:
            pts.append((x, y))
            if y:
                pts.append((x, -y))
    return pts


def make_cases(seed=11, n=1200):
    rng = random.Random(seed)
    circles = [circle_points(r2) for r2 in (25, 625, 5 ** 6, 65 * 65, 1105 * 1105)]
    cases = []
    for it in range(n):
        kind = it % 4
        if kind == 0:   # random integers at a random binary scale
            bits = rng.choice([4, 10, 20, 30, 40, 50])
            pts = [(rng.randint(-2 ** bits, 2 ** bits), rng.randint(-2 ** bits, 2 ** bits)) for _ in range(4)]
            sc = Fraction(2) ** rng.randint(-40, 10)
        elif kind in (1, 2):   # four lattice points of one circle, shifted; kind 2 moves one by a unit
            pts = rng.sample(rng.choice(circles), 4)
            ox, oy = rng.randint(-2 ** 30, 2 ** 30), rng.randint(-2 ** 30, 2 ** 30)
            pts = [(x + ox, y + oy) for x, y in pts]
            if kind == 2:
                i = rng.randrange(4)
                pts[i] = (pts[i][0] + rng.choice([-1, 1]), pts[i][1])
            sc = Fraction(2) ** rng.randint(-60, 0)
        else:   # squares of a lattice snapped at 2^-23 in [0, 1), some corners moved by one unit
            x0, y0, s = rng.randint(0, 2 ** 22), rng.randint(0, 2 ** 22), rng.randint(1, 2000)
            pts = [(x0, y0), (x0 + s, y0), (x0 + s, y0 + s), (x0, y0 + s)]
            if rng.random() < 0.3:
                pts[3] = (pts[3][0], pts[3][1] + 1)
            rng.shuffle(pts)
            sc = Fraction(1, 2 ** 23)
        q = [(Fraction(x) * sc, Fraction(y) * sc) for x, y in pts]
        assert all(float(v) == v for p in q for v in p)   # every coordinate is an exact double
        cases.append(([float(v) for p in q for v in p], 1 if incircle_det(q) == 0 else 0))
    # spreads beyond 60 bits: the test defers to the big integers
    cases.append(([1e10, 0.3, 1e-10, 2.0, 5.0, 7.0, 0.0, 1.0], -1))
    cases.append(([0.0] * 8, 1))
    return [(0, v + [0.0] * 4, z) for v, z in cases]


def _sign(x):
    return (x > 0) - (x < 0)


def rho_value(pts):
    if len(pts) == 2:
        (ax, ay), (bx, by) = pts
        return (ax - bx) ** 2 + (ay - by) ** 2
    (ax, ay), (bx, by), (cx, cy) = pts
    l1 = (ax - bx) ** 2 + (ay - by) ** 2
    l2 = (bx - cx) ** 2 + (by - cy) ** 2
    l3 = (cx - ax) ** 2 + (cy - ay) ** 2
    ar = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
    return l1 * l2 * l3 / (ar * ar)


def make_alpha_cases(seed=12, n=800):
    rng = random.Random(seed)
    cases = []
    unit = Fraction(1, 2 ** 23)
    for it in range(n):
        kind = it % 4
        if kind == 0:   # a lattice triangle and a moved, turned copy (equal), sometimes one corner off
            t = [(rng.randint(0, 3000), rng.randint(0, 3000)) for _ in range(3)]
            if (t[1][0] - t[0][0]) * (t[2][1] - t[0][1]) == (t[1][1] - t[0][1]) * (t[2][0] - t[0][0]):
                continue
            ox, oy = rng.randint(0, 2 ** 21), rng.randint(0, 2 ** 21)
            u = [(ox - y, oy + x) for x, y in t]   # turned by 90 degrees and moved
            rng.shuffle(u)
            if rng.random() < 0.4:
                u[0] = (u[0][0] + 1, u[0][1])
            a, b = t, u
        elif kind == 1:   # a right triangle against its hypotenuse (equal), or against another edge
            x0, y0, p, q = rng.randint(0, 2 ** 21), rng.randint(0, 2 ** 21), rng.randint(1, 999), rng.randint(1, 999)
            a = [(x0, y0), (x0 + p, y0), (x0, y0 + q)]
            b = [(x0 + p, y0), (x0, y0 + q)] if rng.random() < 0.7 else [(x0, y0), (x0 + p, y0)]
        elif kind == 2:   # random pairs of any kinds
            a = [(rng.randint(0, 2 ** 22), rng.randint(0, 2 ** 22)) for _ in range(rng.choice([2, 3]))]
            b = [(rng.randint(0, 2 ** 22), rng.randint(0, 2 ** 22)) for _ in range(rng.choice([2, 3]))]
            if len(a) == 3 and (a[1][0] - a[0][0]) * (a[2][1] - a[0][1]) == (a[1][1] - a[0][1]) * (a[2][0] - a[0][0]):
                continue
            if len(b) == 3 and (b[1][0] - b[0][0]) * (b[2][1] - b[0][1]) == (b[1][1] - b[0][1]) * (b[2][0] - b[0][0]):
                continue
        else:   # attachment: apex on the circle over the edge (a right angle there) or a unit in or out
            x0, y0, p, q = rng.randint(0, 2 ** 21), rng.randint(0, 2 ** 21), rng.randint(1, 999), rng.randint(1, 999)
            i, j, k = (x0 + p, y0), (x0, y0 + q), (x0, y0)
            if rng.random() < 0.6:
                k = (x0 + rng.choice([-1, 1]), y0 + rng.choice([-1, 0, 1]))
            q3 = [(Fraction(x) * unit, Fraction(y) * unit) for x, y in (i, j, k)]
            (ix, iy), (jx, jy), (kx, ky) = q3
            z = 1 if (ix - jx) ** 2 + (iy - jy) ** 2 >= (ix + jx - 2 * kx) ** 2 + (iy + jy - 2 * ky) ** 2 else 0
            cases.append((1, [float(c) for pt in q3 for c in pt] + [0.0] * 6, z))
            continue
        qa = [(Fraction(x) * unit, Fraction(y) * unit) for x, y in a]
        qb = [(Fraction(x) * unit, Fraction(y) * unit) for x, y in b]
        z = 1 if rho_value(qa) == rho_value(qb) else 0
        va = [float(c) for pt in qa for c in pt] + [0.0] * (6 - 2 * len(qa))
        vb = [float(c) for pt in qb for c in pt] + [0.0] * (6 - 2 * len(qb))
        cases.append((2 + (len(qa) == 3) + 2 * (len(qb) == 3), va + vb, z))
    return cases


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_incircle_zero_residues():
    cases = make_cases() + make_alpha_cases()
    tmp = tempfile.mkdtemp()
    try:
        src, exe, data = (os.path.join(tmp, f) for f in ("h.cpp", "h", "cases.txt"))
        with open(src, "w") as f:
            f.write(HARNESS)
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True, capture_output=True)
        with open(data, "w") as f:
            f.write(f"{len(cases)}\n")
            for kind, v, _ in cases:
                f.write(f"{kind} " + " ".join(repr(x) for x in v) + "\n")
        out = subprocess.run([exe, data], check=True, capture_output=True, text=True).stdout.split()
    finally:
        shutil.rmtree(tmp)
    got = [int(x) for x in out]
    assert len(got) == len(cases)
    want = [z for _, _, z in cases]
    bad = [(i, cases[i][0], got[i], want[i]) for i in range(len(cases)) if got[i] != want[i]]
    assert not bad, bad[:5]
    for kind in (0, 1):   # both outcomes of every test are exercised
        assert {z for k, _, z in cases if k == kind} >= {0, 1}
    assert {z for k, _, z in cases if 2 <= k <= 5} == {0, 1}


def _run_harness(cases):
    tmp = tempfile.mkdtemp()
    try:
        src, exe, data = (os.path.join(tmp, f) for f in ("h.cpp", "h", "cases.txt"))
        with open(src, "w") as f:
            f.write(HARNESS)
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True, capture_output=True)
        with open(data, "w") as f:
            f.write(f"{len(cases)}\n")
            for kind, v in cases:
                f.write(f"{kind} " + " ".join(repr(x) for x in v) + "\n")
        return subprocess.run([exe, data], check=True, capture_output=True, text=True).stdout.splitlines()
    finally:
        shutil.rmtree(tmp)


def _entity(rng, kind, sc=1.0):
    """Points of an alpha-value entity as exact doubles: 2 (edge) or 3 (triangle)."""
    n = 3 if rng.random() < 0.6 else 2
    while True:
        if kind == 0:     # lattice cell i/1414 (rounded doubles, like the lattice workload)
            i, j = rng.randint(0, 1412), rng.randint(0, 1412)
            cand = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            rng.shuffle(cand)
            pts = [(a / 1414, b / 1414) for a, b in cand[:n]]
        elif kind == 1:   # uniform doubles in a small window
            x0, y0 = rng.random(), rng.random()
            pts = [(x0 + rng.random() * 1e-3, y0 + rng.random() * 1e-3) for _ in range(n)]
        else:             # integers at a binary scale sc
            pts = [(rng.randint(-2 ** 40, 2 ** 40) * sc, rng.randint(-2 ** 40, 2 ** 40) * sc) for _ in range(n)]
        q = [(Fraction(x), Fraction(y)) for x, y in pts]
        if n == 3 and (q[1][0] - q[0][0]) * (q[2][1] - q[0][1]) == (q[1][1] - q[0][1]) * (q[2][0] - q[0][0]):
            continue
        if n == 2 and q[0] == q[1]:
            continue
        return pts, rho_value(q)


def _defers(entities, to_last=False):
    """True when the fixed-width path must defer: at the entities' common unit 2^t (the lowest set
    bit of any coordinate) a coordinate needs more than 61 bits or a difference within one entity
    more than 58."""
    vals = [Fraction(c) for pts in entities for pt in pts for c in pt]
    t = min((v.numerator & -v.numerator).bit_length() - v.denominator.bit_length() for v in vals if v)
    X = [[(int(Fraction(x) / Fraction(2) ** t), int(Fraction(y) / Fraction(2) ** t)) for x, y in pts]
         for pts in entities]
    if max(abs(c).bit_length() for pts in X for pt in pts for c in pt) > 61:
        return True
    if to_last:   # incircle: differences to the last point only
        D = max(abs(p[k] - pts[-1][k]).bit_length() for pts in X for p in pts for k in (0, 1))
    else:
        D = max(abs(p[k] - q[k]).bit_length() for pts in X for p in pts for q in pts for k in (0, 1))
    return D > 58


def test_fixed_width_alpha_values():
    """rho_cmp_fixed (sign of rho_a - rho_b) and rho_key_fixed (q = floor(rho 2^-E) in [2^255, 2^256))
    against rationals: neighbouring lattice cells (equal and unequal values a few ulps apart),
    close uniform pairs, wide binary scales, and spreads beyond 58 bits that defer."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    rng = random.Random(21)
    cmp_cases, key_cases = [], []
    for it in range(1500):
        kind = it % 3
        sc = 2.0 ** rng.randint(-60, 20)   # one scale for both entities of a pair
        (pa, ra), (pb, rb) = _entity(rng, kind, sc), _entity(rng, kind, sc)
        if it % 5 == 0:   # the same cell moved by whole lattice steps: often the same value
            d = rng.randint(1, 50)
            pb = [(round(x * 1414) / 1414, (round(y * 1414) + d) / 1414) for x, y in pa] if kind == 0 else pa
            rb = rho_value([(Fraction(x), Fraction(y)) for x, y in pb])
        va = [c for pt in pa for c in pt] + [0.0] * (6 - 2 * len(pa))
        vb = [c for pt in pb for c in pt] + [0.0] * (6 - 2 * len(pb))
        cmp_cases.append((11 + (len(pa) == 3) + 2 * (len(pb) == 3), va + vb,
                          2 if _defers([pa, pb]) else _sign(ra - rb)))
        # rho_value is num / den; the alpha value is num / (4 den)
        key_cases.append((10, [float(len(pa) == 3), 0.0] + va + [0.0] * 4, None if _defers([pa]) else ra / 4))
    # spreads beyond 58 bits: both defer
    wide = [1.0, 0.0, 2.0 ** -70, 0.0]
    cmp_cases.append((11, wide + [0.0] * 2 + [0.5, 0.5, 0.0, 0.0, 0.0, 0.0], 2))
    key_cases.append((10, [0.0, 0.0] + wide + [0.0] * 6, None))
    out = _run_harness([(k, v) for k, v, _ in cmp_cases] + [(k, v) for k, v, _ in key_cases])
    got_cmp = [int(x) for x in out[:len(cmp_cases)]]
    want_cmp = [w for _, _, w in cmp_cases]
    bad = [(i, got_cmp[i], want_cmp[i]) for i in range(len(cmp_cases)) if got_cmp[i] != want_cmp[i]]
    assert not bad, bad[:5]
    assert set(want_cmp) == {-1, 0, 1, 2}
    assert sum(w == 2 for w in want_cmp) < len(want_cmp) // 20   # deferral stays rare
    nkey = 0
    for line, (_, v, rho) in zip(out[len(cmp_cases):], key_cases):
        ok, E, *q = (int(x) for x in line.split())
        if rho is None:
            assert ok == 0
            continue
        assert ok == 1, v
        Q = sum(w << (64 * i) for i, w in enumerate(q))
        e = rho.numerator.bit_length() - rho.denominator.bit_length()   # 2^e <= rho < 2^(e + 1), or one less
        if Fraction(2) ** e > rho:
            e -= 1
        assert E == e - 255, (v, E, e)
        assert Q == math.floor(rho / Fraction(2) ** E)
        assert 2 ** 255 <= Q < 2 ** 256
        nkey += 1
    assert nkey > 0.95 * len(key_cases)


DIV_HARNESS = r"""
#include <hip/hip_runtime.h>
#include <cstdio>
#include "dev_common.h"
int main(int argc, char **argv)
{
    FILE *f = fopen(argv[1], "r");
    int n;
    if (!f || fscanf(f, "%d", &n) != 1) return 2;
    for (int c = 0; c < n; c++) {
        uint32_t un[17], vn[8], qd[9];
        for (int i = 0; i < 16; i++)
            if (fscanf(f, "%x", &un[i]) != 1) return 2;
        un[16] = 0;
        for (int i = 0; i < 8; i++)
            if (fscanf(f, "%x", &vn[i]) != 1) return 2;
        mw_div_17_8(un, vn, qd);
        for (int i = 8; i >= 0; i--) printf("%08x", qd[i]);
        printf(" ");
        for (int i = 7; i >= 0; i--) printf("%08x", un[i]);
        printf("\n");
    }
    return 0;
}
"""


def _add_backs(U, V):
    """Quotient digits of U / V (32-bit digits, 17 by 8, V's top bit set) whose two-digit estimate
    still overshoots, so that the division must add back: the rare branch worth pinning."""
    b = 1 << 32
    vn = [(V >> (32 * i)) & (b - 1) for i in range(8)]
    n = 0
    R = U
    for j in range(8, -1, -1):
        un = [(R >> (32 * i)) & (b - 1) for i in range(17)]
        top = (un[j + 8] << 32) | un[j + 7]
        qhat, rhat = divmod(top, vn[7])
        while qhat >= b or qhat * vn[6] > ((rhat << 32) | un[j + 6]):
            qhat -= 1
            rhat += vn[7]
            if rhat >= b:
                break
        qtrue = (R >> (32 * j)) // V
        n += qhat > qtrue
        R -= qtrue * V << (32 * j)
    return n


def test_fixed_division_add_back():
    """mw_div_17_8 (the key's long division) against Python's integers, on digits drawn from edge
    values (0, 1, 2^31 - 1, 2^31, 2^32 - 1) that drive the estimate into its corrections and the
    add-back branch, which random keys almost never reach."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    rng = random.Random(31)
    edge = [0, 1, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 1]
    cases = []
    while len(cases) < 3000:
        vd = [rng.choice(edge) if rng.random() < 0.8 else rng.getrandbits(32) for _ in range(8)]
        vd[7] |= 1 << 31
        V = sum(d << (32 * i) for i, d in enumerate(vd))
        ud = [rng.choice(edge) if rng.random() < 0.8 else rng.getrandbits(32) for _ in range(16)]
        U = sum(d << (32 * i) for i, d in enumerate(ud))
        cases.append((U, V))
    hits = sum(_add_backs(U, V) > 0 for U, V in cases)
    assert hits >= 10   # the add-back branch is exercised
    tmp = tempfile.mkdtemp()
    try:
        src, exe, data = (os.path.join(tmp, f) for f in ("d.cpp", "d", "cases.txt"))
        with open(src, "w") as f:
            f.write(DIV_HARNESS)
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True, capture_output=True)
        with open(data, "w") as f:
            f.write(f"{len(cases)}\n")
            for U, V in cases:
                f.write(" ".join("%x" % ((U >> (32 * i)) & 0xffffffff) for i in range(16)) + " ")
                f.write(" ".join("%x" % ((V >> (32 * i)) & 0xffffffff) for i in range(8)) + "\n")
        out = subprocess.run([exe, data], check=True, capture_output=True, text=True).stdout.split("\n")
    finally:
        shutil.rmtree(tmp)
    for (U, V), line in zip(cases, out):
        q, r = (int(x, 16) for x in line.split())
        assert (q, r) == divmod(U, V), (hex(U), hex(V))


def test_fixed_width_incircle():
    """incircle_fixed (the determinant's sign in 256-bit two's complement) against rationals on the
    co-circular, perturbed and random quadruples of the residue test, plus far translations and
    spreads beyond 58 bits, which defer (2)."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    rng = random.Random(41)
    cases = []
    for _, v, _ in make_cases(seed=13, n=1600):
        pts = [(v[2 * k], v[2 * k + 1]) for k in range(4)]
        if rng.random() < 0.3:   # shift far away: long coordinates, short differences
            sh = rng.randint(1, 2 ** 20) * 2.0 ** 20
            moved = [(x + sh, y - sh) for x, y in pts]
            if any(Fraction(mx) != Fraction(x) + Fraction(sh) or Fraction(my) != Fraction(y) - Fraction(sh)
                   for (x, y), (mx, my) in zip(pts, moved)):
                continue   # the shift was not exact in doubles
            pts = moved
        q = [(Fraction(x), Fraction(y)) for x, y in pts]
        want = 2 if _defers([pts], to_last=True) else _sign(incircle_det(q))
        cases.append((16, [c for p in pts for c in p] + [0.0] * 4, want))
    out = _run_harness([(k, v) for k, v, _ in cases])
    got = [int(x) for x in out]
    bad = [(i, cases[i][1][:8], got[i], cases[i][2]) for i in range(len(cases)) if got[i] != cases[i][2]]
    assert not bad, bad[:5]
    assert {w for _, _, w in cases} == {-1, 0, 1, 2}
    assert sum(w == 2 for _, _, w in cases) < len(cases) // 4


def test_key_equal_residue_bound():
    """rho_equal_res with key_equal (the values share their 256-bit key, so fewer primes or none):
    equal values stay certified equal at every difference length, from lattice-sized cells
    (no prime needed) to cells whose differences take 40-58 bits."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    rng = random.Random(51)
    cases = []
    for it in range(600):
        bits = rng.choice([10, 20, 30, 40, 48, 52])   # coordinates stay exact doubles
        t = [(rng.randint(0, 2 ** bits), rng.randint(0, 2 ** bits)) for _ in range(3)]
        if (t[1][0] - t[0][0]) * (t[2][1] - t[0][1]) == (t[1][1] - t[0][1]) * (t[2][0] - t[0][0]):
            continue
        ox, oy = rng.randint(0, 2 ** 4), rng.randint(0, 2 ** 4)
        u = [(ox - y, oy + x) for x, y in t]   # turned by 90 degrees and moved: the same value
        rng.shuffle(u)
        if rng.random() < 0.3:   # an edge against the same edge turned
            t, u = t[:2], [(ox - y, oy + x) for x, y in t[:2]]
        sc = Fraction(1, 2 ** 60)
        qa = [(Fraction(x) * sc, Fraction(y) * sc) for x, y in t]
        qb = [(Fraction(x) * sc, Fraction(y) * sc) for x, y in u]
        assert rho_value(qa) == rho_value(qb)
        assert all(Fraction(float(c)) == c for pt in qa + qb for c in pt)
        va = [float(c) for pt in qa for c in pt] + [0.0] * (6 - 2 * len(qa))
        vb = [float(c) for pt in qb for c in pt] + [0.0] * (6 - 2 * len(qb))
        want = -1 if _defers([[(float(x), float(y)) for x, y in qa], [(float(x), float(y)) for x, y in qb]]) else 1
        cases.append((21 + (len(qa) == 3) + 2 * (len(qb) == 3), va + vb, want))
    out = _run_harness([(k, v) for k, v, _ in cases])
    got = [int(x) for x in out]
    bad = [(i, got[i], cases[i][2]) for i in range(len(cases)) if got[i] != cases[i][2] and cases[i][2] == 1]
    assert not bad, bad[:5]
    assert sum(w == 1 for _, _, w in cases) > 0.9 * len(cases)


BIGINT_HARNESS = r"""
#include <hip/hip_runtime.h>
#include <cstdio>
#include "dev_common.h"
int main(int argc, char **argv)
{
    FILE *f = fopen(argv[1], "r");
    int n;
    if (!f || fscanf(f, "%d", &n) != 1) return 2;
    for (int c = 0; c < n; c++) {
        int kind;
        double v[8];
        if (fscanf(f, "%d", &kind) != 1) return 2;
        for (int k = 0; k < 8; k++)
            if (fscanf(f, "%lf", &v[k]) != 1) return 2;
        double2 P[4];
        for (int k = 0; k < 4; k++) P[k] = make_double2(v[2 * k], v[2 * k + 1]);
        int ovf = 0;
        const int r = kind == 0 ? orient_exact(P[0], P[1], P[2], &ovf) : incircle_exact(P[0], P[1], P[2], P[3], &ovf);
        printf("%d %d\n", r, ovf);
    }
    return 0;
}
"""


def test_bigint_predicates_host():
    """The big-integer fallbacks orient_exact and incircle_exact (the device's last resort behind
    every filter, residue and fixed-width form), built for the host from the same header, against
    rationals: collinear and co-circular pixel-grid cases at the snapping scales, random doubles,
    and integers over wide binary scales (beyond the fixed-width limits)."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    rng = random.Random(61)
    cases = []
    for it in range(1200):
        T = rng.choice([256, 1024, 4096])
        sc = 2.0 ** -math.ceil(math.log2(T)) * (1 + rng.randint(0, 7) / 8)   # few significant bits
        mode = it % 3
        if mode == 0:     # pixel-grid triples and quadruples, many collinear / co-circular
            a = (rng.randint(0, T), rng.randint(0, T))
            d = (rng.randint(-4, 4), rng.randint(-4, 4))
            pts = [a, (a[0] + d[0], a[1] + d[1]), (a[0] + 2 * d[0], a[1] + 2 * d[1] + rng.choice([0, 0, 1])),
                   (a[0] + d[1], a[1] - d[0])]
            pts = [(x * sc, y * sc) for x, y in pts]
        elif mode == 1:   # random doubles
            pts = [(rng.random(), rng.random()) for _ in range(4)]
        else:             # integers at a wide binary scale: long coordinates
            s2 = 2.0 ** rng.randint(-200, 100)
            pts = [(rng.randint(-2 ** 52, 2 ** 52) * s2, rng.randint(-2 ** 52, 2 ** 52) * s2) for _ in range(4)]
        q = [(Fraction(x), Fraction(y)) for x, y in pts]
        (ax, ay), (bx, by), (cx, cy) = q[:3]
        o = (ax - cx) * (by - cy) - (ay - cy) * (bx - cx)
        v = [c for p in pts for c in p]
        cases.append((0, v, _sign(o)))
        cases.append((1, v, _sign(incircle_det(q))))
    tmp = tempfile.mkdtemp()
    try:
        # the header's device-only inline functions, declared for the host as well
        with open(os.path.join(CSRC, "dev_common.h")) as f:
            hdr = f.read()
        assert "#define GA_INLINE __device__ __forceinline__" in hdr
        with open(os.path.join(tmp, "dev_common.h"), "w") as f:
            f.write(hdr.replace("#define GA_INLINE __device__ __forceinline__",
                                "#define GA_INLINE __host__ __device__ __forceinline__"))
        src, exe, data = (os.path.join(tmp, x) for x in ("b.cpp", "b", "cases.txt"))
        with open(src, "w") as f:
            f.write(BIGINT_HARNESS)
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", tmp, src, "-o", exe], check=True, capture_output=True)
        with open(data, "w") as f:
            f.write(f"{len(cases)}\n")
            for kind, v, _ in cases:
                f.write(f"{kind} " + " ".join(repr(x) for x in v) + "\n")
        out = subprocess.run([exe, data], check=True, capture_output=True, text=True).stdout.splitlines()
    finally:
        shutil.rmtree(tmp)
    bad = []
    for (kind, v, want), line in zip(cases, out):
        r, ovf = (int(x) for x in line.split())
        assert ovf == 0, v
        if r != want:
            bad.append((kind, v, r, want))
    assert not bad, bad[:5]
    for kind in (0, 1):
        assert {w for k, _, w in cases if k == kind} == {-1, 0, 1}
