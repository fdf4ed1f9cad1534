"""The residue zero tests of the device's exact predicates (galpha2d_amd/csrc/dev_common.h),
compiled for the host and checked against exact rational arithmetic:
- `incircle_zero_res`: co-circular quadruples (lattice points on circles with many of them,
  lattice squares at a binary snapping unit) come out zero, perturbed and random ones nonzero;
- `rho_equal_res`: equal alpha values (congruent lattice triangles moved and turned, a right
  triangle against its hypotenuse) come out equal, perturbed and random pairs different;
- `attached_i128` (an exact sign in 128-bit integers): apexes on, inside and outside the circle
  over an edge;
and coordinate spreads beyond their bit limits defer (-1) instead of answering.  A wrong prime, a
dropped term or a bound that lets a nonzero value vanish fails here."""
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
        if (kind == 0) r = incircle_zero_res(P[0], P[1], P[2], P[3]);
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
