"""The residue zero test of the device incircle determinant (`incircle_zero_res` in
galpha2d_amd/csrc/dev_common.h), compiled for the host and checked against exact rational
arithmetic: co-circular quadruples (lattice points on circles with many of them, lattice
squares at a binary snapping unit) must come out zero, perturbed and random ones nonzero, and a
coordinate spread beyond 60 bits must defer (-1) instead of answering.  A wrong prime, a dropped
term of the determinant or a bound that lets a nonzero determinant vanish fails here."""
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
        double v[8];
        for (int k = 0; k < 8; k++)
            if (fscanf(f, "%lf", &v[k]) != 1) return 2;
        printf("%d\n", incircle_zero_res(make_double2(v[0], v[1]), make_double2(v[2], v[3]),
                                         make_double2(v[4], v[5]), make_double2(v[6], v[7])));
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
    return cases


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_incircle_zero_residues():
    cases = make_cases()
    tmp = tempfile.mkdtemp()
    try:
        src, exe, data = (os.path.join(tmp, f) for f in ("h.cpp", "h", "cases.txt"))
        with open(src, "w") as f:
            f.write(HARNESS)
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True, capture_output=True)
        with open(data, "w") as f:
            f.write(f"{len(cases)}\n")
            for v, _ in cases:
                f.write(" ".join(repr(x) for x in v) + "\n")
        out = subprocess.run([exe, data], check=True, capture_output=True, text=True).stdout.split()
    finally:
        shutil.rmtree(tmp)
    got = [int(x) for x in out]
    assert len(got) == len(cases)
    want = [z for _, z in cases]
    assert got == want
    assert sum(1 for z in want if z == 1) > 300   # the co-circular cases are exercised
