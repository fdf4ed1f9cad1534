"""Generates tests/golden/predicates.json: near-degenerate orient2d / incircle cases signed
exactly with rational arithmetic (stdlib fractions).  Run from the repo root:
    python tests/golden/make_predicates.py
The grids follow the classic construction for predicate robustness tests: points on a
line or circle perturbed by a few ulps, where the FP64 filter cannot decide."""
import json
import os
import random
from fractions import Fraction as Fr

random.seed(20260101)


def orient(a, b, c):
    a, b, c = [(Fr(x), Fr(y)) for x, y in (a, b, c)]
    d = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
    return (d > 0) - (d < 0)


def incircle(a, b, c, d):
    rows = []
    for p in (a, b, c):
        dx, dy = Fr(p[0]) - Fr(d[0]), Fr(p[1]) - Fr(d[1])
        rows.append((dx, dy, dx * dx + dy * dy))
    (a1, a2, a3), (b1, b2, b3), (c1, c2, c3) = rows
    det = a1 * (b2 * c3 - b3 * c2) - a2 * (b1 * c3 - b3 * c1) + a3 * (b1 * c2 - b2 * c1)
    return (det > 0) - (det < 0)


def ulps(x, k):
    import math
    for _ in range(abs(k)):
        x = math.nextafter(x, math.inf if k > 0 else -math.inf)
    return x


cases_o, cases_i = [], []
# orient: point near the line through a, b (a, b far apart, c moved by ulps)
for _ in range(400):
    a = (random.uniform(-1, 1), random.uniform(-1, 1))
    b = (random.uniform(10, 20), random.uniform(10, 20))
    t = random.random()
    c = (a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1]))
    c = (ulps(c[0], random.randint(-3, 3)), ulps(c[1], random.randint(-3, 3)))
    cases_o.append([a, b, c, orient(a, b, c)])
# exactly collinear (integer-valued)
for _ in range(50):
    a = (float(random.randint(-100, 100)), float(random.randint(-100, 100)))
    d = (float(random.randint(-5, 5)), float(random.randint(-5, 5)))
    k = random.randint(1, 9)
    cases_o.append([a, (a[0] + d[0], a[1] + d[1]), (a[0] + k * d[0], a[1] + k * d[1]), 0])
# incircle: fourth point on or near the circle through three lattice-like points
for _ in range(400):
    cx, cy, r = random.uniform(-1, 1), random.uniform(-1, 1), random.uniform(0.5, 2)
    import math
    ang = sorted(random.uniform(0, 2 * math.pi) for _ in range(4))
    pts = [(cx + r * math.cos(t), cy + r * math.sin(t)) for t in ang]
    d = (ulps(pts[3][0], random.randint(-2, 2)), ulps(pts[3][1], random.randint(-2, 2)))
    cases_i.append([pts[0], pts[1], pts[2], d, incircle(pts[0], pts[1], pts[2], d)])
# exactly co-circular (integer points on x^2 + y^2 = 25)
circ = [(5.0, 0.0), (4.0, 3.0), (3.0, 4.0), (0.0, 5.0), (-3.0, 4.0), (-4.0, 3.0), (-5.0, 0.0),
        (-4.0, -3.0), (-3.0, -4.0), (0.0, -5.0), (3.0, -4.0), (4.0, -3.0)]
for _ in range(60):
    q = sorted(random.sample(range(len(circ)), 4))
    a, b, c, d = (circ[i] for i in q)
    cases_i.append([a, b, c, d, 0])
out = {"orient": cases_o, "incircle": cases_i}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "predicates.json")
with open(path, "w") as f:
    json.dump(out, f)
print("wrote", path, len(cases_o), len(cases_i))
