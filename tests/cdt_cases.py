"""Constraint sets for the constrained-DT tests: non-crossing by construction."""
import numpy as np


def star_polygon(xy, k, seed, center=(0.5, 0.5)):
    """Edges of a polygon through k random points, taken in angular order around `center`:
    a star-shaped simple polygon, so its edges never cross."""
    rng = np.random.default_rng(seed)
    c = np.asarray(center)
    d = xy - c
    ang = np.arctan2(d[:, 1], d[:, 0])
    sel = rng.choice(len(xy), k, replace=False)
    sel = sel[np.argsort(ang[sel])]
    return np.stack([sel, np.roll(sel, -1)], 1).astype(np.int32)


def cell_polygons(xy, cells, per, seed):
    """One star polygon of up to `per` points inside each cell of a cells x cells grid over the
    bounding box: every edge stays inside its (convex) cell, so polygons of different cells
    never meet."""
    rng = np.random.default_rng(seed)
    lo, hi = xy.min(0), xy.max(0)
    ij = np.minimum(((xy - lo) / (hi - lo + 1e-300) * cells).astype(int), cells - 1)
    key = ij[:, 0] * cells + ij[:, 1]
    out = []
    for cell in np.unique(key):
        idx = np.nonzero(key == cell)[0]
        if len(idx) < 3:
            continue
        pick = rng.choice(idx, min(per, len(idx)), replace=False)
        ctr = xy[pick].mean(0)
        ang = np.arctan2(xy[pick, 1] - ctr[1], xy[pick, 0] - ctr[0])
        pick = pick[np.argsort(ang)]
        out.append(np.stack([pick, np.roll(pick, -1)], 1))
    return np.concatenate(out).astype(np.int32)


def lattice_knights(side, n_lat_sel, k, seed):
    """Knight-move segments ((i, j) -> (i + 1, j + 2)) between selected lattice nodes of a
    side x side lattice scaled to [0, 1)^2.  A knight move passes through no other lattice node;
    segments are kept greedily when they cross none kept before (exact integer test).
    Returns (segments as lattice-node pairs (node = i * side + j), node list)."""
    rng = np.random.default_rng(seed)
    nodes = set(rng.choice(side * side, n_lat_sel, replace=False).tolist())
    cand = []
    for v in nodes:
        i, j = divmod(v, side)
        for di, dj in ((1, 2), (2, 1), (1, -2), (2, -1)):
            w = (i + di) * side + (j + dj)
            if 0 <= i + di < side and 0 <= j + dj < side and w in nodes:
                cand.append((v, w))
    rng.shuffle(cand)

    def pt(v):
        return divmod(v, side)

    def orient(a, b, c):
        return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])

    def cross(s, t):
        if len({*s, *t}) < 4:
            return False
        a, b, c, d = (pt(v) for v in (*s, *t))
        return orient(a, b, c) * orient(a, b, d) < 0 and orient(c, d, a) * orient(c, d, b) < 0

    kept = []
    for s in cand:
        if len(kept) >= k:
            break
        if not any(cross(s, t) for t in kept):
            kept.append(s)
    return kept, sorted(nodes)
