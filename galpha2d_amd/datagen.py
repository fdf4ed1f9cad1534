"""Seeded synthetic point sets for the benchmark configurations (NumPy, host side).

uniform(seed, n): i.i.d. U[0,1)^2 as FP64 via splitmix64 -> (u >> 11) * 2^-53, the same stream
the test oracle's generator produces.
lattice_clusters(...): half the points drawn without replacement from a side x side lattice
scaled to [0,1)^2, half from Gaussian clusters resampled into [0,1)^2.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count):
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + _GAMMA * np.arange(1, count + 1, dtype=np.uint64)
        z = (s ^ (s >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, n):
    u = splitmix64(seed, 2 * n) >> np.uint64(11)
    return (u.astype(np.float64) * 2.0 ** -53).reshape(n, 2)


def lattice_clusters(n, side=1414, nclusters=64, sigma=0.005, seed=3):
    rng = np.random.default_rng(seed)
    n_lat = n // 2
    n_cl = n - n_lat
    sel = rng.choice(side * side, n_lat, replace=False)
    lat = np.stack([(sel // side) / side, (sel % side) / side], axis=1)
    return np.concatenate([lat, gaussian_clusters(rng, n_cl, nclusters, sigma)])


def gaussian_clusters(rng, n, nclusters, sigma):
    """n points from nclusters Gaussian clusters (centres uniform), resampled into [0,1)^2."""
    cent = rng.random((nclusters, 2))
    pts = cent[rng.integers(0, nclusters, n)] + rng.normal(size=(n, 2)) * sigma
    bad = ((pts < 0) | (pts >= 1)).any(axis=1)
    while bad.any():
        k = int(bad.sum())
        pts[bad] = cent[rng.integers(0, nclusters, k)] + rng.normal(size=(k, 2)) * sigma
        bad = ((pts < 0) | (pts >= 1)).any(axis=1)
    return pts


def serving_sizes(nreq, seed=5, lo=1000, hi=2_000_000):
    """Request sizes log-uniform in [lo, hi] and window starts into a pool of hi + 1 points."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), nreq))).astype(np.int64)
    starts = rng.integers(0, hi + 1 - sizes + 1)
    return sizes, starts


def serving_requests(nreq=24, seed=5, lo=1000, hi=2_000_000):
    """Serving workload in arrival order: windows of one seeded pool, uniform(seed, hi + 1), with
    sizes log-uniform in [lo, hi] (serving_sizes)."""
    pool = uniform(seed, hi + 1)
    sizes, starts = serving_sizes(nreq, seed, lo, hi)
    return [pool[s:s + n].copy() for s, n in zip(starts, sizes)]


def batch_sets(nsets=4096, lo=256, hi=65536, seed=4):
    """Batched workload: sizes log-uniform in [lo, hi], each set uniform or an 8-cluster Gaussian
    mix (sigma 0.01, points outside [0,1)^2 resampled), 50/50."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), nsets))).astype(np.int64)
    sets = []
    for i, n in enumerate(sizes):
        if rng.random() < 0.5:
            sets.append(uniform(seed * 1000003 + i, int(n)))
        else:
            sets.append(gaussian_clusters(rng, int(n), 8, 0.01))
    return sets
