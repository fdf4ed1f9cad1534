"""Generates tests/golden/digests.json: oracle digests (canonical triangles and every alpha
array) of the full-size workloads, which the oracle needs minutes for and the GPU tests compare
against.  Run from the repo root after `make -C oracle`:
    python tests/golden/make_digests.py [workload ...]      (default: all)
Single-set workloads (one SHA-256 per output array):
  uniform_1m    1,048,576 i.i.d. U[0,1)^2 points (splitmix64 seed 2)
  lattice_2m    2,000,000 points: 1M from a 1414 x 1414 lattice, 1M from 64 Gaussian clusters
                (sigma 0.005), seed 3
  lattice_400k  400,000 points: 200K from a 632 x 632 lattice, 200K from 16 clusters, seed 3
Many-set workloads (one SHA-256 per set over its triangles and all its alpha arrays, in the
order of set_digest(); stored as the first 16 hex digits per set plus the SHA-256 of all full
set digests in set order):
  batched_4096  the batched workload: datagen.batch_sets(4096, seed=4), 4,096 sets of 256..65,536
                points, uniform or 8-cluster Gaussian
  serving_24    the serving parity workload: datagen.serving_requests(24, seed=5), 24 windows of
                datagen.uniform(5, 2000001) with sizes log-uniform in [1,000, 2,000,000]
The many-set workloads run the oracle on a pool of host processes (one set per task).
"""
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from conftest import digest, set_digest  # noqa: E402
from galpha2d_amd import datagen  # noqa: E402
from oracle import pyoracle  # noqa: E402

SINGLE = {
    "uniform_1m": lambda: datagen.uniform(2, 1048576),
    "lattice_400k": lambda: datagen.lattice_clusters(400000, side=632, nclusters=16, sigma=0.005, seed=3),
    "lattice_2m": lambda: datagen.lattice_clusters(2000000, side=1414, nclusters=64, sigma=0.005, seed=3),
}
MANY = {
    "batched_4096": lambda: datagen.batch_sets(4096, seed=4),
    "serving_24": lambda: datagen.serving_requests(24, seed=5),
}


def _one_set(xy):
    tris, _ = pyoracle.delaunay(xy)
    return set_digest(tris, pyoracle.alpha(xy, tris))


def many_digest(sets, workers):
    # largest sets first, so the pool ends together
    order = sorted(range(len(sets)), key=lambda k: -sets[k].shape[0])
    with ProcessPoolExecutor(workers) as pool:
        got = list(pool.map(_one_set, [sets[k] for k in order], chunksize=4))
    full = [None] * len(sets)
    for k, d in zip(order, got):
        full[k] = d
    return {"sets": len(sets), "points": int(sum(x.shape[0] for x in sets)),
            "all": hashlib.sha256("".join(full).encode()).hexdigest(),
            "per_set16": [d[:16] for d in full]}


def main():
    names = sys.argv[1:] or list(SINGLE) + list(MANY)
    path = os.path.join(HERE, "digests.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    workers = min(8, os.cpu_count() or 1)
    for name in names:
        t0 = time.time()
        if name in SINGLE:
            xy = SINGLE[name]()
            tris, _ = pyoracle.delaunay(xy)
            t1 = time.time()
            a = pyoracle.alpha(xy, tris)
            out[name] = digest(tris, a)
            out[name]["n"] = int(xy.shape[0])
            print(name, "dt %.1fs alpha %.1fs" % (t1 - t0, time.time() - t1), out[name]["nTris"],
                  out[name]["specSize"], flush=True)
        else:
            out[name] = many_digest(MANY[name](), workers)
            print(name, "%d sets, %d points, %.1fs on %d processes" % (out[name]["sets"], out[name]["points"],
                                                                        time.time() - t0, workers), flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
