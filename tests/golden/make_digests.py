"""Generates tests/golden/digests.json: oracle digests (canonical triangles and every alpha
array) of the full-size workloads, which the oracle needs minutes for and the GPU tests compare
against.  Run from the repo root after `make -C oracle`:
    python tests/golden/make_digests.py
Workloads:
  uniform_1m  1,048,576 i.i.d. U[0,1)^2 points (splitmix64 seed 2)
  lattice_2m  2,000,000 points: 1M from a 1414 x 1414 lattice, 1M from 64 Gaussian clusters
              (sigma 0.005), seed 3
  lattice_400k  400,000 points: 200K from a 632 x 632 lattice, 200K from 16 clusters, seed 3
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from conftest import digest  # noqa: E402
from galpha2d_amd import datagen  # noqa: E402
from oracle import pyoracle  # noqa: E402

WORKLOADS = {
    "uniform_1m": lambda: datagen.uniform(2, 1048576),
    "lattice_400k": lambda: datagen.lattice_clusters(400000, side=632, nclusters=16, sigma=0.005, seed=3),
    "lattice_2m": lambda: datagen.lattice_clusters(2000000, side=1414, nclusters=64, sigma=0.005, seed=3),
}


def main():
    out = {}
    for name, gen in WORKLOADS.items():
        xy = gen()
        t0 = time.time()
        tris, _ = pyoracle.delaunay(xy)
        t1 = time.time()
        a = pyoracle.alpha(xy, tris)
        t2 = time.time()
        out[name] = digest(tris, a)
        out[name]["n"] = int(xy.shape[0])
        print(name, "dt %.1fs alpha %.1fs" % (t1 - t0, t2 - t1), out[name]["nTris"], out[name]["specSize"],
              flush=True)
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
