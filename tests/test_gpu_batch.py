"""Batched mode: every set of one batched pass equals the oracle (canonical triangles and every
alpha array, bit-exact), whatever else is in the batch."""
import numpy as np
import pytest

from conftest import ALPHA_FIELDS, lattice_clusters

pytestmark = pytest.mark.gpu


def _mixed_sets(oracle):
    rng = np.random.default_rng(7)
    sets = []
    for k, n in enumerate([3, 7, 50, 300, 1000, 2500, 4000, 700, 120, 64]):
        sets.append(oracle.gen_uniform(100 + k, n))
    # Gaussian clusters, a lattice with ties, exact duplicates, collinear points, tiny sets
    c = rng.random((4, 2))
    sets.append(np.clip(c[rng.integers(0, 4, 3000)] + rng.normal(size=(3000, 2)) * 0.01, 0, 0.999999))
    sets.append(np.array([(i / 17.0, j / 17.0) for i in range(17) for j in range(17)]))
    u = oracle.gen_uniform(55, 400)
    sets.append(np.concatenate([u, u[:40]]))
    sets.append(np.array([[i * 0.1, i * 0.05] for i in range(12)]))
    sets.append(np.zeros((0, 2)))
    sets.append(np.array([[0.5, 0.5], [0.25, 0.75]]))
    sets.append(lattice_clusters(1500, 1500, 60, 4, 0.01, 9))
    return sets


def _check_set(oracle, xy, tris, alpha):
    ref, _ = oracle.delaunay(xy) if xy.shape[0] >= 3 else (np.zeros((0, 9), np.int32), None)
    assert np.array_equal(tris, ref)
    if alpha is None:
        return
    if ref.shape[0] == 0:
        assert len(alpha["spectrum"]) == 1 and np.all(alpha["vert_nu_min"] == -1)
        return
    ra = oracle.alpha(xy, ref)
    for k in ALPHA_FIELDS:
        g, r = np.asarray(alpha[k]), np.asarray(ra[k])
        assert g.shape == r.shape, k
        if k == "spectrum":
            assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), "spectrum differs in ulps"
        else:
            assert np.array_equal(g, r), k


def test_batch_matches_oracle(oracle):
    import galpha2d_amd as G
    print("case: mixed batch of 17 sets, DT + alpha", flush=True)
    sets = _mixed_sets(oracle)
    ctx = G.Context(0)
    tris = ctx.batch(sets, with_alpha=True)
    assert len(tris) == len(sets)
    for k, xy in enumerate(sets):
        _check_set(oracle, xy, tris[k], ctx.batch_alpha(k))
    ctx.close()


def test_batch_explicit_textures_and_reuse(oracle):
    """Explicit texture sides, then a second (smaller) batch on the same context."""
    import galpha2d_amd as G
    print("case: batch with explicit textures, context reuse", flush=True)
    ctx = G.Context(0)
    sets = [oracle.gen_uniform(200 + k, n) for k, n in enumerate([2000, 500, 9000, 30])]
    tris = ctx.batch(sets, fbo_sizes=[512, 256, 1024, 64], with_alpha=True)
    for k, xy in enumerate(sets):
        _check_set(oracle, xy, tris[k], ctx.batch_alpha(k))
    sets2 = sets[1:3]
    tris2 = ctx.batch(sets2, with_alpha=False)
    for k, xy in enumerate(sets2):
        _check_set(oracle, xy, tris2[k], None)
    ctx.close()


def test_batch_equals_plain_loop():
    """The batched workload generator's first sets: batch result == per-set plain API result."""
    import galpha2d_amd as G
    from galpha2d_amd import datagen
    print("case: 96 generated sets, batch vs plain loop", flush=True)
    sets = datagen.batch_sets(96, seed=4)
    ctx = G.Context(0)
    tris = ctx.batch(sets, with_alpha=True)
    for k, xy in enumerate(sets):
        plain = ctx.delaunay(xy)
        assert np.array_equal(tris[k], plain), k
    ctx.close()


def test_deletion_hole_through_diagonal(oracle):
    """A clustered set (8 Gaussian clusters, sigma 0.01, clipped to the unit square) whose snapped
    positions put a deleted vertex exactly on an ear diagonal of its hole; the hole must still be
    triangulated counter-clockwise (plain and batched paths)."""
    import os
    import galpha2d_amd as G
    from conftest import GOLDEN
    print("case: clustered 19976-point set, deletion hole with K on a diagonal", flush=True)
    xy = np.load(os.path.join(GOLDEN, "clustered_19976.npy"))
    ctx = G.Context(0)
    ref, _ = oracle.delaunay(xy)
    assert np.array_equal(ctx.delaunay(xy), ref)
    tris = ctx.batch([xy, oracle.gen_uniform(3, 500)], with_alpha=False)
    assert np.array_equal(tris[0], ref)
    ctx.close()


def test_autotune_picks_candidates_and_keeps_results(oracle):
    import galpha2d_amd as G
    print("case: texture autotuner on a small sample", flush=True)
    sets = [oracle.gen_uniform(500 + k, n) for k, n in enumerate([700, 1500, 5000, 9000, 40000])]
    ctx = G.Context(0)
    dens = ctx.autotune(sets, reps=1)
    assert all(d in (6.0, 9.0, 12.0, 16.0, 24.0, 32.0) for d in dens)
    tris = ctx.batch(sets, with_alpha=False)
    for xy, t in zip(sets, tris):
        ref, _ = oracle.delaunay(xy)
        assert np.array_equal(t, ref)
    for n in (1000, 20000, 100000, 500000):
        T = ctx.texture_for(n)
        assert 256 <= T <= 16384 and (T & (T - 1)) == 0
    ctx.close()


def test_batch_repeated_sets(oracle):
    """The same set several times in one batch (repeated requests): every copy equals the oracle
    (equal values in different sets must not be merged or compared across sets)."""
    import galpha2d_amd as G
    print("case: batch of 6 copies of one set plus a lattice", flush=True)
    u = oracle.gen_uniform(77, 3000)
    g = np.array([(i / 25.0, j / 25.0) for i in range(25) for j in range(25)])
    sets = [u, u, g, u, u, g, u, u]
    ctx = G.Context(0)
    tris = ctx.batch(sets, with_alpha=True)
    for k, xy in enumerate(sets):
        _check_set(oracle, xy, tris[k], ctx.batch_alpha(k))
    ctx.close()


def _digest_sets(ctx, nsets, workers=8):
    """Full set digests (conftest.set_digest) of every set of the context's last batch; the
    copies run on this thread, the hashing on a pool (hashlib releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    from conftest import set_digest
    out = []
    with ThreadPoolExecutor(workers) as pool:
        pending = []
        for k in range(nsets):
            pending.append(pool.submit(set_digest, ctx.batch_triangles(k), ctx.batch_alpha(k)))
            if len(pending) > 4 * workers:   # bound the copied results held in host memory
                out.append(pending.pop(0).result())
        out += [f.result() for f in pending]
    return out


def test_batched_workload_full_size():
    """The batched workload at full size: 4,096 sets of 256..65,536 points (uniform or 8-cluster
    Gaussian, seed 4, 49.5M points), autotuned textures, one device-resident batched pass with
    alpha.  Every set's triangles and alpha arrays equal the oracle's (committed digests)."""
    import hashlib
    import galpha2d_amd as G
    from galpha2d_amd import datagen
    from conftest import golden_digests
    want = golden_digests()["batched_4096"]
    print("case: batched workload, 4096 sets, autotune + galphaBatchDevice", flush=True)
    sets = datagen.batch_sets(4096, seed=4)
    sizes = np.array([x.shape[0] for x in sets], dtype=np.int64)
    assert int(sizes.sum()) == want["points"]
    offs = np.zeros(len(sets) + 1, dtype=np.int32)
    offs[1:] = np.cumsum(sizes)
    ctx = G.Context(0)
    ctx.autotune(sets[:256], reps=1)
    d = G.DeviceBuffer(np.ascontiguousarray(np.concatenate(sets)))
    del sets
    try:
        F = ctx.batch_device(d.ptr, offs, with_alpha=True)
        G.synchronize()
        print(f"batch done: {F} triangles; hashing 4096 sets", flush=True)
        got = _digest_sets(ctx, len(offs) - 1)
    finally:
        d.free()
        ctx.close()
    bad = [k for k, g in enumerate(got) if g[:16] != want["per_set16"][k]]
    assert not bad, f"{len(bad)} sets differ from the oracle, first {bad[:10]}"
    assert hashlib.sha256("".join(got).encode()).hexdigest() == want["all"]
