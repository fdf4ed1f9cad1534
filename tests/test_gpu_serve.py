"""Serving: requests submitted together (continuous batching over several slots, admission by
HBM footprint) give exactly the result each request gets alone; a tight budget still serves
everything; a request larger than the budget is refused."""
import numpy as np
import pytest

from conftest import ALPHA_FIELDS

pytestmark = pytest.mark.gpu


def _requests(oracle):
    sizes = [1000, 35000, 3000, 120000, 800, 15000, 5000, 60000, 2500, 9000, 400, 25000]
    reqs = [oracle.gen_uniform(300 + k, n) for k, n in enumerate(sizes)]
    rng = np.random.default_rng(5)
    c = rng.random((8, 2))
    reqs.append(np.clip(c[rng.integers(0, 8, 20000)] + rng.normal(size=(20000, 2)) * 0.01, 0, 0.999999))
    return reqs


def _same(a, b):
    ta, aa = a
    tb, ab = b
    assert np.array_equal(ta, tb)
    for k in ALPHA_FIELDS:
        x, y = np.asarray(aa[k]), np.asarray(ab[k])
        if k == "spectrum":
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), k
        else:
            assert np.array_equal(x, y), k


def test_mix_invariance(oracle):
    import galpha2d_amd as G
    print("case: serving mix invariance, 13 requests, 4 slots", flush=True)
    reqs = _requests(oracle)
    srv = G.Server(0, slots=4, hbm_budget_bytes=24 << 30, max_group=6)
    # alone: one request at a time
    alone = []
    for xy in reqs:
        t = srv.submit(xy, with_alpha=True)
        srv.wait(t)
        alone.append(srv.result(t))
    # mixed: all in flight together
    tickets = [srv.submit(xy, with_alpha=True) for xy in reqs]
    groups = []
    for t in tickets:
        F, lat, q, g = srv.wait(t)
        groups.append(g)
    mixed = [srv.result(t) for t in tickets]
    st = srv.stats()
    srv.close()
    print("groups", groups, "stats", st, flush=True)
    assert max(groups) > 1, "no request shared a pass"
    for a, b in zip(alone, mixed):
        _same(a, b)
    # and against the oracle for the small ones
    for xy, (tris, al) in zip(reqs, mixed):
        if xy.shape[0] > 10000:
            continue
        ref, _ = oracle.delaunay(xy)
        assert np.array_equal(tris, ref)
        ra = oracle.alpha(xy, ref)
        for k in ALPHA_FIELDS:
            x, y = np.asarray(al[k]), np.asarray(ra[k])
            if k == "spectrum":
                assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), k
            else:
                assert np.array_equal(x, y), k


def test_tight_budget_serves_everything(oracle):
    import galpha2d_amd as G
    print("case: serving with a budget that fits about one large request", flush=True)
    reqs = [oracle.gen_uniform(400 + k, n) for k, n in enumerate([150000, 150000, 2000, 150000, 500])]
    srv = G.Server(0, slots=3, hbm_budget_bytes=900 << 20, max_group=4)
    tickets = [srv.submit(xy, with_alpha=False) for xy in reqs]
    for xy, t in zip(reqs, tickets):
        F, lat, q, g = srv.wait(t)
        tris, _ = srv.result(t, alpha=False)
        assert tris.shape[0] == F > 0
    st = srv.stats()
    srv.close()
    assert st["served"] == len(reqs)


def test_oversize_request_refused(oracle):
    import galpha2d_amd as G
    print("case: serving refuses a request above the budget", flush=True)
    srv = G.Server(0, slots=1, hbm_budget_bytes=8 << 20, max_group=1)
    with pytest.raises(G.GalphaError):
        srv.submit(oracle.gen_uniform(1, 100000))
    srv.close()


def test_group_points_cap(oracle):
    """With a cap on points per pass, no pass holds more than the cap unless it holds a single
    request, and every request's result is the one it gets alone."""
    import galpha2d_amd as G
    print("case: serving with 40,000 points per pass", flush=True)
    reqs = _requests(oracle)
    cap = 40000
    srv = G.Server(0, slots=2, hbm_budget_bytes=24 << 30, max_group=8, max_group_points=cap)
    tickets = [srv.submit(xy, with_alpha=True) for xy in reqs]
    groups = []
    for t in tickets:
        F, lat, q, g = srv.wait(t)
        groups.append(g)
    mixed = [srv.result(t) for t in tickets]
    srv.close()
    for xy, g in zip(reqs, groups):
        if xy.shape[0] > cap:
            assert g == 1, "a request above the cap shares its pass"
    assert max(groups) > 1
    solo = G.Server(0, slots=1, hbm_budget_bytes=24 << 30, max_group=1)
    for xy, m in zip(reqs, mixed):
        t = solo.submit(xy, with_alpha=True)
        solo.wait(t)
        _same(solo.result(t), m)
    solo.close()


def test_idle_slot_gives_memory_back(oracle):
    """An idle server whose only holdings sit in one idle slot: a new request that fits only once
    that slot gives its buffers back is served, whichever slot the submission wakes (a slot that
    holds nothing wakes the idle holders instead of waiting alone)."""
    import threading
    import galpha2d_amd as G
    print("case: serving, request that fits only after an idle slot trims", flush=True)
    n = 100000
    budget = int(1.2 * G.server_footprint(n))
    srv = G.Server(0, slots=2, hbm_budget_bytes=budget, max_group=1)
    reqs = [oracle.gen_uniform(500 + k, n) for k in range(4)]
    for k, xy in enumerate(reqs):
        t = srv.submit(xy, with_alpha=False)
        out = {}
        th = threading.Thread(target=lambda: out.setdefault("r", srv.wait(t)), daemon=True)
        th.start()
        th.join(120)
        assert not th.is_alive(), f"request {k} was never admitted"
        assert out["r"][0] > 0
        tris, _ = srv.result(t, alpha=False)
        if k == 0:
            want, _ = oracle.delaunay(xy)
            assert np.array_equal(tris, want)
        # the server is idle now, and one slot holds about one request's buffers
        assert srv.stats()["held_bytes"] > 0
    srv.close()


def test_serving_workload_full_range():
    """The serving workload's size range: 24 requests, sizes log-uniform in [1,000, 2,000,000]
    (windows of one seeded uniform pool, seed 5), arriving in Poisson order while four 2M-point
    requests hold all 4 slots, then served with a 2M-point cap per pass.  Every result equals the
    oracle's (committed digests) and the same request served alone; a request of 1M points or
    more shares a pass with smaller ones."""
    import hashlib
    import time
    import galpha2d_amd as G
    from galpha2d_amd import datagen
    from conftest import golden_digests, set_digest
    want = golden_digests()["serving_24"]
    reqs = datagen.serving_requests(24, seed=5)
    sizes = [x.shape[0] for x in reqs]
    print("case: serving workload, 24 requests of", min(sizes), "..", max(sizes), "points", flush=True)
    srv = G.Server(0, slots=4, hbm_budget_bytes=192 << 30, max_group=64, max_group_points=2_000_000)
    try:
        busy = datagen.uniform(7, 2_000_000)
        blockers = [srv.submit(busy, with_alpha=True) for _ in range(4)]
        gaps = np.random.default_rng(5).exponential(1.0 / 2000.0, len(reqs))   # Poisson, 2000 per second
        tickets = []
        for xy, g in zip(reqs, gaps):
            time.sleep(g)
            tickets.append(srv.submit(xy, with_alpha=True))
        mixed, groups = [], []
        for t in tickets:
            F, lat, q, grp = srv.wait(t)
            groups.append(grp)
            mixed.append(set_digest(*srv.result(t)))
        for t in blockers:
            srv.wait(t)
            srv.release(t)
        print("groups", groups, flush=True)
        alone = []
        for xy in reqs:
            t = srv.submit(xy, with_alpha=True)
            F, lat, q, grp = srv.wait(t)
            assert grp == 1
            alone.append(set_digest(*srv.result(t)))
    finally:
        srv.close()
    for k in range(len(reqs)):
        assert mixed[k] == alone[k], f"request {k} ({sizes[k]} points) differs between the mix and alone"
        assert mixed[k][:16] == want["per_set16"][k], f"request {k} ({sizes[k]} points) differs from the oracle"
    assert hashlib.sha256("".join(mixed).encode()).hexdigest() == want["all"]
    assert any(g > 1 and n >= 1_000_000 for g, n in zip(groups, sizes)), "no 1M+ request shared a pass"
