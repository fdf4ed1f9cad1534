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
