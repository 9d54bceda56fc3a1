"""The request server (bnsamd_server_*): continuous batching, admission by
HBM budget, overlapping batches, large requests as interleaved pieces.

A request's result must be the same alone and in the mix: every output of a
request served together with others (different sizes, quadrature orders,
precisions, bremsstrahlung models, submitted from several threads) equals,
bit for bit, the same request computed alone through the context API.
"""
import threading

import numpy as np
import pytest

from bns_nurates_amd import workloads as W
from tests.gpu_util import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from bns_nurates_amd import api as a
    return a


def alone(api, ctx, req):
    ctx.set_quadrature(req["nquad"])
    return ctx.compute_m1_opacities(req["pts"], flags=W.FLAGS_ALL, params=req["params"],
                                    precision=req["precision"])


def make_requests(n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        npts = int(rng.choice([1, 512, 4096, 32768, 262144], p=[0.15, 0.35, 0.3, 0.12, 0.08]))
        reqs.append(dict(pts=np.ascontiguousarray(W.make_points(npts, seed=1000 + i)),
                         nquad=int(rng.choice([8, 10, 16])), precision=int(rng.integers(0, 2)),
                         params=W.params_default_all(int(rng.choice([W.BREM_BRT06, W.BREM_HR98])))))
    return reqs


def test_results_same_alone_and_in_the_mix(api):
    reqs = make_requests(36, seed=7)
    say(f"serving: {len(reqs)} requests, {sum(r['pts'].shape[1] for r in reqs)} points, 3 threads")
    srv = api.Server(0, hbm_budget_bytes=256 << 20, streams=4, max_batch_points=1 << 16)
    outs = [np.empty((20, r["pts"].shape[1])) for r in reqs]
    ids = [None] * len(reqs)

    def submitter(k):
        for i in range(k, len(reqs), 3):
            r = reqs[i]
            ids[i] = srv.submit(r["pts"], outs[i], r["nquad"], W.FLAGS_ALL, r["params"],
                                precision=r["precision"])

    th = [threading.Thread(target=submitter, args=(k,)) for k in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    lat = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    say(f"served: {st}; latency max {max(lat) * 1e3:.2f} ms")
    assert st["requests_completed"] == len(reqs)
    ctx = api.Context(0)
    for i, r in enumerate(reqs):
        ref = alone(api, ctx, r)
        assert np.array_equal(outs[i], ref, equal_nan=True), (i, r["nquad"], r["precision"])
    ctx.close()


def test_pending_requests_join_one_batch(api):
    """While one large request runs on the only stream, twenty small
    requests of one kind queue up and leave in a single batch."""
    srv = api.Server(0, hbm_budget_bytes=512 << 20, streams=1, max_batch_points=1 << 20,
                     chunk_points=1 << 20)  # the large request in one piece
    big = np.ascontiguousarray(W.make_points(262144, seed=3))
    small = [np.ascontiguousarray(W.make_points(64, seed=50 + i)) for i in range(20)]
    prm = W.params_default_all(W.BREM_BRT06)
    ob = np.empty((20, big.shape[1]))
    os_ = [np.empty((20, 64)) for _ in small]
    ib = srv.submit(big, ob, 16, W.FLAGS_ALL, prm)
    ids = [srv.submit(p, o, 10, W.FLAGS_ALL, prm) for p, o in zip(small, os_)]
    srv.wait(ib)
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    srv.close()
    say(f"batches {st['batches']} for 21 requests")
    assert st["batches"] <= 3
    ctx = api.Context(0)
    ctx.set_quadrature(10)
    for p, o in zip(small, os_):
        assert np.array_equal(o, ctx.compute_m1_opacities(p, flags=W.FLAGS_ALL, params=prm))
    ctx.close()


def test_small_requests_overtake_a_large_ones_pieces(api):
    """A 64^3 request runs as eight pieces; small requests submitted just
    after it are served before its later pieces (smallest request first), so
    they complete first, and every result equals the one computed alone."""
    srv = api.Server(0, hbm_budget_bytes=512 << 20, streams=1, max_batch_points=1 << 20,
                     chunk_points=32768)
    prm = W.params_default_all(W.BREM_BRT06)
    big = np.ascontiguousarray(W.make_points(262144, seed=4))
    small = [np.ascontiguousarray(W.make_points(n, seed=60 + i))
             for i, n in enumerate([1, 512, 4096, 1, 64])]
    ob = np.empty((20, big.shape[1]))
    os_ = [np.empty((20, p.shape[1])) for p in small]
    ib = srv.submit(big, ob, 16, W.FLAGS_ALL, prm)
    ids = [srv.submit(p, o, 16, W.FLAGS_ALL, prm) for p, o in zip(small, os_)]
    lat_small = [srv.wait(i) for i in ids]
    lat_big = srv.wait(ib)
    st = srv.stats()
    srv.close()
    say(f"pieces: big {lat_big * 1e3:.2f} ms, small max {max(lat_small) * 1e3:.2f} ms; {st}")
    assert max(lat_small) < lat_big
    assert st["batches"] >= 8 and st["points_completed"] == 262144 + sum(p.shape[1] for p in small)
    ctx = api.Context(0)
    ctx.set_quadrature(16)
    assert np.array_equal(ob, ctx.compute_m1_opacities(big, flags=W.FLAGS_ALL, params=prm))
    for p, o in zip(small, os_):
        assert np.array_equal(o, ctx.compute_m1_opacities(p, flags=W.FLAGS_ALL, params=prm))
    ctx.close()


def test_admission_respects_the_hbm_budget(api):
    budget = 3 * 4096 * 41 * 8 + 4096
    srv = api.Server(0, hbm_budget_bytes=budget, streams=4, max_batch_points=4096)
    prm = W.params_default_all(W.BREM_BRT06)
    reqs = [np.ascontiguousarray(W.make_points(4096, seed=200 + i)) for i in range(24)]
    outs = [np.empty((20, 4096)) for _ in reqs]
    ids = [srv.submit(p, o, 8, W.FLAGS_ALL, prm, precision=i % 2)
           for i, (p, o) in enumerate(zip(reqs, outs))]
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    say(f"admission: {st}")
    assert st["requests_completed"] == 24
    assert st["max_inflight_bytes"] <= st["arena_bytes"] <= budget
    assert st["max_concurrent_batches"] <= 3
    # a request above the budget runs as pieces that fit it
    four = np.ascontiguousarray(W.make_points(4 * 4096, seed=230))
    o4 = np.empty((20, 4 * 4096))
    srv.wait(srv.submit(four, o4, 8, W.FLAGS_ALL, prm))
    assert srv.stats()["max_inflight_bytes"] <= budget
    srv.close()
    # a piece larger than the budget is rejected at submission
    one_piece = api.Server(0, hbm_budget_bytes=budget, streams=1, max_batch_points=1 << 20,
                           chunk_points=1 << 20)
    with pytest.raises(api.BnsAmdError):
        one_piece.submit(np.zeros((21, 4 * 4096)), np.empty((20, 4 * 4096)), 8, W.FLAGS_ALL, prm)
    one_piece.close()
    ctx = api.Context(0)
    ctx.set_quadrature(8)
    for i, (p, o) in enumerate(zip(reqs, outs)):
        assert np.array_equal(o, ctx.compute_m1_opacities(p, flags=W.FLAGS_ALL, params=prm,
                                                          precision=i % 2))
    assert np.array_equal(o4, ctx.compute_m1_opacities(four, flags=W.FLAGS_ALL, params=prm))
    ctx.close()


def test_pinned_buffers(api):
    srv = api.Server(0, hbm_budget_bytes=64 << 20, streams=2)
    prm = W.params_default_all(W.BREM_HR98)
    pin_in, pin_out = api.PinnedBuffer((21, 1000)), api.PinnedBuffer((20, 1000))
    pin_in.array[...] = W.make_points(1000, seed=9)
    srv.wait(srv.submit(pin_in.array, pin_out.array, 16, W.FLAGS_ALL, prm))
    ctx = api.Context(0)
    ctx.set_quadrature(16)
    ref = ctx.compute_m1_opacities(np.ascontiguousarray(pin_in.array), flags=W.FLAGS_ALL,
                                   params=prm)
    assert np.array_equal(pin_out.array, ref)
    srv.close()
    ctx.close()
    pin_in.close()
    pin_out.close()
