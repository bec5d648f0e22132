"""Serving loop (sb_server): continuous batching on one GPU.

Every request's result must equal the same request searched alone
(bit-identical: the search arithmetic does not depend on the batch), under a
mix of CMP and CRS requests of uneven fold, length and grid submitted
concurrently; admission respects the HBM budget."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)


def make_requests(sb, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        ns = int(rng.integers(301, 902))
        if rng.random() < 0.7:
            ncdp = int(rng.integers(1, 5))
            cdp_off, geom, samples = sb.synth(0, ncdp, (24, 120), ns, 2000, int(seed * 1000 + i), h_max=1500.0,
                                              h_jitter=5.0, nevents=5)
            nc = int(rng.integers(21, 129))
            # two shapes repeat so that merging happens
            if i % 3 == 0:
                ns, nc = 301, 64
                cdp_off, geom, samples = sb.synth(0, ncdp, (24, 60), ns, 2000, int(seed * 1000 + i), h_max=1500.0)
            prm = sb.cmp_params(1.98e-7, 1.77e-6, nc, 0.002)
            reqs.append(dict(args=(cdp_off, geom, samples, 2000, prm), kw={}))
        else:
            ncdp = int(rng.integers(2, 5))
            cdp_off, geom, samples = sb.synth(1, ncdp, (40, 120), ns, 4000, int(seed * 1000 + i), h_max=800.0,
                                              mid_half_width=20.0, nevents=4)
            nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 30)
            n3 = [int(x) for x in rng.integers(3, 8, 3)]
            prm = sb.crs_params(BOX[0], BOX[1], n3[0], BOX[2], BOX[3], n3[1], BOX[4], BOX[5], n3[2], 0.004)
            reqs.append(dict(args=(cdp_off, geom, samples, 4000, prm), kw=dict(nbr_off=nbr_off, nbr=nbr)))
    return reqs


def alone(sb, ctx, req):
    cdp_off, geom, samples, dt, prm = req["args"]
    g = sb.Gather(ctx, cdp_off, geom, samples, dt, **req["kw"])
    try:
        return sb.search(ctx, g, prm)
    finally:
        g.close()


def test_results_identical_alone_and_in_the_mix(sb, gpu_ctx):
    print("case: serving mix of 24 CMP/CRS requests", flush=True)
    reqs = make_requests(sb, 24, 7)
    ref = [alone(sb, gpu_ctx, r) for r in reqs]
    with sb.Server(0, max_batch=16) as srv:
        tickets = [None] * len(reqs)

        def submitter(lo):
            for i in range(lo, len(reqs), 3):
                tickets[i] = srv.submit(*reqs[i]["args"], **reqs[i]["kw"])

        th = [threading.Thread(target=submitter, args=(k,)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    print("served %d requests in %d launches" % (st["done"], st["launches"]))
    assert st["done"] == len(reqs)
    assert st["launches"] < len(reqs)  # some same-shape requests were merged
    for i, ((out, stats), r) in enumerate(zip(outs, ref)):
        assert out is not None, stats
        assert stats["t_submit_ms"] <= stats["t_admit_ms"] <= stats["t_done_ms"]
        for k in ("index", "a", "b", "c", "coher", "stack"):
            assert np.array_equal(out[k], r[k]), (i, k)


def test_admission_by_budget(sb):
    print("case: serving under a small HBM budget", flush=True)
    cdp_off, geom, samples = sb.synth(0, 8, 60, 1001, 2000, 3, h_max=1500.0)
    prm = sb.cmp_params(1.98e-7, 1.77e-6, 64, 0.002)
    # ~8 x 60 x 1024 x 4 B = 2 MB of samples per request; a 24 MB budget holds few at once
    with sb.Server(0, hbm_budget=24 << 20, max_batch=1) as srv:
        ts = [srv.submit(cdp_off, geom, samples, 2000, prm) for _ in range(12)]
        res = [srv.wait(t) for t in ts]
        big = sb.synth(0, 200, 120, 3001, 2000, 4, h_max=1500.0)
        tb = srv.submit(*big, 2000, prm)
        out, st = srv.wait(tb)
    assert all(o is not None for o, _ in res)
    first = res[0][0]
    for o, _ in res[1:]:
        assert np.array_equal(o["coher"], first["coher"])
    assert out is None and "budget" in st["error"].lower()
