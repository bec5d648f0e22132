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


def _line(sb, kind, seed):
    if kind == "cmp":
        off, geom, s = sb.synth(0, 10, (24, 120), 701, 2000, seed, h_max=1500.0, h_jitter=5.0, nevents=5)
        return dict(cdp_off=off, geom=geom, samples=s, dt_us=2000)
    off, geom, s = sb.synth(1, 9, (40, 100), 601, 4000, seed, h_max=800.0, mid_half_width=20.0, nevents=4)
    no, nb = sb.neighbors(geom[off[:-1]], 30)
    return dict(cdp_off=off, geom=geom, samples=s, dt_us=4000, nbr_off=no, nbr=nb)


@pytest.mark.parametrize("mode", ["fast", "exact"])
@pytest.mark.parametrize("kind", ["cmp", "crs", "crs_de"])
def test_range_search_equals_rows_of_full_search(sb, gpu_ctx, kind, mode):
    """sb_*_run_range over [cb, ce) gives exactly rows cb..ce-1 of the search
    over the whole gather (the work table, keys and DE draws are per CDP)."""
    print("case: range search %s %s" % (kind, mode), flush=True)
    m = sb.MODE_FAST if mode == "fast" else sb.MODE_EXACT
    L = _line(sb, "cmp" if kind == "cmp" else "crs", 11)
    kw = {k: L[k] for k in ("nbr_off", "nbr") if k in L}
    if kind == "cmp":
        prm = sb.cmp_params(1.98e-7, 1.77e-6, 48, 0.002, mode=m)
    elif kind == "crs":
        prm = sb.crs_params(BOX[0], BOX[1], 4, BOX[2], BOX[3], 3, BOX[4], BOX[5], 5, 0.004, mode=m)
    else:
        prm = sb.de_params(*BOX, 0.004, pop=16, ngen=6, seed=5, mode=m)
    g = sb.Gather(gpu_ctx, L["cdp_off"], L["geom"], L["samples"], L["dt_us"], **kw)
    try:
        full = sb.search(gpu_ctx, g, prm)
        for cb, ce in ((0, 1), (3, 7), (g.ncdp - 2, g.ncdp), (0, g.ncdp)):
            part = sb.search(gpu_ctx, g, prm, cdp_range=(cb, ce))
            for k in ("index", "a", "b", "c", "coher", "stack"):
                assert np.array_equal(part[k], full[k][cb:ce]), (cb, ce, k)
        with pytest.raises(sb.SbError):
            sb.search(gpu_ctx, g, prm, cdp_range=(2, 2))
        with pytest.raises(sb.SbError):
            sb.search(gpu_ctx, g, prm, cdp_range=(0, g.ncdp + 1))
    finally:
        g.close()


def test_resident_gathers_alone_and_in_the_mix(sb, gpu_ctx):
    """Range requests on lines kept resident by the server equal the range
    searched alone, while they are merged among themselves and mixed with
    requests that carry their traces."""
    print("case: resident-gather serving mix", flush=True)
    lines = {"cmp": _line(sb, "cmp", 21), "crs": _line(sb, "crs", 22)}
    cprm = sb.cmp_params(1.98e-7, 1.77e-6, 64, 0.002)
    rprm = sb.crs_params(BOX[0], BOX[1], 4, BOX[2], BOX[3], 4, BOX[4], BOX[5], 4, 0.004)
    ref = {}
    for kind, L in lines.items():
        kw = {k: L[k] for k in ("nbr_off", "nbr") if k in L}
        g = sb.Gather(gpu_ctx, L["cdp_off"], L["geom"], L["samples"], L["dt_us"], **kw)
        try:
            ref[kind] = sb.search(gpu_ctx, g, cprm if kind == "cmp" else rprm)
        finally:
            g.close()
    rng = np.random.default_rng(3)
    carried = make_requests(sb, 8, 9)
    with sb.Server(0, max_batch=16) as srv:
        gid = {k: srv.add_gather(L["cdp_off"], L["geom"], L["samples"], L["dt_us"],
                                 **{q: L[q] for q in ("nbr_off", "nbr") if q in L}) for k, L in lines.items()}
        used0 = srv.stats()["hbm_reserved"]
        assert used0 > 0
        jobs = []
        for i in range(40):
            kind = "cmp" if i % 3 else "crs"
            n = lines[kind]["cdp_off"].size - 1
            cb = int(rng.integers(0, n))
            ce = int(rng.integers(cb + 1, min(n, cb + 4) + 1))
            jobs.append((kind, cb, ce, srv.submit_range(gid[kind], cb, ce, cprm if kind == "cmp" else rprm)))
            if i % 5 == 0:
                r = carried[i // 5]
                jobs.append(("carried", i // 5, None, srv.submit(*r["args"], **r["kw"])))
        outs = [(j, srv.wait(j[3])) for j in jobs]
        st = srv.stats()
        with pytest.raises(sb.SbError):
            srv.submit_range(gid["cmp"], 0, 99, cprm)  # range beyond the line
        srv.remove_gather(gid["cmp"])
        with pytest.raises(sb.SbError):
            srv.submit_range(gid["cmp"], 0, 1, cprm)  # removed
        assert srv.stats()["hbm_reserved"] < used0
    print("served %d requests in %d launches" % (st["done"], st["launches"]))
    for (kind, cb, ce, _), (out, stats) in outs:
        assert out is not None, stats
        if kind == "carried":
            want = alone(sb, gpu_ctx, carried[cb])
        else:
            want = {k: v[cb:ce] for k, v in ref[kind].items()}
        for k in ("index", "a", "b", "c", "coher", "stack"):
            assert np.array_equal(out[k], want[k]), (kind, cb, ce, k)


def test_results_beyond_the_initial_staging(sb, gpu_ctx):
    """Range requests whose results (24 bytes a (CDP, t0)) outgrow a slot's
    initial 8 MiB of pinned download staging: 9.2-21.6 MB here, so the
    staging grows (twice over on some slot) while requests run on the other
    slots; every result equals the rows of the line searched alone."""
    print("case: serving results of 0.1-21.6 MB", flush=True)
    cdp_off, geom, samples = sb.synth(0, 300, (6, 10), 3001, 2000, 31, h_max=1500.0, nevents=4)
    prm = sb.cmp_params(1.98e-7, 1.77e-6, 32, 0.002)
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 2000)
    try:
        ref = sb.search(gpu_ctx, g, prm)
    finally:
        g.close()
    ranges = [(0, 128), (10, 300), (5, 7), (0, 300), (100, 250), (0, 300)]
    with sb.Server(0, max_batch=300) as srv:
        gid = srv.add_gather(cdp_off, geom, samples, 2000)
        jobs = [(cb, ce, srv.submit_range(gid, cb, ce, prm)) for cb, ce in ranges]
        outs = [(cb, ce, srv.wait(t)) for cb, ce, t in jobs]
    for cb, ce, (out, st) in outs:
        assert out is not None, st
        for k in ("index", "a", "b", "c", "coher", "stack"):
            assert np.array_equal(out[k], ref[k][cb:ce]), (cb, ce, k)


def test_admission_deadline_sheds_load(sb, gpu_ctx):
    """With an admission deadline, a burst far above the GPU's capacity is
    shed: requests not admitted within the deadline fail with
    SB_ERR_DEADLINE, the admitted ones complete with their alone results, and
    none waited much longer than the deadline before admission.  Without a
    deadline every request of the same burst is served."""
    print("case: serving admission deadline", flush=True)
    line = sb.synth(0, 64, (200, 400), 2001, 2000, 77, h_max=1500.0, h_jitter=5.0, nevents=6)
    prm = sb.cmp_params(1.98e-7, 1.77e-6, 128, 0.002)
    g = sb.Gather(gpu_ctx, *line, 2000)
    alone = sb.search(gpu_ctx, g, prm, (0, 64))
    g.close()
    for deadline in (20.0, 0.0):
        with sb.Server(0, max_batch=1) as srv:
            gid = srv.add_gather(*line, 2000)
            srv.set_deadline(deadline)
            tickets = [srv.submit_range(gid, 0, 64, prm) for _ in range(80)]  # ~80 x 5 ms of GPU work at once
            res = [srv.wait(t) for t in tickets]
        served = [(o, st) for o, st in res if o is not None]
        shed = [st for o, st in res if o is None]
        print("deadline %.0f ms: %d served, %d shed" % (deadline, len(served), len(shed)), flush=True)
        for o, st in served:
            for k in ("index", "coher", "stack", "c"):
                assert np.array_equal(o[k], alone[k]), k
        if deadline > 0:
            assert len(shed) > 0 and len(served) > 0
            assert all(st["status"] == sb.ERR_DEADLINE for st in shed)
            assert all(st["t_admit_ms"] - st["t_submit_ms"] <= deadline + 10.0 for o, st in served)
        else:
            assert not shed
