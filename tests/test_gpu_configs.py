"""The HIP path at the sizes of the CRS grid, CRS-DE and serving workloads
(the CMP production line is in test_gpu_scale.py).

  CRS grid   500 CDPs, 1000 traces each (midpoints +-50 m, half-offsets
             0-2600 m, aph 600 -> about 231 kept), ns 1001, dt 4 ms, apm 50,
             11 x 11 x 11 grid, tau 4 ms; FP32 and FP16 stores.  Whole-line
             properties; the first, middle and last CDP against the
             reference-precision oracle at full fold and ns (FP32: the
             parity bar of tests/parity.py; FP16: semblance within 2e-3
             relative and stack within 2e-3 of the line's amplitude of the
             FP64 evaluation of the reported (A, B, C)).
  CRS-DE     2000 CDPs of fold 200-1000 (aph 600), ns 1001, population 32,
             30 generations, F 0.5, CR 0.9; FP32 and FP16.  Whole-line
             properties; SB_MODE_EXACT on a 3-CDP range bit-identical to the
             CPU twin; fast-mode fitness and stack against the FP64
             evaluation of the reported point (FP32 2e-4, FP16 2e-3 relative);
             agreement of the fast search's best attributes with the twin's
             printed.
  Serving    requests of 1-64 CDPs of fold 24-1000 at ns 501-3001 with
             21-128 C values (CMP, 70 %) and CRS requests of 1-4 CDPs with
             21-24 values per parameter, submitted together to one server:
             every result bit-identical to the same search run alone.
"""
import threading

import numpy as np
import pytest

from parity import check_fast_picks

pytestmark = pytest.mark.gpu

SEED = 0x5EB1A9CE
BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)


def _crs_line(sb, ncdp, fold, seed):
    cdp_off, geom, samples = sb.synth(1, ncdp, fold, 1001, 4000, seed, h_max=2600.0, mid_half_width=50.0,
                                      nevents=6, snr=3.0, aph=600)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 50)
    return cdp_off, geom, samples, nbr_off, nbr


@pytest.fixture(scope="module")
def grid_line(sb):
    return _crs_line(sb, 500, (1000, 1000), SEED + 2)


def _grid_properties(out, n=11):
    idx = out["index"]
    assert idx.min() >= 0 and idx.max() < n ** 3
    coh = out["coher"]
    assert coh.min() >= 0.0 and coh.max() <= 1.0 + 1e-6 and np.isfinite(out["stack"]).all()
    assert np.all(idx[coh == 0] == 0)  # no winner: index 0 (the reference's absolute index 0)
    # parameter outputs as the reference encodes them (double grid, stored float)
    a0, a1, b0, b1, c0, c1 = BOX
    ia, ib, ic = idx // (n * n), (idx // n) % n, idx % n
    assert np.array_equal(out["a"], (a0 + (a1 - a0) * (1.0 / n) * ia).astype(np.float32))
    assert np.array_equal(out["b"], (b0 + (b1 - b0) * (1.0 / n) * ib).astype(np.float32))
    cc0, cc1 = c0 * 1e6, c1 * 1e6
    assert np.array_equal(out["c"], ((cc0 + (cc1 - cc0) * (1.0 / n) * ic) / 1e6).astype(np.float32))
    assert (coh >= 0.3).mean() > 0.02  # the events are found


@pytest.mark.parametrize("store", ["f32", "f16"])
def test_crs_grid_full_line(sb, orc, gpu_ctx, grid_line, store):
    print("case: crs grid 500 CDPs fold 1000 ns 1001 11^3 store", store, flush=True)
    cdp_off, geom, samples, nbr_off, nbr = grid_line
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
    prm = sb.crs_params(BOX[0], BOX[1], 11, BOX[2], BOX[3], 11, BOX[4], BOX[5], 11, 0.004)
    out = sb.search(gpu_ctx, g, prm)
    g.close()
    _grid_properties(out)
    picks = (0, 250, 499)
    if store == "f32":
        for cdp in picks:
            ref = orc.crs(cdp_off, geom, samples, 4000, nbr_off, nbr, BOX[0], BOX[1], 11, BOX[2], BOX[3], 11, BOX[4],
                          BOX[5], 11, 0.004, panel=True, cdp_range=(cdp, cdp + 1))
            mine = {k: v[cdp:cdp + 1] for k, v in out.items()}
            check_fast_picks("crs grid CDP %d" % cdp, ref, mine, None, params=("a", "b", "c"))
    else:
        scale = float(np.max(np.abs(samples)))
        rel, stk = [], []
        rng = np.random.default_rng(3)
        for cdp in picks:
            for t0 in rng.integers(25, 1000, 150):
                t0 = int(t0)
                s64, k64 = orc.crs_eval_f64(cdp, cdp_off, geom, samples, 4000, nbr_off, nbr, float(out["a"][cdp, t0]),
                                            float(out["b"][cdp, t0]), float(out["c"][cdp, t0]), 0.004, t0)
                if s64 >= 0.05:
                    rel.append(abs(out["coher"][cdp, t0] - s64) / s64)
                    stk.append(abs(out["stack"][cdp, t0] - k64) / scale)
        rel, stk = np.array(rel), np.array(stk)
        print("crs grid f16: semblance rel err vs FP64 max %.3g p99 %.3g (n=%d), stack err max %.3g" %
              (rel.max(), np.percentile(rel, 99), len(rel), stk.max()), flush=True)
        assert len(rel) > 60 and rel.max() <= 2e-3 and stk.max() <= 2e-3


@pytest.fixture(scope="module")
def de_line(sb):
    return _crs_line(sb, 2000, (200, 1000), SEED + 3)


def _de_params(sb, mode):
    return sb.de_params(*BOX, 0.004, pop=32, ngen=30, F=0.5, CR=0.9, seed=1, mode=mode)


def test_crs_de_exact_range_matches_twin(sb, orc, gpu_ctx, de_line):
    print("case: crs-de 2000 CDPs exact mode on CDPs 1000-1002 vs the CPU twin", flush=True)
    cdp_off, geom, samples, nbr_off, nbr = de_line
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, nbr_off=nbr_off, nbr=nbr)
    out = sb.search(gpu_ctx, g, _de_params(sb, sb.MODE_EXACT), (1000, 1003))
    g.close()
    ref = orc.crs_de(cdp_off, geom, samples, 4000, nbr_off, nbr, *BOX, 0.004, 32, 30, 0.5, 0.9, 1,
                     cdp_range=(1000, 1003))
    for k in ("a", "b", "c", "coher", "stack"):
        assert np.array_equal(out[k], ref[k]), k


@pytest.mark.parametrize("store,tol", [("f32", 2e-4), ("f16", 2e-3)])
def test_crs_de_full_line(sb, orc, gpu_ctx, de_line, store, tol):
    print("case: crs-de 2000 CDPs fold 200-1000 pop 32 ngen 30 store", store, flush=True)
    cdp_off, geom, samples, nbr_off, nbr = de_line
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
    out = sb.search(gpu_ctx, g, _de_params(sb, sb.MODE_FAST))
    g.close()
    lo = np.array([BOX[0], BOX[2], BOX[4]], np.float32)
    hi = np.array([BOX[1], BOX[3], BOX[5]], np.float32)
    # members stay in the box (A, B exactly; C is reported as the scaled value / 1e6, one more rounding)
    for k, l, h in (("a", lo[0], hi[0]), ("b", lo[1], hi[1])):
        assert out[k].min() >= l and out[k].max() <= h, k
    assert out["c"].min() >= BOX[4] * (1 - 1e-6) and out["c"].max() <= BOX[5] * (1 + 1e-6)
    coh = out["coher"]
    assert coh.min() >= 0.0 and coh.max() <= 1.0 + 1e-6 and np.isfinite(out["stack"]).all()
    assert (out["index"] == -1).all()
    # fitness and stack at the reported point against FP64
    scale = float(np.max(np.abs(samples)))
    rng = np.random.default_rng(4)
    rel, stk = [], []
    for cdp, t0 in zip(rng.integers(0, 2000, 400), rng.integers(25, 1000, 400)):
        cdp, t0 = int(cdp), int(t0)
        s64, k64 = orc.crs_eval_f64(cdp, cdp_off, geom, samples, 4000, nbr_off, nbr, float(out["a"][cdp, t0]),
                                    float(out["b"][cdp, t0]), float(out["c"][cdp, t0]), 0.004, t0)
        if s64 >= 0.05:
            rel.append(abs(coh[cdp, t0] - s64) / s64)
            stk.append(abs(out["stack"][cdp, t0] - k64) / scale)
    rel, stk = np.array(rel), np.array(stk)
    # best attributes against the CPU twin (FP32 in the reference order) on CDPs 1000-1002
    twin = orc.crs_de(cdp_off, geom, samples, 4000, nbr_off, nbr, *BOX, 0.004, 32, 30, 0.5, 0.9, 1,
                      cdp_range=(1000, 1003))
    mine = {k: v[1000:1003] for k, v in out.items()}
    d_fit = mine["coher"] - twin["coher"]
    span = hi - lo
    close = np.ones_like(d_fit, bool)
    for k, sp in zip(("a", "b", "c"), span):
        close &= np.abs(mine[k] - twin[k]) <= 0.01 * sp
    strong = twin["coher"] >= 0.3
    print("crs-de %s: fitness rel err vs FP64 max %.3g p99 %.3g (n=%d), stack err max %.3g; vs the CPU twin on "
          "CDPs 1000-1002: fitness within 1e-3 on %.4f of cells, mean fitness %.5f vs %.5f, attributes within 1%% "
          "of the box on %.4f of cells with fitness >= 0.3" %
          (store, rel.max(), np.percentile(rel, 99), len(rel), stk.max(), float(np.mean(np.abs(d_fit) <= 1e-3)),
           float(mine["coher"].mean()), float(twin["coher"].mean()), float(close[strong].mean())), flush=True)
    assert len(rel) > 40 and rel.max() <= tol and stk.max() <= tol
    assert abs(mine["coher"].mean() - twin["coher"].mean()) <= 0.02 * twin["coher"].mean()


def _serving_requests(sb, rng, n):
    """Requests in the serving workload's ranges, on resident lines."""
    cmp_line = sb.synth(0, 96, (24, 1000), 3001, 2000, SEED + 40, h_max=1500.0, h_jitter=5.0, nevents=8)
    crs_line = sb.synth(1, 24, (24, 200), 1001, 4000, SEED + 41, h_max=800.0, mid_half_width=50.0, nevents=6)
    crs_no, crs_nb = sb.neighbors(crs_line[1][crs_line[0][:-1]], 50)
    lines = {("cmp", ns): (cmp_line[0], cmp_line[1], np.ascontiguousarray(cmp_line[2][:, :ns]), 2000, {})
             for ns in (501, 1777, 3001)}
    lines[("crs", 1001)] = (crs_line[0], crs_line[1], crs_line[2], 4000, dict(nbr_off=crs_no, nbr=crs_nb))
    reqs = []
    for _ in range(n):
        if rng.random() < 0.7:
            k = int(rng.integers(1, 65))
            i0 = int(rng.integers(0, 96 - k + 1))
            key = ("cmp", int(rng.choice([501, 1777, 3001])))
            prm = sb.cmp_params(1.98e-7, 1.77e-6, int(rng.integers(21, 129)), 0.002)
        else:
            k = int(rng.integers(1, 5))
            i0 = int(rng.integers(0, 24 - k + 1))
            key = ("crs", 1001)
            n3 = [int(x) for x in rng.integers(21, 25, 3)]
            prm = sb.crs_params(BOX[0], BOX[1], n3[0], BOX[2], BOX[3], n3[1], BOX[4], BOX[5], n3[2], 0.004)
        reqs.append((key, i0, i0 + k, prm))
    return lines, reqs


@pytest.mark.parametrize("store", ["f32", "f16"])
def test_serving_workload_alone_equals_mix(sb, gpu_ctx, store):
    print("case: serving workload mix, store", store, flush=True)
    rng = np.random.default_rng(SEED + 42)
    lines, reqs = _serving_requests(sb, rng, 40)
    st = sb.STORES[store]
    # alone: each request on its own, through the library directly
    gathers = {key: sb.Gather(gpu_ctx, l[0], l[1], l[2], l[3], store=st, **l[4]) for key, l in lines.items()}
    try:
        ref = [sb.search(gpu_ctx, gathers[key], prm, (b, e)) for key, b, e, prm in reqs]
    finally:
        for g in gathers.values():
            g.close()
    with sb.Server(0, max_batch=64) as srv:
        gid = {key: srv.add_gather(l[0], l[1], l[2], l[3], store=st, **l[4]) for key, l in lines.items()}
        tickets = [None] * len(reqs)

        def submitter(lo):
            for i in range(lo, len(reqs), 4):
                key, b, e, prm = reqs[i]
                tickets[i] = srv.submit_range(gid[key], b, e, prm)

        th = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        outs = [srv.wait(t) for t in tickets]
        stats = srv.stats()
    print("served %d requests in %d launches" % (stats["done"], stats["launches"]), flush=True)
    for i, ((out, rs), r) in enumerate(zip(outs, ref)):
        assert out is not None, rs
        for k in ("index", "a", "b", "c", "coher", "stack"):
            assert np.array_equal(out[k], r[k]), (i, reqs[i][0], k)


def test_device_fp32_fma_rate(sb):
    """The FP32 FMA rate the device sustains without packed FP32: the ceiling
    the search kernels' arithmetic is held to (printed; the published 157.3
    TFLOP/s counts packed FP32)."""
    print("case: fp32 fma rate probe", flush=True)
    t = sb.fp32_fma_rate(0)
    print("sustained FP32 FMA rate: %.1f TFLOP/s" % t, flush=True)
    assert 10.0 < t < 200.0


def test_device_window_rates(sb):
    """Interpolated window samples per second for the four instruction forms
    of the search's inner step (printed).  Every form runs and reports a
    rate of the order of the FMA rate (one to four instructions per sample)."""
    print("case: window arithmetic rate probe", flush=True)
    r = sb.window_rates(0)
    print("window samples/s: " + ", ".join("%s %.3g" % kv for kv in r.items()), flush=True)
    fma = sb.fp32_fma_rate(0) * 1e12 / 2  # FMA per second
    for k, v in r.items():
        assert fma / 8 < v < fma * 2, (k, v, fma)
