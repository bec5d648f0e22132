"""HIP path at production sizes and the reduced-precision stores.

Tolerances (written here, checked below):
  FP32 store, SB_MODE_FAST: winning index = the reference's except where the
      reference's top two semblances are within 1e-5 (tests/parity.py; a
      capped, counted number of picks the FP64 oracle ranks at least as
      high); coherence and stack at every pick bit-identical to the
      reference-precision oracle's values for that pick.
  FP16 store: semblance at the reported (t0, c) within 2e-3 relative of the
      FP64 oracle (semblances >= 0.05), stack within 2e-3 of the CDP's
      amplitude scale (max |stack| of the FP64 oracle).
  BF16 store: semblance within 5e-3 relative (8-bit mantissa: the FP16 bar
      of 2e-3 does not hold for it), stack within 2e-3.
"""
import numpy as np
import pytest

from parity import check_fast_picks

pytestmark = pytest.mark.gpu

SEED = 0x5EB1A9CE
C0, C1 = 1.98e-7, 1.77e-6


def subset(cdp_off, geom, samples, cdps):
    rows = np.concatenate([np.arange(cdp_off[i], cdp_off[i + 1]) for i in cdps])
    off = np.concatenate([[0], np.cumsum(np.diff(cdp_off)[cdps])]).astype(np.int32)
    return off, geom[rows], samples[rows]


@pytest.fixture(scope="module")
def cmp_line(sb):
    """The CMP production line: 2000 CDPs, fold 30-120, ns 1501, dt 2 ms."""
    return sb.synth(0, 2000, (30, 120), 1501, 2000, SEED + 1, h_max=1500.0, h_jitter=5.0, nevents=8, snr=5.0)


def _twelve_cdps(cdp_off):
    folds = np.diff(cdp_off)
    return sorted(set([int(np.argmax(folds)), int(np.argmin(folds)), 0, 1999] +
                      np.random.default_rng(5).choice(2000, 8, replace=False).tolist()))


def test_cmp_production_size_against_oracle(sb, orc, gpu_ctx, cmp_line):
    print("case: cmp 2000 CDPs fold 30-120 ns 1501 nc 256 fast f32", flush=True)
    cdp_off, geom, samples = cmp_line
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 2000)
    out = sb.search(gpu_ctx, g, sb.cmp_params(C0, C1, 256, 0.002, sb.MODE_FAST))
    g.close()
    idx, coh = out["index"], out["coher"]
    assert idx.min() >= -1 and idx.max() < 256
    assert np.all((coh > 0) == (idx >= 0))
    assert coh.max() <= 1.0 + 1e-6 and np.isfinite(out["stack"]).all()
    assert (idx >= 0).mean() > 0.3
    # c written as the reference does: (c0 + inc * idx) / 1e6 in float, 0 for none
    c0 = np.float32(np.float64(np.float32(C0)) * 1e6)
    inc = (np.float32(np.float64(np.float32(C1)) * 1e6) - c0) * (np.float32(1.0) / np.float32(256))
    want = np.where(idx < 0, np.float32(0),
                    ((c0 + inc * idx.astype(np.float32)).astype(np.float64) / 1e6).astype(np.float32))
    assert np.array_equal(out["c"], want)
    pick = _twelve_cdps(cdp_off)
    off, gs, ss = subset(cdp_off, geom, samples, pick)
    ref = orc.cmp(off, gs, ss, 2000, C0, C1, 256, 0.002, panel=True)
    ref64 = orc.cmp_f64(off, gs, ss, 2000, C0, C1, 256, 0.002, panel=True)
    mine = {k: v[pick] for k, v in out.items()}
    ndiff, _ = check_fast_picks("cmp 2000 CDPs fold 30-120 ns 1501 nc 256", ref, mine, ref64, params=("c",))
    assert ndiff <= 1e-3 * mine["index"].size


@pytest.mark.parametrize("store,tol_sem,tol_stk", [("f16", 2e-3, 2e-3), ("bf16", 5e-3, 2e-3)])
def test_production_line_16bit_store_error_vs_fp64(sb, orc, gpu_ctx, cmp_line, store, tol_sem, tol_stk):
    """The 2000-CDP production line stored in FP16 / BF16: on 12 CDPs at full
    fold and ns, the reported semblance within tol_sem (relative, semblance >=
    0.05) of the FP64 oracle's semblance at the reported (t0, c), the stack
    within tol_stk of the CDP's amplitude where the FP64 arg-max agrees.  FP16
    meets the 2e-3 bar; BF16 (8-bit mantissa) is held to 5e-3."""
    print("case: cmp 2000 CDPs fold 30-120 ns 1501 nc 256 store", store, flush=True)
    cdp_off, geom, samples = cmp_line
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 2000, store=sb.STORES[store])
    out = sb.search(gpu_ctx, g, sb.cmp_params(C0, C1, 256, 0.002, sb.MODE_FAST))
    g.close()
    pick = _twelve_cdps(cdp_off)
    off, gs, ss = subset(cdp_off, geom, samples, pick)
    ref = orc.cmp_f64(off, gs, ss, 2000, C0, C1, 256, 0.002, panel=True)
    idx, coh, stk = out["index"][pick], out["coher"][pick], out["stack"][pick]
    ok = idx >= 0
    sem64 = np.take_along_axis(ref["panel_sem"], np.maximum(idx, 0)[..., None], 2)[..., 0]
    big = ok & (sem64 >= 0.05)
    rel = np.abs(coh[big] - sem64[big]) / sem64[big]
    same = ok & (idx == ref["index"])
    scale = np.max(np.abs(ref["stack"]), axis=1, keepdims=True)
    stk_err = np.abs(stk - ref["stack"]) / scale
    print("production line store %s: semblance rel err vs FP64 max %.3g p99 %.3g (n=%d); stack err max %.3g; "
          "index agreement with FP64 %.5f" % (store, rel.max(), np.percentile(rel, 99), big.sum(),
                                              stk_err[same].max(), (idx == ref["index"]).mean()), flush=True)
    assert big.sum() > 1000
    assert rel.max() <= tol_sem and stk_err[same].max() <= tol_stk
    assert (idx == ref["index"]).mean() > 0.99


@pytest.mark.parametrize("store,tol_sem,tol_stk", [("f16", 2e-3, 2e-3), ("bf16", 5e-3, 2e-3)])
def test_reduced_precision_store_error_vs_fp64(sb, orc, gpu_ctx, store, tol_sem, tol_stk):
    print("case: cmp 200 CDPs fold 60 ns 1001 nc 101 store", store, flush=True)
    cdp_off, geom, samples = sb.synth(0, 200, 60, 1001, 2000, SEED, h_max=1500.0, regular_offsets=True,
                                      h_step=25.0, nevents=8, snr=5.0)
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 2000, store=sb.STORES[store])
    out = sb.search(gpu_ctx, g, sb.cmp_params(C0, C1, 101, 0.002, sb.MODE_FAST))
    g.close()
    pick = list(range(0, 200, 10))
    off, gs, ss = subset(cdp_off, geom, samples, pick)
    ref = orc.cmp_f64(off, gs, ss, 2000, C0, C1, 101, 0.002, panel=True)
    idx = out["index"][pick]
    coh = out["coher"][pick]
    stk = out["stack"][pick]
    ok = idx >= 0
    sem64 = np.take_along_axis(ref["panel_sem"], np.maximum(idx, 0)[..., None], 2)[..., 0]
    big = ok & (sem64 >= 0.05)
    rel = np.abs(coh[big] - sem64[big]) / sem64[big]
    # stack of the same (t0, c) in FP64 needs the FP64 oracle's stack at that c:
    # compare where the FP64 arg-max agrees.
    same = ok & (idx == ref["index"])
    scale = np.max(np.abs(ref["stack"]), axis=1, keepdims=True)
    stk_err = np.abs(stk - ref["stack"]) / scale
    print("store %s: semblance rel err max %.3g p99 %.3g; stack err max %.3g; index agreement %.4f" %
          (store, rel.max(), np.percentile(rel, 99), stk_err[same].max(), (idx == ref["index"]).mean()))
    assert rel.max() <= tol_sem
    assert stk_err[same].max() <= tol_stk
    assert (idx == ref["index"]).mean() > 0.9


def test_fp32_fast_error_vs_fp64(sb, orc, gpu_ctx):
    """FP32 store: the reported coherence is the reference-precision value, so
    its error against FP64 is the reference's own (FP32 traveltime rounding);
    it must be no larger than the reference oracle's."""
    cdp_off, geom, samples = sb.synth(0, 40, 60, 1001, 2000, SEED, h_max=1500.0, regular_offsets=True,
                                      h_step=25.0, nevents=8, snr=5.0)
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 2000)
    out = sb.search(gpu_ctx, g, sb.cmp_params(C0, C1, 101, 0.002, sb.MODE_FAST))
    g.close()
    ref64 = orc.cmp_f64(cdp_off, geom, samples, 2000, C0, C1, 101, 0.002, panel=True)
    ref32 = orc.cmp(cdp_off, geom, samples, 2000, C0, C1, 101, 0.002)

    def err(res):
        idx = res["index"]
        sem64 = np.take_along_axis(ref64["panel_sem"], np.maximum(idx, 0)[..., None], 2)[..., 0]
        big = (idx >= 0) & (sem64 >= 0.05)
        return np.abs(res["coher"][big] - sem64[big]) / sem64[big]

    ours, theirs = err(out), err(ref32)
    print("f32 store: semblance rel err vs FP64 max %.3g (reference FP32 path: %.3g)" % (ours.max(), theirs.max()))
    assert ours.max() <= theirs.max() * 1.000001
    assert np.percentile(ours, 99) <= 1e-4


def _de_gather(sb, ncdp=5):
    cdp_off, geom, samples = sb.synth(1, ncdp, (12, 20), 201, 4000, seed=41, h_max=700.0, mid_half_width=20.0,
                                      nevents=4, snr=4.0)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 30)
    return cdp_off, geom, samples, nbr_off, nbr


BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)


@pytest.mark.parametrize("pop,ngen", [(16, 6), (32, 4), (20, 3)])
def test_crs_de_exact_matches_cpu_twin(sb, orc, gpu_ctx, pop, ngen):
    print("case: crs-de exact pop", pop, "ngen", ngen, flush=True)
    cdp_off, geom, samples, nbr_off, nbr = _de_gather(sb)
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, nbr_off=nbr_off, nbr=nbr)
    out = sb.search(gpu_ctx, g, sb.de_params(*BOX, 0.004, pop=pop, ngen=ngen, seed=11, mode=sb.MODE_EXACT))
    g.close()
    ref = orc.crs_de(cdp_off, geom, samples, 4000, nbr_off, nbr, *BOX, 0.004, pop, ngen, 0.5, 0.9, 11)
    for k in ("a", "b", "c", "coher", "stack"):
        assert np.array_equal(out[k], ref[k]), k


def test_crs_de_fast_is_statistically_equivalent(sb, orc, gpu_ctx):
    print("case: crs-de fast vs exact", flush=True)
    cdp_off, geom, samples, nbr_off, nbr = _de_gather(sb)
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, nbr_off=nbr_off, nbr=nbr)
    fast = sb.search(gpu_ctx, g, sb.de_params(*BOX, 0.004, pop=32, ngen=20, seed=5, mode=sb.MODE_FAST))
    exact = sb.search(gpu_ctx, g, sb.de_params(*BOX, 0.004, pop=32, ngen=20, seed=5, mode=sb.MODE_EXACT))
    g.close()
    # Same random stream: trajectories only diverge after near-ties.
    assert abs(fast["coher"].mean() - exact["coher"].mean()) < 0.02 * max(exact["coher"].mean(), 1e-3)
    # The reported fitness is the semblance of the reported parameters.
    rng = np.random.default_rng(0)
    for cdp, t0 in zip(rng.integers(0, 5, 20), rng.integers(10, 190, 20)):
        sem, _ = orc.crs_eval_f64(int(cdp), cdp_off, geom, samples, 4000, nbr_off, nbr, float(fast["a"][cdp, t0]),
                                  float(fast["b"][cdp, t0]), float(fast["c"][cdp, t0]), 0.004, int(t0))
        assert fast["coher"][cdp, t0] == pytest.approx(sem, rel=2e-4, abs=1e-5)


def test_crs_fast_larger_gather(sb, orc, gpu_ctx):
    print("case: crs 12 CDPs fold 200-400 ns 501 11x11x11", flush=True)
    cdp_off, geom, samples = sb.synth(1, 12, (200, 400), 501, 4000, SEED + 2, h_max=2600.0, mid_half_width=50.0,
                                      nevents=6, snr=3.0, aph=600)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 50)
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, nbr_off=nbr_off, nbr=nbr)
    prm = sb.crs_params(-0.7e-3, 0.7e-3, 11, -1e-7, 1e-7, 11, 1.975e-7, 1.384e-6, 11, 0.004, sb.MODE_FAST)
    out = sb.search(gpu_ctx, g, prm)
    g.close()
    pick = [0, 5, 11]
    ref = orc.crs(cdp_off, geom, samples, 4000, nbr_off, nbr, -0.7e-3, 0.7e-3, 11, -1e-7, 1e-7, 11, 1.975e-7,
                  1.384e-6, 11, 0.004, panel=True)
    ref = {k: v[pick] for k, v in ref.items()}
    mine = {k: v[pick] for k, v in out.items()}
    check_fast_picks("crs 12 CDPs 11^3", ref, mine, None, params=("a", "b", "c"))


def _crs_line(sb, ncdp=6):
    cdp_off, geom, samples = sb.synth(1, ncdp, (60, 120), 501, 4000, SEED + 7, h_max=1500.0, mid_half_width=40.0,
                                      nevents=6, snr=3.0)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 40)
    return cdp_off, geom, samples, nbr_off, nbr


def _crs_errors(orc, out, line, cells):
    cdp_off, geom, samples, nbr_off, nbr = line
    rel, stk = [], []
    scale = float(np.max(np.abs(samples)))
    for cdp, t0 in cells:
        sem64, stk64 = orc.crs_eval_f64(int(cdp), cdp_off, geom, samples, 4000, nbr_off, nbr, float(out["a"][cdp, t0]),
                                        float(out["b"][cdp, t0]), float(out["c"][cdp, t0]), 0.004, int(t0))
        if sem64 >= 0.05:
            rel.append(abs(out["coher"][cdp, t0] - sem64) / sem64)
            stk.append(abs(out["stack"][cdp, t0] - stk64) / scale)
    return np.array(rel), np.array(stk)


@pytest.mark.parametrize("store,tol", [("f16", 2e-3), ("bf16", 5e-3)])
def test_crs_reduced_precision_store_error_vs_fp64(sb, orc, gpu_ctx, store, tol):
    """CRS grid search on a 16-bit store: the reported semblance and stack at
    the reported (A, B, C) against the FP64 evaluation of the same point
    (tolerance as for CMP: 2e-3 relative semblance for FP16, stack relative to
    the line's amplitude scale)."""
    print("case: crs store", store, flush=True)
    line = _crs_line(sb)
    cdp_off, geom, samples, nbr_off, nbr = line
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
    out = sb.search(gpu_ctx, g, sb.crs_params(BOX[0], BOX[1], 7, BOX[2], BOX[3], 5, BOX[4], BOX[5], 9, 0.004))
    g.close()
    rng = np.random.default_rng(1)
    cells = list(zip(rng.integers(0, 6, 120), rng.integers(20, 480, 120)))
    rel, stk = _crs_errors(orc, out, line, cells)
    print("crs store %s: semblance rel err vs FP64 max %.3g p99 %.3g (n=%d); stack err max %.3g" %
          (store, rel.max(), np.percentile(rel, 99), len(rel), stk.max()))
    assert len(rel) > 30
    assert rel.max() <= tol and stk.max() <= tol


@pytest.mark.parametrize("store,tol", [("f32", 2e-4), ("f16", 2e-3)])
def test_crs_de_fitness_error_vs_fp64(sb, orc, gpu_ctx, store, tol):
    """CRS-DE (fast evaluator, FP32 accumulation): the reported fitness and
    stack against the FP64 evaluation of the reported (A, B, C)."""
    print("case: crs-de store", store, flush=True)
    line = _crs_line(sb)
    cdp_off, geom, samples, nbr_off, nbr = line
    g = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
    out = sb.search(gpu_ctx, g, sb.de_params(*BOX, 0.004, pop=32, ngen=12, seed=9))
    g.close()
    rng = np.random.default_rng(2)
    cells = list(zip(rng.integers(0, 6, 120), rng.integers(20, 480, 120)))
    rel, stk = _crs_errors(orc, out, line, cells)
    print("crs-de store %s: fitness rel err vs FP64 max %.3g p99 %.3g (n=%d); stack err max %.3g" %
          (store, rel.max(), np.percentile(rel, 99), len(rel), stk.max()))
    assert len(rel) > 30
    assert rel.max() <= tol and stk.max() <= tol
