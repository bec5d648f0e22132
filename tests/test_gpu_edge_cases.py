"""The HIP path on the closed-form and edge cases that pin the oracle
(tests/closed_forms.py): exact mode must equal the oracle's float CMP / double
CRS restatement bit for bit, fast mode must satisfy the closed forms, and the
reference's sentinels, tie rule and err < 2 rule must hold in both modes --
on every store that represents the input exactly (the ramps and constants
here are exact in FP16 after the power-of-two scale)."""
import numpy as np
import pytest

import closed_forms as cf

pytestmark = pytest.mark.gpu

DT = 4000
NS = 201
C0, C1, NC = 1.98e-7, 1.77e-6, 8
TAU = 0.004
KEYS = ("index", "c", "coher", "stack")
MODES = ["fast", "exact"]


def _run(sb, ctx, g, s, mode, store="f32", nc=NC, tau=TAU):
    off = np.array([0, len(g)], np.int32)
    gg = sb.Gather(ctx, off, g, s, DT, store=sb.STORES[store])
    try:
        m = sb.MODE_FAST if mode == "fast" else sb.MODE_EXACT
        return sb.search(ctx, gg, sb.cmp_params(C0, C1, nc, tau, mode=m))
    finally:
        gg.close()


def _oracle(orc, g, s, nc=NC, tau=TAU):
    return orc.cmp(np.array([0, len(g)], np.int32), g, s, DT, C0, C1, nc, tau)


@pytest.mark.parametrize("mode", MODES)
def test_ramp_stack_is_the_moveout_mean(sb, orc, gpu_ctx, mode):
    print("case: ramp", mode, flush=True)
    g, s = cf.line_gather([0.0, 120.0, 333.0, 610.0], NS, "ramp", alpha=0.01, beta=0.25)
    out = _run(sb, gpu_ctx, g, s, mode)
    ref = _oracle(orc, g, s)
    if mode == "exact":
        for k in KEYS:
            assert np.array_equal(out[k], ref[k]), k
    inc = (C1 - C0) / NC
    checked = 0
    for t0 in range(NS):
        ci = out["index"][0, t0]
        if ci < 0:
            continue
        taus = cf.cmp_tau_samples(g, t0, C0 + inc * ci, DT * 1e-6)
        taus = taus[cf.valid_mask(taus, NS, 1)]
        assert out["stack"][0, t0] == pytest.approx(cf.ramp_stack(taus, 0.01, 0.25), rel=2e-5, abs=1e-5)
        checked += 1
    assert checked > NS // 2


@pytest.mark.parametrize("store", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("mode", MODES)
def test_constant_traces_unit_semblance_first_index(sb, orc, gpu_ctx, mode, store):
    print("case: const", mode, store, flush=True)
    g, s = cf.line_gather([0.0, 200.0, 400.0, 800.0, 900.0], NS, "const", amp=0.75)
    out = _run(sb, gpu_ctx, g, s, mode, store)
    ref = _oracle(orc, g, s)
    v = out["index"][0] >= 0
    assert np.array_equal(v, ref["index"][0] >= 0)
    assert v.any()
    assert np.all(out["index"][0][v] == 0)  # all equal: the first parameter wins
    assert np.allclose(out["coher"][0][v], 1.0, rtol=1e-6)
    assert np.allclose(out["stack"][0][v], 0.75, rtol=1e-6)


@pytest.mark.parametrize("mode", MODES)
def test_opposite_polarity_never_wins(sb, orc, gpu_ctx, mode):
    print("case: flip", mode, flush=True)
    g, s = cf.line_gather([0.0, 0.0], NS, "flip", amp=1.0)
    out = _run(sb, gpu_ctx, g, s, mode)
    assert np.all(out["index"] == -1) and np.all(out["coher"] == 0) and np.all(out["stack"] == 0)
    assert np.all(out["c"] == 0)


@pytest.mark.parametrize("mode", MODES)
def test_err_rule(sb, orc, gpu_ctx, mode):
    print("case: err rule", mode, flush=True)
    base = [0.0, 100.0, 200.0]
    g1, s1 = cf.line_gather(base + [5e4], NS, "const", amp=1.0)
    g2, s2 = cf.line_gather(base + [5e4, 6e4], NS, "const", amp=1.0)
    o1 = _run(sb, gpu_ctx, g1, s1, mode)
    o2 = _run(sb, gpu_ctx, g2, s2, mode)
    assert np.isclose(o1["coher"][0], 1.0, rtol=1e-6).any()
    assert np.all(o2["index"] == -1) and np.all(o2["coher"] == 0)


@pytest.mark.parametrize("mode", MODES)
def test_single_trace_single_parameter(sb, orc, gpu_ctx, mode):
    print("case: single trace", mode, flush=True)
    rng = np.random.default_rng(3)
    g, _ = cf.line_gather([150.0], NS)
    s = rng.standard_normal((1, NS)).astype(np.float32)
    out = _run(sb, gpu_ctx, g, s, mode, nc=1)
    ref = _oracle(orc, g, s, nc=1)
    v = ref["index"][0] >= 0
    assert np.array_equal(out["index"][0] >= 0, v)
    assert np.all(out["index"][0][v] == 0)
    assert np.allclose(out["coher"][0][v], 1.0, rtol=1e-6)
    for k in KEYS:  # the winner is re-evaluated in reference arithmetic in both modes
        assert np.array_equal(out[k], ref[k]), k


@pytest.mark.parametrize("tau", [0.0, 0.004, 0.016, 0.028])
@pytest.mark.parametrize("mode", MODES)
def test_window_lengths_up_to_the_maximum(sb, orc, gpu_ctx, mode, tau):
    """w = 1, 3, 9, 15 (= SB_MAX_W): fast mode runs w > 7 in reference
    arithmetic, so both modes equal the oracle bit for bit there."""
    print("case: window", mode, tau, flush=True)
    cdp_off, geom, samples = sb.synth(0, 3, (8, 16), 301, DT, seed=12, h_max=500.0, nevents=4)
    gg = sb.Gather(gpu_ctx, cdp_off, geom, samples, DT)
    try:
        m = sb.MODE_FAST if mode == "fast" else sb.MODE_EXACT
        out = sb.search(gpu_ctx, gg, sb.cmp_params(C0, C1, 16, tau, mode=m))
    finally:
        gg.close()
    ref = orc.cmp(cdp_off, geom, samples, DT, C0, C1, 16, tau)
    w = 2 * int(np.float32(tau) * np.float32(1.0 / np.float32(DT / 1e6))) + 1
    if mode == "exact" or w > 7:
        for k in KEYS:
            assert np.array_equal(out[k], ref[k]), (k, w)
    else:
        same = out["index"] == ref["index"]
        assert same.mean() > 0.97
        for k in KEYS:
            assert np.array_equal(out[k][same], ref[k][same]), (k, w)


@pytest.mark.parametrize("ns", [1, 2, 3, 4])
@pytest.mark.parametrize("mode", MODES)
def test_traces_shorter_than_a_window(sb, orc, gpu_ctx, mode, ns):
    """ns <= w + 1: no window start is valid anywhere, every output is the
    CMP sentinel."""
    print("case: short traces", mode, ns, flush=True)
    g, s = cf.line_gather([0.0, 50.0, 90.0], ns, "const", amp=1.0)
    out = _run(sb, gpu_ctx, g, s, mode)
    ref = _oracle(orc, g, s)
    for k in KEYS:
        assert np.array_equal(out[k], ref[k]), k


@pytest.mark.parametrize("mode", MODES)
def test_crs_no_winner_reports_first_parameter(sb, orc, gpu_ctx, mode):
    """CRS quirk: index 0 (a0, b0, c0), coherence 0, stack of (t0 = 0, par 0)."""
    print("case: crs no winner", mode, flush=True)
    g, s = cf.line_gather([0.0, 0.0], NS, "flip")
    off = np.array([0, 2], np.int32)
    no, nb = np.array([0, 1], np.int32), np.array([0], np.int32)
    gg = sb.Gather(gpu_ctx, off, g, s, DT, nbr_off=no, nbr=nb)
    try:
        m = sb.MODE_FAST if mode == "fast" else sb.MODE_EXACT
        out = sb.search(gpu_ctx, gg, sb.crs_params(-1e-3, 1e-3, 2, -1e-7, 1e-7, 2, 2e-7, 1e-6, 2, TAU, mode=m))
    finally:
        gg.close()
    ref = orc.crs(off, g, s, DT, no, nb, -1e-3, 1e-3, 2, -1e-7, 1e-7, 2, 2e-7, 1e-6, 2, TAU)
    for k in ("index", "a", "b", "c", "coher", "stack"):
        assert np.array_equal(out[k], ref[k]), k
    assert np.all(out["index"] == 0) and np.all(out["coher"] == 0)
