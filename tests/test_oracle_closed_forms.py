"""Pins the oracle routines beyond the reference fixtures: the float CMP and
double CRS restatements (also pinned byte-for-byte by tests/golden), the FP64
CMP oracle, the FP64 CRS point evaluation and the CRS-DE twin, against closed
forms (tests/closed_forms.py), invariants and brute force."""
import numpy as np
import pytest

import closed_forms as cf

DT = 4000
NS = 201
C0, C1, NC = 1.98e-7, 1.77e-6, 8
TAU = 0.004  # w = 3 at 4 ms


def _cmp(orc, g, s, fn="cmp", nc=NC, tau=TAU, **kw):
    off = np.array([0, len(g)], np.int32)
    return getattr(orc, fn)(off, g, s, DT, C0, C1, nc, tau, **kw)


@pytest.mark.parametrize("fn", ["cmp", "cmp_f64"])
def test_cmp_ramp_stack_is_the_moveout_mean(orc, fn):
    g, s = cf.line_gather([0.0, 120.0, 333.0, 610.0], NS, "ramp", alpha=0.01, beta=0.25)
    out = _cmp(orc, g, s, fn, panel=True)
    dt = DT * 1e-6
    inc = (C1 - C0) / NC
    checked = 0
    for t0 in range(NS):
        ci = out["index"][0, t0]
        if ci < 0:
            continue
        taus = cf.cmp_tau_samples(g, t0, C0 + inc * ci, dt)
        assert cf.valid_mask(taus, NS, 1).sum() >= len(taus) - 1
        taus = taus[cf.valid_mask(taus, NS, 1)]
        assert out["stack"][0, t0] == pytest.approx(cf.ramp_stack(taus, 0.01, 0.25), rel=2e-5, abs=1e-5)
        checked += 1
    assert checked > NS // 2


@pytest.mark.parametrize("fn", ["cmp", "cmp_f64"])
def test_cmp_constant_traces_have_unit_semblance(orc, fn):
    g, s = cf.line_gather([0.0, 200.0, 400.0, 800.0, 900.0], NS, "const", amp=0.75)
    out = _cmp(orc, g, s, fn, panel=True)
    ps = out["panel_sem"][0]
    valid = ps > -1
    assert valid.any()
    assert np.allclose(ps[valid], 1.0, rtol=1e-6)
    # Equal semblance for every c: the first maximum (index 0) wins.
    assert np.all(out["index"][0][valid.any(axis=1)] == 0)
    assert np.allclose(out["stack"][0][valid.any(axis=1)], 0.75, rtol=1e-6)


def test_cmp_opposite_polarity_never_wins(orc):
    g, s = cf.line_gather([0.0, 0.0], NS, "flip", amp=1.0)
    out = _cmp(orc, g, s, panel=True)
    assert np.all(out["panel_sem"][0][out["panel_sem"][0] > -1] == 0)
    assert np.all(out["index"] == -1) and np.all(out["coher"] == 0) and np.all(out["stack"] == 0)
    assert np.all(out["c"] == 0)


def test_cmp_err_rule(orc):
    # Trace 3 and 4 leave the trace at every t0 (h = 50 km): with one of them
    # the result stays valid from the other traces, with both it is invalid.
    base = [0.0, 100.0, 200.0]
    g1, s1 = cf.line_gather(base + [5e4], NS, "const", amp=1.0)
    g2, s2 = cf.line_gather(base + [5e4, 6e4], NS, "const", amp=1.0)
    o1 = _cmp(orc, g1, s1, panel=True)
    o2 = _cmp(orc, g2, s2, panel=True)
    assert np.isclose(o1["panel_sem"][0], 1.0, rtol=1e-6).any()
    assert np.all(o2["panel_sem"][0] == -1)
    assert np.all(o2["index"] == -1)


def test_cmp_single_trace_and_single_param(orc):
    rng = np.random.default_rng(3)
    g, _ = cf.line_gather([150.0], NS)
    s = rng.standard_normal((1, NS)).astype(np.float32)
    out = _cmp(orc, g, s, nc=1, panel=True)
    ps = out["panel_sem"][0, :, 0]
    v = ps > -1
    assert np.allclose(ps[v], 1.0, rtol=1e-6)
    assert np.all(out["index"][0][v] == 0)


def test_cmp_f64_agrees_with_float_restatement(orc, sb):
    cdp_off, geom, samples = sb.synth(0, 6, (10, 20), 501, 2000, seed=9, h_max=600.0, h_jitter=5.0)
    a = orc.cmp(cdp_off, geom, samples, 2000, C0, C1, 24, 0.002, panel=True)
    b = orc.cmp_f64(cdp_off, geom, samples, 2000, C0, C1, 24, 0.002, panel=True)
    v = (a["panel_sem"] > 0) & (b["panel_sem"] > 0)
    assert v.mean() > 0.5
    assert np.max(np.abs(a["panel_sem"][v] - b["panel_sem"][v])) < 1e-4
    agree = (a["index"] == b["index"]).mean()
    assert agree > 0.97


def test_crs_ramp_stack_matches_moveout(orc):
    # Two CDPs 20 m apart sharing one supergather; unequal midpoints.
    mids = [(0.0, 0.0), (3.0, 1.0), (-6.0, 2.5), (20.0, 0.0), (24.0, -3.0)]
    hs = [0.0, 150.0, 300.0, 80.0, 420.0]
    g, s = cf.line_gather(hs, NS, "ramp", alpha=0.02, beta=-0.5, mids=mids)
    g["cdp"][3:] = 2
    off = np.array([0, 3, 5], np.int32)
    nbr_off = np.array([0, 2, 4], np.int32)
    nbr = np.array([0, 1, 0, 1], np.int32)
    a0, a1, b0, b1, c0, c1 = -0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6
    out = orc.crs(off, g, s, DT, nbr_off, nbr, a0, a1, 3, b0, b1, 2, c0, c1, 3, TAU, panel=True)
    dt = DT * 1e-6
    for cdp, ref in ((0, g[0]), (1, g[3])):
        for t0 in range(0, NS, 7):
            idx = out["index"][cdp, t0]
            if out["coher"][cdp, t0] <= 0:
                continue
            ia, ib, ic = idx // 6, (idx // 3) % 2, idx % 3
            A = a0 + (a1 - a0) / 3 * ia
            B = b0 + (b1 - b0) / 2 * ib
            Cc = c0 + (c1 - c0) / 3 * ic
            taus = cf.crs_tau_samples(g, ref, t0, A, B, Cc, dt)
            taus = taus[cf.valid_mask(taus, NS, 1)]
            assert out["stack"][cdp, t0] == pytest.approx(cf.ramp_stack(taus, 0.02, -0.5), rel=1e-6, abs=1e-6)
            sem64, stk64 = orc.crs_eval_f64(cdp, off, g, s, DT, nbr_off, nbr, A, B, Cc, TAU, t0)
            assert stk64 == pytest.approx(out["stack"][cdp, t0], rel=1e-6, abs=1e-6)
            assert sem64 == pytest.approx(out["coher"][cdp, t0], rel=1e-6)


@pytest.mark.parametrize("az", [90.0, 30.0, 0.0])
def test_crs_projected_midpoint_ramp_stack(orc, az):
    """The azimuth extension (orc_crs_range_az): with m the signed projection
    of the midpoint displacement onto the azimuth, the ramp stack is the mean
    moveout of the projected traveltimes.  Midpoints on both sides of the
    reference trace: a sign or sine/cosine mistake moves the taus.  At
    azimuth 0 (north) the x displacements project to zero."""
    mids = [(0.0, 0.0), (-12.0, 1.0), (9.0, -2.5), (25.0, 3.0), (-30.0, -1.0)]
    hs = [40.0, 150.0, 300.0, 80.0, 420.0]
    g, s = cf.line_gather(hs, NS, "ramp", alpha=0.02, beta=-0.5, mids=mids)
    off = np.array([0, 5], np.int32)
    nbr_off, nbr = np.array([0, 1], np.int32), np.array([0], np.int32)
    a0, a1, b0, b1, c0, c1 = -0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6
    out = orc.crs(off, g, s, DT, nbr_off, nbr, a0, a1, 4, b0, b1, 2, c0, c1, 3, TAU, az_mdeg=int(az * 1000))
    dt = DT * 1e-6
    checked = 0
    for t0 in range(0, NS, 5):
        idx = out["index"][0, t0]
        if out["coher"][0, t0] <= 0:
            continue
        ia, ib, ic = idx // 6, (idx // 3) % 2, idx % 3
        A = a0 + (a1 - a0) / 4 * ia
        B = b0 + (b1 - b0) / 2 * ib
        Cc = c0 + (c1 - c0) / 3 * ic
        taus = cf.crs_tau_samples_projected(g, g[0], t0, A, B, Cc, dt, az)
        taus = taus[cf.valid_mask(taus, NS, 1)]
        assert out["stack"][0, t0] == pytest.approx(cf.ramp_stack(taus, 0.02, -0.5), rel=1e-6, abs=1e-6), (t0, idx)
        checked += 1
    assert checked > 25  # of the 41 t0 visited
    if az != 0.0:  # the signed projection differs from the distance somewhere
        ref = orc.crs(off, g, s, DT, nbr_off, nbr, a0, a1, 4, b0, b1, 2, c0, c1, 3, TAU)
        assert not np.array_equal(ref["stack"], out["stack"])


def test_crs_no_winner_reports_first_parameter(orc):
    g, s = cf.line_gather([0.0, 0.0], NS, "flip")
    off = np.array([0, 2], np.int32)
    out = orc.crs(off, g, s, DT, np.array([0, 1], np.int32), np.array([0], np.int32), -1e-3, 1e-3, 2, -1e-7, 1e-7, 2,
                  2e-7, 1e-6, 2, TAU)
    assert np.all(out["index"] == 0) and np.all(out["coher"] == 0)
    assert np.all(out["a"] == np.float32(-1e-3)) and np.all(out["b"] == np.float32(-1e-7))
    assert np.all(out["c"] == np.float32(2e-7))


def test_de_hash_is_a_bijection_of_its_counter(orc):
    seen = set()
    for t0 in range(8):
        for gen in range(4):
            for m in range(8):
                for d in range(7):
                    seen.add(orc.de_hash(1, 3, t0, gen, m, d))
    assert len(seen) == 8 * 4 * 8 * 7  # no collisions in this range
    assert orc.de_hash(1, 3, 4, 5, 6, 7) == orc.de_hash(1, 3, 4, 5, 6, 7)
    assert orc.de_hash(2, 3, 4, 5, 6, 7) != orc.de_hash(1, 3, 4, 5, 6, 7)


def _de_gather(sb):
    cdp_off, geom, samples = sb.synth(1, 4, (10, 14), 151, 4000, seed=31, h_max=600.0, mid_half_width=20.0,
                                      nevents=3, snr=4.0)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 30)
    return cdp_off, geom, samples, nbr_off, nbr


BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)


def test_de_ngen0_is_best_of_initial_population(orc, sb):
    cdp_off, geom, samples, nbr_off, nbr = _de_gather(sb)
    out = orc.crs_de(cdp_off, geom, samples, 4000, nbr_off, nbr, *BOX, TAU, 8, 0, 0.5, 0.9, 7)
    # Re-evaluate every initial member in FP64: the reported best is the max.
    lo = np.array([BOX[0], BOX[2], BOX[4] * 1e6], np.float32)
    hi = np.array([BOX[1], BOX[3], BOX[5] * 1e6], np.float32)
    for cdp in range(4):
        for t0 in (20, 60, 100):
            fits = []
            for i in range(8):
                u = np.array([(orc.de_hash(7, cdp, t0, 0, i, d) >> 8) for d in range(3)], np.float32) / np.float32(
                    16777216.0)
                x = lo + u * (hi - lo)
                sem, _ = orc.crs_eval_f64(cdp, cdp_off, geom, samples, 4000, nbr_off, nbr, float(x[0]), float(x[1]),
                                          float(x[2]) / 1e6, TAU, t0)
                fits.append(sem)
            assert out["coher"][cdp, t0] == pytest.approx(max(fits), rel=1e-4, abs=1e-6)


def test_de_improves_with_generations(orc, sb):
    cdp_off, geom, samples, nbr_off, nbr = _de_gather(sb)
    f0 = orc.crs_de(cdp_off, geom, samples, 4000, nbr_off, nbr, *BOX, TAU, 16, 0, 0.5, 0.9, 3)["coher"]
    f1 = orc.crs_de(cdp_off, geom, samples, 4000, nbr_off, nbr, *BOX, TAU, 16, 15, 0.5, 0.9, 3)["coher"]
    assert np.all(f1 >= f0)  # ">=" selection never loses the best
    assert f1.mean() > f0.mean()
