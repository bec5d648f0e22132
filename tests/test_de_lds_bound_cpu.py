"""CPU replay of the LDS-staged CRS-DE evaluator's window-range bound.

The staged evaluator (k_crs_de_lds in csrc/de.hip) stages, for every trace
of a batch, the window starts e = floor(sqrt(tt)) - tau that any candidate
of its workgroup can reach, where tt = fma(K2, m2, fma(K1, m, fma(c, h,
T0^2))) with K1 = 2 T0 a, K2 = a^2 + b is the FP32 traveltime of
de_eval_fast, for every t0 of the workgroup's span and every (a, b, c) in
the box -- the bound holds for every generation because every trial is
clipped to the box.  A window outside the staged range would be read from
the wrong LDS bytes, so this replays both computations in FP32 (numpy,
fused multiply-adds through float64, the hardware square root pushed one ulp
either way) over random and extreme points and checks that every start lies
inside the bound (before the bound is clipped to the valid starts), and that
a trace whose whole bound lies at or after the first valid start -- which the
kernel runs without gating and without clamping t^2 before the square root --
has a positive t^2 at every point.

The planted-mistake check: dropping the 2-sample margin, or taking the
bound at T0lo only, must make the check fail on the same points.
"""
import numpy as np

F = np.float32


def fma(a, b, c):
    return F(np.float64(a) * np.float64(b) + np.float64(c))


def kernel_traveltime(T0, pa, pb, pc, m, m2, h):
    """tt of de_eval_fast / k_crs_de_lds for arrays of points."""
    T0 = F(T0)
    T0sq = F(T0 * T0)
    K1 = F(F(F(2.0) * T0) * F(pa))
    K2 = fma(pa, pa, pb)
    return fma(K2, m2, fma(K1, m, fma(pc, h, T0sq)))


def kernel_starts(T0, pa, pb, pc, m, m2, h, tau):
    """e of de_eval_fast / k_crs_de_lds for arrays of points, with the square
    root exact and one ulp below / above."""
    tt = kernel_traveltime(T0, pa, pb, pc, m, m2, h)
    ts = np.sqrt(np.maximum(tt, F(0)), dtype=np.float32)
    out = []
    for s in (ts, np.nextafter(ts, F(0)), np.nextafter(ts, F(np.inf))):
        out.append(s.astype(np.int64) - tau)
    return out


def plan_bound(T0lo, T0hi, lo, hi, m, m2, h, tau, margin=2, t0_hi=True):
    """de_plan's range of window starts for one trace (unclipped)."""
    m, m2, h = F(m), F(m2), F(h)
    am0, am1 = F(F(lo[0]) * m), F(F(hi[0]) * m)
    ulo = F(F(T0lo) + min(am0, am1))
    uhi = F(F(T0hi if t0_hi else T0lo) + max(am0, am1))
    u2lo, u2hi = F(ulo * ulo), F(uhi * uhi)
    usqmin = F(0) if (ulo <= 0 <= uhi) else min(u2lo, u2hi)
    usqmax = max(u2lo, u2hi)
    bm0, bm1 = F(F(lo[1]) * m2), F(F(hi[1]) * m2)
    ch0, ch1 = F(F(lo[2]) * h), F(F(hi[2]) * h)
    ttlo = F(F(usqmin + min(bm0, bm1)) + min(ch0, ch1))
    tthi = F(F(usqmax + max(bm0, bm1)) + max(ch0, ch1))
    elo = int(np.sqrt(max(ttlo, F(0)), dtype=np.float32)) - tau - margin
    ehi = int(np.sqrt(max(tthi, F(0)), dtype=np.float32)) - tau + margin
    return elo, ehi


def _violations(rng, ntraces, npts, margin=2, t0_hi=True, signed=False):
    """Config-3 box (c scaled by 1e6; geometry in samples at dt = 4 ms):
    count staged-range violations over random traces, tiles and points."""
    idt = 250.0
    lo = [F(-0.7e-3), F(-1e-7), F(1.975e-7 * 1e6)]
    hi = [F(0.7e-3), F(1e-7), F(1.384e-6 * 1e6)]
    bad = 0
    for _ in range(ntraces):
        dm = rng.uniform(-60.0, 60.0) if signed else rng.uniform(0.0, 60.0)
        half = rng.choice([0.0, rng.uniform(0.0, 2600.0)])
        m = F(dm * idt)
        m2 = F(dm * dm * idt * idt)
        h = F(half * half / 1e6 * idt * idt)
        nt = int(rng.choice([8, 16, 32, 64]))
        T0lo = int(rng.integers(0, 3001 - nt))
        T0hi = T0lo + nt - 1
        tau = int(rng.choice([0, 1, 2, 3]))
        elo, ehi = plan_bound(T0lo, T0hi, lo, hi, m, m2, h, tau, margin, t0_hi)
        T0 = rng.integers(T0lo, T0hi + 1, npts).astype(np.float32)
        pts = [rng.uniform(lo[k], hi[k], npts).astype(np.float32) for k in range(3)]
        # the box's corners at both ends of the span
        cor = np.array(np.meshgrid([T0lo, T0hi], *[[lo[k], hi[k]] for k in range(3)])).reshape(4, -1)
        T0 = np.concatenate([T0, cor[0].astype(np.float32)])
        pts = [np.concatenate([pts[k], cor[k + 1].astype(np.float32)]) for k in range(3)]
        for e in kernel_starts(T0, pts[0], pts[1], pts[2], m, m2, h, tau):
            bad += int(np.sum((e < elo) | (e > ehi)))
        if elo >= 0:
            # an ungated trace (its whole range valid) runs without the
            # clamp before the square root: t^2 must be positive
            tt = kernel_traveltime(T0, pts[0], pts[1], pts[2], m, m2, h)
            bad += int(np.sum(tt < F(1)))
    return bad


def test_every_start_lies_in_the_staged_range():
    rng = np.random.default_rng(20261204)
    assert _violations(rng, 400, 300) == 0


def test_signed_midpoint_projections_stay_in_range():
    rng = np.random.default_rng(7)
    assert _violations(rng, 300, 300, signed=True) == 0


def test_a_planted_mistake_fails():
    """The check has teeth: a bound taken at the span's first t0 only, or
    one without the margin, is violated on the same kind of points."""
    rng = np.random.default_rng(20261204)
    assert _violations(rng, 200, 300, t0_hi=False) > 0
    rng = np.random.default_rng(3)
    assert _violations(rng, 400, 300, margin=-1) > 0
