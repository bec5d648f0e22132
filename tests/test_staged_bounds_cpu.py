"""CPU replay of the window-range bounds of the LDS-staged CRS-DE evaluator
and CRS grid search.

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

The CRS grid search (k_search_crs_lds in csrc/crs_lds.hip) bounds the same
traveltime over a workgroup's box: the (a, b) values of its run of grid pairs
and every c value, formed from the grid's end points with the kernels' FMAs;
the replay checks that every grid point of the run lies in that box and that
every window start of the run's 64 t0 lies in the staged range.
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


# ---- the LDS-staged CRS grid search (k_search_crs_lds in csrc/crs_lds.hip)

def _grid_value(f0, finc, i):
    """A grid value as both CRS fast kernels form it: one FMA."""
    return fma(F(finc), F(i), F(f0))


def crs_block_box(pair0, pair1, na, nb, nc, g):
    """The workgroup's (a, b, c) box from its run of (a, b) pairs [pair0,
    pair1), as k_search_crs_lds forms it."""
    fa0, finc_a, fb0, finc_b, fc0, finc_c = g
    ia0, ia1 = pair0 // nb, (pair1 - 1) // nb
    ib0, ib1 = (pair0 % nb, (pair1 - 1) % nb) if ia0 == ia1 else (0, nb - 1)
    ea = (_grid_value(fa0, finc_a, ia0), _grid_value(fa0, finc_a, ia1))
    eb = (_grid_value(fb0, finc_b, ib0), _grid_value(fb0, finc_b, ib1))
    ec = (_grid_value(fc0, finc_c, 0), _grid_value(fc0, finc_c, nc - 1))
    return [min(ea), min(eb), min(ec)], [max(ea), max(eb), max(ec)]


def test_crs_block_box_holds_every_parameter_of_its_pairs():
    """Every grid point of a workgroup's pairs lies in the box it stages for
    (both grid directions, runs within one a value and across several), and
    the staged range holds every such point's window start."""
    rng = np.random.default_rng(11)
    idt = 250.0
    bad = 0
    for _ in range(300):
        na, nb, nc = (int(x) for x in rng.integers(1, 12, 3))
        a0, a1, b0, b1 = -0.7e-3, 0.7e-3, -1e-7, 1e-7
        c0, c1 = 1.975e-7 * 1e6, 1.384e-6 * 1e6
        if rng.random() < 0.5:
            a0, a1, c0, c1 = a1, a0, c1, c0
        g = (a0, (a1 - a0) / na, b0, (b1 - b0) / nb, c0, (c1 - c0) / nc)
        npairs = na * nb
        per = int(rng.integers(1, 9))
        pair0 = int(rng.integers(0, npairs))
        pair1 = min(npairs, pair0 + per)
        lo, hi = crs_block_box(pair0, pair1, na, nb, nc, g)
        for pair in range(pair0, pair1):
            pa = _grid_value(g[0], g[1], pair // nb)
            pb = _grid_value(g[2], g[3], pair % nb)
            bad += int(not (lo[0] <= pa <= hi[0] and lo[1] <= pb <= hi[1]))
            for ic in range(nc):
                pc = _grid_value(g[4], g[5], ic)
                bad += int(not lo[2] <= pc <= hi[2])
        # window starts of the pairs' grid points against the box's bound
        dm = rng.uniform(-60.0, 60.0)
        half = rng.uniform(0.0, 2600.0)
        m, m2, h = F(dm * idt), F(dm * dm * idt * idt), F(half * half / 1e6 * idt * idt)
        T0lo = int(rng.integers(0, 1000)) // 64 * 64
        tau = int(rng.choice([0, 1, 2, 3]))
        elo, ehi = plan_bound(T0lo, T0lo + 63, lo, hi, m, m2, h, tau)
        T0 = np.arange(T0lo, T0lo + 64, dtype=np.float32)
        for pair in range(pair0, pair1):
            pa = _grid_value(g[0], g[1], pair // nb)
            pb = _grid_value(g[2], g[3], pair % nb)
            for ic in range(nc):
                pc = _grid_value(g[4], g[5], ic)
                for e in kernel_starts(T0, pa, pb, pc, m, m2, h, tau):
                    bad += int(np.sum((e < elo) | (e > ehi)))
    assert bad == 0


def test_crs_block_box_planted_mistake_fails():
    """Taking b's range from the first and last pair even when the run crosses
    an a value (instead of the whole b axis) misses grid points."""
    rng = np.random.default_rng(5)
    misses = 0
    for _ in range(200):
        na, nb, nc = 3, int(rng.integers(3, 9)), 4
        g = (-0.7e-3, 1.4e-3 / na, -1e-7, 2e-7 / nb, 0.1975, 1.19 / nc)
        pair0 = nb - 2  # the run crosses from a = 0 to a = 1
        pair1 = pair0 + 4
        ib0, ib1 = pair0 % nb, (pair1 - 1) % nb
        eb = (_grid_value(g[2], g[3], ib0), _grid_value(g[2], g[3], ib1))
        for pair in range(pair0, pair1):
            pb = _grid_value(g[2], g[3], pair % nb)
            misses += int(not min(eb) <= pb <= max(eb))
    assert misses > 0
