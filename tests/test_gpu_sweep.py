"""Seeded random shapes through every launch path, against the oracle.

Each case draws a line -- CDP count, ragged fold, trace length, sample
interval, window, half-offset and midpoint spreads, events and noise -- and a
parameter grid (count, range, direction), then runs

  - SB_MODE_EXACT: every output bit-identical to the oracle's restatement of
    the reference (CMP float, CRS double, CRS-DE the FP32 twin);
  - SB_MODE_FAST on a drawn launch shape (any global-memory or LDS-staged
    variant -- sample, paired or mixed slabs --, an automatic or explicit
    parameter chunk, per-item or persistent grid): the parity bar of tests/parity.py against the reference-precision
    oracle (picks differ only at top-two gaps within 1e-5, or, counted and
    capped, where the FP64 oracle ranks our pick at least as high; coherence
    and stack bit-identical at every agreeing pick and equal to the oracle's
    panel value at every differing one).

The draws are fixed by the case number, and every case prints its shape
before it runs.
"""
import numpy as np
import pytest

from parity import check_fast_picks

pytestmark = pytest.mark.gpu

KEYS_CMP = ("index", "c", "coher", "stack")
KEYS_CRS = ("index", "a", "b", "c", "coher", "stack")


def _cmp_case(sb, i, nvariants):
    r = np.random.default_rng(1000 + i)
    dt_us = int(r.choice([1000, 2000, 4000]))
    fmin = int(r.integers(1, 41))
    fold = (fmin, int(r.integers(fmin, 91)))
    ns = int(r.integers(40, 801))
    ncdp = int(r.integers(1, 11))
    k = int(r.integers(0, 4))  # window half-length in samples: w = 2k + 1 <= 7
    tau = (k + 0.25) * dt_us * 1e-6
    nc = int(r.integers(1, 65))
    c0, c1 = sorted(r.uniform(1e-7, 2e-6, 2))
    if r.random() < 0.3:
        c0, c1 = c1, c0  # decreasing grid
    # half-offsets mostly within reach: the far traces' moveout c h^2 up to
    # about the trace length squared (a fifth of the cases reach past it)
    T = ns * dt_us * 1e-6
    reach = np.sqrt(T * T / max(c0, c1)) * (2.0 if r.random() < 0.2 else 0.7)
    line = dict(h_max=float(np.clip(reach * r.uniform(0.5, 1.0), 30.0, 3000.0)),
                h_jitter=float(r.choice([0.0, 5.0])), nevents=int(r.integers(1, 9)),
                snr=float(r.choice([2.0, 5.0, 100.0])))
    variant = int(r.integers(0, nvariants))
    pchunk = int(r.choice([0, 0, 1, 7, 16, 64]))
    persistent = int(variant < 4 and r.random() < 0.3)
    off, geom, s = sb.synth(0, ncdp, fold, ns, dt_us, 2000 + i, **line)
    desc = "cmp case %d: ncdp %d fold %s ns %d dt %d tau %.4g nc %d c %.3g..%.3g variant %d pchunk %d persistent %d" % (
        i, ncdp, fold, ns, dt_us, tau, nc, c0, c1, variant, pchunk, persistent)
    return desc, (off, geom, s, dt_us), (c0, c1, nc, tau), (variant, pchunk, persistent)


def _crs_line(sb, r, i):
    dt_us = int(r.choice([2000, 4000]))
    ns = int(r.integers(40, 401)) if r.random() < 0.2 else int(r.integers(200, 401))
    ncdp = int(r.integers(2, 9))
    fmin = int(r.integers(5, 31))
    fold = (fmin, int(r.integers(fmin, 61)))
    mid = float(r.uniform(5, 50))
    apm = int(r.integers(15, 61))
    k = int(r.integers(0, 3))
    tau = (k + 0.25) * dt_us * 1e-6
    T = ns * dt_us * 1e-6
    h_max = float(np.clip(np.sqrt(T * T / 1.6e-6) * r.uniform(0.4, 0.9), 30.0, 2000.0))
    off, geom, s = sb.synth(1, ncdp, fold, ns, dt_us, 3000 + i, h_max=h_max,
                            mid_half_width=mid, nevents=int(r.integers(1, 6)), snr=float(r.choice([3.0, 100.0])))
    nbr_off, nbr = sb.neighbors(geom[off[:-1]], apm)
    desc = "ncdp %d fold %s ns %d dt %d tau %.4g apm %d mid %.1f" % (ncdp, fold, ns, dt_us, tau, apm, mid)
    return desc, (off, geom, s, dt_us, nbr_off, nbr), tau


def _box(r):
    a = float(r.uniform(0.2e-3, 1e-3))
    b = float(r.uniform(0.3e-7, 2e-7))
    c0 = float(r.uniform(1.5e-7, 4e-7))
    return (-a, a, -b, b, c0, float(c0 + r.uniform(3e-7, 1.4e-6)))


@pytest.mark.parametrize("i", range(16))
def test_cmp_sweep(sb, orc, gpu_ctx, i):
    desc, (off, geom, s, dt_us), (c0, c1, nc, tau), (variant, pchunk, persistent) = _cmp_case(sb, i, gpu_ctx.tune_get()[2])
    print(desc, flush=True)
    g = sb.Gather(gpu_ctx, off, geom, s, dt_us)
    try:
        exact = sb.search(gpu_ctx, g, sb.cmp_params(c0, c1, nc, tau, sb.MODE_EXACT))
        gpu_ctx.tune_set(variant, pchunk)
        gpu_ctx.set_option(sb.OPT_PERSISTENT, persistent)
        fast = sb.search(gpu_ctx, g, sb.cmp_params(c0, c1, nc, tau))
    finally:
        gpu_ctx.set_option(sb.OPT_PERSISTENT, 0)
        gpu_ctx.tune_set(-1, 0)
        g.close()
    ref = orc.cmp(off, geom, s, dt_us, c0, c1, nc, tau, panel=True)
    for k in KEYS_CMP:
        assert np.array_equal(exact[k], ref[k]), (desc, "exact", k)
    ref64 = orc.cmp_f64(off, geom, s, dt_us, c0, c1, nc, tau, panel=True)
    check_fast_picks(desc, ref, fast, ref64, params=("c",))


@pytest.mark.parametrize("i", range(8))
def test_crs_sweep(sb, orc, gpu_ctx, i):
    r = np.random.default_rng(4000 + i)
    desc, (off, geom, s, dt_us, nbr_off, nbr), tau = _crs_line(sb, r, i)
    box = _box(r)
    na, nb, nc = (int(v) for v in r.integers(1, 6, 3))
    variant = int(r.integers(0, 4))
    desc = "crs case %d: %s grid %dx%dx%d variant %d" % (i, desc, na, nb, nc, variant)
    print(desc, flush=True)
    grid = (box[0], box[1], na, box[2], box[3], nb, box[4], box[5], nc, tau)
    g = sb.Gather(gpu_ctx, off, geom, s, dt_us, nbr_off=nbr_off, nbr=nbr)
    try:
        exact = sb.search(gpu_ctx, g, sb.crs_params(*grid, sb.MODE_EXACT))
        gpu_ctx.tune_set(variant, 0)
        fast = sb.search(gpu_ctx, g, sb.crs_params(*grid))
    finally:
        gpu_ctx.tune_set(-1, 0)
        g.close()
    ref = orc.crs(off, geom, s, dt_us, nbr_off, nbr, *grid, panel=True)
    for k in KEYS_CRS:
        assert np.array_equal(exact[k], ref[k]), (desc, "exact", k)
    check_fast_picks(desc, ref, fast, None, params=("a", "b", "c"))


@pytest.mark.parametrize("i", range(4))
def test_crs_de_sweep(sb, orc, gpu_ctx, i):
    r = np.random.default_rng(5000 + i)
    desc, (off, geom, s, dt_us, nbr_off, nbr), tau = _crs_line(sb, r, i)
    box = _box(r)
    pop = int(r.choice([4, 7, 16, 32, 50]))
    ngen = int(r.integers(0, 9))
    seed = int(r.integers(1, 2 ** 63))
    desc = "crs-de case %d: %s pop %d ngen %d" % (i, desc, pop, ngen)
    print(desc, flush=True)
    g = sb.Gather(gpu_ctx, off, geom, s, dt_us, nbr_off=nbr_off, nbr=nbr)
    try:
        got = sb.search(gpu_ctx, g, sb.de_params(*box, tau, pop=pop, ngen=ngen, seed=seed, mode=sb.MODE_EXACT))
    finally:
        g.close()
    want = orc.crs_de(off, geom, s, dt_us, nbr_off, nbr, *box, tau, pop, ngen, 0.5, 0.9, seed)
    for k in ("a", "b", "c", "coher", "stack"):
        assert np.array_equal(got[k], want[k]), (desc, k)
