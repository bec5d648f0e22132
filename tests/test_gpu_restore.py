"""Transient FP32 copies of a 16-bit store (SB_OPT_RESTORE_FP32).

A fast search of an FP16 or BF16 store restores the rows it reads -- CMP: the
rows of its CDP range; CRS grid and CRS-DE: the rows of the range's
supergathers -- to FP32 and runs the FP32 kernels on the copy (CRS-DE with a
copy of its slot table addressing it).  With the option off the 16-bit kernels
restore every window after its load.  Both evaluate the same FP32 values with
the same arithmetic, so every output must be identical either way: on whole
lines, on CDP ranges (the copy then starts at the range's first row), and when
the copy does not fit the searching context's budget (the 16-bit kernels run
instead).  CRS-DE's outputs come from its FP32 evaluator, so on a 16-bit store
they must also equal CRS-DE's on an FP32 store of the restored samples."""
import numpy as np
import pytest

from test_gpu_store import restored

pytestmark = pytest.mark.gpu

KEYS = ("index", "a", "b", "c", "coher", "stack")
BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)


def _both(sb, ctx, g, prm, rng=None):
    """The search with the option on and off: (on, off) outputs."""
    out = []
    try:
        for flag in (1, 0):
            ctx.set_option(sb.OPT_RESTORE_FP32, flag)
            out.append(sb.search(ctx, g, prm, cdp_range=rng))
            assert ctx.timing()["restored"] == flag, (flag, rng)
    finally:
        ctx.set_option(sb.OPT_RESTORE_FP32, 1)
    return out


def _same(x, y):
    return all(np.array_equal(x[k], y[k]) for k in KEYS)


@pytest.mark.parametrize("store", ["f16", "bf16"])
def test_cmp_restore_on_and_off_agree(sb, gpu_ctx, store):
    print("case: cmp restore on/off", store, flush=True)
    off, geom, s = sb.synth(0, 30, (20, 110), 901, 2000, 61, h_max=1500.0, h_jitter=5.0, nevents=6, snr=5.0)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000, store=sb.STORES[store])
    try:
        prm = sb.cmp_params(1.98e-7, 1.77e-6, 96, 0.002)
        on, off_ = _both(sb, gpu_ctx, g, prm)
        assert _same(on, off_)
        assert (on["index"] >= 0).mean() > 0.3
        for rng in ((0, 5), (7, 19), (25, 30)):
            a, b = _both(sb, gpu_ctx, g, prm, rng)
            assert _same(a, b), rng
            assert all(np.array_equal(a[k], on[k][rng[0]:rng[1]]) for k in KEYS), rng
        # the global-memory shapes too
        gpu_ctx.tune_set(1, 16)
        try:
            a, b = _both(sb, gpu_ctx, g, prm, (3, 12))
        finally:
            gpu_ctx.tune_set(-1, 0)
        assert _same(a, b)
        assert np.array_equal(a["index"], on["index"][3:12])
    finally:
        g.close()


def _crs_line(sb, ctx, store, ncdp=10):
    off, geom, s = sb.synth(1, ncdp, (40, 90), 501, 4000, 62, h_max=1500.0, mid_half_width=40.0, nevents=6, snr=3.0)
    nbr_off, nbr = sb.neighbors(geom[off[:-1]], 40)
    return off, geom, s, nbr_off, nbr


@pytest.mark.parametrize("store", ["f16", "bf16"])
def test_crs_restore_on_and_off_agree(sb, gpu_ctx, store):
    print("case: crs restore on/off", store, flush=True)
    off, geom, s, nbr_off, nbr = _crs_line(sb, gpu_ctx, store)
    g = sb.Gather(gpu_ctx, off, geom, s, 4000, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
    try:
        prm = sb.crs_params(BOX[0], BOX[1], 5, BOX[2], BOX[3], 3, BOX[4], BOX[5], 7, 0.004)
        on, off_ = _both(sb, gpu_ctx, g, prm)
        assert _same(on, off_)
        for rng in ((3, 6), (0, 2), (8, 10)):
            a, b = _both(sb, gpu_ctx, g, prm, rng)
            assert _same(a, b), rng
            assert all(np.array_equal(a[k], on[k][rng[0]:rng[1]]) for k in KEYS), rng
    finally:
        g.close()


@pytest.mark.parametrize("store", ["f16", "bf16"])
def test_crs_de_restore_equals_f32_store_on_restored_samples(sb, gpu_ctx, store):
    print("case: crs-de restore", store, flush=True)
    off, geom, s, nbr_off, nbr = _crs_line(sb, gpu_ctx, store)
    prm = sb.de_params(*BOX, 0.004, pop=16, ngen=6, seed=11)
    g16 = sb.Gather(gpu_ctx, off, geom, s, 4000, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
    g32 = sb.Gather(gpu_ctx, off, geom, restored(s, store), 4000, nbr_off=nbr_off, nbr=nbr)
    try:
        want = sb.search(gpu_ctx, g32, prm)
        assert gpu_ctx.timing()["restored"] == 0
        on, off_ = _both(sb, gpu_ctx, g16, prm)
        assert _same(on, want) and _same(off_, want)
        for rng in ((2, 6), (9, 10)):
            a, b = _both(sb, gpu_ctx, g16, prm, rng)
            assert _same(a, b), rng
            assert all(np.array_equal(a[k], want[k][rng[0]:rng[1]]) for k in KEYS), rng
        assert (want["coher"] > 0).mean() > 0.5
    finally:
        g16.close()
        g32.close()


def test_restore_falls_back_when_the_copy_does_not_fit(sb, gpu_ctx):
    """A context whose budget holds the results and the work table but not
    the 80 MB copy: the 16-bit kernels run, with the same results."""
    print("case: restore budget fallback", flush=True)
    off, geom, s = sb.synth(0, 40, (240, 260), 2001, 1000, 63, h_max=1500.0, nevents=5, snr=5.0)
    g = sb.Gather(gpu_ctx, off, geom, s, 1000, store=sb.STORE_F16)
    prm = sb.cmp_params(1.98e-7, 1.77e-6, 32, 0.002)
    try:
        want = sb.search(gpu_ctx, g, prm)
        assert gpu_ctx.timing()["restored"] == 1
        with sb.Context(0, hbm_budget=24 << 20) as small:
            got = sb.search(small, g, prm)
            assert small.timing()["restored"] == 0
            used, budget = small.memory()
            assert used <= budget
        assert _same(got, want)
    finally:
        g.close()
