"""LDS-staged CRS-DE evaluator (SB_OPT_DE_LDS) against the global-memory one.

The staged evaluator walks each supergather in batches of consecutive traces
whose window ranges (over the workgroup's t0 span and the whole (A, B, C)
box) it stages in LDS as (sample, difference) pairs.  It makes the same
roundings in the same order as the global-memory evaluator, so every output
-- A, B, C, coherence, stack -- must be bit-identical for every slab size
(one resident batch, a few batches, one trace per batch), every workgroup
size, window, population, store and CDP range.  SB_MODE_FAST throughout.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)
KEYS = ("index", "a", "b", "c", "coher", "stack")


def _line(sb, ncdp, fold, ns, seed, mid=20.0, h_max=700.0):
    cdp_off, geom, samples = sb.synth(1, ncdp, fold, ns, 4000, seed=seed, h_max=h_max, mid_half_width=mid,
                                      nevents=4, snr=4.0)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 30)
    return cdp_off, geom, samples, nbr_off, nbr


def _run(sb, ctx, line, params, lds, waves=8, store="f32", cdp_range=None, pairs=1):
    cdp_off, geom, samples, nbr_off, nbr = line
    ctx.set_option(sb.OPT_DE_LDS, lds)
    ctx.set_option(sb.OPT_DE_WAVES, waves)
    ctx.set_option(sb.OPT_DE_PAIRS, pairs)
    try:
        g = sb.Gather(ctx, cdp_off, geom, samples, 4000, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
        out = sb.search(ctx, g, params, cdp_range=cdp_range)
        t = ctx.timing()
        g.close()
    finally:
        ctx.set_option(sb.OPT_DE_LDS, -1)
        ctx.set_option(sb.OPT_DE_WAVES, 8)
        ctx.set_option(sb.OPT_DE_PAIRS, 1)
    return out, t


def _same(ref, out, what):
    for k in KEYS:
        a, b = ref[k], out[k]
        same = np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b)
        assert same, "%s: %s differs at %d cells" % (what, k, int((a != b).sum()))


@pytest.mark.parametrize("pairs", [1, 0])
@pytest.mark.parametrize("lds", [-1, 4096, 12288, 65536])
@pytest.mark.parametrize("waves", [4, 8, 16])
def test_de_lds_equals_global(sb, gpu_ctx, lds, waves, pairs):
    print("case: de lds", lds, "waves", waves, "pairs", pairs, flush=True)
    line = _line(sb, 6, (12, 24), 301, 41)
    prm = sb.de_params(*BOX, 0.004, pop=32, ngen=5, seed=7)
    ref, t0 = _run(sb, gpu_ctx, line, prm, 0)
    assert t0["lds_staged"] == 0
    out, t = _run(sb, gpu_ctx, line, prm, lds, waves, pairs=pairs)
    assert t["lds_staged"] == 1 and t["block_threads"] == 64 * waves
    _same(ref, out, "lds %d waves %d pairs %d" % (lds, waves, pairs))


@pytest.mark.parametrize("pop,tau", [(4, 0.0), (16, 0.004), (20, 0.008), (32, 0.012), (64, 0.004), (8, 0.002)])
def test_de_lds_windows_and_populations(sb, gpu_ctx, pop, tau):
    print("case: de lds pop", pop, "tau", tau, flush=True)
    line = _line(sb, 5, (10, 30), 257, 43 + pop)
    prm = sb.de_params(*BOX, tau, pop=pop, ngen=4, seed=3 + pop)
    ref, _ = _run(sb, gpu_ctx, line, prm, 0)
    for lds, waves, pairs in ((-1, 8, 1), (8192, 4, 1), (5120, 16, 1), (-1, 8, 0), (4096, 4, 0)):
        out, t = _run(sb, gpu_ctx, line, prm, lds, waves, pairs=pairs)
        assert t["lds_staged"] == 1
        _same(ref, out, "pop %d tau %g lds %d waves %d pairs %d" % (pop, tau, lds, waves, pairs))


@pytest.mark.parametrize("store", ["f16", "bf16"])
def test_de_lds_restored_16bit_store(sb, gpu_ctx, store):
    """16-bit stores run the staged evaluator on their restored FP32 copy:
    the same bits as the global-memory evaluator reading the store directly
    (each window restored after its load)."""
    print("case: de lds store", store, flush=True)
    line = _line(sb, 6, (12, 24), 301, 47)
    prm = sb.de_params(*BOX, 0.004, pop=32, ngen=5, seed=8)
    gpu_ctx.set_option(sb.OPT_RESTORE_FP32, 0)
    try:
        ref, t0 = _run(sb, gpu_ctx, line, prm, 0, store=store)
    finally:
        gpu_ctx.set_option(sb.OPT_RESTORE_FP32, 1)
    assert t0["restored"] == 0 and t0["lds_staged"] == 0
    out, t = _run(sb, gpu_ctx, line, prm, -1, store=store)
    assert t["restored"] == 1 and t["lds_staged"] == 1
    _same(ref, out, "store " + store)


def test_de_lds_cdp_range_far_offsets_long_traces(sb, gpu_ctx):
    """A CDP range of a line with far offsets (wide window ranges: several
    batches per generation even at 64 KiB) and long traces."""
    print("case: de lds cdp range, far offsets, ns 1501", flush=True)
    line = _line(sb, 8, (30, 60), 1501, 53, mid=40.0, h_max=2500.0)
    prm = sb.de_params(*BOX, 0.004, pop=32, ngen=3, seed=12)
    for rng_ in ((2, 7), (0, 8)):
        ref, _ = _run(sb, gpu_ctx, line, prm, 0, cdp_range=rng_)
        for lds, pairs in ((-1, 1), (65536, 1), (16384, 1), (-1, 0), (12288, 0)):
            out, t = _run(sb, gpu_ctx, line, prm, lds, cdp_range=rng_, pairs=pairs)
            assert t["lds_staged"] == 1
            _same(ref, out, "range %s lds %d pairs %d" % (rng_, lds, pairs))


def test_de_lds_azimuth_signed_offsets(sb, gpu_ctx):
    """Signed midpoint projections (SB_OPT_DM_AZIMUTH): the range bound
    takes min/max over both signs of a m."""
    print("case: de lds azimuth", flush=True)
    line = _line(sb, 6, (12, 24), 301, 59, mid=40.0)
    prm = sb.de_params(*BOX, 0.004, pop=32, ngen=4, seed=13)
    gpu_ctx.set_option(sb.OPT_DM_AZIMUTH, 30000)
    try:
        ref, _ = _run(sb, gpu_ctx, line, prm, 0)
        out, t = _run(sb, gpu_ctx, line, prm, -1)
    finally:
        gpu_ctx.set_option(sb.OPT_DM_AZIMUTH, -1)
    assert t["lds_staged"] == 1
    _same(ref, out, "azimuth")


def test_de_lds_falls_back_when_the_slab_is_too_small(sb, gpu_ctx):
    """A slab below one trace's longest range: the global-memory evaluator
    runs (lds_staged 0), same results."""
    print("case: de lds too small", flush=True)
    line = _line(sb, 3, (10, 14), 1501, 61)
    prm = sb.de_params(*BOX, 0.004, pop=16, ngen=2, seed=1)
    ref, _ = _run(sb, gpu_ctx, line, prm, 0)
    out, t = _run(sb, gpu_ctx, line, prm, 4096)
    assert t["lds_staged"] == 0
    _same(ref, out, "fallback")


def test_de_lds_option_validation(sb, gpu_ctx):
    for bad in (1, 4095, 65537, 5000 + 3):
        with pytest.raises(sb.SbError):
            gpu_ctx.set_option(sb.OPT_DE_LDS, bad)
    with pytest.raises(sb.SbError):
        gpu_ctx.set_option(sb.OPT_DE_WAVES, 2)
    assert gpu_ctx.get_option(sb.OPT_DE_LDS) == -1
    assert gpu_ctx.get_option(sb.OPT_DE_WAVES) == 8
    assert gpu_ctx.get_option(sb.OPT_DE_PAIRS) == 1
