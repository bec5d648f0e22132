"""LDS-staged CRS grid search (SB_OPT_CRS_LDS) against the global-memory one.

The staged search gives each wave NC c values of one (a, b) pair for its 64
t0 and walks the supergather in batches of consecutive traces whose window
ranges over the workgroup's parameter box it stages in LDS.  It makes the
same roundings in the same trace order as k_search_fast, so every output --
index, A, B, C, coherence, stack -- must be bit-identical for every slab
size, c values per wave, slab layout, window, grid shape and direction,
store and CDP range.  SB_MODE_FAST throughout.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)
KEYS = ("index", "a", "b", "c", "coher", "stack")


def _line(sb, ncdp, fold, ns, seed, mid=20.0, h_max=700.0, dt=4000):
    cdp_off, geom, samples = sb.synth(1, ncdp, fold, ns, dt, seed=seed, h_max=h_max, mid_half_width=mid,
                                      nevents=4, snr=4.0)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 30)
    return cdp_off, geom, samples, nbr_off, nbr, dt


def _run(sb, ctx, line, params, lds, cpw=0, pairs=1, store="f32", cdp_range=None):
    cdp_off, geom, samples, nbr_off, nbr, dt = line
    ctx.set_option(sb.OPT_CRS_LDS, lds)
    ctx.set_option(sb.OPT_CRS_CPW, cpw)
    ctx.set_option(sb.OPT_CRS_PAIRS, pairs)
    try:
        g = sb.Gather(ctx, cdp_off, geom, samples, dt, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
        out = sb.search(ctx, g, params, cdp_range=cdp_range)
        t = ctx.timing()
        g.close()
    finally:
        ctx.set_option(sb.OPT_CRS_LDS, -1)
        ctx.set_option(sb.OPT_CRS_CPW, 0)
        ctx.set_option(sb.OPT_CRS_PAIRS, 1)
    return out, t


def _same(ref, out, what):
    for k in KEYS:
        a, b = ref[k], out[k]
        same = np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b)
        assert same, "%s: %s differs at %d cells" % (what, k, int((a != b).sum()))


def _prm(sb, na, nb, nc, tau=0.004, flip=False):
    a0, a1, b0, b1, c0, c1 = BOX
    if flip:
        a0, a1, c0, c1 = a1, a0, c1, c0
    return sb.crs_params(a0, a1, na, b0, b1, nb, c0, c1, nc, tau, sb.MODE_FAST)


@pytest.mark.parametrize("pairs", [1, 0])
@pytest.mark.parametrize("cpw", [6, 12])
@pytest.mark.parametrize("lds", [-1, 8192, 65536])
def test_crs_lds_equals_global(sb, gpu_ctx, lds, cpw, pairs):
    print("case: crs lds", lds, "cpw", cpw, "pairs", pairs, flush=True)
    line = _line(sb, 6, (12, 24), 301, 71)
    prm = _prm(sb, 5, 4, 11)
    ref, t0 = _run(sb, gpu_ctx, line, prm, 0)
    assert t0["lds_staged"] == 0
    out, t = _run(sb, gpu_ctx, line, prm, lds, cpw, pairs)
    assert t["lds_staged"] == 1 and t["block_threads"] == 512
    _same(ref, out, "lds %d cpw %d pairs %d" % (lds, cpw, pairs))


@pytest.mark.parametrize("tau", [0.0, 0.004, 0.008, 0.012])
@pytest.mark.parametrize("grid", [(1, 1, 1), (3, 2, 5), (2, 3, 6), (4, 1, 7), (1, 5, 12), (2, 2, 13), (1, 1, 48),
                                  (3, 3, 11)])
def test_crs_lds_windows_and_grids(sb, gpu_ctx, tau, grid):
    print("case: crs lds tau", tau, "grid", grid, flush=True)
    line = _line(sb, 5, (10, 30), 257, 73)
    for flip in (False, True):
        prm = _prm(sb, *grid, tau=tau, flip=flip)
        ref, _ = _run(sb, gpu_ctx, line, prm, 0)
        for lds, cpw, pairs in ((-1, 0, 1), (6144, 12, 1), (-1, 6, 0)):
            out, t = _run(sb, gpu_ctx, line, prm, lds, cpw, pairs)
            assert t["lds_staged"] == 1
            _same(ref, out, "tau %g grid %s flip %s lds %d cpw %d pairs %d" % (tau, grid, flip, lds, cpw, pairs))


@pytest.mark.parametrize("store", ["f16", "bf16"])
def test_crs_lds_restored_16bit_store(sb, gpu_ctx, store):
    """16-bit stores run the staged search on their restored FP32 copy: the
    same bits as the global-memory search reading the store directly."""
    print("case: crs lds store", store, flush=True)
    line = _line(sb, 6, (12, 24), 301, 77)
    prm = _prm(sb, 4, 3, 9)
    gpu_ctx.set_option(sb.OPT_RESTORE_FP32, 0)
    try:
        ref, t0 = _run(sb, gpu_ctx, line, prm, 0, store=store)
    finally:
        gpu_ctx.set_option(sb.OPT_RESTORE_FP32, 1)
    assert t0["restored"] == 0 and t0["lds_staged"] == 0
    out, t = _run(sb, gpu_ctx, line, prm, -1, store=store)
    assert t["restored"] == 1 and t["lds_staged"] == 1
    _same(ref, out, "store " + store)


def test_crs_lds_cdp_range_far_offsets_long_traces(sb, gpu_ctx):
    """CDP ranges of a line with far offsets (wide window ranges: several
    batches per workgroup) and long traces."""
    print("case: crs lds cdp range, far offsets, ns 1501", flush=True)
    line = _line(sb, 8, (30, 60), 1501, 79, mid=40.0, h_max=2500.0)
    prm = _prm(sb, 3, 3, 11)
    for rng_ in ((2, 7), (0, 8)):
        ref, _ = _run(sb, gpu_ctx, line, prm, 0, cdp_range=rng_)
        for lds, pairs in ((-1, 1), (65536, 1), (16384, 1), (-1, 0), (12288, 0)):
            out, t = _run(sb, gpu_ctx, line, prm, lds, pairs=pairs, cdp_range=rng_)
            assert t["lds_staged"] == 1
            _same(ref, out, "range %s lds %d pairs %d" % (rng_, lds, pairs))


def test_crs_lds_azimuth_signed_offsets(sb, gpu_ctx):
    print("case: crs lds azimuth", flush=True)
    line = _line(sb, 6, (12, 24), 301, 83, mid=40.0)
    prm = _prm(sb, 5, 3, 7)
    gpu_ctx.set_option(sb.OPT_DM_AZIMUTH, 30000)
    try:
        ref, _ = _run(sb, gpu_ctx, line, prm, 0)
        out, t = _run(sb, gpu_ctx, line, prm, -1)
    finally:
        gpu_ctx.set_option(sb.OPT_DM_AZIMUTH, -1)
    assert t["lds_staged"] == 1
    _same(ref, out, "azimuth")


def test_crs_lds_falls_back(sb, gpu_ctx):
    """More c values than the waves hold (49) or a slab below one trace's
    longest range: the global-memory search runs, same results."""
    print("case: crs lds fallback", flush=True)
    line = _line(sb, 3, (10, 14), 1501, 89)
    for prm, lds in ((_prm(sb, 1, 1, 49), -1), (_prm(sb, 2, 2, 5), 4096)):
        ref, _ = _run(sb, gpu_ctx, line, prm, 0)
        out, t = _run(sb, gpu_ctx, line, prm, lds)
        assert t["lds_staged"] == 0
        _same(ref, out, "fallback")


def test_crs_lds_option_validation(sb, gpu_ctx):
    for bad in (1, 4095, 65537, 5003):
        with pytest.raises(sb.SbError):
            gpu_ctx.set_option(sb.OPT_CRS_LDS, bad)
    with pytest.raises(sb.SbError):
        gpu_ctx.set_option(sb.OPT_CRS_CPW, 7)
    assert gpu_ctx.get_option(sb.OPT_CRS_LDS) == -1
    assert gpu_ctx.get_option(sb.OPT_CRS_CPW) == 0
    assert gpu_ctx.get_option(sb.OPT_CRS_PAIRS) == 1
