"""LDS-staged CMP search (k_search_lds) against the global-memory kernel.

Both kernels do the same arithmetic on the same stored bits in the same trace
order, so every output must be bit-identical -- for every store, window
length, launch shape, parameter-grid direction, with one staging shared by
several parameter groups per wave, and on gathers whose window ranges do not
all fit the LDS (wide offset range, long traces, many parameters per
workgroup), where workgroups stage their first traces and read the other
traces' windows from global memory."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("index", "c", "coher", "stack")


def _search(sb, ctx, g, prm, variant, pchunk=0):
    ctx.tune_set(variant, pchunk)
    out = sb.search(ctx, g, prm)
    return out, ctx.timing()


def _lds_variants(sb, ctx):
    n = ctx.tune_get()[2]
    return [v for v in range(n) if v >= 4]


@pytest.mark.parametrize("store", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("tau,dt_us", [(0.002, 2000), (0.0, 2000), (0.004, 2000), (0.006, 2000)])
def test_lds_equals_global(sb, gpu_ctx, store, tau, dt_us):
    print("case: lds vs global store=%s tau=%g" % (store, tau), flush=True)
    st = {"f32": sb.STORE_F32, "f16": sb.STORE_F16, "bf16": sb.STORE_BF16}[store]
    off, geom, s = sb.synth(0, 7, (20, 150), 733, dt_us, 17, h_max=1500.0, h_jitter=5.0, nevents=6)
    g = sb.Gather(gpu_ctx, off, geom, s, dt_us, store=st)
    try:
        for c0, c1 in ((1.98e-7, 1.77e-6), (1.77e-6, 1.98e-7)):
            prm = sb.cmp_params(c0, c1, 57, tau)
            ref, t0 = _search(sb, gpu_ctx, g, prm, 0)
            assert t0["lds_staged"] == 0
            for v in _lds_variants(sb, gpu_ctx):
                for pc in (0, 57):
                    out, t = _search(sb, gpu_ctx, g, prm, v, pc)
                    assert t["lds_staged"] == 1, (v, t)
                    for k in KEYS:
                        assert np.array_equal(out[k], ref[k]), (v, pc, c0, k)
    finally:
        gpu_ctx.tune_set(-1, 0)
        g.close()


@pytest.mark.parametrize("store", ["f32", "f16", "bf16"])
def test_lds_ranges_beyond_lds(sb, gpu_ctx, store):
    """Half-offsets to 2.5 km, a wide c range and w = 7: with 200 parameters
    per workgroup the window ranges do not fit the LDS and those workgroups
    read their windows from global memory; the results must not change."""
    print("case: lds ranges beyond lds %s" % store, flush=True)
    st = {"f32": sb.STORE_F32, "f16": sb.STORE_F16, "bf16": sb.STORE_BF16}[store]
    off, geom, s = sb.synth(0, 3, (90, 120), 2001, 2000, 5, h_max=2500.0, nevents=5)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000, store=st)
    try:
        prm = sb.cmp_params(1.0e-7, 1.5e-6, 200, 0.006)
        ref, _ = _search(sb, gpu_ctx, g, prm, 0)
        for v in _lds_variants(sb, gpu_ctx):
            for pc in (0, 200):  # 200: the wide staging does not fit, narrow rounds run
                out, t = _search(sb, gpu_ctx, g, prm, v, pc)
                assert t["lds_staged"] == 1, (v, t)
                for k in KEYS:
                    assert np.array_equal(out[k], ref[k]), (v, pc, k)
    finally:
        gpu_ctx.tune_set(-1, 0)
        g.close()


@pytest.mark.parametrize("store", ["f32", "f16"])
def test_lds_partial_staging(sb, gpu_ctx, store):
    """128 and 256 parameters per workgroup on the headline line's geometry
    (ns 1501, nc 256, w 3): most workgroups stage only their first traces and
    read the others' windows from global memory; every output must equal the
    global-memory kernel's."""
    print("case: lds partial staging %s" % store, flush=True)
    st = sb.STORE_F32 if store == "f32" else sb.STORE_F16
    off, geom, s = sb.synth(0, 5, (60, 120), 1501, 2000, 31, h_max=1500.0, h_jitter=5.0, nevents=8)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000, store=st)
    try:
        prm = sb.cmp_params(1.98e-7, 1.77e-6, 256, 0.002)
        ref, _ = _search(sb, gpu_ctx, g, prm, 0)
        for v in _lds_variants(sb, gpu_ctx):
            for pc in (128, 256):
                out, t = _search(sb, gpu_ctx, g, prm, v, pc)
                assert t["lds_staged"] == 1, (v, t)
                for k in KEYS:
                    assert np.array_equal(out[k], ref[k]), (v, pc, k)
    finally:
        gpu_ctx.tune_set(-1, 0)
        g.close()


def test_lds_far_offsets_stage_nothing(sb, gpu_ctx):
    """Half-offsets of 3-4 km, 20001 samples at 0.5 ms and C from 1e-8 to
    5e-6: many workgroups cannot stage even their nearest trace's range (K =
    0, every window from global memory; tests/lds_replay.py counts them on the
    CPU), the others stage some; results equal the global-memory kernel's."""
    import lds_replay as lr
    print("case: lds far offsets (K = 0)", flush=True)
    off, geom, s = lr.far_offset_line()
    g = sb.Gather(gpu_ctx, off, geom, s, 500)
    try:
        prm = sb.cmp_params(1e-8, 5e-6, 16, 0.001)
        ref, _ = _search(sb, gpu_ctx, g, prm, 1)
        assert (ref["index"] >= 0).mean() > 0.05
        for v in _lds_variants(sb, gpu_ctx):
            out, t = _search(sb, gpu_ctx, g, prm, v, 16)
            assert t["lds_staged"] == 1, (v, t)
            for k in KEYS:
                assert np.array_equal(out[k], ref[k]), (v, k)
    finally:
        gpu_ctx.tune_set(-1, 0)
        g.close()


def test_lds_falls_back_when_it_cannot_stage(sb, gpu_ctx):
    """A CDP with more traces than the slab table holds runs on the global
    kernel (same results), and CRS never uses the LDS variants."""
    print("case: lds fallback", flush=True)
    off, geom, s = sb.synth(0, 2, (2100, 2100), 301, 2000, 9, h_max=1500.0, nevents=3)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000)
    try:
        prm = sb.cmp_params(1.98e-7, 1.77e-6, 16, 0.002)
        ref, _ = _search(sb, gpu_ctx, g, prm, 0)
        out, t = _search(sb, gpu_ctx, g, prm, 4)
        assert t["lds_staged"] == 0 and t["variant"] == 1, t
        for k in KEYS:
            assert np.array_equal(out[k], ref[k]), k
    finally:
        gpu_ctx.tune_set(-1, 0)
        g.close()


def test_lds_range_search(sb, gpu_ctx):
    print("case: lds range search", flush=True)
    off, geom, s = sb.synth(0, 9, (30, 90), 801, 2000, 23, h_max=1500.0, nevents=5)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000, store=sb.STORE_F16)
    try:
        prm = sb.cmp_params(1.98e-7, 1.77e-6, 64, 0.002)
        gpu_ctx.tune_set(-1, 0)
        full = sb.search(gpu_ctx, g, prm)
        part = sb.search(gpu_ctx, g, prm, cdp_range=(2, 6))
        assert gpu_ctx.timing()["lds_staged"] == 1
        for k in KEYS:
            assert np.array_equal(part[k], full[k][2:6]), k
    finally:
        gpu_ctx.tune_set(-1, 0)
        g.close()
