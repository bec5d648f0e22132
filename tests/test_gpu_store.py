"""16-bit trace stores: the fast search evaluates the restored samples
(stored value times the trace's power-of-two scale, exact in FP32) with the
FP32 arithmetic, so a search on an FP16 or BF16 store must pick exactly what
the FP32-store search picks on the same samples quantised and restored on the
host.  The host quantiser restates k_quantize's rule: scale = 2^e with
max|s| = f 2^e, f in [0.5, 1) (frexp), value = s / scale rounded to nearest
even in the 16-bit format.  Both ways of running a 16-bit search are checked:
the FP32 kernels on a transient restored copy (SB_OPT_RESTORE_FP32 = 1, the
default) and the 16-bit kernels reading the store (0)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)


def restored(samples, store):
    s = np.ascontiguousarray(samples, np.float32)
    mx = np.abs(s).max(axis=1)
    _, e = np.frexp(mx)
    sc = np.where(mx > 0, np.ldexp(np.float32(1.0), e), np.float32(1.0)).astype(np.float32)
    v = s * (np.float32(1.0) / sc)[:, None]
    if store == "f16":
        q = v.astype(np.float16).astype(np.float32)
    else:  # bfloat16, round to nearest even on the upper 16 bits
        u = v.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        q = u.astype(np.uint32).view(np.float32)
    return (q * sc[:, None]).astype(np.float32)


@pytest.mark.parametrize("restore", [1, 0])
@pytest.mark.parametrize("store", ["f16", "bf16"])
@pytest.mark.parametrize("variant,pchunk", [(5, 0), (7, 128), (9, 128), (12, 64), (1, 16), (0, 0)])
def test_cmp_16bit_store_equals_f32_on_restored_samples(sb, gpu_ctx, store, variant, pchunk, restore):
    print("case: cmp store", store, "variant", variant, "pchunk", pchunk, "restore", restore, flush=True)
    cdp_off, geom, samples = sb.synth(0, 60, (30, 120), 1001, 2000, 91, h_max=1500.0, h_jitter=5.0, nevents=6,
                                      snr=5.0)
    prm = sb.cmp_params(1.98e-7, 1.77e-6, 128, 0.002)
    gpu_ctx.tune_set(variant, pchunk)
    gpu_ctx.set_option(sb.OPT_RESTORE_FP32, restore)
    try:
        g16 = sb.Gather(gpu_ctx, cdp_off, geom, samples, 2000, store=sb.STORES[store])
        a = sb.search(gpu_ctx, g16, prm)
        assert gpu_ctx.timing()["restored"] == restore
        g16.close()
        g32 = sb.Gather(gpu_ctx, cdp_off, geom, restored(samples, store), 2000)
        b = sb.search(gpu_ctx, g32, prm)
        assert gpu_ctx.timing()["restored"] == 0
        g32.close()
    finally:
        gpu_ctx.tune_set(-1, 0)
        gpu_ctx.set_option(sb.OPT_RESTORE_FP32, 1)
    assert np.array_equal(a["index"], b["index"])
    assert (a["index"] >= 0).mean() > 0.3


@pytest.mark.parametrize("restore", [1, 0])
@pytest.mark.parametrize("store", ["f16", "bf16"])
def test_crs_16bit_store_equals_f32_on_restored_samples(sb, gpu_ctx, store, restore):
    print("case: crs store", store, "restore", restore, flush=True)
    cdp_off, geom, samples = sb.synth(1, 8, (60, 120), 501, 4000, 92, h_max=1500.0, mid_half_width=40.0, nevents=6,
                                      snr=3.0)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 40)
    prm = sb.crs_params(BOX[0], BOX[1], 7, BOX[2], BOX[3], 5, BOX[4], BOX[5], 9, 0.004)
    gpu_ctx.set_option(sb.OPT_RESTORE_FP32, restore)
    try:
        g16 = sb.Gather(gpu_ctx, cdp_off, geom, samples, 4000, store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
        a = sb.search(gpu_ctx, g16, prm)
        assert gpu_ctx.timing()["restored"] == restore
        g16.close()
    finally:
        gpu_ctx.set_option(sb.OPT_RESTORE_FP32, 1)
    g32 = sb.Gather(gpu_ctx, cdp_off, geom, restored(samples, store), 4000, nbr_off=nbr_off, nbr=nbr)
    b = sb.search(gpu_ctx, g32, prm)
    g32.close()
    assert np.array_equal(a["index"], b["index"])
