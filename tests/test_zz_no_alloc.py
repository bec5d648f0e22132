"""sb_cmp_run_noalloc (the allocation-free CMP search, scratch in a static
device array) against sb_cmp_run_range on the same gather: bit-identical
outputs on two calls in a row, the second smaller, and no device memory
taken by the calling context.  Output arrays are the caller's (allocated here
with hipMalloc through ctypes, outside the library)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("index", "a", "b", "c", "coher", "stack")


class Hip:
    def __init__(self, sb):
        # The HIP runtime the library already uses (by soname; torch, when
        # imported, may have brought its own copy of it): never a second one.
        sb.lib()
        self.L = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_NOW)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), n) == 0
        assert self.L.hipMemset(p, 0xff, n) == 0  # stale values: every element must be written
        assert self.L.hipDeviceSynchronize() == 0  # the library's streams do not wait for the null stream
        return p

    def download(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out


def _noalloc(sb, hip, ctx, g, prm, cb, ce, bufs):
    n = (ce - cb) * g.ns
    arr = sb.DeviceArrays(*[bufs[k] for k in KEYS])
    sb.check(sb.lib().sb_cmp_run_noalloc(ctx.h, g.h, C.byref(prm), cb, ce, C.byref(arr)))
    ctx.sync()
    return {k: hip.download(bufs[k], n, np.int32 if k == "index" else np.float32).reshape(ce - cb, g.ns)
            for k in KEYS}


@pytest.mark.parametrize("store", ["f32", "f16"])
def test_noalloc_equals_run_range_on_two_calls(sb, store):
    print("case: noalloc cmp store", store, flush=True)
    hip = Hip(sb)
    big = sb.synth(0, 48, (30, 90), 1001, 2000, seed=71, h_max=1500.0, h_jitter=5.0, nevents=6, snr=5.0)
    small = sb.synth(0, 9, (24, 60), 601, 2000, seed=72, h_max=1200.0, h_jitter=5.0, nevents=4, snr=5.0)
    prm_big = sb.cmp_params(1.98e-7, 1.77e-6, 101, 0.002)
    prm_small = sb.cmp_params(1.98e-7, 1.77e-6, 64, 0.004)
    owner = sb.Context(0)
    caller = sb.Context(0)  # the context the allocation-free calls run on
    try:
        gb = sb.Gather(owner, *big, 2000, store=sb.STORES[store])
        gs = sb.Gather(owner, *small, 2000, store=sb.STORES[store])
        want_big = sb.search(owner, gb, prm_big)
        want_small = sb.search(owner, gs, prm_small, (2, 7))
        nmax = 48 * 1001
        bufs = {k: hip.malloc(nmax * 4) for k in KEYS}
        try:
            used0 = caller.memory()[0]
            got_big = _noalloc(sb, hip, caller, gb, prm_big, 0, 48, bufs)
            got_small = _noalloc(sb, hip, caller, gs, prm_small, 2, 7, bufs)
            assert caller.memory()[0] == used0 == 0
        finally:
            for p in bufs.values():
                hip.L.hipFree(p)
        for k in KEYS:
            assert np.array_equal(got_big[k], want_big[k]), ("big", k)
            assert np.array_equal(got_small[k], want_small[k]), ("small", k)
        assert (want_big["index"] >= 0).mean() > 0.3
        gb.close()
        gs.close()
    finally:
        caller.close()
        owner.close()


def test_noalloc_rejects_searches_beyond_the_scratch(sb):
    print("case: noalloc size limit", flush=True)
    line = sb.synth(0, 300, 30, 1501, 2000, seed=73, h_max=1500.0, nevents=2, snr=5.0)
    with sb.Context(0) as ctx:
        g = sb.Gather(ctx, *line, 2000)
        arr = sb.DeviceArrays(*([None] * 6))
        prm = sb.cmp_params(1.98e-7, 1.77e-6, 32, 0.002)
        # 300 x 1501 (CDP, t0) keys exceed the scratch's 393 216: refused before any launch
        rc = sb.lib().sb_cmp_run_noalloc(ctx.h, g.h, C.byref(prm), 0, 300, C.byref(arr))
        assert rc == 1  # null outputs: SB_ERR_INVALID, checked first
        hip = Hip(sb)
        bufs = {k: hip.malloc(300 * 1501 * 4) for k in KEYS}
        try:
            arr = sb.DeviceArrays(*[bufs[k] for k in KEYS])
            rc = sb.lib().sb_cmp_run_noalloc(ctx.h, g.h, C.byref(prm), 0, 300, C.byref(arr))
            assert rc == 7 and b"static scratch" in sb.lib().sb_last_error()  # SB_ERR_UNSUPPORTED
            rc = sb.lib().sb_cmp_run_noalloc(ctx.h, g.h, C.byref(prm), 0, 200, C.byref(arr))  # 300 200 keys: fits
            assert rc == 0
            ctx.sync()
        finally:
            for p in bufs.values():
                hip.L.hipFree(p)
        g.close()
