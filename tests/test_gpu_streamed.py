"""Streamed one-shot search (sb_search_streamed): the line is uploaded in
CDP batches on a copy stream while completed batches are searched.  Results
must equal sb_gather_create + one search bit for bit, for every method and
store, with many small batches (CRS: CDPs searched only once all their
neighbours have arrived), and the kept gather must serve later searches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)
KEYS = ("index", "a", "b", "c", "coher", "stack")


def _eager(sb, ctx, line, prm, store):
    kw = {k: line[k] for k in ("nbr_off", "nbr") if k in line}
    g = sb.Gather(ctx, line["cdp_off"], line["geom"], line["samples"], line["dt_us"], store=store, **kw)
    try:
        return sb.search(ctx, g, prm)
    finally:
        g.close()


def _lines(sb):
    off, geom, s = sb.synth(0, 14, (20, 130), 701, 2000, 51, h_max=1500.0, h_jitter=5.0, nevents=6)
    cmp_line = dict(cdp_off=off, geom=geom, samples=s, dt_us=2000)
    off, geom, s = sb.synth(1, 10, (30, 90), 501, 4000, 52, h_max=800.0, mid_half_width=30.0, nevents=4)
    no, nb = sb.neighbors(geom[off[:-1]], 40)
    crs_line = dict(cdp_off=off, geom=geom, samples=s, dt_us=4000, nbr_off=no, nbr=nb)
    return cmp_line, crs_line


@pytest.mark.parametrize("rows", [0, 64, 37, 1])
@pytest.mark.parametrize("store", ["f32", "f16"])
def test_streamed_equals_gather_then_search(sb, gpu_ctx, store, rows):
    print("case: streamed store=%s rows=%d" % (store, rows), flush=True)
    st = sb.STORES[store]
    cmp_line, crs_line = _lines(sb)
    cases = [(cmp_line, sb.cmp_params(1.98e-7, 1.77e-6, 64, 0.002)),
             (cmp_line, sb.cmp_params(1.98e-7, 1.77e-6, 40, 0.004, mode=sb.MODE_EXACT)),
             (crs_line, sb.crs_params(BOX[0], BOX[1], 4, BOX[2], BOX[3], 3, BOX[4], BOX[5], 5, 0.004)),
             (crs_line, sb.de_params(*BOX, 0.004, pop=16, ngen=4, seed=7))]
    gpu_ctx.set_option(sb.OPT_STREAM_ROWS, rows)
    try:
        for line, prm in cases:
            want = _eager(sb, gpu_ctx, line, prm, st)
            kw = {k: line[k] for k in ("nbr_off", "nbr") if k in line}
            got = sb.search_streamed(gpu_ctx, prm, line["cdp_off"], line["geom"], line["samples"], line["dt_us"],
                                     store=st, **kw)
            for k in KEYS:
                assert np.array_equal(got[k], want[k]), (type(prm).__name__, k)
    finally:
        gpu_ctx.set_option(sb.OPT_STREAM_ROWS, 0)


def test_streamed_keeps_gather(sb, gpu_ctx):
    print("case: streamed keep", flush=True)
    cmp_line, _ = _lines(sb)
    p1 = sb.cmp_params(1.98e-7, 1.77e-6, 64, 0.002)
    p2 = sb.cmp_params(1.5e-7, 1.9e-6, 33, 0.002)
    gpu_ctx.set_option(sb.OPT_STREAM_ROWS, 100)
    try:
        first, g = sb.search_streamed(gpu_ctx, p1, cmp_line["cdp_off"], cmp_line["geom"], cmp_line["samples"],
                                      cmp_line["dt_us"], keep=True)
        try:
            again = sb.search(gpu_ctx, g, p1)
            other = sb.search(gpu_ctx, g, p2)
        finally:
            g.close()
        assert all(np.array_equal(first[k], again[k]) for k in KEYS)
        assert all(np.array_equal(other[k], _eager(sb, gpu_ctx, cmp_line, p2, sb.STORE_F32)[k]) for k in KEYS)
    finally:
        gpu_ctx.set_option(sb.OPT_STREAM_ROWS, 0)


def test_streamed_after_queued_searches(sb, gpu_ctx):
    """Device blocks freed while work that reads them is still queued on the
    ctx stream are handed out again only after that work has run (the
    allocator records an event on the stream at each free and the streamed
    upload writes on its own copy stream): a streamed search issued right
    after unsynchronised searches -- results never downloaded and destroyed,
    work tables of several CDP ranges built and replaced, their gather
    destroyed -- returns what the same streamed search returns on an idle
    ctx, bit for bit."""
    print("case: streamed after queued searches", flush=True)
    cmp_line, _ = _lines(sb)
    p = sb.cmp_params(1.98e-7, 1.77e-6, 64, 0.002)
    args = (cmp_line["cdp_off"], cmp_line["geom"], cmp_line["samples"], cmp_line["dt_us"])
    want = sb.search_streamed(gpu_ctx, p, *args)
    gpu_ctx.sync()
    off, geom, s = sb.synth(0, 160, (60, 120), 1501, 2000, 53, h_max=1500.0, h_jitter=5.0, nevents=6)
    pb = sb.cmp_params(1.98e-7, 1.77e-6, 256, 0.002)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000)
    for cb, ce in ((0, 160), (0, 40), (7, 151), (0, 160), (90, 100)):
        r = sb.DeviceResult(gpu_ctx, ce - cb, 1501)
        sb.run(gpu_ctx, g, pb, r, cdp_range=(cb, ce))
        r.close()  # freed with its search still queued
    g.close()
    got = sb.search_streamed(gpu_ctx, p, *args)
    for k in KEYS:
        assert np.array_equal(got[k], want[k]), k
