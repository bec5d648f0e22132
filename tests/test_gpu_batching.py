"""Batched launches: the persistent grid and HIP-graph replay.

The persistent grid (SB_OPT_PERSISTENT) runs the same work items as the
one-workgroup-per-item launch, so every output is bit-identical.  A captured
graph replays key clear, search and winner re-evaluation; each replay,
interleaved with eager searches on the same context (other parameters,
other gathers) and synchronisations, must equal the eager result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)
KEYS = ("index", "a", "b", "c", "coher", "stack")


def _cmp_line(sb, ctx, store, seed=31):
    off, geom, s = sb.synth(0, 24, (20, 140), 901, 2000, seed, h_max=1500.0, h_jitter=5.0, nevents=6)
    return sb.Gather(ctx, off, geom, s, 2000, store=store)


def _crs_line(sb, ctx, store, seed=32):
    off, geom, s = sb.synth(1, 8, (30, 90), 601, 4000, seed, h_max=800.0, mid_half_width=20.0, nevents=4)
    no, nb = sb.neighbors(geom[off[:-1]], 30)
    return sb.Gather(ctx, off, geom, s, 4000, store=store, nbr_off=no, nbr=nb)


def _same(x, y):
    return all(np.array_equal(x[k], y[k]) for k in KEYS)


@pytest.mark.parametrize("store", ["f32", "f16"])
def test_persistent_grid_equals_per_item_launch(sb, gpu_ctx, store):
    print("case: persistent grid %s" % store, flush=True)
    st = sb.STORES[store]
    g, gc = _cmp_line(sb, gpu_ctx, st), _crs_line(sb, gpu_ctx, st)
    try:
        for gg, prm in ((g, sb.cmp_params(1.98e-7, 1.77e-6, 96, 0.002)),
                        (gc, sb.crs_params(BOX[0], BOX[1], 5, BOX[2], BOX[3], 4, BOX[4], BOX[5], 6, 0.004))):
            for v in range(4):
                gpu_ctx.tune_set(v, 0)
                gpu_ctx.set_option(sb.OPT_PERSISTENT, 0)
                ref = sb.search(gpu_ctx, gg, prm)
                gpu_ctx.set_option(sb.OPT_PERSISTENT, 1)
                out = sb.search(gpu_ctx, gg, prm)
                out2 = sb.search(gpu_ctx, gg, prm, cdp_range=(1, 5))
                assert _same(out, ref), v
                assert all(np.array_equal(out2[k], ref[k][1:5]) for k in KEYS), v
    finally:
        gpu_ctx.set_option(sb.OPT_PERSISTENT, 0)
        gpu_ctx.tune_set(-1, 0)
        g.close()
        gc.close()


@pytest.mark.parametrize("persistent", [0, 1])
def test_graph_replays_equal_eager(sb, gpu_ctx, persistent):
    print("case: graph replay persistent=%d" % persistent, flush=True)
    gpu_ctx.set_option(sb.OPT_PERSISTENT, persistent)
    g = _cmp_line(sb, gpu_ctx, sb.STORE_F32)
    g16 = _cmp_line(sb, gpu_ctx, sb.STORE_F16)
    gc = _crs_line(sb, gpu_ctx, sb.STORE_F16)
    try:
        # the 16-bit cases replay a restored FP32 copy held by the graph
        # (SB_OPT_RESTORE_FP32), rewritten by every replay
        cases = [
            (g, sb.cmp_params(1.98e-7, 1.77e-6, 80, 0.002), None),
            (g, sb.cmp_params(1.98e-7, 1.77e-6, 33, 0.004, mode=sb.MODE_EXACT), (3, 11)),
            (g16, sb.cmp_params(1.98e-7, 1.77e-6, 64, 0.002), (2, 9)),
            (gc, sb.crs_params(BOX[0], BOX[1], 4, BOX[2], BOX[3], 3, BOX[4], BOX[5], 5, 0.004), None),
            (gc, sb.de_params(*BOX, 0.004, pop=16, ngen=5, seed=3), (2, 6)),
        ]
        other = sb.cmp_params(1.5e-7, 1.9e-6, 41, 0.002)
        for gg, prm, rng in cases:
            want = sb.search(gpu_ctx, gg, prm, cdp_range=rng)
            n = gg.ncdp if rng is None else rng[1] - rng[0]
            r = sb.DeviceResult(gpu_ctx, n, gg.ns)
            try:
                with sb.Graph(gpu_ctx, gg, prm, r, cdp_range=rng) as gr:
                    for rep in range(4):
                        gr.launch()
                        got = r.download()
                        assert _same(got, want), (prm, rep)
                        # one copy into a host buffer laid out as the device's
                        assert _same(r.download(contiguous=True), want), (prm, rep)
                        t = gpu_ctx.timing()
                        assert t["total_ms"] > 0
                        # eager work on the same ctx between replays
                        mid = sb.search(gpu_ctx, g, other)
                        assert mid["index"].shape == (g.ncdp, g.ns)
            finally:
                r.close()
    finally:
        gpu_ctx.set_option(sb.OPT_PERSISTENT, 0)
        g.close()
        g16.close()
        gc.close()
