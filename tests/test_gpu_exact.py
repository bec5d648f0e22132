"""SB_MODE_EXACT equals the oracle (the reference's arithmetic and order)
bit for bit, on both of its kernels.

On an FP32 store every window (1-15 samples) runs the tiled search (64-t0
tiles, several parameters per trace pass, first maxima as packed keys for
CMP or merged in the workgroup for CRS, then the winner's re-evaluation);
16-bit stores run one thread per (CDP, t0).
Lines with uneven fold, shuffled offsets, late windows off the trace end, a
grid with negative C (invalid traveltimes) and several chunkings of the
parameter range."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("index", "c", "coher", "stack")


def _line(sb, seed, ns=601, ncdp=10, fold=(2, 70), h_max=2500.0):
    off, geom, s = sb.synth(0, ncdp, fold, ns, 2000, seed, h_max=h_max, h_jitter=5.0, nevents=5)
    rng = np.random.default_rng(seed)
    perm = np.concatenate([off[i] + rng.permutation(off[i + 1] - off[i]) for i in range(len(off) - 1)])
    return off, geom[perm], np.ascontiguousarray(s[perm])


@pytest.mark.parametrize("tau", [0.0, 0.002, 0.004, 0.006, 0.008, 0.014])
def test_exact_equals_oracle(sb, orc, gpu_ctx, tau):
    print("case: exact cmp tau=%g" % tau, flush=True)
    off, geom, s = _line(sb, 61)
    c0, c1, nc = 1.0e-7, 2.5e-6, 45
    ref = orc.cmp(off, geom, s, 2000, c0, c1, nc, tau)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000)
    try:
        out = sb.search(gpu_ctx, g, sb.cmp_params(c0, c1, nc, tau, sb.MODE_EXACT))
        part = sb.search(gpu_ctx, g, sb.cmp_params(c0, c1, nc, tau, sb.MODE_EXACT), cdp_range=(3, 7))
    finally:
        g.close()
    assert (ref["index"] >= 0).any() and (ref["index"] < 0).any()
    for k in KEYS:
        assert np.array_equal(out[k], ref[k]), k
        assert np.array_equal(part[k], ref[k][3:7]), k


def test_exact_negative_c_grid(sb, orc, gpu_ctx):
    print("case: exact cmp negative c", flush=True)
    off, geom, s = _line(sb, 63, ncdp=6)
    c0, c1, nc, tau = -1.0e-6, 1.5e-6, 40, 0.002
    ref = orc.cmp(off, geom, s, 2000, c0, c1, nc, tau)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000)
    try:
        fast = sb.search(gpu_ctx, g, sb.cmp_params(c0, c1, nc, tau, sb.MODE_FAST))  # runs exact (c < 0)
        ex = sb.search(gpu_ctx, g, sb.cmp_params(c0, c1, nc, tau, sb.MODE_EXACT))
    finally:
        g.close()
    for k in KEYS:
        assert np.array_equal(ex[k], ref[k]), k
        assert np.array_equal(fast[k], ref[k]), k


@pytest.mark.parametrize("store", ["f16", "bf16"])
def test_exact_on_16bit_store_matches_restored_input(sb, orc, gpu_ctx, store):
    """16-bit stores run the per-(CDP, t0) kernel on the restored amplitudes:
    equal to the oracle on the same amplitudes rounded as the store does
    (power-of-two scale, round to nearest)."""
    print("case: exact cmp store=%s" % store, flush=True)
    off, geom, s = _line(sb, 65, ncdp=5)
    scale = np.ones(len(s), np.float32)
    mx = np.abs(s).max(axis=1)
    nz = mx > 0
    scale[nz] = np.ldexp(np.float32(1.0), np.frexp(mx[nz])[1]).astype(np.float32)
    q = s / scale[:, None]
    if store == "f16":
        q = q.astype(np.float16).astype(np.float32)
    else:
        u = q.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000  # round to nearest even, as __float2bfloat16
        q = u.astype(np.uint32).view(np.float32)
    restored = (q * scale[:, None]).astype(np.float32)
    c0, c1, nc, tau = 1.98e-7, 1.77e-6, 33, 0.002
    ref = orc.cmp(off, geom, restored, 2000, c0, c1, nc, tau)
    g = sb.Gather(gpu_ctx, off, geom, s, 2000, store=sb.STORES[store])
    try:
        out = sb.search(gpu_ctx, g, sb.cmp_params(c0, c1, nc, tau, sb.MODE_EXACT))
    finally:
        g.close()
    for k in KEYS:
        assert np.array_equal(out[k], ref[k]), k


def _crs_line(sb, seed, ncdp=6):
    off, geom, s = sb.synth(1, ncdp, (20, 60), 401, 4000, seed, h_max=800.0, mid_half_width=25.0, nevents=4)
    no, nb = sb.neighbors(geom[off[:-1]], 30)
    return off, geom, s, no, nb


@pytest.mark.parametrize("tau", [0.0, 0.004, 0.008, 0.012, 0.016, 0.028])
def test_crs_exact_equals_oracle(sb, orc, gpu_ctx, tau):
    """CRS in the reference's double arithmetic: the tiled search (one
    workgroup per CDP and tile, first maximum merged across waves) equals
    the oracle bit for bit for windows of 1-15 samples, including t0 with no
    positive semblance (index 0 and the stack of (t0 = 0, parameter 0))."""
    print("case: exact crs tau=%g" % tau, flush=True)
    off, geom, s, no, nb = _crs_line(sb, 71)
    box = (-0.7e-3, 0.7e-3, 5, -1e-7, 1e-7, 4, 1.975e-7, 1.384e-6, 6)
    ref = orc.crs(off, geom, s, 4000, no, nb, *box, tau)
    g = sb.Gather(gpu_ctx, off, geom, s, 4000, nbr_off=no, nbr=nb)
    try:
        prm = sb.crs_params(box[0], box[1], box[2], box[3], box[4], box[5], box[6], box[7], box[8], tau,
                            sb.MODE_EXACT)
        out = sb.search(gpu_ctx, g, prm)
        part = sb.search(gpu_ctx, g, prm, cdp_range=(2, 5))
    finally:
        g.close()
    for k in ("index", "a", "b", "c", "coher", "stack"):
        assert np.array_equal(out[k], ref[k]), k
        assert np.array_equal(part[k], ref[k][2:5]), k
