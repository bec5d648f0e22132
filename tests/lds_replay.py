"""CPU replay of the LDS-staged CMP search's staging decision (k_search_lds):
for every work item (CDP, 64-t0 tile, parameter chunk), how many of the
CDP's traces, in slot order (increasing h), the workgroup stages in LDS --
all of them (full), the first K < S (partial), none (K = 0: every window read
from global memory) -- or whether it runs nothing at all (dead: past tau,
the smallest c already puts the second-to-last trace out of range).

Restates the kernel's arithmetic in FP32 (numpy's square root is correctly
rounded; the hardware's is within 1 ulp, which moves a range end by at most
one sample and the counts negligibly), and the LDS layout of an 8-wave
workgroup: 64 zero bytes, the scan scratch, TraceLds[S], s_pre[S + 1],
s_lo[S], then the slabs: 16 bytes per quad of samples, or 32 for the shapes
that stage (sample, difference) pairs.  The mixed shapes stage pairs for the
first Kp traces and samples for traces Kp .. K - 1: every trace as pairs
when that fits; else, when the samples of every trace fit, every trace with
as many pairs as the rest of the slab holds; else pairs in at most share/16
of the slab bytes and samples after them.  Test infrastructure only.
"""
import numpy as np

f32 = np.float32


def _fma(a, b, c):
    return (np.float64(a) * np.float64(b) + np.float64(c)).astype(np.float32)


def slab_offset(S, nw=8):
    tr = 64 + ((4 * nw + 15) & ~15)
    return tr + ((8 * S + 4 * (2 * S + 1) + 15) & ~15)


def slot_h(cdp_off, geom, dt_us):
    """Per CDP, the slot table's scaled h (FP32, increasing, ties by row)."""
    s = np.where(geom["scalco"] == 0, 1.0, np.where(geom["scalco"] > 0, geom["scalco"], 1.0 / geom["scalco"]))
    hx = (geom["gx"].astype(np.float64) - geom["sx"]) * s
    hy = (geom["gy"].astype(np.float64) - geom["sy"]) * s
    h0 = 0.25 * (hx * hx + hy * hy) / 1e6
    idt = 1.0 / (dt_us / 1e6)
    h = (h0 * idt * idt).astype(np.float32)
    return [np.sort(h[cdp_off[i]:cdp_off[i + 1]], kind="stable") for i in range(len(cdp_off) - 1)]


def _largest_le(pre, x):
    """The largest k with pre[k] <= x (pre nondecreasing from pre[0] = 0 <= x)."""
    return int(np.searchsorted(pre, x, side="right") - 1)


def mixed_split(pre, cap16, share):
    """(Kp, K) of a mixed-slab workgroup: pre = the exclusive prefix sums of
    the traces' quads (length S + 1), cap16 = the slab bytes / 16."""
    S = len(pre) - 1
    total = int(pre[S])
    if 2 * total <= cap16:
        return S, S
    if total <= cap16:
        return _largest_le(pre, cap16 - total), S
    kp = _largest_le(pre, (share * cap16) >> 5)
    return kp, _largest_le(pre, cap16 - int(pre[kp]))


def staging_cases(cdp_off, geom, dt_us, ns, tau_s, c0, c1, nc, pchunk, lds_kb, w_compiled=None, quad_bytes=16,
                  share=None):
    """Counts of work items by staging case: dict(full, partial, none, dead);
    with share (mixed slabs) also by the traces staged as pairs: pairs_all
    (Kp = S), mixed_full (0 < Kp < K = S), mixed_partial (0 < Kp <= K < S),
    pairs_none (Kp = 0)."""
    dt = f32(f32(dt_us) / f32(1e6))
    tau = max(int(f32(tau_s) * (f32(1.0) / dt)), 0)
    W = 2 * tau + 1
    cf0 = f32(np.float64(f32(c0)) * 1e6)
    cf1 = f32(np.float64(f32(c1)) * 1e6)
    inc = f32((cf1 - cf0) * f32(f32(1.0) / f32(nc)))
    hi = ns - tau - 1
    hi2 = f32(hi * hi * (1.0 - 2.0 ** -20)) if hi > 0 else f32(0)
    out = dict(full=0, partial=0, none=0, dead=0)
    if share is not None:
        out.update(pairs_all=0, mixed_full=0, mixed_partial=0, pairs_none=0)
    for hs in slot_h(cdp_off, geom, dt_us):
        S = len(hs)
        cap16 = (lds_kb * 1024 - slab_offset(S)) // 16
        capQ = cap16 // (quad_bytes // 16)
        h2 = hs[S - 2] if S >= 2 else f32(0)
        for t0min in range(0, ns, 64):
            sqlo = f32(f32(t0min) * f32(t0min))
            sqhi = f32(f32(t0min + 63) * f32(t0min + 63))
            prefix = t0min > tau
            for p0 in range(0, nc, pchunk):
                p1 = min(nc, p0 + pchunk)
                cA = f32(cf0 + f32(inc * f32(p0)))
                cB = f32(cf0 + f32(inc * f32(p1 - 1)))
                cmin, cmax = min(cA, cB), max(cA, cB)
                if prefix and S >= 2 and not (_fma(cmin, h2, sqlo) < hi2):
                    out["dead"] += 1
                    continue
                elo = np.maximum(np.sqrt(_fma(cmin, hs, sqlo)).astype(np.int64) - tau - 2, 0)
                ehi = np.sqrt(_fma(cmax, hs, sqhi)).astype(np.int64) - tau + 2
                lo = elo >> 2
                length = np.where(ehi >= elo, ((ehi + W) >> 2) - lo + 1, 0)
                pre = np.concatenate([[0], np.cumsum(length)])
                if share is not None:
                    Kp, K = mixed_split(pre, cap16, share)
                    out["pairs_all" if Kp == S else "pairs_none" if Kp == 0 else
                        "mixed_full" if K == S else "mixed_partial"] += 1
                elif pre[S] <= capQ:
                    K = S
                else:
                    K = _largest_le(pre, capQ)  # pre[K] <= capQ < pre[K + 1]
                out["full" if K == S else "partial" if K > 0 else "none"] += 1
    return out


def far_offset_line(ncdp=3, fold=12, ns=20001, seed=0):
    """CMP gathers whose every half-offset is 3-4 km (coordinates in cm,
    scalco -100) with long traces of noise and one flat event: with dt 0.5 ms
    and a C range of 1e-8 to 5e-6 s^2/m^2 even the nearest trace's window range
    over a 64-t0 tile exceeds what an LDS workgroup holds (K = 0)."""
    rng = np.random.default_rng(seed)
    dtype = np.dtype([("cdp", "<i4"), ("scalco", "<i4"), ("sx", "<i4"), ("sy", "<i4"), ("gx", "<i4"), ("gy", "<i4")])
    geom = np.zeros(ncdp * fold, dtype)
    for i in range(ncdp):
        for k in range(fold):
            h = rng.uniform(3000.0, 4000.0)
            mx = 12.5 * i
            r = i * fold + k
            geom[r] = (i + 1, -100, round((mx - h) * 100), 0, round((mx + h) * 100), 0)
    samples = rng.standard_normal((ncdp * fold, ns)).astype(np.float32)
    samples[:, 3000:3010] += 5.0
    cdp_off = (np.arange(ncdp + 1) * fold).astype(np.int32)
    return cdp_off, geom, samples
