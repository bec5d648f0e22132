"""The LDS-staged CMP search reads every window from its workgroup's staged
sample ranges (CPU check of the bound, no GPU).

k_search_lds stages, per trace of a (CDP, 64-t0 tile, parameter chunk), the
16-byte quads that hold samples max(elo, 0) .. ehi + W, where

    elo = floor(sqrt(fma(cmin, h, t0min^2))) - tau - 2
    ehi = floor(sqrt(fma(cmax, h, (t0min + 63)^2))) - tau + 2

(FP32, the square root of the hardware, accurate to 1 ulp).  A lane reads the
window starting at e = floor(sqrt(fma(c, h, T0^2))) - tau.  In the prefix
path (t0min > tau) every lane reads its window without a validity test --
also lanes whose result is discarded and lanes past the end of the trace --
so each such window must lie inside the staged quads; in the general path
only valid windows (0 <= e < ns - w and t^2 < hi2) are read.  This test
evaluates both sides in FP32 with the square roots pushed 1 ulp the wrong
way (the bound's inward, the lanes' outward) on random lines, parameter
grids and both grid directions, including the last, partial tile.  Shapes
that stage (sample, difference) pairs stage the same samples and read the W
pairs of samples e .. e + W - 1, a subset of the same window."""
import numpy as np

f32 = np.float32


def _fma(a, b, c):
    # float32 fused multiply-add through float64 (the product is exact)
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def _sqrt(x, direction):
    r = np.sqrt(np.maximum(x, f32(0))).astype(np.float32)
    return np.nextafter(r, np.float32(np.inf if direction > 0 else -np.inf)).astype(np.float32)


def _check(rng, ns, tau, c0, c1, nc, chunk, S, hmax):
    w = 2 * tau + 1
    nwin = ns - w
    hi = ns - tau - 1
    hi2 = f32(hi * hi * (1.0 - 2.0 ** -20))
    inc = f32((f32(c1) - f32(c0)) * f32(1.0 / nc))
    h = np.sort(rng.uniform(0, hmax, S).astype(np.float32))
    lanes = np.arange(64)
    bad = 0
    for tile in range(0, (ns + 63) // 64):
        t0min = tile * 64
        T0 = (t0min + lanes).astype(np.float32)
        T0sq = (T0 * T0).astype(np.float32)
        sqlo = f32(f32(t0min) * f32(t0min))
        sqhi = f32(f32(t0min + 63) * f32(t0min + 63))
        prefix = t0min > tau
        for q0 in range(0, nc, chunk):
            q1 = min(q0 + chunk, nc)
            cA = f32(f32(c0) + f32(inc * f32(q0)))
            cB = f32(f32(c0) + f32(inc * f32(q1 - 1)))
            cmin, cmax = min(cA, cB), max(cA, cB)
            q = np.arange(q0, q1)
            pc = (f32(c0) + (inc * q.astype(np.float32)).astype(np.float32)).astype(np.float32)
            # the kernel's per-trace range, square roots pushed inward
            elo = np.maximum(_sqrt(_fma(np.full(S, cmin, np.float32), h, np.full(S, sqlo, np.float32)), +1)
                             .astype(np.int64) - tau - 2, 0)
            ehi = _sqrt(_fma(np.full(S, cmax, np.float32), h, np.full(S, sqhi, np.float32)), -1) \
                .astype(np.int64) - tau + 2
            first = (elo >> 2) * 4                       # first staged sample
            last = ((ehi + w) >> 2) * 4 + 3              # last staged sample
            # every lane x parameter x trace, square roots pushed outward
            tt = _fma(pc[None, :, None], h[None, None, :], T0sq[:, None, None])
            for d in (-1, +1):
                e = _sqrt(tt, d).astype(np.int64) - tau
                if prefix:
                    read = np.ones_like(e, dtype=bool)
                else:
                    read = (e >= 0) & (e < nwin) & (tt < hi2)
                out = read & ((e < first[None, None, :]) | (e + w > last[None, None, :]) |
                              (ehi[None, None, :] < elo[None, None, :]))
                bad += int(out.sum())
    return bad


def test_lds_staged_ranges_cover_every_window():
    rng = np.random.default_rng(2024)
    cases = [
        # ns, tau, c0, c1, nc, chunk, S, hmax (h in samples^2 per unit of c)
        (1501, 1, 0.198, 1.77, 256, 32, 120, 5.6e5),   # the headline line's scales
        (733, 0, 0.198, 1.77, 57, 57, 150, 5.6e5),
        (733, 3, 1.77, 0.198, 57, 32, 90, 5.6e5),      # decreasing grid
        (301, 2, 0.1, 3.0, 72, 40, 60, 2.0e6),          # far offsets, short traces
        (2001, 3, 0.1, 1.5, 200, 64, 40, 1.6e6),
    ]
    for case in cases:
        assert _check(rng, *case) == 0, case


def test_staging_cases_reached_by_the_gpu_tests(sb):
    """The staging decision replayed on the CPU for the gathers and launch
    shapes of tests/test_gpu_lds.py: between them the GPU tests run
    workgroups that stage every trace, workgroups that stage only their first
    traces and workgroups that stage none (K = 0, every window from global
    memory; the far-offset case)."""
    import lds_replay as lr
    seen = dict(full=0, partial=0, none=0, dead=0)
    # test_lds_partial_staging: headline geometry, 128 and 256 parameters per workgroup, 40-64 KiB
    # of sample slabs and 40-64 KiB of (sample, difference) slabs
    off, geom, _ = sb.synth(0, 5, (60, 120), 1501, 2000, 31, h_max=1500.0, h_jitter=5.0, nevents=8)
    paired = dict(full=0, partial=0, none=0, dead=0)
    for kb, qb in ((40, 16), (48, 16), (52, 16), (64, 16), (64, 32), (52, 32), (40, 32)):
        for pc in (128, 256):
            c = lr.staging_cases(off, geom, 2000, 1501, 0.002, 1.98e-7, 1.77e-6, 256, pc, kb, quad_bytes=qb)
            for k in seen:
                (seen if qb == 16 else paired)[k] += c[k]
    part = dict(seen)
    # test_lds_far_offsets_stage_nothing
    off, geom, _ = lr.far_offset_line()
    c = lr.staging_cases(off, geom, 500, 20001, 0.001, 1e-8, 5e-6, 16, 16, 40)
    print("partial-staging gathers: %s, paired slabs %s; far-offset gather: %s" % (part, paired, c))
    assert part["full"] > 0 and part["partial"] > 0
    assert paired["full"] > 0 and paired["partial"] > 0
    assert c["none"] > 0


def test_mixed_split_fits_and_is_maximal():
    """The mixed-slab decision (pairs for the first Kp traces, samples for
    traces Kp .. K - 1) on random quad counts: the slabs never exceed the
    capacity, Kp <= K, every trace is staged when its samples fit, every
    trace is paired when its pairs fit, and K is as large as the bytes left
    after the pairs allow."""
    import lds_replay as lr
    rng = np.random.default_rng(7)
    for _ in range(3000):
        S = int(rng.integers(1, 60))
        length = rng.integers(0, 40, S)
        pre = np.concatenate([[0], np.cumsum(length)])
        cap16 = int(rng.integers(0, 3 * pre[S] + 2))
        share = int(rng.choice([4, 8, 16]))
        Kp, K = lr.mixed_split(pre, cap16, share)
        assert 0 <= Kp <= K <= S
        assert 32 * pre[Kp] + 16 * (pre[K] - pre[Kp]) <= 16 * cap16
        if 2 * pre[S] <= cap16:
            assert Kp == S
        elif pre[S] <= cap16:
            assert K == S and (Kp == S or pre[Kp + 1] + pre[S] > cap16)
        else:
            assert 32 * pre[Kp] <= (share * cap16 >> 5) * 32
            assert K == S or pre[K + 1] + pre[Kp] > cap16


def test_mixed_slab_cases_reached_by_the_gpu_tests(sb):
    """For the mixed-slab shapes (52 KiB, pairs in at most 16/16 or 8/16 of
    the slab bytes), the gathers of tests/test_gpu_lds.py run workgroups
    that pair every trace, that stage every trace with pairs for the first
    ones, that stage pairs, samples and leave the rest in global memory, and
    that stage samples only."""
    import lds_replay as lr
    for share in (16, 8):
        seen = {}
        # test_lds_partial_staging: 128 and 256 parameters per workgroup
        off, geom, _ = sb.synth(0, 5, (60, 120), 1501, 2000, 31, h_max=1500.0, h_jitter=5.0, nevents=8)
        for pc in (128, 256):
            c = lr.staging_cases(off, geom, 2000, 1501, 0.002, 1.98e-7, 1.77e-6, 256, pc, 52, share=share)
            for k, v in c.items():
                seen[k] = seen.get(k, 0) + v
        # test_lds_ranges_beyond_lds: the automatic split (32 parameters on 256 CUs) and 200
        off, geom, _ = sb.synth(0, 3, (90, 120), 2001, 2000, 5, h_max=2500.0, nevents=5)
        for pc in (32, 200):
            c = lr.staging_cases(off, geom, 2000, 2001, 0.006, 1.0e-7, 1.5e-6, 200, pc, 52, share=share)
            for k, v in c.items():
                seen[k] = seen.get(k, 0) + v
        print("mixed slabs, share %d/16: %s" % (share, seen))
        for k in ("pairs_all", "mixed_full", "mixed_partial", "pairs_none"):
            assert seen[k] > 0, (share, k)
