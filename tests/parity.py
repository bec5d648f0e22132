"""The fast-mode parity bar, shared by the GPU parity tests.

A fast search (FP32, contracted, hardware square root) picks the reference's
parameter at every (CDP, t0) except where the reference's own top two
semblances are within 1e-5 of each other (semblance lies in [0, 1]; the gap is
absolute): there the two orders of rounding may rank them either way.  The
winner is re-evaluated in the reference's precision and order, so at every
pick -- agreeing or not -- the reported coherence and stack are the
reference-precision oracle's values for that parameter, bit for bit.

One more case is admitted, counted and capped: the FP64 oracle ranks our pick
at least as high as the reference's (the reference's FP32 traveltime rounding,
about 1e-4 samples at late t0, moves its semblances by a few 1e-5 and decides
the pick).  At most `cap` (default 1e-3) of all picks may be such cases; the
count is printed.
"""
import numpy as np

TIE = 1e-5


def check_fast_picks(label, ref, mine, ref64=None, tol=TIE, cap=1e-3, params=("c",)):
    """ref: oracle outputs with panel=True (panel_sem, panel_stk [ncdp][ns][npar]);
    mine: the search's outputs (index, coher, stack and the parameter arrays).
    Returns (number of differing picks, number of FP64-ranked exceptions)."""
    ri, mi = ref["index"], mine["index"]
    diff = np.argwhere(mi != ri)
    gaps, fp64_cases = [], 0
    for cdp, t0 in diff:
        row = ref["panel_sem"][cdp, t0]
        p, q = int(mi[cdp, t0]), int(ri[cdp, t0])
        top = float(np.max(row))
        if p < 0:  # nothing beat 0 in our search: the reference's best must be a near-tie of 0
            assert top <= tol, (label, cdp, t0, top)
            continue
        gap = top - float(row[p])
        gaps.append(gap)
        # the reported values are the reference-precision values of our pick
        assert mine["coher"][cdp, t0] == row[p], (label, cdp, t0, mine["coher"][cdp, t0], row[p])
        assert mine["stack"][cdp, t0] == ref["panel_stk"][cdp, t0, p], (label, cdp, t0)
        if gap <= tol:
            continue
        assert ref64 is not None and q >= 0, (label, "pick differs beyond a near-tie", cdp, t0, p, q, gap)
        row64 = ref64["panel_sem"][cdp, t0]
        assert row64[p] >= row64[q] - 1e-9, (label, "FP64 ranks the reference's pick higher", cdp, t0, gap)
        fp64_cases += 1
    print("%s: %d of %d picks differ (top-two gaps %s), %d decided by the reference's FP32 rounding (cap %d)" %
          (label, len(diff), ri.size, np.round(sorted(gaps, reverse=True)[:8], 9).tolist(), fp64_cases,
           int(cap * ri.size)), flush=True)
    assert fp64_cases <= cap * ri.size, (label, fp64_cases)
    same = mi == ri
    for k in ("coher", "stack") + tuple(params):
        assert np.array_equal(mine[k][same], ref[k][same]), (label, k)
    return len(diff), fp64_cases
