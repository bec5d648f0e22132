"""HIP path (through the C ABI) against the reference programs' output files
and the oracle, on the golden fixtures.

SB_MODE_EXACT must reproduce the reference output files byte for byte.
SB_MODE_FAST must pick the reference's parameter except where the top two
semblances of the oracle panel are within 1e-5 (relative), and then report
the reference's coherence and stack bit for bit (the winner is re-evaluated in
the reference's precision).
"""
import filecmp
import os

import numpy as np
import pytest

from conftest import GOLD, cmp_args, crs_args

pytestmark = pytest.mark.gpu

CMP_CASES = ["cmp_w3", "cmp_uneven_w5", "cmp_aph_shuffled", "cmp_scalco_pos_w1"]
CRS_CASES = ["crs_w3", "crs_w5"]


def _print_case(*a):
    print("case:", *a, flush=True)


def _cmp_gather(sb, case, p):
    return sb.SuGather(os.path.join(GOLD, case, "input.su"), sb.METHOD_CMP, p["aph"])


def _crs_gather(sb, case, p):
    return sb.SuGather(os.path.join(GOLD, case, "input.su"), sb.METHOD_CRS, p["aph"], p["apm"])


def near_tie(row, pick, tol=1e-5):
    """Top-two semblances (values in [0, 1]) within tol of each other."""
    return pick >= 0 and float(np.max(row)) - float(row[pick]) <= tol


def explained(row32, row64, ref_pick, our_pick, tol=1e-5):
    """A differing pick is a near-tie of the reference's own values, or the
    FP64 oracle ranks it at least as high as the reference's pick (the
    reference's FP32 rounding decided)."""
    if our_pick < 0 or ref_pick < 0:
        return False
    return near_tie(row32, our_pick, tol) or row64[our_pick] >= row64[ref_pick] - 1e-9


@pytest.mark.parametrize("case", CMP_CASES)
def test_cmp_exact_reproduces_reference_files(sb, gpu_ctx, manifest, tmp_path, case):
    _print_case("cmp exact", case)
    p = cmp_args(manifest[case])
    G = _cmp_gather(sb, case, p)
    dg = sb.Gather(gpu_ctx, G.cdp_off, G.geom, G.samples, G.dt_us)
    out = sb.search(gpu_ctx, dg, sb.cmp_params(p["c0"], p["c1"], p["nc"], p["tau"], sb.MODE_EXACT))
    dg.close()
    for key, fname in (("c", "cmp.c.su"), ("coher", "cmp.coher.su"), ("stack", "cmp.stack.su")):
        G.write_section(str(tmp_path / fname), out[key], key == "c")
        assert filecmp.cmp(str(tmp_path / fname), os.path.join(GOLD, case, fname), shallow=False), (case, fname)


@pytest.mark.parametrize("case", CRS_CASES)
def test_crs_exact_reproduces_reference_files(sb, gpu_ctx, manifest, tmp_path, case):
    _print_case("crs exact", case)
    p = crs_args(manifest[case])
    G = _crs_gather(sb, case, p)
    dg = sb.Gather(gpu_ctx, G.cdp_off, G.geom, G.samples, G.dt_us, nbr_off=G.nbr_off, nbr=G.nbr)
    prm = sb.crs_params(p["a0"], p["a1"], p["na"], p["b0"], p["b1"], p["nb"], p["c0"], p["c1"], p["nc"], p["tau"],
                        sb.MODE_EXACT)
    out = sb.search(gpu_ctx, dg, prm)
    dg.close()
    for key, fname in (("a", "crs.a.su"), ("b", "crs.b.su"), ("c", "crs.c.su"), ("coher", "crs.coher.su"),
                       ("stack", "crs.stack.su")):
        G.write_section(str(tmp_path / fname), out[key], key in ("a", "b", "c"))
        assert filecmp.cmp(str(tmp_path / fname), os.path.join(GOLD, case, fname), shallow=False), (case, fname)


@pytest.mark.parametrize("case", CMP_CASES)
def test_cmp_fast_matches_reference(sb, orc, gpu_ctx, manifest, case):
    _print_case("cmp fast", case)
    p = cmp_args(manifest[case])
    G = _cmp_gather(sb, case, p)
    ref = orc.cmp(G.cdp_off, G.geom, G.samples, G.dt_us, p["c0"], p["c1"], p["nc"], p["tau"], panel=True)
    ref64 = orc.cmp_f64(G.cdp_off, G.geom, G.samples, G.dt_us, p["c0"], p["c1"], p["nc"], p["tau"], panel=True)
    dg = sb.Gather(gpu_ctx, G.cdp_off, G.geom, G.samples, G.dt_us)
    out = sb.search(gpu_ctx, dg, sb.cmp_params(p["c0"], p["c1"], p["nc"], p["tau"], sb.MODE_FAST))
    dg.close()
    diff = np.argwhere(out["index"] != ref["index"])
    for cdp, t0 in diff:
        assert explained(ref["panel_sem"][cdp, t0], ref64["panel_sem"][cdp, t0], ref["index"][cdp, t0],
                         out["index"][cdp, t0]), (case, cdp, t0)
    same = out["index"] == ref["index"]
    for k in ("c", "coher", "stack"):
        assert np.array_equal(out[k][same], ref[k][same]), (case, k)


@pytest.mark.parametrize("case", CRS_CASES)
def test_crs_fast_matches_reference(sb, orc, gpu_ctx, manifest, case):
    _print_case("crs fast", case)
    p = crs_args(manifest[case])
    G = _crs_gather(sb, case, p)
    ref = orc.crs(G.cdp_off, G.geom, G.samples, G.dt_us, G.nbr_off, G.nbr, p["a0"], p["a1"], p["na"], p["b0"],
                  p["b1"], p["nb"], p["c0"], p["c1"], p["nc"], p["tau"], panel=True)
    dg = sb.Gather(gpu_ctx, G.cdp_off, G.geom, G.samples, G.dt_us, nbr_off=G.nbr_off, nbr=G.nbr)
    prm = sb.crs_params(p["a0"], p["a1"], p["na"], p["b0"], p["b1"], p["nb"], p["c0"], p["c1"], p["nc"], p["tau"],
                        sb.MODE_FAST)
    out = sb.search(gpu_ctx, dg, prm)
    dg.close()
    # CRS reference values are double: a differing pick must be a near-tie of them.
    diff = np.argwhere(out["index"] != ref["index"])
    for cdp, t0 in diff:
        assert near_tie(ref["panel_sem"][cdp, t0], out["index"][cdp, t0]), (case, cdp, t0)
    same = out["index"] == ref["index"]
    for k in ("a", "b", "c", "coher", "stack"):
        assert np.array_equal(out[k][same], ref[k][same]), (case, k)
