"""HIP path (through the C ABI) against the reference programs' output files
and the oracle, on the golden fixtures.

SB_MODE_EXACT must reproduce the reference output files byte for byte.
SB_MODE_FAST must pick the reference's parameter except where the reference's
top two semblances are within 1e-5, and report the reference-precision
coherence and stack of its pick bit for bit (the winner is re-evaluated in the
reference's precision); tests/parity.py states the bar.
"""
import filecmp
import os

import numpy as np
import pytest

from conftest import GOLD, cmp_args, crs_args
from parity import check_fast_picks

pytestmark = pytest.mark.gpu

CMP_CASES = ["cmp_w3", "cmp_uneven_w5", "cmp_aph_shuffled", "cmp_scalco_pos_w1"]
CRS_CASES = ["crs_w3", "crs_w5"]


def _print_case(*a):
    print("case:", *a, flush=True)


def _cmp_gather(sb, case, p):
    return sb.SuGather(os.path.join(GOLD, case, "input.su"), sb.METHOD_CMP, p["aph"])


def _crs_gather(sb, case, p):
    return sb.SuGather(os.path.join(GOLD, case, "input.su"), sb.METHOD_CRS, p["aph"], p["apm"])


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
    check_fast_picks("cmp fast " + case, ref, out, ref64, params=("c",))


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
    # CRS reference values are double: a differing pick must be a near-tie of them (no FP64 exception).
    check_fast_picks("crs fast " + case, ref, out, None, params=("a", "b", "c"))
