"""The oracle and the host-side SU path against the reference programs' own
outputs (tests/golden, written by oracle/_ref/{cmp,crs}-omp2).

Pins, byte for byte: the SU reader, aph filter, CDP binning and order
(CMP fold-descending std::sort, CRS key order + apm neighbour lists), the
output header rewrite, and the oracle's arithmetic (CMP float, CRS double).
"""
import filecmp
import os

import numpy as np
import pytest

from conftest import GOLD, cmp_args, crs_args

CMP_CASES = ["cmp_w3", "cmp_uneven_w5", "cmp_aph_shuffled", "cmp_scalco_pos_w1"]
CRS_CASES = ["crs_w3", "crs_w5"]


def _cmp_outputs(orc, G, p):
    return orc.cmp(G.cdp_off, G.geom, G.samples, G.dt_us, p["c0"], p["c1"], p["nc"], p["tau"])


@pytest.mark.parametrize("case", CMP_CASES)
def test_cmp_oracle_matches_reference_files(sb, orc, manifest, tmp_path, case):
    p = cmp_args(manifest[case])
    G = sb.SuGather(os.path.join(GOLD, case, "input.su"), sb.METHOD_CMP, p["aph"])
    out = _cmp_outputs(orc, G, p)
    for key, fname in (("c", "cmp.c.su"), ("coher", "cmp.coher.su"), ("stack", "cmp.stack.su")):
        G.write_section(str(tmp_path / fname), out[key], key == "c")
        assert filecmp.cmp(str(tmp_path / fname), os.path.join(GOLD, case, fname), shallow=False), (case, fname)


@pytest.mark.parametrize("case", CRS_CASES)
def test_crs_oracle_matches_reference_files(sb, orc, manifest, tmp_path, case):
    p = crs_args(manifest[case])
    G = sb.SuGather(os.path.join(GOLD, case, "input.su"), sb.METHOD_CRS, p["aph"], p["apm"])
    out = orc.crs(G.cdp_off, G.geom, G.samples, G.dt_us, G.nbr_off, G.nbr, p["a0"], p["a1"], p["na"], p["b0"],
                  p["b1"], p["nb"], p["c0"], p["c1"], p["nc"], p["tau"])
    for key, fname in (("a", "crs.a.su"), ("b", "crs.b.su"), ("c", "crs.c.su"), ("coher", "crs.coher.su"),
                       ("stack", "crs.stack.su")):
        G.write_section(str(tmp_path / fname), out[key], key in ("a", "b", "c"))
        assert filecmp.cmp(str(tmp_path / fname), os.path.join(GOLD, case, fname), shallow=False), (case, fname)


def test_golden_cases_exercise_the_edges(sb, orc, manifest):
    """The fixtures cover what they claim: repeated folds, dropped traces,
    gaps in neighbour lists, invalid (-1 / 0) outputs and a tie-free arg-max."""
    p = cmp_args(manifest["cmp_uneven_w5"])
    G = sb.SuGather(os.path.join(GOLD, "cmp_uneven_w5", "input.su"), sb.METHOD_CMP, p["aph"])
    folds = np.diff(G.cdp_off)
    assert len(set(folds.tolist())) < len(folds), "needs CDPs of equal fold"
    assert np.all(folds[:-1] >= folds[1:])
    hdr, _ = orc.read_su(os.path.join(GOLD, "cmp_aph_shuffled", "input.su"))
    G2 = sb.SuGather(os.path.join(GOLD, "cmp_aph_shuffled", "input.su"), sb.METHOD_CMP, 650)
    assert G2.ntr < len(hdr), "aph must drop traces"
    q = crs_args(manifest["crs_w3"])
    G3 = sb.SuGather(os.path.join(GOLD, "crs_w3", "input.su"), sb.METHOD_CRS, q["aph"], q["apm"])
    gaps = 0
    for i in range(G3.ncdp):
        lst = G3.nbr[G3.nbr_off[i]:G3.nbr_off[i + 1]]
        gaps += int(len(lst) and (lst[-1] - lst[0] + 1) != len(lst))
        assert i in lst.tolist()
    assert gaps > 0, "needs non-contiguous neighbour lists"
    _, coh = orc.read_su(os.path.join(GOLD, "cmp_w3", "cmp.coher.su"))
    assert (coh == 0).any() and (coh > 0).any()
