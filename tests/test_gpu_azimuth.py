"""The azimuth extension (SB_OPT_DM_AZIMUTH, crs-de -azimuth with
SB_AZIMUTH=project): CRS supergathers whose midpoint displacement is the
signed projection onto an azimuth.

  - SB_MODE_EXACT CRS grid: bit-identical to the oracle's projected search
    (orc_crs_range_az, pinned on the CPU by a closed form);
  - SB_MODE_FAST CRS grid: the parity bar of tests/parity.py against it;
  - CRS-DE on a noise-free line of one dipping event (the synthetic's events
    use the signed inline displacement): projecting onto the line direction
    finds at least the coherence the unsigned distance does.
"""
import numpy as np
import pytest

from parity import check_fast_picks

pytestmark = pytest.mark.gpu

BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)


def _line(sb, snr=3.0, nevents=4, seed=61):
    cdp_off, geom, samples = sb.synth(1, 8, (40, 90), 401, 4000, seed, h_max=900.0, mid_half_width=40.0,
                                      nevents=nevents, snr=snr)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 40)
    return cdp_off, geom, samples, nbr_off, nbr


def _gather(sb, az_mdeg, line, store=0):
    ctx = sb.Context(0)
    ctx.set_option(sb.OPT_DM_AZIMUTH, az_mdeg)
    g = sb.Gather(ctx, *line[:3], 4000, store=store, nbr_off=line[3], nbr=line[4])
    return ctx, g


@pytest.mark.parametrize("az", [90000, 33000])
def test_crs_grid_exact_equals_oracle_projection(sb, orc, az):
    print("case: crs grid exact, azimuth", az, flush=True)
    line = _line(sb)
    ctx, g = _gather(sb, az, line)
    try:
        assert ctx.get_option(sb.OPT_DM_AZIMUTH) == az
        prm = sb.crs_params(BOX[0], BOX[1], 5, BOX[2], BOX[3], 3, BOX[4], BOX[5], 4, 0.004, sb.MODE_EXACT)
        out = sb.search(ctx, g, prm)
        fast = sb.search(ctx, g, sb.crs_params(BOX[0], BOX[1], 5, BOX[2], BOX[3], 3, BOX[4], BOX[5], 4, 0.004))
    finally:
        g.close()
        ctx.close()
    ref = orc.crs(*line[:3], 4000, line[3], line[4], BOX[0], BOX[1], 5, BOX[2], BOX[3], 3, BOX[4], BOX[5], 4, 0.004,
                  panel=True, az_mdeg=az)
    for k in ("index", "a", "b", "c", "coher", "stack"):
        assert np.array_equal(out[k], ref[k]), k
    check_fast_picks("crs grid fast, azimuth %d mdeg" % az, ref, fast, None, params=("a", "b", "c"))
    dist = orc.crs(*line[:3], 4000, line[3], line[4], BOX[0], BOX[1], 5, BOX[2], BOX[3], 3, BOX[4], BOX[5], 4, 0.004)
    assert not np.array_equal(dist["coher"], ref["coher"])  # the projection changed the search


def test_crs_de_projection_fits_a_dipping_event(sb):
    print("case: crs-de azimuth projection on a dipping event", flush=True)
    line = _line(sb, snr=0.0, nevents=1, seed=62)
    prm = sb.de_params(*BOX, 0.004, pop=32, ngen=30, seed=3)
    fit = {}
    for az in (-1, 90000):
        ctx, g = _gather(sb, az, line)
        try:
            fit[az] = sb.search(ctx, g, prm)["coher"]
        finally:
            g.close()
            ctx.close()
    ev = np.maximum(fit[-1], fit[90000]) >= 0.5
    print("event cells %d: mean fitness unsigned distance %.4f, projected onto the line %.4f" %
          (ev.sum(), fit[-1][ev].mean(), fit[90000][ev].mean()), flush=True)
    assert ev.sum() > 50
    assert fit[90000][ev].mean() >= fit[-1][ev].mean() - 1e-3
