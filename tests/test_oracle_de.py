"""Pins the CRS-DE twin (oracle/oracle.c, orc_crs_de) beyond "it improves".

The reference ships no DE code (README.md:82-131 only names crs-ocl-de and its
flags), so the twin's DE/rand/1/bin (Storn & Price 1997) is the specification
and the GPU kernel is checked against it bit for bit.  Each test below fixes a
setting where the algorithm's behaviour follows from its definition alone,
and reads the twin's whole trajectory at one (CDP, t0) (orc_de_history):

  F = 0, CR = 1   every trial is a copy of another member (x_r1, r1 != i),
                  so the best fitness never changes
  CR = 0          every trial differs from its target in exactly one
                  coordinate, the forced one (jr)
  pop = 4, CR = 1 every trial is x_a + F (x_b - x_c) for a permutation
                  (a, b, c) of the three other members of the generation's
                  population (synchronous update), clipped to the box
  F = 2           mutants that leave the box are clipped onto its faces
  all fitness 0   a trial replaces its target on a tie (">=" selection)
  noise-free      DE with pop 32 and 30 generations reaches the 11^3 grid's
  single event    best semblance (minus 1e-3) on at least 95 % of the
                  event's (CDP, t0) cells

Mutations of the twin each of these fails (checked by planting them in a copy
of oracle.c on the CPU): the target x_i as base vector (F = 0 test), an
inverted crossover test u >= CR (CR = 0 test), no forced coordinate (CR = 0
test), r1/r2/r3 drawn with the target included or not distinct (pop = 4
test), no clipping (F = 2 test), strict ">" selection (tie test), in-place
(asynchronous) generation updates (pop = 4 test).
"""
import itertools

import numpy as np
import pytest

BOX = (-0.7e-3, 0.7e-3, -1e-7, 1e-7, 1.975e-7, 1.384e-6)
LO = np.array([BOX[0], BOX[2], BOX[4] * 1e6], np.float32)
HI = np.array([BOX[1], BOX[3], BOX[5] * 1e6], np.float32)
TAU = 0.004
NS = 151


@pytest.fixture(scope="module")
def line(sb):
    cdp_off, geom, samples = sb.synth(1, 4, (10, 14), NS, 4000, seed=31, h_max=600.0, mid_half_width=20.0,
                                      nevents=3, snr=4.0)
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 30)
    return cdp_off, geom, samples, nbr_off, nbr


def _hist(orc, line, pop, ngen, F, CR, seed, cdp, t0):
    return orc.de_history(*line[:3], 4000, *line[3:], *BOX, TAU, pop, ngen, F, CR, seed, cdp, t0)


CELLS = [(0, 40), (1, 75), (2, 100), (3, 60)]


def test_history_reports_what_the_search_reports(orc, line):
    out = orc.crs_de(*line[:3], 4000, *line[3:], *BOX, TAU, 12, 5, 0.5, 0.9, 3)
    for cdp, t0 in CELLS:
        h = _hist(orc, line, 12, 5, 0.5, 0.9, 3, cdp, t0)
        fin_x, fin_f = h["pop_x"][-1], h["pop_f"][-1]
        b = int(np.argmax(fin_f))  # the first member of maximal fitness
        assert h["best"][3] == fin_f[b] == out["coher"][cdp, t0]
        assert np.array_equal(h["best"][:3], fin_x[b])
        assert out["a"][cdp, t0] == fin_x[b, 0] and out["b"][cdp, t0] == fin_x[b, 1]
        assert out["c"][cdp, t0] == np.float32(np.float64(fin_x[b, 2]) / 1e6)
        # the initial population is uniform in the box, from the shared hash
        u = np.array([[orc.de_hash(3, cdp, t0, 0, i, d) >> 8 for d in range(3)] for i in range(12)], np.float32)
        assert np.all((h["pop_x"][0] >= LO) & (h["pop_x"][0] <= HI))
        assert np.array_equal(h["pop_x"][0], LO + (u * np.float32(1.0 / 16777216.0)) * (HI - LO))


def test_f0_cr1_trials_copy_another_member(orc, line):
    """F = 0, CR = 1: the mutant x_r1 + 0 (x_r2 - x_r3) is x_r1 in every
    coordinate, r1 != i: each trial is another member's vector, so the
    population only ever holds copies of its initial members and the best
    fitness is the initial best exactly."""
    for cdp, t0 in CELLS:
        h = _hist(orc, line, 10, 6, 0.0, 1.0, 5, cdp, t0)
        for g in range(6):
            P, T = h["pop_x"][g], h["trial_x"][g]
            for i in range(10):
                same = np.all(P == T[i], axis=1)
                assert same.any(), (cdp, t0, g, i)
                assert same[h["picks"][g, i, 0]] and h["picks"][g, i, 0] != i
                assert not np.all(P[i] == T[i]) or np.sum(same) > 1  # a copy of another member, not of itself
        assert h["pop_f"][-1].max() == h["pop_f"][0].max()
    out0 = orc.crs_de(*line[:3], 4000, *line[3:], *BOX, TAU, 10, 0, 0.0, 1.0, 5)
    outn = orc.crs_de(*line[:3], 4000, *line[3:], *BOX, TAU, 10, 9, 0.0, 1.0, 5)
    assert np.array_equal(out0["coher"], outn["coher"])


def test_cr0_changes_exactly_the_forced_coordinate(orc, line):
    changed = np.zeros(3, int)
    total = single = 0
    for cdp, t0 in CELLS:
        h = _hist(orc, line, 16, 5, 0.5, 0.0, 9, cdp, t0)
        for g in range(5):
            d = h["trial_x"][g] != h["pop_x"][g]  # [pop][3]
            n = d.sum(axis=1)
            assert n.max() <= 1, (cdp, t0, g)
            total += n.size
            single += int((n == 1).sum())
            for i in np.nonzero(n == 1)[0]:
                k = int(np.nonzero(d[i])[0][0])
                assert k == h["picks"][g, i, 3]
                changed[k] += 1
    assert single >= 0.95 * total  # the rest: a mutant that equals the target (clipped onto the same face)
    assert np.all(changed > 0.2 * single)  # jr covers the three coordinates


def test_pop4_mutant_uses_the_three_other_members(orc, line):
    """pop = 4: r1, r2, r3 are distinct and differ from i, so they are the
    other three members in some order, taken from the generation's population
    (not one already updated by this generation's selection)."""
    F = np.float32(0.5)
    for cdp, t0 in CELLS:
        h = _hist(orc, line, 4, 8, 0.5, 1.0, 13, cdp, t0)
        for g in range(8):
            P = h["pop_x"][g]
            for i in range(4):
                others = [j for j in range(4) if j != i]
                match = []
                for a, b, c in itertools.permutations(others):
                    v = np.minimum(np.maximum(P[a] + F * (P[b] - P[c]), LO), HI)
                    if np.array_equal(v, h["trial_x"][g, i]):
                        match.append((a, b, c))
                assert match, (cdp, t0, g, i)
                assert tuple(h["picks"][g, i, :3]) in match


def test_f2_mutants_are_clipped_to_the_box(orc, line):
    on_face = 0
    for cdp, t0 in CELLS:
        h = _hist(orc, line, 16, 6, 2.0, 1.0, 17, cdp, t0)
        T = h["trial_x"]
        assert np.all((T >= LO) & (T <= HI))
        on_face += int(np.sum((T == LO) | (T == HI)))
    assert on_face > 50  # F = 2 steps leave the box often: those are clipped onto a face


def test_ties_replace_the_target(orc, line):
    """At t0 = ns - 2 every window runs off the trace, every fitness is 0 and
    every trial ties with its target: ">=" selection takes the trial."""
    for cdp in range(4):
        h = _hist(orc, line, 8, 4, 0.5, 0.9, 19, cdp, NS - 2)
        assert np.all(h["pop_f"] == 0) and np.all(h["trial_f"] == 0)
        for g in range(4):
            assert np.array_equal(h["pop_x"][g + 1], h["trial_x"][g])
            assert not np.array_equal(h["pop_x"][g + 1], h["pop_x"][g])


def test_selection_keeps_the_better_of_target_and_trial(orc, line):
    for cdp, t0 in CELLS:
        h = _hist(orc, line, 24, 8, 0.5, 0.9, 23, cdp, t0)
        for g in range(8):
            better = h["trial_f"][g] >= h["pop_f"][g]
            assert np.array_equal(h["pop_f"][g + 1], np.where(better, h["trial_f"][g], h["pop_f"][g]))
            assert np.array_equal(h["pop_x"][g + 1], np.where(better[:, None], h["trial_x"][g], h["pop_x"][g]))
        assert np.all(np.diff(h["pop_f"], axis=0) >= 0)


def test_de_reaches_the_grid_optimum_on_a_single_event(orc, sb):
    """Noise-free line with one CRS event: on every cell where the 11^3 grid
    finds the event (semblance >= 0.5), DE with population 32 and 30
    generations finds at least the grid's semblance minus 1e-3 on >= 95 % of
    them (the grid's 1331 evaluations against DE's 992)."""
    ok = total = 0
    for seed in (1, 2, 3):
        cdp_off, geom, samples = sb.synth(1, 5, (40, 60), 301, 4000, seed=seed, h_max=1200.0, mid_half_width=50.0,
                                          nevents=1, snr=0.0, aph=600)
        nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], 50)
        grid = orc.crs(cdp_off, geom, samples, 4000, nbr_off, nbr, BOX[0], BOX[1], 11, BOX[2], BOX[3], 11, BOX[4],
                       BOX[5], 11, TAU)
        de = orc.crs_de(cdp_off, geom, samples, 4000, nbr_off, nbr, *BOX, TAU, 32, 30, 0.5, 0.9, 5)
        ev = grid["coher"] >= 0.5
        total += int(ev.sum())
        ok += int(np.sum(de["coher"][ev] >= grid["coher"][ev] - 1e-3))
    print("DE >= grid - 1e-3 on %d of %d event cells" % (ok, total))
    assert total > 50 and ok >= 0.95 * total
