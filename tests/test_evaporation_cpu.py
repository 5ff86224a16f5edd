"""Solvent evaporation, host side (no GPU): the plan, its refusals and both ladders against the
reference's own known answers (tests/soup-evaporation.test.ts:39-140); the insertion sampler
(the library's soup_sample_insertion) against the oracle's loop restatement bit for bit and both
against a brute-force replay of the candidate stream; the returning beads' velocity and RNG draws;
and createSoup's resolution of an evaporating cycle and of a live count below the census."""
import math

import numpy as np
import pytest

import protocell_soup as ps
from oracle import evaporate as oev
from protocell_soup.box_scale import computeDryBox, deriveCycleConfig
from protocell_soup.evaporate import evaporationLadder, planEvaporation, rehydrationLadder, sampleInsertionPositions
from protocell_soup.init_state import Mulberry32
from protocell_soup.rules import loadParams, loadSoup, wcaCutoff

BOX30 = (30.0, 30.0, 30.0)
START30 = {"C": 1860, "O": 7440, "H": 1860, "M": 124, "W": 21600}
BOX24 = (24.0, 24.0, 24.0)
START24 = {"C": 952, "O": 3808, "H": 952, "M": 63, "W": 11059}


def test_plan_reference_known_answers():
    soup, p = loadSoup(), loadParams()
    dwc = soup["dryWetCycle"]
    plan = planEvaporation(soup, p, BOX30, START30)
    # (1) the residual fraction is 1/1400: 15 of 21,600 solvent beads stay
    assert dwc["residualSolventFraction"] == pytest.approx(1 / 1400, abs=1e-12)
    assert plan.wetSolventCount == 21600 and plan.drySolventCount == 15
    assert plan.organicCount == 1860 + 7440 + 1860 + 124 == plan.solventBlockStart
    assert plan.solventKind == [m["id"] for m in soup["monomers"]].index("W")
    # (2) the dry box realises targetDryDensity for what is left
    n_dry = plan.organicCount + plan.drySolventCount
    assert n_dry / float(np.prod(plan.dryBox)) == pytest.approx(dwc["targetDryDensity"], abs=1e-9)
    assert plan.dryBox[0] == pytest.approx(computeDryBox(BOX30, n_dry, dwc["targetDryDensity"])[0], abs=1e-12)
    # (3) the concentration factor between its two bounds
    rho_org = plan.organicCount / 27000.0
    ceiling = math.sqrt(2) / rho_org
    previous = dwc["targetDryDensity"] / (32884 / 27000)
    assert 3.1 < plan.concentrationFactor < ceiling
    assert plan.concentrationFactor / previous > 2.9
    # (4) the minimiser budget: a bead placed at the floor lifted to the stiffest solvent pair's WCA
    # radius plus the mean spacing at the wet density; the descent's displacement bound covers it
    stiffest = wcaCutoff(p["sigma"] * (1.0 + max(m["radiusSigma"] for m in soup["monomers"])) * 0.5)
    d0 = soup["coldStartRelax"]["maxDisplacementSigma"] * p["sigma"]
    need = stiffest - dwc["insertionMinSeparationSigma"] * p["sigma"] + float(np.cbrt(27000 / (plan.organicCount + 21600)))
    assert plan.relaxIterations == math.ceil(2 * need / d0)
    assert d0 * (plan.relaxIterations + 1) / 2 >= need
    # (5) the floor stays below random-sequential-addition saturation for the whole wet pool
    phi = (plan.organicCount + 21600) * (math.pi / 6) * dwc["insertionMinSeparationSigma"] ** 3 / 27000
    assert phi < 0.3841


def test_plan_refusals():
    soup, p = loadSoup(), loadParams()
    with pytest.raises(ValueError, match="last non-empty block"):
        planEvaporation(soup, p, BOX30, dict(START30, K=10))
    bad = dict(soup, dryWetCycle=dict(soup["dryWetCycle"], residualSolventFraction=0.999999))
    with pytest.raises(ValueError, match="must remove the solvent"):
        planEvaporation(bad, p, BOX30, START30)
    with pytest.raises(ValueError, match="contains no solvent"):
        planEvaporation(soup, p, BOX30, dict(START30, W=0))
    missing = dict(soup, dryWetCycle={k: v for k, v in soup["dryWetCycle"].items() if k != "evaporationRampSteps"})
    with pytest.raises(ValueError, match="does not contain its fields"):
        planEvaporation(missing, p, BOX30, START30)


def test_ladders_uniform_monotone_and_exact_at_the_ends():
    soup, p = loadSoup(), loadParams()
    dwc = soup["dryWetCycle"]
    plan = planEvaporation(soup, p, BOX30, START30)
    down = evaporationLadder(plan, dwc["targetDryDensity"])
    assert len(down) == dwc["evaporationRampSteps"]
    assert down[-1]["box"][0] == plan.dryBox[0] and down[-1]["solvent"] == plan.drySolventCount
    prev_box, prev_sol, ln_steps = BOX30[0], plan.wetSolventCount, []
    for rung in down:
        assert rung["box"][0] < prev_box and rung["solvent"] <= prev_sol
        ln_steps.append(math.log(prev_box / rung["box"][0]))
        prev_box, prev_sol = rung["box"][0], rung["solvent"]
    assert max(ln_steps) - min(ln_steps) < 1e-9
    assert 1 - math.exp(-max(ln_steps)) <= 0.032 + 1e-9
    # the solvent follows the density ramp: every intermediate rung holds rho(f) * V(box) beads
    rho_wet = (plan.organicCount + plan.wetSolventCount) / 27000.0
    for s, rung in enumerate(down[:-1], start=1):
        f = s / len(down)
        rho = rho_wet * (dwc["targetDryDensity"] / rho_wet) ** f
        n = (plan.organicCount + rung["solvent"]) / float(np.prod(rung["box"]))
        assert n == pytest.approx(rho, abs=1.0 / float(np.prod(rung["box"])))
    up = rehydrationLadder(plan)
    assert len(up) == dwc["evaporationRampSteps"]
    assert up[-1]["box"] == BOX30 and up[-1]["solvent"] == plan.wetSolventCount
    assert all(r["solvent"] == plan.drySolventCount for r in up[:-1])
    # the same boxes walked back: rung s of the rehydration is rung (R - s) of the evaporation
    for s in range(1, len(up)):
        assert up[s - 1]["box"][0] == pytest.approx(down[len(down) - 1 - s]["box"][0], rel=1e-12)


def _brute_replay(active, box, count, min_sep, state):
    """sampleInsertionPositions by brute force: every candidate of the stream scored against every
    point already present, no cell list. Within min_sep both searches see the same points (a cell is
    at least min_sep wide), so the chosen beads, the shortfall count and the stream agree."""
    rng = oev.Mulberry32(state)
    pts = [tuple(float(v) for v in active.reshape(-1, 4)[i, :3]) for i in range(active.size // 4)]
    out, short, worst = [], 0, math.inf
    L = np.asarray(box, np.float64)

    def dist(x):
        if not pts:
            return math.inf
        d = np.asarray(x)[None, :] - np.asarray(pts)
        r = np.floor(d / L)
        m = d - (r + ((d / L) - r >= 0.5)) * L
        return float(np.sqrt(((m[:, 0] * m[:, 0] + m[:, 1] * m[:, 1]) + m[:, 2] * m[:, 2]).min()))

    for _ in range(count):
        best, bd = (0.0, 0.0, 0.0), -1.0
        for _t in range(oev.INSERTION_CANDIDATES):
            x = (rng() * box[0], rng() * box[1], rng() * box[2])
            d = dist(x)
            if d > bd:
                bd, best = d, x
            if d >= min_sep:
                break
        if bd < min_sep:
            short += 1
            worst = min(worst, bd)
        out.append(best)
        pts.append(best)
    return np.asarray(out, np.float32), short, worst, rng.a


@pytest.mark.parametrize("n_active,count,box,min_sep,state", [
    (400, 60, (8.0, 9.5, 7.25), 0.75, 12345),     # the floor met almost everywhere
    (900, 120, (8.0, 8.0, 8.0), 0.9, 0xDEADBEEF),  # dense: candidates fall short, best-of-64 kicks in
    (30, 25, (2.0, 5.0, 1.6), 0.75, 7),           # fewer than 3 cells on two axes (repeated cell visits)
    (0, 10, (5.0, 5.0, 5.0), 0.75, 99),           # nothing present yet
    (50, 0, (5.0, 5.0, 5.0), 0.75, 1),            # nothing to place
])
def test_sample_insertion_matches_the_oracle_and_brute_force(n_active, count, box, min_sep, state):
    rng = np.random.default_rng(n_active + count)
    act = np.zeros((n_active, 4), np.float32)
    act[:, :3] = (rng.random((n_active, 3)) * np.asarray(box)).astype(np.float32)
    act[:, 3] = 4
    got = sampleInsertionPositions(act, n_active, box, count, min_sep, state)
    orng = oev.Mulberry32(state)
    exp = oev.sample_insertion_positions(act, n_active, box, count, min_sep, orng)
    assert np.array_equal(got["positions"], exp["positions"])
    assert got["shortOfFloor"] == exp["shortOfFloor"] and got["minAchieved"] == exp["minAchieved"]
    assert got["rng"] == orng.a
    pos, short, worst, a = _brute_replay(act, box, count, min_sep, state)
    assert np.array_equal(got["positions"].reshape(-1, 4)[:, :3], pos.reshape(-1, 3))
    assert got["shortOfFloor"] == short and got["rng"] == a
    if short:
        assert got["minAchieved"] == pytest.approx(worst, rel=0, abs=0)
    if n_active == 900:
        assert short > 0, "the dense case must exercise the best-of-64 fallback"
    if count:
        # every returned coordinate lies in the box
        p = got["positions"].reshape(-1, 4)[:, :3]
        assert (p >= 0).all() and (p <= np.asarray(box, np.float32)).all()


def test_sample_insertion_rejects_bad_input():
    act = np.zeros(8, np.float32)
    with pytest.raises(ValueError):
        sampleInsertionPositions(act, 2, (5.0, 5.0, 5.0), 3, 0.0, 1)
    with pytest.raises(ValueError):
        sampleInsertionPositions(act, 2, (5.0, -1.0, 5.0), 3, 0.75, 1)
    act[1] = np.nan
    with pytest.raises(ValueError):
        sampleInsertionPositions(act, 2, (5.0, 5.0, 5.0), 3, 0.75, 1)


def test_returning_bead_draws_continue_the_sampler_stream():
    # the oracle's statement-by-statement draws against the product's vectorised ones
    rng = np.random.default_rng(3)
    act = np.zeros((200, 4), np.float32)
    act[:, :3] = (rng.random((200, 3)) * 7.0).astype(np.float32)
    seed, gs, count = 19, 23456, 40
    exp = oev.returning_beads(seed, gs, act, 200, (7.0, 7.0, 7.0), count, 0.75, 1.1, 4)
    a = oev.rehydration_seed(seed, gs)
    got = sampleInsertionPositions(act, 200, (7.0, 7.0, 7.0), count, 0.75, a)
    assert np.array_equal(got["positions"].reshape(-1, 4)[:, :3], exp["positions"].reshape(-1, 4)[:, :3])
    from protocell_soup.init_state import _bm

    u = Mulberry32(got["rng"]).take(6 * count).reshape(count, 3, 2)
    vel = (math.sqrt(1.1) * _bm(u[:, :, 0], u[:, :, 1])).astype(np.float32)
    ev = exp["velocities"].reshape(-1, 4)[:, :3]
    # numpy's log/cos against the C library's: the f32 results agree (at most a rare last-bit tie)
    assert np.count_nonzero(vel != ev) <= 1 and np.abs(vel - ev).max() <= np.spacing(np.float32(4.0))
    # the known answer of the seed mix: Math.imul(23457, 0x9e3779b9) ^ 19
    assert a == ((23457 * 0x9E3779B9) & 0xFFFFFFFF) ^ 19


def test_resolve_evaporating_cycle_and_a_live_count_below_the_census():
    soup = loadSoup()
    opts = {"box": list(BOX24), "seed": 19, "kT": 1.1, "start": START24, "dryWetCycle": True, "evaporateSolvent": True,
            "clay": False}
    R = ps.resolve(opts)
    plan = planEvaporation(soup, loadParams(), BOX24, START24)
    assert R.evap == plan
    # the dry box is the one for what is left, not for the wet census
    assert R.dry_box == plan.dryBox
    wet_dry = computeDryBox(BOX24, sum(START24.values()), soup["dryWetCycle"]["targetDryDensity"])
    assert R.dry_box[0] < wet_dry[0]
    assert R.N == R.capacity_n == R.cfg.n == R.cfg.capacity == sum(START24.values())
    # the dry grid of 24 sigma is the smallest with the minimum cell count (soup-evaporation.test.ts:29-33)
    assert 16.2 < plan.dryBox[0] < 16.3
    _, dry = deriveCycleConfig(soup, opts, BOX24, R.capacity_n, R.start, plan.dryBox)
    assert dry == plan.dryBox
    # a resume mid dry phase: the live census is smaller than the creation one
    n = plan.organicCount + 100
    st = {"globalStep": 12000, "liveBox": list(plan.dryBox), "activeCounts": dict(START24, W=100),
          "positions": np.zeros(4 * n, np.float32), "velocities": np.zeros(4 * n, np.float32),
          "bondSlots": np.full(3 * n, 0xFFFFFFFF, np.uint32), "centerLink": np.full(n, 0xFFFFFFFF, np.uint32),
          "centerHeldSteps": np.zeros(n, np.uint32), "desorbEvents": np.zeros(2, np.uint32),
          "bondRng": np.zeros(n, np.uint32), "thermoRng": np.zeros(n, np.uint32), "events": {}}
    R2 = ps.resolve(dict(opts, resume=st))
    assert R2.N == R2.cfg.n == n and R2.cfg.capacity == R.capacity_n
    assert (R2.cycle_phase, R2.cycle_index) == ("dry", 1)
    too_many = dict(st, activeCounts=dict(START24, W=20000))
    with pytest.raises(ValueError, match="larger than the starting one"):
        ps.resolve(dict(opts, resume=too_many))
