"""The initial-state and constant-pH random streams against known answers computed in V8.

tests/golden/rng_vectors.json is written by tests/golden/make_rng_vectors.js under node (the
reference's JavaScript runtime: its Math.imul, >>> and Math.log / Math.cos / Math.pow), from the
published definitions of mulberry32, Box-Muller and the PCG output step. Neither the host's vectorised
numpy code nor the oracle's scalar C produced these numbers, so a wrong shift, a wrong constant, a
swapped Box-Muller draw or a wrong draw offset in the lattice fails here.

The C library's cbrt is not exact on perfect cubes (cbrt(27) = 3 + 1 ulp), which once gave the
oracle's lattice an nx^3 = 64-site grid for 27 particles; the lattice case N = 27 is that regression.
"""
import json
import os

import numpy as np

from oracle import oracle as orc
from protocell_soup.electrostatics import PcgState, hendersonAlpha, initialCharges, makeEsBasis, pcgNext
from protocell_soup.init_state import _bm, lattice_state, mulberry32_draws
from protocell_soup.rules import loadParams, loadSoup

HERE = os.path.dirname(os.path.abspath(__file__))
V = json.load(open(os.path.join(HERE, "golden", "rng_vectors.json")))


def test_mulberry32_known_answers():
    for case in V["mulberry32"]:
        host = (mulberry32_draws(case["seed"], len(case["u32"])) * 4294967296.0).astype(np.int64).tolist()
        assert host == case["u32"], case["seed"]


def test_gaussian_within_one_ulp_of_v8():
    # the host's log and cos are numpy's, V8's are its own: the doubles agree to one ulp
    for case in V["gaussian"]:
        u = mulberry32_draws(case["seed"], 2 * len(case["values"]))
        g = _bm(u[0::2], u[1::2])
        ref = np.array(case["values"])
        assert np.all(np.abs(g - ref) <= np.spacing(np.abs(ref))), case["seed"]


def test_lattice_velocities_equal_v8_in_f32():
    for case in V["lattice_velocities"]:
        ref = np.array(case["velocities"]).astype(np.float32)
        host = lattice_state((10.0, 10.0, 10.0), [case["n"]], case["seed"], case["kT"])
        assert np.array_equal(host["velocities"].reshape(-1, 4)[:, :3].reshape(-1), ref), case
        ora = orc.lattice((10.0, 10.0, 10.0), [case["n"]], case["seed"], case["kT"])
        assert np.array_equal(ora["velocities"].reshape(-1, 4)[:, :3].reshape(-1), ref), case


def test_lattice_grid_is_the_smallest_cube_holding_n():
    # 27 = 3^3 sits on a 3 x 3 x 3 grid: site spacing box / 3, one particle per site
    for n in (27, 216, 729, 28):
        lat = lattice_state((12.0, 12.0, 12.0), [n], 5, 1.0)
        ora = orc.lattice((12.0, 12.0, 12.0), [n], 5, 1.0)
        nx = round(n ** (1 / 3))
        nx = nx if nx ** 3 >= n else nx + 1
        pos = lat["positions"].reshape(-1, 4)[:, :3]
        site = np.floor(pos / (12.0 / nx)).astype(int)
        assert len({tuple(s) for s in site}) == n
        assert site.max() == nx - 1
        assert np.array_equal(lat["positions"], ora["positions"]), n


def test_pcg_next_and_initial_charges_known_answers():
    for case in V["pcg_next"]:
        s = PcgState(case["start"])
        host = [int(pcgNext(s) * 4294967296.0) for _ in case["u32"]]
        assert host == case["u32"] and s.state == case["end_state"]
        st, ora = case["start"], []
        for _ in case["u32"]:
            v, st = orc.pcg_next(st)
            ora.append(int(v * 4294967296.0))
        assert ora == case["u32"] and st == case["end_state"]
    for case in V["henderson_alpha"]:
        assert hendersonAlpha(case["pH"], case["pKa"]) == case["alpha"]
    # initialCharges (soup/src/soup-protonation.ts:38-49): head h is deprotonated iff the h-th draw
    # of the stream seeded seed ^ 0x9e3779b9 is below the HH alpha
    soup, p = loadSoup(), loadParams()
    case = V["pcg_next"][0]
    es = makeEsBasis(soup, p, {"enabled": True, "pH": 5.5, "ionicStrengthMolar": 0.01})
    alpha = next(a["alpha"] for a in V["henderson_alpha"] if a["pH"] == 5.5 and a["pKa"] == es.pKaIntrinsic)
    n = len(case["u32"])
    pos = np.zeros((n + 3, 4), np.float32)
    pos[:, 3] = [es.chargedKind] * n + [0, 0, 0]  # the heads first, three other beads after
    ch, rng = initialCharges(pos.reshape(-1), es, case["seed"])
    expect = [np.float32(es.chargeDeprotonated) if u / 4294967296.0 < alpha else 0 for u in case["u32"]] + [0, 0, 0]
    assert np.array_equal(ch, np.array(expect, np.float32))
    assert rng.state == case["end_state"]
    assert 0 < (ch != 0).sum() < n  # both outcomes occur at alpha = 0.80
