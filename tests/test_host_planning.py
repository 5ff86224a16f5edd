"""Host-side planning against the reference's own known answers (CPU only).

Known answers are the ones the reference's tests assert: tests/verlet-capacity.test.ts:28-80,
tests/soup-electrostatics.test.ts:39-122 and :460-495, tests/soup-rules.test.ts:4-32.
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from protocell_soup import native
from protocell_soup.electrostatics import (PcgState, apparentPKa, esPairEnergy, esPairForceMag, hendersonAlpha,
                                           makeEsBasis, pcgNext)
from protocell_soup.plan import (deriveGridGeometry, deriveListCapacity, densestDensityOf, planSoupGrid,
                                 verletListBytes)
from protocell_soup.rules import (acceptanceProbability, acidSoapScaleOf, attemptProbability, attractionScaleTable,
                                  forwardBackwardRatio, loadParams, loadSoup, resolveRules, wcaCutoff)
from protocell_soup.sim import resolve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOUP = loadSoup()
P = loadParams()
LIST_RANGE = wcaCutoff(P["sigma"] * max(m["radiusSigma"] for m in SOUP["monomers"])) + P["attraction"]["wc"] + SOUP["verletList"]["skin"]


def test_list_range_and_capacity_known_answers():
    assert LIST_RANGE == pytest.approx(4.446954, abs=5e-7)
    n54, box54 = 173353, (54, 54, 54)
    rho_dry = densestDensityOf(SOUP, n54, box54, n54, box54, True)
    assert rho_dry == pytest.approx(1.34, abs=1e-10)
    cap = deriveListCapacity(SOUP, LIST_RANGE, n54, rho_dry)
    assert cap["derived"] and cap["uniform"] == pytest.approx(493.607, abs=5e-4) and cap["capacity"] == 1284
    rho_wet = densestDensityOf(SOUP, n54, box54, n54, box54, False)
    assert rho_wet == pytest.approx(1.10091, abs=5e-6)
    assert deriveListCapacity(SOUP, LIST_RANGE, n54, rho_wet)["capacity"] == 1126
    assert math.ceil(2.6 * (4 * math.pi / 3) * LIST_RANGE ** 3 * 0.8) == 767
    cap_w = deriveListCapacity(SOUP, LIST_RANGE, 21600, 0.8)
    assert cap_w["capacity"] == 1126
    assert 4294967292 // (cap_w["capacity"] * 4) == 953589
    assert verletListBytes(483268, 1284) == 2482064448


def test_campaign_configuration_sizes():
    opts = {"box": [76, 76, 76], "seed": 19, "kT": 1.1, "clay": False,
            "start": {"C": 55988, "O": 18659, "H": 55988, "M": 1452, "W": 351181},
            "electrostatics": {"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.01}}
    R = resolve(opts)
    assert R.N == 483268 and tuple(R.plan.dims) == (25, 25, 25) and R.list_cap == 1126
    assert R.es.cutoff == pytest.approx(15.2) and R.es_heads == 18659 and R.cfg.es_cap == 1327


def test_grid_plan_validity_and_messages():
    bad = planSoupGrid((12, 12, 12))
    assert not bad.valid and bad.reason.startswith("neighbour grid: box=[12,12,12] gives cellSize=2.9470, dims=[4,4,4]")
    assert "it needs dims>=5 (2*effectiveWalkRadius+1)" in bad.reason
    assert planSoupGrid((14.74, 14.74, 14.74)).valid
    assert not planSoupGrid((14.73, 14.73, 14.73)).valid
    with pytest.raises(ValueError, match=r"Verlet list: 2\*rebuildEvery\*dt\*vBound=.* exceeds skin=1.5 at kT=20"):
        deriveGridGeometry(SOUP, P, 20.0)
    g = deriveGridGeometry(SOUP, P, 1.1)
    assert g.effectiveWalkRadius == 2 and g.interactionRange == pytest.approx(2.946954, abs=5e-7)


def test_es_basis_known_answers():
    p = dict(P, thermostat=dict(P["thermostat"], kT=1.0))
    b = makeEsBasis(SOUP, p, {"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.01, "minBoxSigma": 30.0})
    assert b.coeffA == pytest.approx(0.8875, abs=1e-12)
    assert b.debyeSigma == pytest.approx(3.8, abs=1e-12) and b.kappa == pytest.approx(1 / 3.8)
    assert b.cutoff == pytest.approx(13.5)  # min(4 lambda_D = 15.2, 0.45 * 30)
    for I in (0.01, 0.1):
        bi = makeEsBasis(SOUP, p, {"enabled": True, "pH": 7, "ionicStrengthMolar": I})
        assert bi.debyeLengthsSpanned == pytest.approx(4.0, abs=1e-9) and bi.cutoff > bi.nbCutoff
    old = makeEsBasis(SOUP, p, {"enabled": True, "pH": 7, "ionicStrengthMolar": 0.01, "cutoffSigma": 2.7224620})
    assert old.discardedIntegratedFraction > 0.8
    off = makeEsBasis(SOUP, p)
    assert off.coeffA == 0 and esPairEnergy(0.95, 1, off) == 0 and esPairForceMag(0.95, 1, off) == 0
    assert hendersonAlpha(7.0, 4.9) == pytest.approx(0.99212, abs=5e-6)
    with pytest.raises(ValueError, match="violates the minimum image convention"):
        makeEsBasis(SOUP, p, {"enabled": True, "ionicStrengthMolar": 0.01, "minBoxSigma": 20.0, "cutoffSigma": 10.5})


def test_es_pair_force_is_minus_dU_and_continuous():
    b = makeEsBasis(SOUP, P, {"enabled": True, "pH": 7})
    h = 1e-6
    worst = 0.0
    for r in [0.6, 0.8, 0.95, 1.2, 1.5, 2.0, 2.4, 2.7, 2.9469545, 3.2, 3.8, 4.4, 4.8]:
        num = -(esPairEnergy(r + h, 1, b) - esPairEnergy(r - h, 1, b)) / (2 * h)
        ana = esPairForceMag(r, 1, b)
        if ana == 0:
            continue
        worst = max(worst, abs(num - ana) / max(1e-12, abs(ana)))
    assert worst < 1e-5
    assert abs(esPairEnergy(b.cutoff - 1e-7, 1, b)) < 1e-6 and abs(esPairForceMag(b.cutoff - 1e-7, 1, b)) < 1e-6
    assert esPairEnergy(b.cutoff, 1, b) == 0 and esPairForceMag(b.cutoff, 1, b) == 0


def test_apparent_pka_and_pcg_stream():
    for alpha, pH in [(0.5, 7), (0.1, 7), (0.9, 7)]:
        assert hendersonAlpha(pH, apparentPKa(alpha, pH)) == pytest.approx(alpha, abs=1e-12)
    assert apparentPKa(0.1, 7) > 7 > apparentPKa(0.9, 7)
    rng = PcgState(1)
    u = np.array([pcgNext(rng) for _ in range(100000)])
    assert 0.49 < u.mean() < 0.51 and u.min() >= 0.0 and u.max() < 1.0


def test_rule_probabilities():
    rules, cat = resolveRules(SOUP)
    assert [r.bond["id"] for r in rules] == ["cc_bond", "co_bond"] and SOUP["monomers"][cat]["id"] == "M"
    for r in rules:
        for kT in (1.0, 1.1):
            ratio = acceptanceProbability(r.brk, kT) / acceptanceProbability(r.bond, kT)
            assert ratio == pytest.approx(math.exp(-r.bond["energyKT"] / kT))
        assert forwardBackwardRatio(r.bond) == pytest.approx(math.exp(-r.bond["energyKT"]))
    bond_dt = P["integrator"]["dt"] * SOUP["bondAttemptInterval"]["steps"]
    assert attemptProbability(rules[0].bond, bond_dt) == pytest.approx(0.01)
    assert attemptProbability(rules[1].bond, bond_dt) == pytest.approx(0.5125338224)
    assert acceptanceProbability(rules[0].brk, 1.1) == pytest.approx(0.00428, abs=5e-6)
    assert acceptanceProbability(rules[1].brk, 1.1) == pytest.approx(0.000694, abs=5e-7)
    assert (rules[0].slotRoleA, rules[0].slotRoleB, rules[1].slotRoleA, rules[1].slotRoleB) == (0, 0, 1, 2)


def test_attraction_table_and_acid_soap():
    t = attractionScaleTable(SOUP)
    assert t[0][0] == 1.0 and t[1][1] == 0.0 and t[1][0] == 0.0
    assert t[2][2] == pytest.approx(5 / 3.5) and t[2][1] == pytest.approx(4 / 3.5) and t[2][0] == pytest.approx(2 / 3.5)
    assert t[3][3] == pytest.approx(5.6 / 3.5) and t[3][1] == pytest.approx(5 / 3.5)
    assert acidSoapScaleOf(SOUP) == pytest.approx(2.0)


def test_config_struct_layout_matches_c(tmp_path):
    src = tmp_path / "off.c"
    fields = ["n", "box", "dims", "list_range", "kT", "b_tt", "wca_factor", "radius", "cls", "attr", "acid_soap",
              "es_enabled", "es_Frc", "n_heads", "ph_A", "ph_q", "as_eps", "sweep_every", "n_rules", "kindA", "roleB",
              "catalyst_kind", "bond_interval", "capacity", "reserved"]
    body = "\n".join(f'printf("{f} %zu\\n", offsetof(SoupConfig, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "soup_abi.h"\nint main(){' + body +
                   'printf("size %zu\\n", sizeof(SoupConfig)); printf("init %zu\\n", sizeof(SoupInitState)); return 0;}')
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for f in fields:
        assert int(out[f]) == getattr(native.SoupConfig, f).offset, f
    import ctypes

    assert int(out["size"]) == ctypes.sizeof(native.SoupConfig)
    assert int(out["init"]) == ctypes.sizeof(native.SoupInitState)


def _declared_entry_points():
    text = open(os.path.join(ROOT, "include", "soup_abi.h")).read()
    return sorted(set(re.findall(r"^(?:int|size_t|const char\*)\s+(soup_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib_path = native.LIB_PATH
    if not os.path.exists(lib_path):
        from protocell_soup.build import build_library

        build_library()
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT\s+(soup_\w+)$", out, flags=re.M))
    declared = _declared_entry_points()
    assert declared and set(declared) <= exported, set(declared) - exported
    assert set(native.EXPORTS) == set(declared)
    import ctypes

    lib = ctypes.CDLL(lib_path)  # loads without a GPU; no compute is called
    for name in declared:
        getattr(lib, name)
