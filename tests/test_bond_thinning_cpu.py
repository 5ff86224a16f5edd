"""Poisson thinning of the bond Monte Carlo in the oracle (soup/src/soup-buffers.ts:345-356).

Bond attempts run every bondAttemptInterval steps with attemptProbability(rule, dt * interval) =
attemptRate * dt * interval (soup/src/rules.ts:576-578), each candidate with its own attempt draw,
then acceptanceProbability = min(1, exp(-dE / kT)) (:588-591). The oracle's draw counters give, per
rule, how many candidates reach each draw and how many pass it; every pass fraction must equal the
probability computed here from data/soup.json and data/params.json within five binomial standard
deviations. A dropped interval factor (x20), an attempt probability taken per step, one draw shared
by a particle's candidates, or a break acceptance without the kT scaling fails.

The configuration is never integrated (positions fixed, 400 bond phases on one dense soup at kT 4,
where the C-O break acceptance exp(-8/4) = 0.135 is measurable); bonds form and break between phases.
"""
import math

import numpy as np

import protocell_soup as ps
from oracle import oracle as orc
from protocell_soup.rules import loadParams, loadSoup
from soup_fixtures import amphiphile_opts

KT = 4.0


def _within(passed, n, p):
    sd = math.sqrt(n * p * (1.0 - p))
    return abs(passed - n * p) <= 5.0 * sd + 1.0


def test_bond_draws_follow_the_thinned_rates():
    soup, params = loadSoup(), loadParams()
    rules = {r["id"]: r for r in soup["rules"]}
    interval = soup["bondAttemptInterval"]["steps"]
    dt = params["integrator"]["dt"]
    opts = amphiphile_opts(15.0, 400, 4, {"C": 400, "O": 300, "M": 80, "W": 700}, seed=3, kT=KT)
    R = ps.resolve(opts)
    o = orc.OracleState(R.cfg, R.init, R.charges0)
    o.rebuild()
    orc.bond_draws_reset()
    for _ in range(400):
        o.bond_form()
        o.bond_break()
    draws = orc.bond_draws()
    for r, rule in enumerate(R.rules):
        form, brk = rules[rule.bond["id"]], rules[rule.brk["id"]]
        p_form = form["attemptRate"] * dt * interval
        p_break = brk["attemptRate"] * dt * interval
        acc_break = min(1.0, math.exp(-brk["energyKT"] / KT))
        n0, a0, c0, n3, a3, c3 = (int(x) for x in draws[r])
        assert n0 > 50_000, (form["id"], n0)  # enough formation candidates to resolve the rate
        assert _within(a0, n0, p_form), (form["id"], a0, n0, p_form)
        assert c0 == a0  # formation is downhill: acceptance 1
        assert n3 > 100_000, (brk["id"], n3)
        if p_break == 0.0:
            assert a3 == 0 and c3 == 0
        else:
            assert _within(a3, n3, p_break), (brk["id"], a3, n3, p_break)
            assert _within(c3, a3, acc_break), (brk["id"], c3, a3, acc_break)
            assert c3 > 50
    # the two formation rules' attempts are resolved to better than 4 %
    assert np.all(draws[: len(R.rules), 1] > 600)
