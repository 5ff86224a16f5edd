"""Shared system builders and comparisons for the parity tests."""
import numpy as np

import protocell_soup as ps
from protocell_soup.replicas import amphiphileSoupOpts
from oracle import oracle as orc

# small broth: catalysts, free carbons and heads, water (tests/soup-forces.test.ts:31-45 style)
BROTH = {"box": [16, 16, 16], "seed": 7, "kT": 1.0, "clay": False,
         "start": {"C": 600, "O": 100, "H": 500, "M": 20, "W": 0}}


def charged(box=20.0, seed=11, pH=5.0, ionic=0.01, start=None):
    return {"box": [box, box, box], "seed": seed, "kT": 1.0, "clay": False,
            "start": start or {"C": 700, "O": 700, "H": 300, "M": 30, "W": 1000},
            "electrostatics": {"enabled": True, "pH": pH, "ionicStrengthMolar": ionic}}


def amphiphile_opts(box, n_amph, tail, extra_ids, seed=1, kT=1.0, electrostatics=None):
    """createSoup opts for a soup with pre-bonded amphiphiles passed through ``resume``."""
    return amphiphileSoupOpts(box, n_amph, tail, extra_ids, seed=seed, kT=kT, electrostatics=electrostatics)


def oracle_from_system(sys_, R):
    """Oracle state holding exactly the system's current state (reference layout)."""
    init = {"positions": sys_.particles(), "velocities": sys_.velocities(), "bondSlots": sys_.bondSlots(),
            "centerLink": sys_.centerLinks(), "centerHeldSteps": sys_.centerHeldSteps(), "frozen": sys_.frozen(),
            "bondRng": sys_.rngState()["bond"], "thermoRng": sys_.rngState()["thermo"]}
    ev = sys_.readDEBUG("EVENTS")
    de = sys_.readDEBUG("DESORB")
    o = orc.OracleState(R.cfg, init, sys_.charges(), global_step=sys_.steps, prng=sys_.protonationRngState(),
                        events=ev, desorb=de)
    return o


def force_rel_err(f32, f64):
    """Per-particle max-component error over max(|F64_i|, median_j |F64_j|)."""
    f32 = np.asarray(f32, dtype=np.float64).reshape(-1, 3)
    mag = np.abs(f64).max(axis=1)
    den = np.maximum(mag, np.median(mag))
    return float((np.abs(f32 - f64).max(axis=1) / den).max())


def state_of(s):
    """Every per-particle array and counter a step advances, in particle order."""
    rng = s.rngState()
    return {"step": np.array([s.steps]), "pos": s.particles(), "vel": s.velocities(), "force": s.forcesNoRebuildDEBUG(),
            "bonds": s.bondSlots(), "clink": s.centerLinks(), "cheld": s.centerHeldSteps(), "brng": rng["bond"],
            "trng": rng["thermo"], "events": s.readDEBUG("EVENTS"), "charge": s.charges(),
            "prng": np.array([s.protonationRngState()])}


def assert_same_state(a, b, what):
    sa, sb = state_of(a), state_of(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), f"{what}: {k} differs"
