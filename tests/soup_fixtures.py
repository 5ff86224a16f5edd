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


def bilayer_patch_opts(n_lipids=2087, tail=4, L=36.84, seed=5, pH=7.0, ionic=0.01, density=0.8):
    """A pre-assembled bilayer patch of ``n_lipids`` amphiphiles (head + ``tail`` carbons) normal to z
    in a cubic box of edge L, with water outside the slab bringing the whole box to ``density``: the
    construction of tests/water-bilayer-area-move.test.ts:75-194 (two leaflets on a jittered square
    lattice, tails stacked at the bond lengths, the head outermost, water on a jittered lattice
    skipping the slab, zero velocities, bondRng i + 1, thermoRng i + 1000003) with longer tails and
    charged heads. The DH list is sized for the heads' slab (densityVolumeSigma3 = L^2 x the slab
    thickness)."""
    import math

    from protocell_soup.init_state import Mulberry32, NONE_U32, link_chain
    from protocell_soup.rules import loadParams, loadSoup

    soup, p = loadSoup(), loadParams()
    mons = soup["monomers"]
    carbon = next(k for k, m in enumerate(mons) if m["kind"] == "carbon")
    head = next(k for k, m in enumerate(mons) if m["kind"] == "head")
    water = next(k for k, m in enumerate(mons) if m["id"] == soup["solvent"]["waterId"])
    b_tt = p["sigma"] * mons[carbon]["radiusSigma"]
    b_ht = p["sigma"] * (mons[head]["radiusSigma"] + mons[carbon]["radiusSigma"]) / 2
    b_wt = p["sigma"] * (mons[water]["radiusSigma"] + mons[carbon]["radiusSigma"]) / 2
    z0, gap = L / 2, b_tt / 2
    half = gap + (tail - 1) * b_tt + b_ht
    per = math.ceil(n_lipids / 2)
    n_side = max(1, math.ceil(math.sqrt(per)))
    s_l = L / n_side
    jit = 0.15
    rng = Mulberry32(seed)
    r = lambda: float(rng.take(1)[0])  # noqa: E731
    n_water = round(density * L ** 3) - n_lipids * (1 + tail)
    cap = n_lipids * (1 + tail) + n_water
    pos = np.zeros((cap, 4), np.float32)
    bs = np.full(3 * cap, NONE_U32, np.uint32)
    k = 0
    for i in range(n_lipids):
        leaf = 1 if i < per else -1
        idx = i if i < per else i - per
        a, bb = idx % n_side, idx // n_side
        x = ((a + 0.5) * s_l + (r() * 2 - 1) * jit * s_l + L) % L
        y = ((bb + 0.5) * s_l + (r() * 2 - 1) * jit * s_l + L) % L
        zt = [z0 + leaf * (gap + (tail - 1 - m) * b_tt) for m in range(tail)]  # t1 (outer) .. tk (inner)
        zh = zt[0] + leaf * b_ht
        pos[k] = [x, y, zh, head]
        for m in range(tail):
            pos[k + 1 + m] = [x, y, zt[m], carbon]
        link_chain(bs, k, list(range(k + 1, k + 1 + tail)))
        k += 1 + tail
    z_lo, z_hi = z0 - half - 1.5 * b_wt, z0 + half + 1.5 * b_wt
    # the lattice fine enough that the planes outside the slab hold every water bead
    wn = max(1, math.ceil((n_water * L / (L - (z_hi - z_lo))) ** (1 / 3)))
    sw = L / wn
    placed = 0
    for iz in range(wn):
        z_raw = (iz + 0.5) * sw
        if z_lo < z_raw < z_hi:
            continue
        for iy in range(wn):
            for ix in range(wn):
                if k >= cap:
                    break
                x = ((ix + 0.5) * sw + (r() * 2 - 1) * jit * sw + L) % L
                y = ((iy + 0.5) * sw + (r() * 2 - 1) * jit * sw + L) % L
                z = min(L - 1e-4, max(0.0, z_raw + (r() * 2 - 1) * jit * sw))
                pos[k] = [x, y, z, water]
                k += 1
                placed += 1
    n = k
    start = {m["id"]: 0 for m in mons}
    start[mons[head]["id"]], start[mons[carbon]["id"]], start[mons[water]["id"]] = n_lipids, tail * n_lipids, placed
    resume = {"globalStep": 0, "liveBox": [L, L, L], "positions": pos[:n].reshape(-1),
              "velocities": np.zeros(4 * n, np.float32), "bondSlots": bs[: 3 * n],
              "centerLink": np.full(n, NONE_U32, np.uint32), "centerHeldSteps": np.zeros(n, np.uint32),
              "desorbEvents": np.zeros(2, np.uint32), "bondRng": np.arange(1, n + 1, dtype=np.uint32),
              "thermoRng": np.arange(1000003, 1000003 + n, dtype=np.uint32), "events": {}}
    es = {"enabled": True, "pH": pH, "ionicStrengthMolar": ionic, "densityVolumeSigma3": L * L * 2 * (half + 1.0)}
    return {"box": [L, L, L], "seed": seed, "kT": p["thermostat"]["kT"], "clay": False, "start": start, "resume": resume,
            "electrostatics": es}
