"""The benchmark configurations beyond the campaign box, each on the GPU.

- Charged soup (30 sigma, 21,600 beads at density 0.80, 1,500 five-bead amphiphiles, pH 7.0, 0.01 M,
  sweeps every 100 steps, bond Monte Carlo on). Chain growth and head termination need free carbons,
  heads and catalysts, which a soup of pre-built amphiphiles in water lacks, so part of the water is
  the reference broth at its ratio C:O:H:M = 1500:500:1500:100 (data/soup.json's start): census
  {C 6,750 (6,000 in chains + 750 free), O 1,750 (1,500 + 250), H 750, M 50, W 12,300}. One snapshot
  against the oracle (cell assignment, every Verlet row, every DH row, the force within 1e-5 of FP64,
  one constant-pH sweep, one bond phase), then 10,000 steps in chunks of 100 with census, charge,
  valence and finiteness checks.
- Batched sweep: the 64 replicas of sweepReplicaParams(64) (edges 14.7-40 sigma at density 0.80, pH
  5-9, ionic strength 0.001-0.1 M, supply 1.5-4.8 thresholds, one seed each) stepped together for a
  1,000-step chunk and past it; the smallest, the median and the largest replica must equal the same
  replica stepped alone bit for bit, and every replica must keep its census, finite state, charges
  and list guards.
"""
import numpy as np
import pytest

import protocell_soup as ps
from protocell_soup.replicas import sweepReplicaOpts, sweepReplicaParams
from soup_fixtures import amphiphile_opts, assert_same_state, force_rel_err, oracle_from_system

pytestmark = pytest.mark.gpu

C2_BROTH = {"C": 750, "O": 250, "H": 750, "M": 50, "W": 12300}


def _c2_opts():
    return amphiphile_opts(30.0, 1500, 4, C2_BROTH, seed=2, kT=1.0,
                           electrostatics={"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.01, "sweepEverySteps": 100})


def _valence_ok(s):
    soup = s._R.soup
    kinds = np.rint(s.particles().reshape(-1, 4)[:, 3]).astype(np.int64)
    kind_of = np.array([m["kind"] for m in soup["monomers"]])[kinds]
    slots = s.bondSlots().reshape(-1, 3)
    edges = s.bonds().reshape(-1, 2).astype(np.int64)
    cc = np.zeros(kinds.size, int)
    co = np.zeros(kinds.size, int)
    hd = np.zeros(kinds.size, int)
    for i, j in edges:
        a, b = kind_of[i], kind_of[j]
        if a == b == "carbon":
            cc[i] += 1
            cc[j] += 1
        elif {a, b} == {"carbon", "head"}:
            c, h = (i, j) if a == "carbon" else (j, i)
            co[c] += 1
            hd[h] += 1
        else:
            return False
    sym = all(i in slots[j] for i, j in edges) and all(j in slots[i] for i, j in edges)
    return sym and cc.max() <= 2 and co.max() <= 1 and hd.max() <= soup["headPlacement"]["chainCapacity"]


def test_charged_soup_30_sigma():
    opts = _c2_opts()
    s = ps.createSoup(opts)
    try:
        R = s._R
        assert s.N == 21600 and R.cfg.es_enabled == 1 and R.sweep_every == 100
        census = s.invariants()["monomers"]
        assert census == {"C": 6750, "O": 1750, "H": 750, "M": 50, "W": 12300, "K": 0}
        s.relaxColdStart()
        s.step(37)
        s.forces()
        o = oracle_from_system(s, R)
        o.rebuild()
        assert np.array_equal(s.readDEBUG("CELL_OF"), o.cell_of)
        assert np.array_equal(s.verletCount(), o.vcount)
        assert np.array_equal(s.verletList(), o.verlet_rows())
        nh = o.s.nh
        assert nh == 1750
        assert np.array_equal(s.readDEBUG("ES_COUNT"), o.es_count[:nh])
        assert np.array_equal(s.readDEBUG("ES_LIST").reshape(nh, -1), o.es_list.reshape(-1, R.cfg.es_cap)[:nh])
        f = s.forcesNoRebuildDEBUG().reshape(-1, 4)[:, :3]
        assert force_rel_err(f, o.force_f64(use_list=True)) < 1e-5
        r_gpu, r_ref = s.protonationSweepDEBUG(), o.protonation_sweep()
        assert np.array_equal(s.charges(), o.charge) and s.protonationRngState() == o.s.prng
        assert (r_gpu["accepted"], r_gpu["deprotonated"]) == (r_ref["accepted"], r_ref["deprotonated"])
        s.runPhasesDEBUG("BONDS")
        o.bond_form()
        o.bond_break()
        assert np.array_equal(s.bondSlots(), o.bond_slots)
        assert np.array_equal(s.centerLinks(), o.center_link)
        assert np.array_equal(s.rngState()["bond"], o.bond_rng)
        assert np.array_equal(s.readDEBUG("EVENTS"), o.events[: 2 * R.cfg.n_rules])

        sweeps0 = s.electrostatics()["sweeps"]
        for _ in range(100):
            s.step(100)
        es = s.electrostatics()
        ev = s.events()
        q = s.charges()
        print(f"C2 steps={s.steps} sweeps={es['sweeps'] - sweeps0} alpha={es['last']['alpha']:.4f} events={ev} "
              f"bonds={s.bonds().size // 2}", flush=True)
        assert es["sweeps"] - sweeps0 == 100
        assert s.invariants()["monomers"] == census
        assert s.nonFiniteCount() == {"pos": 0, "vel": 0}
        assert set(np.unique(q).tolist()) <= {0.0, float(np.float32(R.cfg.ph_q))}
        heads = np.rint(s.particles().reshape(-1, 4)[:, 3]) == R.cfg.charged_kind
        assert np.all(q[~heads] == 0)
        assert 0.5 < es["last"]["alpha"] < 1.0  # pH 7 is above the intrinsic pKa 4.9; the field lowers it
        assert ev["cc_bond"] > 0 and ev["co_bond"] > 0  # chains grew and heads terminated them
        assert _valence_ok(s)
    finally:
        s.dispose()


def test_batched_sweep_of_64_replicas():
    params = sweepReplicaParams(64)
    order = np.argsort([p["N"] for p in params], kind="stable")
    picks = {int(order[0]): "smallest", int(order[len(order) // 2]): "median", int(order[-1]): "largest"}
    batch = [ps.createSoup(sweepReplicaOpts(p)) for p in params]
    solo = {r: ps.createSoup(sweepReplicaOpts(params[r])) for r in picks}
    try:
        census = [s.invariants()["monomers"] for s in batch]
        for s in batch + list(solo.values()):
            s.relaxColdStart()
        ps.stepReplicas(batch, 1000)
        ps.stepReplicas(batch, 230)
        for r, s in solo.items():
            s.step(1000)
            s.step(230)
            assert_same_state(batch[r], s, f"{picks[r]} replica {r} (N={s.N})")
        for r, s in enumerate(batch):
            assert s.steps == 1230
            assert s.invariants()["monomers"] == census[r], r
            assert s.nonFiniteCount() == {"pos": 0, "vel": 0}, r
            q = s.charges()
            assert set(np.unique(q).tolist()) <= {0.0, float(np.float32(s._R.cfg.ph_q))}, r
            assert s.electrostatics()["sweeps"] == 1, r
            s._call(s._lib.soup_check_lists)
        print("C4 N", min(p["N"] for p in params), "..", max(p["N"] for p in params), "total", sum(s.N for s in batch), flush=True)
    finally:
        for s in batch + list(solo.values()):
            s.dispose()
