"""``createSoup`` / ``SoupSystem``: the soup API over libsoup_hip.

Host-side restatement of soup/src/sim.ts:95-600 (``createSoup``) for the step path: it resolves the
options exactly as the reference does (rules, plan, grid geometry and its guards, list capacity from
the densest density, the Debye-Hueckel basis, the initial lattice or a resume, the
Henderson-Hasselbalch charges), packs the flat ``SoupConfig`` and hands everything to the HIP
library, which owns the device state and runs the priming pass. ``SoupSystem`` mirrors the
reference's method names (soup/src/soup-types.ts:181-442) for the methods on this path.

Out of scope here (they raise): the clay platelet and parcel confinement.

Capacity and live count (soup/src/sim.ts:139-151): the device buffers hold the creation census
(``capacity``); the live count ``N`` is smaller on a system resumed mid dry phase or after solvent
evaporation, and grows back on rehydration (protocell_soup.evaporate).
"""
from __future__ import annotations

import ctypes as C
import math
import threading
from typing import Any, Dict, Optional, Sequence

import numpy as np

from . import native
from .electrostatics import EsBasis, initialCharges, makeEsBasis
from .init_state import NONE_U32, lattice_state
from .native import FIELD, SoupError
from .box_scale import applyBoxScale, cyclePhaseAt, deriveCycleConfig, growBoxTo, stepCycled
from .evaporate import planEvaporation
from .plan import deriveGridGeometry, deriveListCapacity, densestDensityOf, planSoupGrid, verletListBytes
from .rules import (
    acceptanceProbability,
    acidSoapScaleOf,
    attemptProbability,
    attractionScaleTable,
    loadParams,
    loadSoup,
    resolveRules,
    speciesClassOf,
)

F32 = np.float32


def _f32(x: float) -> float:
    return float(F32(x))


def _u32_array(a, n: int) -> np.ndarray:
    out = np.ascontiguousarray(np.asarray(a, dtype=np.uint32))
    if out.size != n:
        raise ValueError(f"expected {n} entries, got {out.size}")
    return out


def _f32_array(a, n: int) -> np.ndarray:
    out = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if out.size != n:
        raise ValueError(f"expected {n} entries, got {out.size}")
    return out


class Resolved:
    """Everything createSoup derives before touching the device."""

    pass


def resolve(opts: Dict[str, Any]) -> Resolved:
    """soup/src/sim.ts:96-322 restricted to unconfined, clay-free systems."""
    soup = loadSoup()
    if len(soup["monomers"]) > 8:
        raise ValueError(f"data/soup.json: {len(soup['monomers'])} monomer kinds, but the shader holds at most 8")
    base = loadParams()
    p = dict(base)
    p["thermostat"] = dict(base["thermostat"], kT=opts["kT"])
    rules, catalyst_kind = resolveRules(soup)
    event_rule_ids = [(r.bond["id"], r.brk["id"]) for r in rules]
    start = dict(soup["start"])
    start.update(opts.get("start") or {})
    if opts.get("catalystCount") is not None:
        cat = next(m for m in soup["monomers"] if m["kind"] == "catalyst")
        start[cat["id"]] = opts["catalystCount"]
    clay = soup.get("clay")
    if clay and opts.get("clay", clay.get("enabled", False)):
        raise ValueError("createSoup: the clay platelet is not part of this build; pass clay: False")
    if clay:
        start[clay["mineralId"]] = 0
    if opts.get("confine"):
        raise ValueError("createSoup: parcel confinement is not part of this build")
    cycling = bool(opts.get("dryWetCycle", soup["dryWetCycle"]["enabled"]))
    evap_requested = cycling and bool(opts.get("evaporateSolvent", soup["dryWetCycle"].get("evaporateSolvent", False)))
    counts_by_kind = [int(start.get(m["id"], 0)) for m in soup["monomers"]]
    capacity_n = sum(counts_by_kind)
    resume = opts.get("resume")
    active = dict(resume["activeCounts"]) if resume and resume.get("activeCounts") else dict(start)
    N = sum(int(active.get(m["id"], 0)) for m in soup["monomers"])
    if capacity_n == 0:
        raise ValueError("createSoup: the starting composition is empty")
    if N > capacity_n:
        raise ValueError(f"createSoup: the live composition ({N}) is larger than the starting one ({capacity_n}) -- buffers are allocated for the starting one")
    geo = deriveGridGeometry(soup, p, opts["kT"])
    box = tuple(float(x) for x in opts["box"])
    initial_step = int(resume["globalStep"]) if resume else 0
    live_box = tuple(float(x) for x in resume["liveBox"]) if resume and resume.get("liveBox") is not None else box
    plan = planSoupGrid(live_box, start)
    if not plan.valid:
        raise ValueError(plan.reason)
    # the dry box of an evaporating cycle holds what is left (planEvaporation), so it is resolved first
    evap = planEvaporation(soup, p, box, start) if evap_requested else None
    cycle_cfg, dry_box = deriveCycleConfig(soup, opts, box, capacity_n, start, evap.dryBox if evap else None)
    min_box = min(*live_box, *dry_box)
    densest = densestDensityOf(soup, capacity_n, box, N, live_box, cycle_cfg is not None)
    cap = deriveListCapacity(soup, geo.listRange, capacity_n, densest)
    vo = opts.get("verletOverride") or {}
    if vo.get("enabled") is False:
        raise ValueError("createSoup: the list-free cell-walk force path is not part of this build (verletList.enabled must stay true)")
    list_cap = int(vo.get("listCapacity", cap["capacity"]))
    rebuild = resolveRebuildSchedule(soup, p, opts["kT"], vo)
    es_over = dict(opts.get("electrostatics") or {})
    es_max_heads = int(start.get(soup["electrostatics"]["chargedKind"], 0)) if soup.get("electrostatics") else 0
    es = makeEsBasis(soup, p, {**es_over, "minBoxSigma": min_box, "maxHeads": es_max_heads,
                               "acidSoapScale": opts.get("acidSoapScaleOverride"),
                               "densityVolumeSigma3": es_over.get("densityVolumeSigma3")})
    sweep_every = int(es_over.get("sweepEverySteps", es.sweepEverySteps))
    seed = int(opts["seed"]) & 0xFFFFFFFF
    if resume:
        init = {
            "positions": _f32_array(resume["positions"], 4 * N),
            "velocities": _f32_array(resume["velocities"], 4 * N),
            "bondSlots": _u32_array(resume["bondSlots"], 3 * N),
            "centerLink": _u32_array(resume["centerLink"], N),
            "centerHeldSteps": _u32_array(resume["centerHeldSteps"], N),
            "frozen": _u32_array(resume.get("frozen", np.zeros(N)), N),
            "desorbEvents": _u32_array(resume["desorbEvents"], 2),
            "bondRng": _u32_array(resume["bondRng"], N),
            "thermoRng": _u32_array(resume["thermoRng"], N),
        }
    else:
        init = lattice_state(box, counts_by_kind, seed, opts["kT"])
    charges_fresh, prng = initialCharges(init["positions"], es, seed)
    charges0 = charges_fresh
    if resume and resume.get("charges") is not None:
        charges0 = _f32_array(resume["charges"], N)
    prng_state = int(resume["protonationRng"]) if resume and resume.get("protonationRng") is not None else prng.state
    events0 = np.zeros(2 * len(rules), dtype=np.uint32)
    if resume:
        ev = resume.get("events") or {}
        for ri, (bid, kid) in enumerate(event_rule_ids):
            events0[2 * ri] = int(ev.get(bid, 0))
            events0[2 * ri + 1] = int(ev.get(kid, 0))
    es_heads = es_max_heads if es.enabled else 0

    R = Resolved()
    R.soup, R.p, R.rules, R.catalyst_kind, R.event_rule_ids = soup, p, rules, catalyst_kind, event_rule_ids
    R.start, R.counts_by_kind, R.N, R.box, R.live_box, R.plan, R.geo = start, counts_by_kind, N, box, live_box, plan, geo
    R.capacity_n, R.evap = capacity_n, evap
    R.list_cap, R.es, R.es_heads, R.sweep_every, R.seed = list_cap, es, es_heads, sweep_every, seed
    R.opts = dict(opts)
    R.init, R.charges0, R.prng, R.events0, R.initial_step = init, charges0, prng_state, events0, initial_step
    R.next_sweep_at = initial_step + sweep_every if es.enabled else 2 ** 64 - 1
    R.verlet = dict(soup["verletList"], listCapacity=list_cap, rebuild=rebuild["mode"])
    if rebuild["mode"] == "drift":
        R.verlet.update(driftCheckEvery=rebuild["checkEvery"], driftTrigger=rebuild["trigger"])
    R.rebuild = rebuild
    R.cycle_cfg, R.dry_box = cycle_cfg, dry_box
    ph = cyclePhaseAt(initial_step, cycle_cfg) if cycle_cfg else {"phase": "none", "cycleIndex": 0}
    R.cycle_phase, R.cycle_index = ph["phase"], ph["cycleIndex"]
    R.cfg = build_config(R, opts)
    return R


def resolveRebuildSchedule(soup: dict, p: dict, kT: float, vo: Dict[str, Any]) -> Dict[str, Any]:
    """The Verlet rebuild schedule: the reference's fixed cadence (verletList.rebuildEvery,
    soup/src/soup-integrate.ts:149), or, with verletOverride.rebuild = "drift", a rebuild whenever the
    measured drift maximum exceeds a trigger, checked every driftCheckEvery steps (default: the
    autotuner's choice for the campaign, else 5). The default trigger
    leaves the reference's own per-step displacement estimate, dt * 2 sqrt(3 kT) (the drift condition
    of data/soup.json's verletList basis), for every step until the next check: skin/2 - checkEvery *
    dt * 2 sqrt(3 kT). The tuned interval was measured at the campaign's temperature and step; where
    it leaves no margin (a hotter or coarser soup), the largest interval up to it that does is taken.
    An interval the caller names is kept as given."""
    mode = vo.get("rebuild", "cadence")
    if mode not in ("cadence", "drift"):
        raise ValueError(f"verletOverride.rebuild={mode!r}: the schedule is 'cadence' (the reference's) or 'drift'")
    if mode == "cadence":
        return {"mode": "cadence"}
    skin = soup["verletList"]["skin"]
    from .autotune import tuned_drift_check_every

    v_bound = 2.0 * math.sqrt(3.0 * kT)
    step_drift = p["integrator"]["dt"] * v_bound
    if "driftCheckEvery" in vo:
        every = int(vo["driftCheckEvery"])
    else:
        every = tuned_drift_check_every()
        while every > 1 and skin / 2 - every * step_drift <= 0.0:
            every -= 1
    if not 1 <= every <= 1000:
        raise ValueError(f"verletOverride.driftCheckEvery={every} must be 1..1000")
    trigger = float(vo.get("driftTrigger", skin / 2 - every * step_drift))
    if not 0.0 < trigger <= skin / 2:
        raise ValueError(f"verletOverride.driftTrigger={trigger:.4f} must lie in (0, skin/2={skin / 2}] "
                         f"(at driftCheckEvery={every} the default margin leaves none)")
    return {"mode": "drift", "checkEvery": every, "trigger": trigger}


def build_config(R: Resolved, opts: Dict[str, Any]) -> native.SoupConfig:
    """Packs the uniforms the reference uploads (soup/src/soup-buffers.ts:291-414) into SoupConfig."""
    soup, p, es = R.soup, R.p, R.es
    c = native.SoupConfig()
    c.n = R.N
    c.capacity = R.capacity_n
    c.n_kinds = len(soup["monomers"])
    for a in range(3):
        c.box[a] = R.live_box[a]
        c.dims[a] = R.plan.dims[a]
    c.walk_radius = R.geo.effectiveWalkRadius
    c.list_range = R.geo.listRange
    c.list_cap = R.list_cap
    c.rebuild_every = soup["verletList"]["rebuildEvery"]
    c.skin = soup["verletList"]["skin"]
    c.sigma, c.epsilon = p["sigma"], p["epsilon"]
    c.k_fene, c.r_inf = p["fene"]["k"], p["fene"]["rInf"]
    c.k_bend, c.r_bend = p["bend"]["k"], p["bend"]["r0"]
    c.wc, c.kT, c.gamma, c.dt = p["attraction"]["wc"], p["thermostat"]["kT"], p["thermostat"]["gamma"], p["integrator"]["dt"]
    c.b_tt = p["beadSizes"]["tail_tail"]
    c.wca_factor = math.pow(2.0, 1.0 / 6.0)
    for k, m in enumerate(soup["monomers"]):
        c.radius[k] = m["radiusSigma"]
        c.cls[k] = speciesClassOf(m)
    table = attractionScaleTable(soup, opts.get("solventAttractionScaleOverride"), opts.get("claySurfaceChemistry"))
    for i in range(4):
        for j in range(4):
            c.attr[i][j] = table[i][j]
    c.acid_soap = acidSoapScaleOf(soup, opts.get("acidSoapScaleOverride"), opts.get("solventAttractionScaleOverride"))
    c.es_enabled = 1 if es.enabled else 0
    c.es_A, c.es_kappa, c.es_rc, c.es_Frc = es.coeffA, es.kappa, es.cutoff, es.shiftF
    c.es_split, c.es_list_range, c.es_cap = es.splitRadius, es.listRange, max(0, int(es.listCapacity))
    c.charged_kind = max(0, es.chargedKind)
    c.n_heads = R.es_heads
    c.ph_A, c.ph_kappa, c.ph_cutoff, c.ph_shiftU, c.ph_shiftF = es.coeffA, es.kappa, es.cutoff, es.shiftU, es.shiftF
    c.ph_pKa, c.ph_pH, c.ph_kT, c.ph_q = es.pKaIntrinsic, es.pH, es.kT, es.chargeDeprotonated
    c.as_scale, c.as_rc, c.as_wc, c.as_eps = es.acidSoap.scale, es.acidSoap.rcAttr, es.acidSoap.wc, es.acidSoap.epsilon
    c.sweep_every = R.sweep_every
    interval = soup["bondAttemptInterval"]["steps"]
    bond_dt = p["integrator"]["dt"] * interval
    kT = p["thermostat"]["kT"]
    c.n_rules = len(R.rules)
    for ri, r in enumerate(R.rules):
        c.kindA[ri], c.kindB[ri] = r.kindA, r.kindB
        c.att_form[ri] = attemptProbability(r.bond, bond_dt)
        c.att_break[ri] = attemptProbability(r.brk, bond_dt)
        c.acc_form[ri] = acceptanceProbability(r.bond, kT)
        c.acc_break[ri] = acceptanceProbability(r.brk, kT)
        c.req_cat[ri] = 1.0 if r.bond.get("requiresCatalyst") else 0.0
        c.roleA[ri], c.roleB[ri] = r.slotRoleA, r.slotRoleB
    c.catalyst_kind = R.catalyst_kind
    c.head_terminal_only = 1.0 if soup["headPlacement"]["terminalOnly"] else 0.0
    c.head_chain_cap = soup["headPlacement"]["chainCapacity"]
    c.max_hold_steps = soup["adsorption"]["maxHoldSteps"]
    c.bond_interval_f = interval
    c.bond_interval = interval
    stage = int((opts.get("verletOverride") or {}).get("stageCapacity", 0))
    if stage < 0:
        raise ValueError(f"verletOverride.stageCapacity={stage} must be >= 0 (0: the library's limit)")
    c.list_stage = stage
    if R.rebuild["mode"] == "drift":
        c.rebuild_mode = native.REBUILD_DRIFT
        c.drift_check_every = R.rebuild["checkEvery"]
        c.drift_trigger = R.rebuild["trigger"]
    return c


def _sweep_dict(r: "native.SoupSweepResult") -> Dict[str, Any]:
    return {"attempts": int(r.attempts), "accepted": int(r.accepted), "heads": int(r.heads),
            "deprotonated": int(r.deprotonated), "alpha": float(r.alpha), "dEsMeanKT": float(r.dEsMeanKT)}


class SoupSystem:
    """The reference's SoupSystem (soup/src/soup-types.ts:181-442) for the step path."""

    def __init__(self, R: Resolved, handle, lib):
        self._R = R
        self._h = handle
        self._lib = lib
        self._lock = threading.Lock()
        self.N = R.N

    # -- plumbing --------------------------------------------------------------------------------
    def _call(self, fn, *args):
        if self._h is None:
            raise RuntimeError("SoupSystem: disposed")
        if not self._lock.acquire(blocking=False):
            raise RuntimeError("SoupSystem: concurrent calls on one system are not allowed")
        try:
            native.check(self._lib, self._h, fn(self._h, *args))
        finally:
            self._lock.release()

    def _read(self, field: str, dtype) -> np.ndarray:
        nbytes = self._lib.soup_field_bytes(self._h, FIELD[field])
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        if nbytes:
            self._call(self._lib.soup_read, FIELD[field], out.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes))
        return out

    def _write(self, field: str, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        self._call(self._lib.soup_write, FIELD[field], arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes))

    # -- stepping --------------------------------------------------------------------------------
    def step(self, n: int) -> None:
        """n integration steps (soup/src/soup-integrate.ts:136-175)."""
        self._call(self._lib.soup_step, C.c_uint32(int(n)))

    def useGraphs(self, enable: bool = True) -> None:
        """Replay whole step periods from captured HIP graphs (on by default). The results are
        bit-identical either way; off, every step is launched kernel by kernel."""
        self._call(self._lib.soup_use_graphs, C.c_int(1 if enable else 0))

    @property
    def steps(self) -> int:
        v = C.c_uint64()
        native.check(self._lib, self._h, self._lib.soup_global_step(self._h, C.byref(v)))
        return int(v.value)

    @property
    def box(self):
        return list(self._R.live_box)

    @property
    def cyclePhase(self) -> str:
        """'wet' / 'dry' on a dry-wet cycling system, 'none' otherwise (soup-types.ts:387-404)."""
        return self._R.cycle_phase

    @property
    def cycleIndex(self) -> int:
        """1-based cycle of the current phase; 0 when cycling is off or over."""
        return self._R.cycle_index

    def stepCycled(self, n: int) -> None:
        """n steps through the dry-wet schedule (soup-box-scale.ts:478-532); a plain step(n) when the
        system does not cycle."""
        stepCycled(self, n)

    def scaleBoxTo(self, targetBox, rampSteps: int, rampRelaxSteps: int, cohesion: Optional[Dict[str, Any]] = None) -> None:
        """A ramped rigid-molecule box change to targetBox (soup-box-scale.ts:266-275); `cohesion` =
        {cutoff, memberIndices?} makes proximity-connected clusters scale as one rigid body."""
        applyBoxScale(self, targetBox, int(rampSteps), int(rampRelaxSteps), cohesion)

    def growBoxTo(self, targetBox, rampSteps: int, rampRelaxSteps: int) -> None:
        """Grows the box to targetBox without moving any particle (soup-box-scale.ts:318-373)."""
        growBoxTo(self, targetBox, int(rampSteps), int(rampRelaxSteps))

    def evaporateDEBUG(self, targetSolvent: int) -> None:
        """Solvent removal to targetSolvent live solvent beads at the current box (sim.ts:569-572)."""
        from .evaporate import evaporateSolventTo

        if self._R.evap is None:
            raise ValueError("evaporateDEBUG: the system was created without solvent evaporation (CreateSoupOpts.evaporateSolvent)")
        evaporateSolventTo(self, self._R.evap, int(targetSolvent))

    def rehydrateDEBUG(self, targetSolvent: int) -> Dict[str, Any]:
        """Solvent insertion back to targetSolvent at the current box, with its report (sim.ts:573-576)."""
        from .evaporate import rehydrateSolventTo

        if self._R.evap is None:
            raise ValueError("rehydrateDEBUG: the system was created without solvent evaporation (CreateSoupOpts.evaporateSolvent)")
        return rehydrateSolventTo(self, self._R.evap, int(targetSolvent), self._R.live_box, self._R.seed)

    # -- readbacks (soup/src/soup-readback.ts) ----------------------------------------------------
    def particles(self) -> np.ndarray:
        return self._read("POS", np.float32)

    def velocities(self) -> np.ndarray:
        return self._read("VEL", np.float32)

    def forces(self) -> np.ndarray:
        self._call(self._lib.soup_forces, native.FORCES_REBUILD)
        return self._read("FORCE", np.float32)

    def forcesBruteForce(self) -> np.ndarray:
        self._call(self._lib.soup_forces, native.FORCES_BRUTE)
        return self._read("FORCE", np.float32)

    def forcesNoRebuildDEBUG(self) -> np.ndarray:
        return self._read("FORCE", np.float32)

    def bondSlots(self) -> np.ndarray:
        return self._read("BOND_SLOTS", np.uint32)

    def bonds(self) -> np.ndarray:
        """soup-readback.ts:72-82: (i, j) pairs with j > i in (i, slot) order."""
        s = self.bondSlots().reshape(-1, 3).astype(np.int64)
        i = np.repeat(np.arange(s.shape[0]), 3)
        j = s.reshape(-1)
        keep = (j != NONE_U32) & (j > i)
        return np.stack([i[keep], j[keep]], axis=1).reshape(-1).astype(np.uint32)

    def centerLinks(self) -> np.ndarray:
        return self._read("CENTER_LINK", np.uint32)

    def centerHeldSteps(self) -> np.ndarray:
        return self._read("CENTER_HELD", np.uint32)

    def frozen(self) -> np.ndarray:
        return self._read("FROZEN", np.uint32)

    def charges(self) -> np.ndarray:
        return self._read("CHARGE", np.float32)

    def rngState(self) -> Dict[str, np.ndarray]:
        return {"bond": self._read("BOND_RNG", np.uint32), "thermo": self._read("THERMO_RNG", np.uint32)}

    def events(self) -> Dict[str, int]:
        ev = self._read("EVENTS", np.uint32)
        out = {}
        for ri, (bid, kid) in enumerate(self._R.event_rule_ids):
            out[bid] = int(ev[2 * ri])
            out[kid] = int(ev[2 * ri + 1])
        return out

    def desorbEvents(self) -> Dict[str, int]:
        d = self._read("DESORB", np.uint32)
        return {"stretch": int(d[0]), "timeout": int(d[1])}

    def invariants(self) -> Dict[str, Any]:
        pos = self.particles().reshape(-1, 4)
        soup = self._R.soup
        kinds = np.rint(pos[:, 3]).astype(np.int64)
        monomers = {m["id"]: int(np.count_nonzero(kinds == k)) for k, m in enumerate(soup["monomers"])}
        charge = 0.0
        for k, m in enumerate(soup["monomers"]):
            charge += float(m.get("charge", 0)) * monomers[m["id"]]
        return {"monomers": monomers, "bonds": int(self.bonds().size // 2), "charge": charge}

    def nonFiniteCount(self) -> Dict[str, int]:
        out = (C.c_uint32 * 2)()
        self._call(self._lib.soup_nonfinite, out)
        return {"pos": int(out[0]), "vel": int(out[1])}

    # -- electrostatics -----------------------------------------------------------------------------
    def electrostatics(self) -> Dict[str, Any]:
        d = dict(vars(self._R.es))
        d["acidSoap"] = dict(vars(self._R.es.acidSoap))
        n = C.c_uint32()
        r = native.SoupSweepResult()
        native.check(self._lib, self._h, self._lib.soup_last_sweep(self._h, C.byref(n), C.byref(r)))
        d["sweeps"] = int(n.value)
        d["last"] = _sweep_dict(r) if n.value else None
        return d

    def protonationSweepDEBUG(self) -> Optional[Dict[str, Any]]:
        if not self._R.es.enabled:
            return None
        r = native.SoupSweepResult()
        self._call(self._lib.soup_protonation_sweep, C.byref(r))
        return _sweep_dict(r)

    def protonationRngState(self) -> int:
        v = C.c_uint32()
        native.check(self._lib, self._h, self._lib.soup_protonation_rng(self._h, C.byref(v)))
        return int(v.value)

    # -- lists ----------------------------------------------------------------------------------
    def verletConfig(self) -> Dict[str, Any]:
        return dict(self._R.verlet, listRange=self._R.geo.listRange, walkRadius=self._R.geo.effectiveWalkRadius,
                    listBytes=verletListBytes(self._R.capacity_n, self._R.list_cap))

    def potentialBasis(self, centerLink: bool = True, charges: bool = True) -> native.SoupPotentialBasis:
        """makePotentialBasis for this system's species, parameters and DH basis."""
        from .potential import makePotentialBasis

        o = self._R.opts
        return makePotentialBasis(self._R.soup, self._R.p, self._R.es, o.get("solventAttractionScaleOverride"),
                                  o.get("claySurfaceChemistry"), o.get("acidSoapScaleOverride"), centerLink, charges)

    def potential(self, basis: Optional[native.SoupPotentialBasis] = None, centerLink: bool = True,
                  charges: bool = True) -> Dict[str, float]:
        """soupPotential (soup/src/soup-potential.ts:145-301) of the current configuration, on the GPU:
        {nonbonded, fene, bend, tether, total} in FP64."""
        b = basis if basis is not None else self.potentialBasis(centerLink, charges)
        out = (C.c_double * 5)()
        self._call(self._lib.soup_potential, C.byref(b), out)
        return {"nonbonded": out[0], "fene": out[1], "bend": out[2], "tether": out[3], "total": out[4]}

    def listOccupancyDEBUG(self) -> Dict[str, Any]:
        s = native.SoupListStats()
        self._call(self._lib.soup_list_stats, C.byref(s))
        return {"verletMean": s.verlet_mean, "verletMax": s.verlet_max, "verletCap": s.verlet_cap,
                "esMean": s.es_mean, "esMax": s.es_max, "esCap": s.es_cap,
                "verletPairs": int(s.verlet_pairs), "esPairs": int(s.es_pairs)}

    def rebuildCount(self) -> int:
        """Verlet rebuilds the steps of this system made since creation (forced ones not counted)."""
        return int(self._read("REBUILDS", np.uint64)[0])

    def verletList(self) -> np.ndarray:
        return self._read("VERLET_LIST", np.uint32).reshape(self.N, self._R.list_cap)

    def verletCount(self) -> np.ndarray:
        return self._read("VERLET_COUNT", np.uint32)

    # -- zero-tension area move (soup/src/soup-area-move.ts:167-279) -------------------------------
    def areaMove(self, trials: int, opts: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
        """`trials` Metropolis box proposals (mode from data/soup.json's areaMove unless opts['mode']);
        the trial energies are soupPotential on the GPU, the accepted box is applied once at the end.
        Returns {trials, accepted, acceptedFraction, box, lateralTrajectory, energyStart, energyEnd}."""
        from .area_move import SoupAreaMove

        if getattr(self, "_area", None) is None:
            self._area = SoupAreaMove(self)
        return self._area(trials, opts)

    @property
    def headCharge(self) -> float:
        """chargeDeprotonated: the charge of a deprotonated head (every charge is 0 or this)."""
        return float(self._R.cfg.ph_q)

    def energyAdjoint(self, basis: Optional[native.SoupPotentialBasis] = None, w: Optional[np.ndarray] = None) -> Dict[str, Any]:
        """Gradients of soupPotential at the current configuration (protocell_soup.adjoint)."""
        from .adjoint import energyAdjoint

        return energyAdjoint(self, basis, w)

    def potentialAt(self, positions: np.ndarray, box, basis: Optional[native.SoupPotentialBasis] = None) -> Dict[str, float]:
        """soupPotential of an arbitrary configuration of this system (positions in particle order, its
        own periodic box), with the system's bonds, centre links and charges, on the GPU."""
        b = basis if basis is not None else self.potentialBasis()
        out = (C.c_double * 5)()
        bx = (C.c_double * 3)(*[float(x) for x in box])
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        if pos.size != 4 * self.N:
            raise ValueError(f"potentialAt: positions hold {pos.size // 4} particles, the system {self.N}")
        self._call(self._lib.soup_potential_at, C.byref(b), pos.ctypes.data_as(C.c_void_p), bx, out)
        return {"nonbonded": out[0], "fene": out[1], "bend": out[2], "tether": out[3], "total": out[4]}

    # -- cold start (soup/src/soup-relax.ts:141-185) ---------------------------------------------
    def relaxColdStart(self, opts: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
        opts = opts or {}
        cfg = self._R.soup.get("coldStartRelax")
        if not cfg and (opts.get("iterations") is None or opts.get("maxDisplacementSigma") is None):
            raise ValueError("relaxColdStart: data/soup.json has no coldStartRelax section, and the call did not set iterations and maxDisplacementSigma explicitly")
        iterations = opts.get("iterations", cfg["iterations"] if cfg else None)
        d_max = opts.get("maxDisplacementSigma", cfg["maxDisplacementSigma"] if cfg else None)
        if not isinstance(iterations, int) or iterations < 1:
            raise ValueError(f"relaxColdStart: iterations={iterations} must be an integer >= 1")
        if not (d_max > 0):
            raise ValueError(f"relaxColdStart: maxDisplacementSigma={d_max} must be positive")
        if self.steps != 0:
            raise ValueError(
                f"relaxColdStart: the system is already at step {self.steps} -- minimisation is allowed only before the first step, "
                "otherwise it lands inside the trajectory on which the measurements are taken"
            )
        d0 = d_max * self._R.p["sigma"]

        def stats(f):
            f3 = f.reshape(-1, 4)[:, :3]
            fin = np.isfinite(f3)
            return float(np.abs(f3[fin]).max(initial=0.0)), int((~fin).sum())

        before = stats(self.forces())
        self._call(self._lib.soup_relax, C.c_uint32(iterations), C.c_double(d0))
        after = stats(self.forces())
        return {"iterations": iterations, "maxDisplacementStart": d0, "displacementBound": d0 * (iterations + 1) / 2,
                "maxForceBefore": before[0], "maxForceAfter": after[0], "nonFiniteBefore": before[1], "nonFiniteAfter": after[1]}

    # -- profiling -----------------------------------------------------------------------------------
    def profile(self, enable: bool, phases=None) -> None:
        """Starts (resets) or stops per-phase device timing; ``phases`` limits it to the named phases."""
        if not enable:
            mask = 0
        elif phases is None:
            mask = -1
        else:
            mask = 0
            for ph in phases:
                mask |= 1 << native.PROF_PHASES.index(ph)
        self._call(self._lib.soup_profile, C.c_int(mask))

    def profileRead(self) -> Dict[str, Dict[str, float]]:
        ms = (C.c_double * 16)()
        cnt = (C.c_uint64 * 16)()
        self._call(self._lib.soup_profile_read, ms, cnt)
        return {name: {"ms": float(ms[k]), "count": int(cnt[k])} for k, name in enumerate(native.PROF_PHASES)}

    # -- test hooks --------------------------------------------------------------------------------
    def runPhasesDEBUG(self, *phases: str) -> None:
        mask = 0
        for ph in phases:
            mask |= native.PHASE[ph]
        self._call(self._lib.soup_run_phases, C.c_uint32(mask), C.c_uint32(0))

    def stepPhasesDEBUG(self, n: int) -> Dict[str, float]:
        """Device milliseconds per iteration of each phase group run n times (soup-readback.ts:191-239):
        full (a step with the rebuild and the bond phase), gridBuild (the rebuild: sort, DH list,
        Verlet list), force (the list force), bondAttempts, integration. The step counter does not
        move; the lists and F(x) are rebuilt at the end."""
        ms = (C.c_double * 5)()
        self._call(self._lib.soup_step_phases, C.c_uint32(int(n)), ms)
        return dict(zip(("full", "gridBuild", "force", "bondAttempts", "integration"), (float(v) for v in ms)))

    def forceCandidateStatsDEBUG(self) -> Dict[str, Any]:
        """The cell walk's candidate count and the pairs within range at the current positions
        (soup-readback.ts:246-263); the state is not touched."""
        out = (C.c_uint64 * 2)()
        self._call(self._lib.soup_force_candidate_stats, out)
        cand, pairs = int(out[0]), int(out[1])
        return {"candidatesExamined": cand, "pairsWithinRange": pairs, "ratio": cand / max(1, pairs)}

    def minimiseNowDEBUG(self, iterations: int) -> Dict[str, Any]:
        """The cold-start minimiser applied mid-run at the current box (sim.ts:518-528), without
        relaxColdStart's step-0 refusal."""
        d0 = (self._R.soup.get("coldStartRelax") or {}).get("maxDisplacementSigma", 0.1) * self._R.p["sigma"]
        n = self.N

        def mx(f):
            f3 = f.reshape(-1, 4)[:n, :3]
            return float(np.abs(f3).max(initial=0.0))  # NaN propagates, as Math.max does

        before = mx(self.forces())
        self._call(self._lib.soup_relax, C.c_uint32(int(iterations)), C.c_double(d0))
        after = mx(self.forces())
        return {"iterations": int(iterations), "globalStep": self.steps, "maxForceBefore": before, "maxForceAfter": after,
                "displacementBound": d0 * (int(iterations) + 1) / 2}

    def runUntil(self, stage: str, opts: Dict[str, int], log=None) -> Dict[str, Any]:
        """sim.ts:452-484: stepCycled in batches of sampleEvery, detectStage after each batch; stops at
        the first batch whose stage is `stage`, else after maxSteps. `steps` counts this call's
        batches; each trace entry also carries the cycle phase and index."""
        from .analysis import AnalysisContext, detectStage

        trace = []
        steps, reached = 0, False
        max_steps, every = int(opts["maxSteps"]), int(opts["sampleEvery"])
        ctx = AnalysisContext()
        try:
            while steps < max_steps:
                chunk = min(every, max_steps - steps)
                self.stepCycled(chunk)
                steps += chunk
                current, evidence = detectStage(self, ctx)
                entry = {"steps": steps, "stage": current, "evidence": evidence, "cyclePhase": self.cyclePhase,
                         "cycleIndex": self.cycleIndex}
                trace.append(entry)
                if log is not None:
                    log(entry)
                if current == stage:
                    reached = True
                    break
        finally:
            ctx.close()
        return {"reached": reached, "steps": steps, "trace": trace}

    def writeDEBUG(self, field: str, arr: np.ndarray) -> None:
        self._write(field, arr)

    def readDEBUG(self, field: str, dtype=np.uint32) -> np.ndarray:
        return self._read(field, dtype)

    def dispose(self) -> None:
        if self._h is not None:
            self._lib.soup_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass


def stepReplicas(systems: Sequence[SoupSystem], n: int) -> None:
    """n steps of every system, queued together on their own streams (soup_step_batch). Each
    system ends exactly where system.step(n) would leave it; this is the batched form of a replica
    sweep's per-replica loop. On a guard failure the error names the failing replica."""
    systems = list(systems)
    if not systems:
        return
    lib = systems[0]._lib
    held = []
    try:
        for s in systems:
            if s._h is None:
                raise RuntimeError("stepReplicas: a system is disposed")
            if not s._lock.acquire(blocking=False):
                raise RuntimeError("stepReplicas: concurrent calls on one system are not allowed (or a system is listed twice)")
            held.append(s)
        hs = (C.c_void_p * len(systems))(*[s._h for s in systems])
        failed = C.c_uint32(0)
        r = lib.soup_step_batch(hs, C.c_uint32(len(systems)), C.c_uint32(int(n)), C.byref(failed))
        if r != 0:
            k = int(failed.value)
            msg = lib.soup_last_error(systems[k]._h) if k < len(systems) else None
            raise SoupError(f"replica {k}: " + (msg.decode() if msg else f"libsoup_hip error {r}"), r)
    finally:
        for s in held:
            s._lock.release()


def createSoup(opts: Dict[str, Any]) -> SoupSystem:
    """soup/src/sim.ts:95 createSoup(opts) -> SoupSystem (synchronous here)."""
    R = resolve(opts)
    lib = native.load_library()
    init = R.init
    keep = {k: np.ascontiguousarray(v) for k, v in init.items()}
    charges = np.ascontiguousarray(R.charges0, dtype=np.float32)
    events = np.ascontiguousarray(R.events0, dtype=np.uint32)
    st = native.SoupInitState()
    st.positions = keep["positions"].ctypes.data
    st.velocities = keep["velocities"].ctypes.data
    st.bond_slots = keep["bondSlots"].ctypes.data
    st.center_link = keep["centerLink"].ctypes.data
    st.center_held = keep["centerHeldSteps"].ctypes.data
    st.frozen = keep["frozen"].ctypes.data
    st.bond_rng = keep["bondRng"].ctypes.data
    st.thermo_rng = keep["thermoRng"].ctypes.data
    st.charges = charges.ctypes.data
    st.events = events.ctypes.data
    st.desorb_events = keep["desorbEvents"].ctypes.data
    st.global_step = R.initial_step
    st.next_sweep_at = min(R.next_sweep_at, 2 ** 64 - 1)
    st.protonation_rng = R.prng & 0xFFFFFFFF
    h = C.c_void_p()
    code = lib.soup_create(C.byref(R.cfg), C.byref(st), C.byref(h))
    if code != 0:
        msg = lib.soup_last_error(None)
        raise SoupError(msg.decode() if msg else f"soup_create failed ({code})", code)
    return SoupSystem(R, h, lib)
