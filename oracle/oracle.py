"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker and by
bench.py's cpu_baseline leg. The product (protocell-genesis_amd) never imports this module.

``OracleState`` holds the reference-layout state (particle-index order, Verlet list rows of
list_cap entries) as numpy arrays and exposes one method per restated reference function; each
C function names the reference file:line it follows (oracle/soup_oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
NONE = 0xFFFFFFFF


def build() -> str:
    """Compiles liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.or_check_lists.restype = C.c_int
        _lib.or_step.restype = C.c_int
        _lib.or_relax.restype = C.c_int
        _lib.or_energy_f64.restype = C.c_double
        _lib.or_pcg_next.restype = C.c_double
        _lib.or_pcg_soup.restype = C.c_uint32
        _lib.or_pcg_soup.argtypes = [C.c_uint32]
        _lib.or_cell_of.restype = C.c_uint32
        _lib.or_initial_charges.restype = C.c_uint32
        _lib.or_verlet_build.restype = C.c_uint32
        _lib.or_es_build.restype = C.c_uint32
        _lib.or_threads.restype = C.c_int
        _lib.or_relax_step.argtypes = [C.c_void_p, C.c_void_p, C.c_float]
        _lib.or_relax.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_double]
    return _lib


class _OrState(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "pos", "vel", "force", "charge", "bond_slots", "center_link", "center_held", "frozen", "bond_rng", "thermo_rng",
        "events", "desorb", "cell_of", "cell_start", "cell_idx", "vl", "vcount", "pos_at_rebuild", "head_idx", "es_list",
        "es_count")] + [
        ("nh", C.c_uint32), ("v_overflow", C.c_uint32), ("es_overflow", C.c_uint32), ("max_drift_bits", C.c_uint32),
        ("prng", C.c_uint32), ("pad", C.c_uint32), ("global_step", C.c_uint64), ("next_sweep_at", C.c_uint64)]


class SweepResult(C.Structure):
    _fields_ = [("attempts", C.c_uint32), ("accepted", C.c_uint32), ("heads", C.c_uint32), ("deprotonated", C.c_uint32),
                ("alpha", C.c_double), ("dEsMeanKT", C.c_double)]


class OracleState:
    """Reference-layout state + the restated kernels. ``cfg`` is the product's SoupConfig (the ABI)."""

    def __init__(self, cfg, init: Dict[str, np.ndarray], charges: Optional[np.ndarray] = None, global_step: int = 0,
                 next_sweep_at: int = 2 ** 64 - 1, prng: int = 0, events=None, desorb=None):
        self.cfg = cfg
        n = int(cfg.n)
        self.n = n
        ncells = int(cfg.dims[0] * cfg.dims[1] * cfg.dims[2])
        f = lambda a, cnt: np.ascontiguousarray(np.array(a, dtype=np.float32).reshape(-1)[:cnt])  # noqa: E731
        u = lambda a, cnt: np.ascontiguousarray(np.array(a, dtype=np.uint32).reshape(-1)[:cnt])  # noqa: E731
        self.pos = f(init["positions"], 4 * n)
        self.vel = f(init["velocities"], 4 * n)
        self.force = np.zeros(4 * n, dtype=np.float32)
        self.charge = f(charges if charges is not None else np.zeros(n), n)
        self.bond_slots = u(init["bondSlots"], 3 * n)
        self.center_link = u(init["centerLink"], n)
        self.center_held = u(init["centerHeldSteps"], n)
        self.frozen = u(init.get("frozen", np.zeros(n)), n)
        self.bond_rng = u(init["bondRng"], n)
        self.thermo_rng = u(init["thermoRng"], n)
        self.events = np.zeros(8, dtype=np.uint32)
        if events is not None:
            self.events[: len(events)] = events
        self.desorb = np.zeros(2, dtype=np.uint32)
        if desorb is not None:
            self.desorb[:] = desorb
        self.cell_of = np.zeros(n, dtype=np.uint32)
        self.cell_start = np.zeros(ncells + 1, dtype=np.uint32)
        self.cell_idx = np.zeros(n, dtype=np.uint32)
        self.vl = np.full(n * int(cfg.list_cap), NONE, dtype=np.uint32)
        self.vcount = np.zeros(n, dtype=np.uint32)
        self.pos_at_rebuild = self.pos.copy()
        nh = max(1, int(cfg.n_heads))
        self.head_idx = np.zeros(nh, dtype=np.uint32)
        self.es_list = np.full(nh * max(1, int(cfg.es_cap)), NONE, dtype=np.uint32)
        self.es_count = np.zeros(nh, dtype=np.uint32)
        self.s = _OrState()
        for name, _ in _OrState._fields_[:21]:
            setattr(self.s, name, getattr(self, name).ctypes.data)
        self.s.global_step = global_step
        self.s.next_sweep_at = next_sweep_at
        self.s.prng = prng & 0xFFFFFFFF
        self.last_sweep = SweepResult()
        self.sweeps = C.c_uint32(0)

    # -- pointers ------------------------------------------------------------------------------
    @property
    def _c(self):
        return C.byref(self.cfg)

    @property
    def _s(self):
        return C.byref(self.s)

    @property
    def global_step(self) -> int:
        return int(self.s.global_step)

    # -- restated functions ----------------------------------------------------------------------
    def grid_build(self):
        lib().or_grid_build(self._c, self._s)

    def verlet_build(self) -> int:
        return int(lib().or_verlet_build(self._c, self._s))

    def es_build(self) -> int:
        return int(lib().or_es_build(self._c, self._s))

    def rebuild(self):
        lib().or_rebuild(self._c, self._s)

    def force_list(self):
        lib().or_force_list(self._c, self._s)

    def force_brute(self):
        lib().or_force_brute(self._c, self._s)

    def force_candidate_stats(self):
        """(candidates examined, pairs within range) of the cell walk at the current positions."""
        out = (C.c_uint64 * 2)()
        lib().or_force_candidate_stats(self._c, self._s, out)
        return int(out[0]), int(out[1])

    def force_f64(self, use_list: bool = False) -> np.ndarray:
        out = np.zeros(3 * self.n, dtype=np.float64)
        lib().or_force_f64(self._c, self._s, C.c_int(1 if use_list else 0), out.ctypes.data_as(C.c_void_p))
        return out.reshape(-1, 3)

    def energy_f64(self) -> float:
        return float(lib().or_energy_f64(self._c, self._s))

    def kick_drift(self):
        lib().or_kick_drift(self._c, self._s)

    def kick_thermo(self):
        lib().or_kick_thermo(self._c, self._s)

    def max_drift(self):
        lib().or_max_drift(self._c, self._s)

    def bond_form(self):
        lib().or_bond_form(self._c, self._s)

    def bond_break(self):
        lib().or_bond_break(self._c, self._s)

    def one_step(self):
        lib().or_one_step(self._c, self._s)

    def step(self, nsteps: int) -> int:
        return int(lib().or_step(self._c, self._s, C.c_uint32(nsteps), C.byref(self.last_sweep), C.byref(self.sweeps)))

    def relax(self, iterations: int, d0: float) -> int:
        return int(lib().or_relax(self._c, self._s, C.c_uint32(iterations), C.c_double(d0)))

    def nonfinite(self):
        out = (C.c_uint32 * 2)()
        lib().or_nonfinite(self._c, self._s, out)
        return int(out[0]), int(out[1])

    def check_lists(self) -> int:
        return int(lib().or_check_lists(self._c, self._s))

    def protonation_sweep(self) -> dict:
        r = SweepResult()
        st = C.c_uint32(self.s.prng)
        lib().or_protonation_sweep(self._c, self.pos.ctypes.data_as(C.c_void_p), self.charge.ctypes.data_as(C.c_void_p),
                                   C.c_uint32(self.n), C.byref(st), C.byref(r))
        self.s.prng = st.value
        return {"attempts": r.attempts, "accepted": r.accepted, "heads": r.heads, "deprotonated": r.deprotonated,
                "alpha": r.alpha, "dEsMeanKT": r.dEsMeanKT}

    def max_drift_sq(self) -> float:
        return float(np.array([self.s.max_drift_bits], dtype=np.uint32).view(np.float32)[0])

    def verlet_rows(self) -> np.ndarray:
        return self.vl.reshape(self.n, -1)

    def load(self, positions, bond_slots, center_link, charges):
        """Replaces positions and topology in place (particle order) and rebuilds the lists."""
        self.pos[:] = np.asarray(positions, dtype=np.float32).reshape(-1)
        self.bond_slots[:] = np.asarray(bond_slots, dtype=np.uint32).reshape(-1)
        self.center_link[:] = np.asarray(center_link, dtype=np.uint32).reshape(-1)
        self.charge[:] = np.asarray(charges, dtype=np.float32).reshape(-1)
        self.rebuild()

    def energy_adjoint(self, basis, box, q_head: float, w=None, strain=None) -> dict:
        """FP64 gradient, Hessian-vector product (direction w plus diag(strain) r per pair term) and
        parameter derivatives of soupPotential over the state's current lists (soup_oracle.c
        or_energy_adjoint): {grad, hw, dU, dFw, We}."""
        n = self.n
        g = np.zeros(3 * n)
        hw = np.zeros(3 * n)
        dU = np.zeros(14)
        dF = np.zeros(14)
        we = C.c_double(0.0)
        bx = np.ascontiguousarray(box, dtype=np.float64)
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        st = None if strain is None else np.ascontiguousarray(strain, dtype=np.float64)
        lib().or_energy_adjoint(self._s, C.c_uint32(n), C.c_uint32(int(self.cfg.list_cap)), C.c_uint32(int(self.cfg.es_cap)),
                                C.byref(basis), bx.ctypes.data_as(C.c_void_p), C.c_double(q_head),
                                None if wa is None else wa.ctypes.data_as(C.c_void_p),
                                None if st is None else st.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p),
                                hw.ctypes.data_as(C.c_void_p), dU.ctypes.data_as(C.c_void_p), dF.ctypes.data_as(C.c_void_p),
                                C.byref(we))
        return {"grad": g.reshape(n, 3), "hw": hw.reshape(n, 3), "dU": dU, "dFw": dF, "We": we.value}


def segment_gradient_f64(state: "OracleState", basis, box, q_head: float, dt: float, gamma: float, start, records, swept,
                         bond_interval: int, start_step: int, objective: str = "energy", strain=None, scale: float = 1.0):
    """CPU restatement, in FP64, of the product's trajectory-segment adjoint (soup_segment_adjoint) over a
    recorded trajectory: ``start`` = (positions, bond slots, centre links, charges) before the first
    step, ``records[s]`` the same after step s (charges after a sweep where swept[s]), K = len(records).
    The reverse sweep through the velocity-Verlet Langevin step takes one or_energy_adjoint per force
    evaluation; objective "energy" is U of the last record, "virial" the mean over records of
    scale * W_e(strain). Returns {J, dJ_dtheta, dJ_dx0, dJ_dv0}. Test infrastructure: the check of the
    product's gradient, not a restatement of a reference function."""
    K = len(records)
    n = state.n
    a = 1.0 - gamma * dt
    st0 = np.zeros(3)
    obj = None if objective == "energy" else np.asarray(strain, dtype=np.float64) * scale / K

    def vjp(P, T, w, stv):
        state.load(P[0], T[1], T[2], T[3])
        return state.energy_adjoint(basis, box, q_head, w=w, strain=stv)

    lx = np.zeros((n, 3))
    lv = np.zeros((n, 3))
    lth = np.zeros(14)
    J = 0.0
    if objective == "energy":
        r = vjp(records[-1], records[-1], None, None)
        J = soup_potential(basis, records[-1][0], records[-1][1], records[-1][2], records[-1][3], box)[4]
        lx += r["grad"]
        lth += r["dU"]
    for s in range(K - 1, -1, -1):
        P = records[s]
        T0 = start if s == 0 else records[s - 1]
        last = s == K - 1
        cw = a * dt / 2 + (0.0 if last else dt / 2)
        done = obj is None
        if not last and swept[s]:
            r = vjp(P, P, (dt / 2) * lv, obj if obj is not None else st0)
            lx -= r["hw"] + (obj if obj is not None else st0) * r["grad"]
            lth += r["dFw"]
            J += r["We"]
            cw = a * dt / 2
            done = True
        if not done and (start_step + s) % bond_interval == 0:
            r = vjp(P, P, None, obj)
            lx -= r["hw"] + obj * r["grad"]
            lth += r["dFw"]
            J += r["We"]
            done = True
        stv = st0 if done else obj
        r = vjp(P, T0, cw * lv, stv)
        lx -= r["hw"] + stv * r["grad"]
        lth += r["dFw"]
        J += r["We"]
        lv = a * lv + dt * lx
    r = vjp(start, start, (dt / 2) * lv, st0)
    lx -= r["hw"]
    lth += r["dFw"]
    return {"J": J, "dJ_dtheta": lth, "dJ_dx0": lx, "dJ_dv0": lv}


def pcg_soup(v: int) -> int:
    return int(lib().or_pcg_soup(C.c_uint32(v)))


def pcg_next(state: int):
    """(value, new_state) of the CPU protonation stream (soup/src/electrostatics.ts:308-315)."""
    s = C.c_uint32(state)
    v = lib().or_pcg_next(C.byref(s))
    return float(v), int(s.value)


def lattice(box, counts, seed: int, kT: float) -> Dict[str, np.ndarray]:
    """The oracle's scalar restatement of buildInitialState (soup/src/soup-init-state.ts:58-307)."""
    n = int(sum(counts))
    pos = np.zeros(4 * n, dtype=np.float32)
    vel = np.zeros(4 * n, dtype=np.float32)
    bs = np.zeros(3 * n, dtype=np.uint32)
    cl = np.zeros(n, dtype=np.uint32)
    br = np.zeros(n, dtype=np.uint32)
    tr = np.zeros(n, dtype=np.uint32)
    b = (C.c_double * 3)(*box)
    cnt = (C.c_uint32 * len(counts))(*counts)
    lib().or_lattice(b, cnt, C.c_uint32(len(counts)), C.c_uint32(seed), C.c_double(kT), pos.ctypes.data_as(C.c_void_p),
                     vel.ctypes.data_as(C.c_void_p), bs.ctypes.data_as(C.c_void_p), cl.ctypes.data_as(C.c_void_p),
                     br.ctypes.data_as(C.c_void_p), tr.ctypes.data_as(C.c_void_p))
    return {"positions": pos, "velocities": vel, "bondSlots": bs, "centerLink": cl, "bondRng": br, "thermoRng": tr}


def initial_charges(cfg, positions: np.ndarray, seed: int):
    n = positions.size // 4
    ch = np.zeros(n, dtype=np.float32)
    rng = lib().or_initial_charges(C.byref(cfg), np.ascontiguousarray(positions, dtype=np.float32).ctypes.data_as(C.c_void_p),
                                   C.c_uint32(n), C.c_uint32(seed), ch.ctypes.data_as(C.c_void_p))
    return ch, int(rng)


def soup_potential(basis, positions, bond_slots, center_link, charges, box):
    """soupPotential (soup/src/soup-potential.ts:145-301) in FP64: (nonbonded, fene, bend, tether, total).
    ``basis`` is the product's SoupPotentialBasis (the ABI struct)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    n = pos.size // 4
    bs = np.ascontiguousarray(bond_slots, dtype=np.uint32).reshape(-1)
    cl = None if center_link is None else np.ascontiguousarray(center_link, dtype=np.uint32)
    ch = None if charges is None else np.ascontiguousarray(charges, dtype=np.float32)
    bx = np.ascontiguousarray(box, dtype=np.float64)  # the reference's box is a JS number
    out = (C.c_double * 5)()
    lib().or_soup_potential(C.byref(basis), C.c_uint32(n), pos.ctypes.data_as(C.c_void_p), bs.ctypes.data_as(C.c_void_p),
                            None if cl is None else cl.ctypes.data_as(C.c_void_p),
                            None if ch is None else ch.ctypes.data_as(C.c_void_p), bx.ctypes.data_as(C.c_void_p), out)
    return tuple(out)


def soup_potential_d(basis, positions, bond_slots, center_link, charges, box):
    """soup_potential with FP64 positions (x, y, z, kind per particle): for strain derivatives, whose
    scaled coordinates f32 would round."""
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
    n = pos.size // 4
    bs = np.ascontiguousarray(bond_slots, dtype=np.uint32).reshape(-1)
    cl = None if center_link is None else np.ascontiguousarray(center_link, dtype=np.uint32)
    ch = None if charges is None else np.ascontiguousarray(charges, dtype=np.float32)
    bx = np.ascontiguousarray(box, dtype=np.float64)
    out = (C.c_double * 5)()
    lib().or_soup_potential_d(C.byref(basis), C.c_uint32(n), pos.ctypes.data_as(C.c_void_p), bs.ctypes.data_as(C.c_void_p),
                              None if cl is None else cl.ctypes.data_as(C.c_void_p),
                              None if ch is None else ch.ctypes.data_as(C.c_void_p), bx.ctypes.data_as(C.c_void_p), out)
    return tuple(out)


def bond_draws() -> np.ndarray:
    """Per-rule bond Monte Carlo draw counters since the last reset: formation candidates reaching the
    attempt draw, attempts passed, accepts passed; bonds reaching the break draw, break attempts
    passed, break accepts passed."""
    out = np.zeros(4 * 6, dtype=np.uint64)
    lib().or_bond_draws(out.ctypes.data_as(C.c_void_p))
    return out.reshape(4, 6)


def bond_draws_reset() -> None:
    lib().or_bond_draws_reset()


def set_threads(t: int) -> None:
    lib().or_set_threads(C.c_int(t))


def threads() -> int:
    return int(lib().or_threads())
