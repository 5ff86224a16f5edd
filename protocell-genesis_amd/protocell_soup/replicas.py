"""Replica sweeps: many independent soups of uneven size, stepped together with stepReplicas.

A sweep replica is an amphiphile soup at bulk density 0.80: pre-bonded five-bead amphiphiles
(one O head, four C tail beads) in water, charged heads at the replica's pH and ionic strength, and
its own seed. The reference runs such sweeps one replica at a time (one page per system); here
every replica keeps its own handle and stream and soup_step_batch queues all of them together.

The amphiphile supply is a multiple of the closure threshold taken as a concentration: the
smallest closed vesicle needs 930 amphiphiles (docs/soup-to-vesicle-verdict.md:268), which in the
campaign's 76 sigma box is 930 / 76^3 amphiphiles per sigma^3. A replica of edge L at supply s holds
round(s * 930 * (L / 76)^3) amphiphiles. The reference's own threshold is box-independent, so its
multiples cannot be placed in the small boxes of a sweep; the concentration form keeps the same
ratio at the campaign box.
"""
from __future__ import annotations

import math
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

from .init_state import Mulberry32, amphiphile_soup

CLOSURE_THRESHOLD_AMPHIPHILES = 930
CAMPAIGN_EDGE = 76.0
MIN_EDGE = 14.735  # 5 cells of the 2.947 sigma interaction range: the smallest box the grid plans


def amphiphileSoupOpts(box: float, n_amph: int, tail: int, extra: Dict[str, int], seed: int = 1, kT: float = 1.0,
                       electrostatics: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
    """createSoup opts for a cubic soup of ``n_amph`` pre-bonded amphiphiles (1 head + ``tail``
    carbons) plus free beads ``extra`` (monomer id -> count), passed as a resume state."""
    from .rules import loadSoup

    soup = loadSoup()
    ids = [m["id"] for m in soup["monomers"]]
    head, carbon = ids.index("O"), ids.index("C")
    extra_k = {ids.index(k): v for k, v in extra.items()}
    st = amphiphile_soup([box] * 3, n_amph, tail, extra_k, head, carbon, len(ids), seed, kT)
    start = {k: 0 for k in ids}
    start["O"] += n_amph
    start["C"] += n_amph * tail
    for k, v in extra.items():
        start[k] += v
    resume = {"globalStep": 0, "liveBox": [box] * 3, "positions": st["positions"], "velocities": st["velocities"],
              "bondSlots": st["bondSlots"], "centerLink": st["centerLink"], "centerHeldSteps": st["centerHeldSteps"],
              "desorbEvents": st["desorbEvents"], "bondRng": st["bondRng"], "thermoRng": st["thermoRng"], "events": {}}
    opts = {"box": [box] * 3, "seed": seed, "kT": kT, "clay": False, "start": start, "resume": resume}
    if electrostatics:
        opts["electrostatics"] = electrostatics
    return opts


def sweepReplicaParams(count: int = 64, seed: int = 1000, edge: Sequence[float] = (MIN_EDGE, 40.0),
                       pH: Sequence[float] = (5.0, 9.0), ionic: Sequence[float] = (0.001, 0.1),
                       supply: Sequence[float] = (1.5, 4.8), density: float = 0.80) -> List[Dict[str, Any]]:
    """The replicas of a sweep: edge ~ U(edge), pH ~ U(pH), ionic strength log-uniform in ``ionic``,
    supply ~ U(supply) closure thresholds, replica r seeded with seed + r. Draws come from one
    mulberry32 stream seeded with ``seed``, four per replica."""
    rng = Mulberry32(seed & 0xFFFFFFFF)
    out = []
    for r in range(count):
        u = rng.take(4).astype(np.float64)
        L = edge[0] + (edge[1] - edge[0]) * float(u[0])
        ph = pH[0] + (pH[1] - pH[0]) * float(u[1])
        ion = math.exp(math.log(ionic[0]) + (math.log(ionic[1]) - math.log(ionic[0])) * float(u[2]))
        s = supply[0] + (supply[1] - supply[0]) * float(u[3])
        n = int(round(density * L ** 3))
        n_amph = int(round(s * CLOSURE_THRESHOLD_AMPHIPHILES * (L / CAMPAIGN_EDGE) ** 3))
        out.append({"edge": L, "N": n, "pH": ph, "ionicStrengthMolar": ion, "supply": s, "amphiphiles": n_amph,
                    "water": n - 5 * n_amph, "seed": seed + r})
    return out


def sweepReplicaOpts(params: Dict[str, Any], kT: float = 1.0) -> Dict[str, Any]:
    """createSoup opts of one replica from sweepReplicaParams. The long-range DH list is sized for
    the heads gathered into an eighth of the box (densityVolumeSigma3 = L^3 / 8): the amphiphiles
    aggregate within a few hundred steps, and the capacity from a uniform head density then
    overflows (367 heads in a 34.7 sigma box at rc_es = 9.1 sigma outgrew the uniform capacity of 71)."""
    L = params["edge"]
    return amphiphileSoupOpts(L, params["amphiphiles"], 4, {"W": params["water"]}, seed=params["seed"], kT=kT,
                              electrostatics={"enabled": True, "pH": params["pH"],
                                              "ionicStrengthMolar": params["ionicStrengthMolar"],
                                              "densityVolumeSigma3": L ** 3 / 8.0})
