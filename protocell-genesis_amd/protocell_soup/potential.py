"""soupPotential's basis (soup/src/soup-potential.ts:91-122), packed for the device.

The energy itself is evaluated on the GPU by ``SoupSystem.potential`` (soup_potential in the C-ABI)
over the system's current configuration.
"""
from __future__ import annotations

import math
from typing import Any, Dict, Optional

from . import native
from .rules import acidSoapScaleOf, attractionScaleTable, speciesClassOf, wcaCutoff


def makePotentialBasis(soup: Dict[str, Any], p: Dict[str, Any], es=None, attractionOverride: Optional[float] = None,
                       claySurfaceChemistry: Optional[str] = None, acidSoapOverride: Optional[float] = None,
                       centerLink: bool = True, charges: bool = True) -> native.SoupPotentialBasis:
    """makePotentialBasis (soup-potential.ts:91-122) in FP64; ``es`` is the system's EsBasis (None or
    disabled: no screened-Coulomb term). ``centerLink`` / ``charges`` say whether soupPotential's
    optional arguments are given."""
    b = native.SoupPotentialBasis()
    radii = [p["sigma"] * m["radiusSigma"] for m in soup["monomers"]]
    rc_attr = wcaCutoff(p["sigma"] * p["beadSizes"]["tail_tail"])
    max_wca = max(wcaCutoff((ra + rb) / 2) for ra in radii for rb in radii)
    b.epsilon = p["epsilon"]
    b.fene_k, b.fene_r_inf = p["fene"]["k"], p["fene"]["rInf"]
    b.bend_k, b.bend_r0 = p["bend"]["k"], p["bend"]["r0"]
    b.wc = p["attraction"]["wc"]
    b.rc_attr = rc_attr
    b.cutoff = max(max_wca, rc_attr + p["attraction"]["wc"])
    b.wca_factor = math.pow(2.0, 1.0 / 6.0)
    b.acid_soap = acidSoapScaleOf(soup, acidSoapOverride, attractionOverride)
    for k, r in enumerate(radii):
        b.b_radius[k] = r
    for k, m in enumerate(soup["monomers"]):
        b.cls[k] = speciesClassOf(m)
    table = attractionScaleTable(soup, attractionOverride, claySurfaceChemistry)
    for i in range(4):
        for j in range(4):
            b.attr[i][j] = table[i][j]
    if es is not None and es.enabled and es.coeffA != 0:
        b.es_A, b.es_kappa, b.es_cutoff, b.es_shiftU, b.es_shiftF = es.coeffA, es.kappa, es.cutoff, es.shiftU, es.shiftF
        b.charged_kind = max(0, es.chargedKind)
    b.use_center_link = 1 if centerLink else 0
    b.use_charges = 1 if charges else 0
    return b
