"""Debye-Hueckel basis and the CPU-side constant-pH helpers.

Restates soup/src/electrostatics.ts: ``makeEsBasis`` (:166-247), ``esUniform``/``esUniform2``
(:253-274), the pair functions ``esPairEnergy``/``esPairForceMag`` (:277-289), ``hendersonAlpha``
(:293-295), ``apparentPKa`` (:300-303) and ``pcgNext`` (:308-315); plus ``initialCharges``
(soup/src/soup-protonation.ts:38-49). The sweep itself runs on the GPU (libsoup_hip).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Dict, Optional

import numpy as np

from .rules import acidSoapScaleOf, wcaCutoff


@dataclass
class AcidSoapWell:
    scale: float = 0.0
    rcAttr: float = 0.0
    wc: float = 0.0
    epsilon: float = 0.0


@dataclass
class EsBasis:
    enabled: bool
    chargedKind: int
    chargeDeprotonated: float
    coeffA: float
    kappa: float
    cutoff: float
    nbCutoff: float
    splitRadius: float
    listRange: float
    listCapacity: int
    debyeLengthsSpanned: float
    discardedIntegratedFraction: float
    discardedContactFraction: float
    targetCutoff: float
    imageCap: float
    shiftU: float
    shiftF: float
    pKaIntrinsic: float
    pH: float
    ionicStrengthMolar: float
    sigmaNm: float
    debyeSigma: float
    sweepEverySteps: int
    kT: float
    acidSoap: AcidSoapWell = field(default_factory=AcidSoapWell)


def makeEsBasis(soup: dict, p: dict, over: Optional[dict] = None) -> EsBasis:
    over = over or {}
    es = soup.get("electrostatics")
    kT = p["thermostat"]["kT"]
    cutoff = wcaCutoff(p["sigma"] * p["beadSizes"]["tail_tail"]) + p["attraction"]["wc"]
    interaction_range = wcaCutoff(p["sigma"] * max(m["radiusSigma"] for m in soup["monomers"])) + p["attraction"]["wc"]
    if not es:
        return EsBasis(False, -1, 0.0, 0.0, 0.0, cutoff, cutoff, interaction_range, cutoff, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, kT)
    enabled = over.get("enabled", es["enabled"])
    pH = over.get("pH", es["pH"])
    ionic = over.get("ionicStrengthMolar", es["ionicStrengthMolar"])
    charged_kind = next((i for i, m in enumerate(soup["monomers"]) if m["id"] == es["chargedKind"]), -1)
    if charged_kind < 0:
        raise ValueError(f"electrostatics.chargedKind='{es['chargedKind']}' is not among monomers")
    if not (ionic > 0):
        raise ValueError(f"electrostatics: ionicStrengthMolar must be > 0, given {ionic}")
    if not (es["sigmaToNm"] > 0):
        raise ValueError(f"electrostatics: sigmaToNm must be > 0, given {es['sigmaToNm']}")
    sigma_nm = es["sigmaToNm"]
    debye_sigma = es["debyeLengthNmAtUnitMolar"] / math.sqrt(ionic) / sigma_nm
    coeff_a = kT * (es["bjerrumLengthNm"] / sigma_nm) if enabled else 0.0
    kappa = 1 / debye_sigma
    target = es["longRangeDebyeLengths"] * debye_sigma
    min_box = over.get("minBoxSigma")
    image_cap = es["longRangeMaxBoxFraction"] * min_box if min_box is not None else math.inf
    es_cutoff = over.get("cutoffSigma", max(cutoff, min(target, image_cap)))
    if min_box is not None and es_cutoff > min_box / 2:
        raise ValueError(
            f"electrostatics: the cutoff rc_es={es_cutoff:.6f} violates the minimum image convention "
            f"at min(box)={min_box:.4f} (need rc_es <= {min_box / 2:.4f}) -- "
            f"the box is too small for {es['longRangeDebyeLengths']} Debye lengths at I={ionic} M"
        )
    list_range = es_cutoff + soup["verletList"]["skin"]
    list_cap = es["longRangeListCapacity"]
    max_heads = over.get("maxHeads")
    if max_heads is not None and max_heads > 0 and min_box is not None:
        dv = over.get("densityVolumeSigma3")
        denom = dv if (dv is not None and dv > 0) else min_box ** 3
        uniform = (4 * math.pi / 3) * list_range ** 3 * (max_heads / denom)
        list_cap = max(64, min(max_heads, math.ceil(es["longRangeListSafetyFactor"] * uniform)))
    shift_u = coeff_a * math.exp(-kappa * es_cutoff) / es_cutoff
    shift_f = shift_u * (1 / es_cutoff + kappa)
    x = es_cutoff / debye_sigma
    return EsBasis(
        enabled, charged_kind, es["chargeDeprotonated"], coeff_a, kappa, es_cutoff, cutoff, interaction_range, list_range,
        int(list_cap), x, math.exp(-x) * (1 + x), math.exp(-x), target, image_cap, shift_u, shift_f, es["pKaIntrinsic"], pH,
        ionic, sigma_nm, debye_sigma, es["sweepEverySteps"], kT,
        AcidSoapWell(acidSoapScaleOf(soup, over.get("acidSoapScale")), wcaCutoff(p["sigma"] * p["beadSizes"]["tail_tail"]),
                     p["attraction"]["wc"], p["epsilon"]),
    )


def esUniform(b: EsBasis) -> np.ndarray:
    return np.array([b.coeffA, b.kappa, b.cutoff, b.shiftF], dtype=np.float32)


def esUniform2(b: EsBasis) -> np.ndarray:
    return np.array([b.splitRadius, b.listRange, b.listCapacity, b.chargedKind], dtype=np.float32)


def esPairEnergy(r: float, qq: float, b: EsBasis) -> float:
    if b.coeffA == 0 or qq == 0 or r >= b.cutoff or r <= 0:
        return 0.0
    u1 = (b.coeffA * math.exp(-b.kappa * r)) / r
    return qq * (u1 - b.shiftU + b.shiftF * (r - b.cutoff))


def esPairForceMag(r: float, qq: float, b: EsBasis) -> float:
    if b.coeffA == 0 or qq == 0 or r >= b.cutoff or r <= 0:
        return 0.0
    u1 = (b.coeffA * math.exp(-b.kappa * r)) / r
    return qq * (u1 * (1 / r + b.kappa) - b.shiftF)


def hendersonAlpha(pH: float, pKa: float) -> float:
    return 1 / (1 + math.pow(10, pKa - pH))


def apparentPKa(alpha: float, pH: float) -> float:
    if alpha <= 0 or alpha >= 1:
        return math.nan
    return pH - math.log10(alpha / (1 - alpha))


class PcgState:
    __slots__ = ("state",)

    def __init__(self, state: int):
        self.state = state & 0xFFFFFFFF


def pcgNext(s: PcgState) -> float:
    """electrostatics.ts:308-315: LCG state step, RXS-M-XS output permutation, / 2^32 in double."""
    v = (s.state * 747796405 + 2891336453) & 0xFFFFFFFF
    s.state = v
    v = ((((v >> ((v >> 28) + 4)) ^ v) * 277803737) & 0xFFFFFFFF)
    v = ((v >> 22) ^ v) & 0xFFFFFFFF
    return v / 4294967296.0


def headIndices(positions: np.ndarray, chargedKind: int) -> np.ndarray:
    kinds = np.trunc(positions.reshape(-1, 4)[:, 3]).astype(np.int64)
    return np.nonzero(kinds == chargedKind)[0].astype(np.int32)


def _js_round(x: np.ndarray) -> np.ndarray:
    r = np.floor(x)
    return r + (x - r >= 0.5)


def pairingStats(positions: np.ndarray, charges: np.ndarray, box, b: EsBasis, contactRadius: float) -> Dict[str, Any]:
    """electrostatics.ts:520-565: head-head contacts inside contactRadius (FP64 minimum image of the
    f32 coordinates), split into unlike (one charged, one neutral) and like pairs seen from each
    head, and the heads whose nearest head neighbour in contact has the unlike charge (paired).
    The reference walks a head cell list of side >= contactRadius; the candidate pairs here come
    from a periodic k-d tree queried slightly wider and are then filtered by the same distance."""
    from scipy.spatial import cKDTree

    pos = np.asarray(positions, dtype=np.float32).reshape(-1, 4)
    q = np.asarray(charges, dtype=np.float32).reshape(-1)
    heads = headIndices(pos.reshape(-1), b.chargedKind).astype(np.int64)
    nh = heads.size
    charged_h = q[heads] != 0
    alpha = float(charged_h.sum()) / nh if nh > 0 else 0.0
    unlike = like = with_nb = paired = 0
    if nh > 1:
        L = np.asarray(box, np.float64)
        x = pos[heads, :3].astype(np.float64)
        tree = cKDTree(np.mod(x, L) % L, boxsize=L)
        pr = tree.query_pairs(r=contactRadius * (1 + 1e-9) + 1e-9, output_type="ndarray")
        if pr.size:
            a, c = pr[:, 0], pr[:, 1]
            d = x[a] - x[c]
            d = d - _js_round(d / L) * L
            r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
            keep = r <= contactRadius
            a, c, r = a[keep], c[keep], r[keep]
            # each unordered contact is seen from both of its heads
            src = np.concatenate([a, c])
            dst = np.concatenate([c, a])
            rr = np.concatenate([r, r])
            diff = charged_h[src] != charged_h[dst]
            unlike, like = int(diff.sum()), int((~diff).sum())
            with_nb = int(np.unique(src).size)
            order = np.lexsort((rr, src))
            first = np.ones(order.size, bool)
            first[1:] = src[order][1:] != src[order][:-1]
            paired = int(diff[order][first].sum())
    contacts = unlike + like
    return {"heads": nh, "withNeighbour": with_nb, "unlikeContacts": unlike, "likeContacts": like,
            "unlikeFraction": unlike / contacts if contacts > 0 else 0, "unlikeFractionRandom": 2 * alpha * (1 - alpha),
            "paired": paired, "pairedFraction": paired / nh if nh > 0 else 0, "contactRadius": contactRadius,
            "alpha": alpha}


def initialCharges(positions: np.ndarray, es: EsBasis, seed: int):
    """soup/src/soup-protonation.ts:38-49 -> (charges f32[n], PcgState)."""
    n = positions.size // 4
    charges = np.zeros(n, dtype=np.float32)
    rng = PcgState((seed ^ 0x9E3779B9) & 0xFFFFFFFF)
    if not es.enabled:
        return charges, rng
    alpha = hendersonAlpha(es.pH, es.pKaIntrinsic)
    q = np.float32(es.chargeDeprotonated)
    for h in headIndices(positions, es.chargedKind):
        if pcgNext(rng) < alpha:
            charges[h] = q
    return charges, rng
