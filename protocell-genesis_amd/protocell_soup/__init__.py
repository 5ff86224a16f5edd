"""protocell_soup -- the protocell-genesis soup Langevin step on MI355X (gfx950).

The public surface mirrors the reference's soup API: ``createSoup(opts) -> SoupSystem``
(soup/src/sim.ts:95) with the SoupSystem methods of soup/src/soup-types.ts:181-442 that belong to
the step path. Every device computation runs in libsoup_hip.so through the C-ABI of
include/soup_abi.h; this package only plans, packs and translates.
"""
from .sim import SoupSystem, createSoup, resolve, stepReplicas  # noqa: F401
from .plan import planSoupGrid, deriveGridGeometry, deriveListCapacity  # noqa: F401
from .electrostatics import makeEsBasis, hendersonAlpha, apparentPKa, pcgNext, PcgState  # noqa: F401
from .native import SoupError, load_library  # noqa: F401

__all__ = ["createSoup", "SoupSystem", "stepReplicas", "planSoupGrid", "makeEsBasis", "SoupError"]
