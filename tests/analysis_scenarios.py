"""Synthetic snapshots of the reference's own analysis tests, rebuilt with seeded generators:
tests/soup-aggregates.test.ts:60-235 (micelles, flat bilayer patch, hollow vesicle shell, solid
blob) and tests/water-closure.test.ts:14-273 (closed shell with water, solid blob, holed shell)."""
import math

import numpy as np

from protocell_soup.rules import loadParams, loadSoup, wcaCutoff

SOUP = loadSoup()
MONOMERS = SOUP["monomers"]
C_KIND = next(k for k, m in enumerate(MONOMERS) if m["kind"] == "carbon")
O_KIND = next(k for k, m in enumerate(MONOMERS) if m["kind"] == "head")
P = loadParams()
CUTOFF = wcaCutoff(P["sigma"] * max(m["radiusSigma"] for m in MONOMERS if m["kind"] in ("carbon", "head"))) + P["attraction"]["wc"]


class Build:
    def __init__(self):
        self.p, self.b = [], []

    def amphiphile(self, head, tail):
        h = len(self.p)
        self.p.append([head[0], head[1], head[2], O_KIND])
        self.p.append([tail[0], tail[1], tail[2], C_KIND])
        self.b += [h, h + 1]

    def arrays(self):
        return np.array(self.p, np.float32).reshape(-1), np.array(self.b, np.uint32)


def _random_dir(rng):
    u = rng.random() * 2 - 1
    phi = rng.random() * 2 * math.pi
    s = math.sqrt(max(0.0, 1 - u * u))
    return [s * math.cos(phi), s * math.sin(phi), u]


def _fibonacci(n):
    golden = math.pi * (3 - math.sqrt(5))
    out = []
    for i in range(n):
        y = 1 - (i / (n - 1)) * 2
        r = math.sqrt(1 - y * y)
        out.append([math.cos(golden * i) * r, y, math.sin(golden * i) * r])
    return out


def micelles():
    """Scenario A: three 20-amphiphile spheres, box 100."""
    bld = Build()
    for c in ([20, 20, 20], [70, 20, 20], [20, 70, 20]):
        for d in _fibonacci(20):
            bld.amphiphile([c[k] + d[k] * 1.2 for k in range(3)], [c[k] + d[k] * 0.5 for k in range(3)])
    return (*bld.arrays(), [100.0, 100.0, 100.0])


def bilayer_patch():
    """Scenario B: a 13 x 13 two-leaflet square patch, box 100."""
    bld = Build()
    for ix in range(13):
        for iy in range(13):
            x, y = 30 + ix * 1.4, 30 + iy * 1.4
            bld.amphiphile([x, y, 30 + 2.2], [x, y, 30 + 1.0])
            bld.amphiphile([x, y, 30 - 2.2], [x, y, 30 - 1.0])
    return (*bld.arrays(), [100.0, 100.0, 100.0])


def vesicle_shell(seed=5, n_outer=8000, n_inner=4000):
    """Scenario C: heads on two shells (radius 8 and 12.4), tails filling the wall, box 60."""
    rng = np.random.default_rng(seed)
    bld = Build()
    c = [30.0, 30.0, 30.0]

    def tail():
        d = _random_dir(rng)
        r = 8 + rng.random() * (12.4 - 8)
        return [c[k] + d[k] * r for k in range(3)]

    for r0, count in ((12.4, n_outer), (8.0, n_inner)):
        for _ in range(count):
            d = _random_dir(rng)
            rh = r0 + (rng.random() - 0.5) * 1.0
            bld.amphiphile([c[k] + d[k] * rh for k in range(3)], tail())
    return (*bld.arrays(), [60.0, 60.0, 60.0])


def blob():
    """Scenario D: a dense cubic-lattice ball of amphiphiles, box 60."""
    bld = Build()
    for ix in range(-6, 7):
        for iy in range(-6, 7):
            for iz in range(-6, 7):
                x, y, z = ix * 1.8, iy * 1.8, iz * 1.8
                if math.sqrt(x * x + y * y + z * z) > 9:
                    continue
                bld.amphiphile([30 + x, 30 + y, 30 + z], [30 + x + 0.3, 30 + y, 30 + z])
    return (*bld.arrays(), [60.0, 60.0, 60.0])


# ---- water closure (tests/water-closure.test.ts) -----------------------------------------------
SYN_BOX = [50.0, 50.0, 50.0]
R_IN, R_OUT = 12.7715, 17.2285
BLOB_DENSITY = 150_000 / ((4 / 3) * math.pi * 15 ** 3)
MEMBER_COUNT = math.ceil(BLOB_DENSITY * ((4 / 3) * math.pi * (R_OUT ** 3 - R_IN ** 3)))


def _angle_deg(a, b):
    return math.degrees(math.acos(max(-1.0, min(1.0, a[0] * b[0] + a[1] * b[1] + a[2] * b[2]))))


def _shell_members(rng, centre, count, hole=None):
    out = []
    for _ in range(count):
        d = _random_dir(rng)
        while hole is not None and _angle_deg(d, hole[0]) <= hole[1]:
            d = _random_dir(rng)
        r = (R_IN ** 3 + rng.random() * (R_OUT ** 3 - R_IN ** 3)) ** (1 / 3)
        out.append([centre[k] + d[k] * r for k in range(3)])
    return out


def _fill_ball(rng, centre, radius, count):
    out = []
    for _ in range(count):
        d = _random_dir(rng)
        r = radius * rng.random() ** (1 / 3)
        out.append([centre[k] + d[k] * r for k in range(3)])
    return out


def _fill_box_excluding(rng, centre, exclude, count):
    out = []
    while len(out) < count:
        p = [rng.random() * SYN_BOX[k] for k in range(3)]
        d = [p[k] - centre[k] for k in range(3)]
        d = [d[k] - round(d[k] / SYN_BOX[k]) * SYN_BOX[k] for k in range(3)]
        if sum(x * x for x in d) >= exclude * exclude:
            out.append(p)
    return out


def _assemble(members, water):
    pos = np.zeros((len(members) + len(water), 4), np.float32)
    pos[: len(members), :3] = members
    pos[len(members):, :3] = water
    pos[len(members):, 3] = 1.0  # water beads are kind 1 in these fixtures; members kind 0
    return pos.reshape(-1), list(range(len(members))), list(range(len(members), len(members) + len(water)))


def shift_and_wrap(pos, shift):
    p = np.asarray(pos, np.float32).reshape(-1, 4).copy()
    for k in range(3):
        v = p[:, k].astype(np.float64) + shift[k]
        p[:, k] = (v - np.floor(v / SYN_BOX[k]) * SYN_BOX[k]).astype(np.float32)
    return p.reshape(-1)


def water_shell(seed, hole=None, members_count=MEMBER_COUNT):
    rng = np.random.default_rng(seed)
    centre = [25.0, 25.0, 25.0]
    members = _shell_members(rng, centre, members_count, hole)
    inner = _fill_ball(rng, centre, R_IN - 1, 900)
    outer = _fill_box_excluding(rng, centre, R_OUT + 1, 6000)
    return _assemble(members, inner + outer)


def water_blob(seed):
    rng = np.random.default_rng(seed)
    centre = [25.0, 25.0, 25.0]
    members = _fill_ball(rng, centre, R_OUT, 150_000)
    water = _fill_box_excluding(rng, centre, R_OUT + 1, 6000)
    return _assemble(members, water)
