"""Long-run gate: a bilayer patch in explicit water with the zero-tension area move, measuring area per
lipid and bilayer thickness at the model's own zero tension.

Follows tests/water-bilayer-area-move.test.ts (the reference's gate):
- construction: 400 three-bead lipids (head + 2 carbons) in two leaflets on a jittered square
  lattice at area 1.3 per lipid, water at 0.8 outside the patch, box [L, L, 30], seed 31, zero
  velocities, bondRng i+1, thermoRng i+1000003, clay-free;
- protocol: 2000 fixed-box steps, then chunks of 500 steps + 10 area-move trials, up to 900 chunks,
  settled when a 150-chunk window of ln(area) has a 6-block OLS drift with |t| < 2,
  |drift| * 150 < 0.01 and a span < 0.02.
The reference's published measurement is area per lipid 1.1734 sigma^2 (window 1.1696-1.1765) and
thickness 4.4689 sigma, inside its corridors [1.1, 1.5] and [4, 6].

Writes one JSON object to stdout (and to the path given as the first argument), with a progress line
per 20 chunks on stderr. Run on a GPU: python tests/gates/water_bilayer_area_move.py [--seed N] out.json
(the seed drives the lattice jitter, the thermostat and bond streams and the area-move stream).
"""
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "protocell-genesis_amd"))

import numpy as np  # noqa: E402

import protocell_soup as ps  # noqa: E402
from protocell_soup.analysis import AnalysisContext, clusterComponents, findAmphiphiles, memberIndicesOf  # noqa: E402
from protocell_soup.init_state import Mulberry32, NONE_U32  # noqa: E402
from protocell_soup.rules import loadParams, loadSoup, wcaCutoff  # noqa: E402

N_LIPIDS, WATER_DENSITY, BOX_Z = 400, 0.8, 30.0
SEED = 31  # the reference's; --seed N for the run-to-run spread
EQUILIBRATE, CHUNK, TRIALS, CAP, TAIL, BLOCKS = 2000, 500, 10, 900, 150, 6


def build():
    """The reference's construction (water-bilayer-area-move.test.ts:75-194)."""
    soup, p = loadSoup(), loadParams()
    mons = soup["monomers"]
    carbon = next(k for k, m in enumerate(mons) if m["kind"] == "carbon")
    head = next(k for k, m in enumerate(mons) if m["kind"] == "head")
    water = next(k for k, m in enumerate(mons) if m["id"] == soup["solvent"]["waterId"])
    lit = json.load(open(os.path.join(ROOT, "protocell-genesis_amd", "data", "literature.json")))
    area = next(g for g in lit["gates"] if g["id"] == "area-per-lipid")["target"]
    area_start = (area["min"] + area["max"]) / 2
    b_tt = p["sigma"] * (mons[carbon]["radiusSigma"] * 2) / 2
    b_ht = p["sigma"] * (mons[head]["radiusSigma"] + mons[carbon]["radiusSigma"]) / 2
    b_wt = p["sigma"] * (mons[water]["radiusSigma"] + mons[carbon]["radiusSigma"]) / 2
    L = math.sqrt((N_LIPIDS / 2) * area_start)
    box = [L, L, BOX_Z]
    z0, gap = BOX_Z / 2, b_tt / 2
    half = gap + b_tt + b_ht
    per = math.ceil(N_LIPIDS / 2)
    n_side = max(1, math.ceil(math.sqrt(per)))
    sx_l, sy_l = L / n_side, L / n_side
    jit = 0.15
    rng = Mulberry32(SEED)
    r = lambda: float(rng.take(1)[0])  # noqa: E731
    N = N_LIPIDS * 3 + math.ceil(WATER_DENSITY * L * L * BOX_Z)
    pos = np.zeros((N, 4), np.float32)
    bs = np.full((N, 3), NONE_U32, np.uint32)
    k = 0
    for i in range(N_LIPIDS):
        leaf = 1 if i < per else -1
        idx = i if i < per else i - per
        a, b = idx % n_side, idx // n_side
        x = ((a + 0.5) * sx_l + (r() * 2 - 1) * jit * sx_l + L) % L
        y = ((b + 0.5) * sy_l + (r() * 2 - 1) * jit * sy_l + L) % L
        zt2 = z0 + leaf * gap
        zt1 = zt2 + leaf * b_tt
        zh = zt1 + leaf * b_ht
        pos[k] = [x, y, zh, head]
        pos[k + 1] = [x, y, zt1, carbon]
        pos[k + 2] = [x, y, zt2, carbon]
        bs[k, 0] = k + 1
        bs[k + 1, 2] = k
        bs[k + 1, 0] = k + 2
        bs[k + 2, 0] = k + 1
        k += 3
    margin = b_wt * 1.5
    z_lo, z_hi = z0 - half - margin, z0 + half + margin
    spacing = (1 / WATER_DENSITY) ** (1 / 3)
    wnx = max(1, round(L / spacing))
    wnz = max(1, round(BOX_Z / spacing))
    sx, sz = L / wnx, BOX_Z / wnz
    placed = 0
    for iz in range(wnz):
        if k >= N:
            break
        z_raw = (iz + 0.5) * sz
        if z_lo < z_raw < z_hi:
            continue
        for iy in range(wnx):
            if k >= N:
                break
            for ix in range(wnx):
                if k >= N:
                    break
                x = ((ix + 0.5) * sx + (r() * 2 - 1) * jit * sx + L) % L
                y = ((iy + 0.5) * sx + (r() * 2 - 1) * jit * sx + L) % L
                z = min(BOX_Z - 1e-4, max(0.0, z_raw + (r() * 2 - 1) * jit * sz))
                pos[k] = [x, y, z, water]
                k += 1
                placed += 1
    n = k
    start = {m["id"]: 0 for m in mons}
    start[mons[head]["id"]], start[mons[carbon]["id"]], start[mons[water]["id"]] = N_LIPIDS, 2 * N_LIPIDS, placed
    resume = {"globalStep": 0, "liveBox": box, "positions": pos[:n].reshape(-1), "velocities": np.zeros(4 * n, np.float32),
              "bondSlots": bs[:n].reshape(-1), "centerLink": np.full(n, NONE_U32, np.uint32),
              "centerHeldSteps": np.zeros(n, np.uint32), "desorbEvents": np.zeros(2, np.uint32),
              "bondRng": np.arange(1, n + 1, dtype=np.uint32), "thermoRng": np.arange(1000003, 1000003 + n, dtype=np.uint32),
              "events": {}}
    opts = {"box": box, "seed": SEED, "kT": p["thermostat"]["kT"], "clay": False, "start": start, "resume": resume}
    return opts, dict(soup=soup, p=p, head=head, carbon=carbon, water=water, placed=placed, n=n, area=area, lit=lit)


def area_per_lipid(box):
    return box[0] * box[1] / (N_LIPIDS / 2)  # engine/src/metrics.ts:8-10


def density_profile_z(member_pos, box, bins):
    """engine/src/metrics.ts:21-43 (w = 0 heads, nonzero tails)."""
    dz = box[2] / bins
    head, tail = np.zeros(bins), np.zeros(bins)
    for x in member_pos:
        z = float(x[2])
        if not (0 <= z < box[2]):
            raise ValueError(f"densityProfileZ: bead with z={z} outside [0, {box[2]})")
        b = min(bins - 1, int(math.floor(z / dz)))
        (head if x[3] == 0 else tail)[b] += 1
    return [(i + 0.5) * dz for i in range(bins)], head, tail


def bilayer_peaks(z, head):
    """engine/src/metrics.ts:118-157: global maximum, a +-1 sigma mask, the next maximum, each refined
    by a three-point parabola."""
    a = int(np.argmax(head))
    if head[a] == 0:
        raise ValueError("bilayerPeaks: the head profile is empty")
    masked = np.array([0 if abs(z[i] - z[a]) <= 1.0 else v for i, v in enumerate(head)])
    b = int(np.argmax(masked))
    if masked[b] == 0:
        raise ValueError("bilayerPeaks: second peak not found")

    def sub(i):
        if i == 0 or i == len(head) - 1:
            return z[i]
        dzv = z[1] - z[0]
        den = head[i - 1] - 2 * head[i] + head[i + 1]
        if den >= 0:
            return z[i]
        off = 0.5 * (head[i - 1] - head[i + 1]) / den
        return z[i] + max(-0.5, min(0.5, off)) * dzv

    za, zb = sub(a), sub(b)
    return (za, zb) if za <= zb else (zb, za)


def block_drift(ys, blocks):
    """water-bilayer-area-move.test.ts:255-279: OLS slope over block means."""
    per = len(ys) // blocks
    xs = [b * per + (per - 1) / 2 for b in range(blocks)]
    ms = [sum(ys[b * per:(b + 1) * per]) / per for b in range(blocks)]
    xm, ym = sum(xs) / blocks, sum(ms) / blocks
    sxy = sum((xs[b] - xm) * (ms[b] - ym) for b in range(blocks))
    sxx = sum((xs[b] - xm) ** 2 for b in range(blocks))
    slope = sxy / sxx
    sse = sum((ms[b] - ym - slope * (xs[b] - xm)) ** 2 for b in range(blocks))
    se = math.sqrt(sse / (blocks - 2) / sxx)
    return slope, se, slope / se if se > 0 else 0.0


def measure(s, ctx, info):
    soup, p = info["soup"], info["p"]
    parts = s.particles().reshape(-1, 4)
    box = s.box
    amph = findAmphiphiles(parts, s.bonds(), soup["monomers"], ctx)
    members = memberIndicesOf(amph)
    mp = parts[members].copy()
    mp[:, 3] = [0 if soup["monomers"][int(round(k))].get("polar") else 1 for k in parts[members, 3]]
    z, head, _ = density_profile_z(mp, box, 200)
    try:
        lower, upper = bilayer_peaks(z, head)
        thick = upper - lower
    except ValueError:
        lower = upper = thick = None
    radii = [m["radiusSigma"] for m in soup["monomers"] if m["kind"] in ("carbon", "head")]
    cutoff = wcaCutoff(p["sigma"] * max(radii)) + p["attraction"]["wc"]
    comps = clusterComponents(ctx, members, box, cutoff, False)
    cluster_fraction = (len(comps[0]) / len(members)) if comps else 0.0
    wic = tot_w = buried = heads = 0
    if lower is not None:
        wk = parts[:, 3].round().astype(int) == info["water"]
        zw = parts[wk, 2]
        tot_w, wic = int(wk.sum()), int(((zw > lower) & (zw < upper)).sum())
        for i in members:
            if soup["monomers"][int(round(parts[i, 3]))]["kind"] != "head":
                continue
            heads += 1
            zi = float(parts[i, 2])
            buried += 1 if abs(zi - lower) > 1 and abs(zi - upper) > 1 else 0
    tails = [t for a in amph for t in a["tailLengths"]]
    return {"box": list(box), "areaPerLipid": area_per_lipid(box), "thickness": thick, "clusterFraction": cluster_fraction,
            "waterInCore": wic, "totalWater": tot_w, "headBuriedFraction": buried / heads if heads else None,
            "amphiphileCount": len(amph), "meanTail": sum(tails) / len(tails) if tails else None}


def main():
    global SEED
    args = sys.argv[1:]
    if "--seed" in args:
        k = args.index("--seed")
        SEED = int(args[k + 1])
        del args[k:k + 2]
    out_path = args[0] if args else None
    opts, info = build()
    s = ps.createSoup(opts)
    ctx = AnalysisContext()
    start_box = list(s.box)
    t0 = time.perf_counter()
    s.step(EQUILIBRATE)
    area_eq = area_per_lipid(s.box)
    series, traj = [], []
    acc = trials = 0
    chunks_used, settled, checks = CAP, False, 0
    t_run = time.perf_counter()
    for i in range(CAP):
        s.step(CHUNK)
        mv = s.areaMove(TRIALS)
        acc += mv["accepted"]
        trials += mv["trials"]
        series.append(area_per_lipid(s.box))
        idx = i + 1
        if idx % 20 == 0:
            snap = measure(s, ctx, info)
            traj.append({"chunk": idx, **{k: snap[k] for k in ("areaPerLipid", "thickness", "headBuriedFraction", "clusterFraction")},
                         "acceptedFraction": acc / trials})
            print(f"[gate] chunk {idx} area {snap['areaPerLipid']:.4f} thick {snap['thickness']} "
                  f"buried {snap['headBuriedFraction']} acc {acc / trials:.3f} ({time.perf_counter() - t0:.0f} s)",
                  file=sys.stderr, flush=True)
        if len(series) >= TAIL and idx % TAIL == 0:
            checks += 1
            w = series[-TAIL:]
            slope, _, t = block_drift([math.log(a) for a in w], BLOCKS)
            span = math.log(max(w) / min(w))
            if abs(t) < 2 and abs(slope) * TAIL < 0.01 and span < 0.02:
                settled, chunks_used = True, idx
                break
    wall = time.perf_counter() - t_run
    tail = series[-TAIL:]
    slope, _, t = block_drift([math.log(a) for a in tail], BLOCKS)
    final = measure(s, ctx, info)
    events = s.events()
    s.dispose()
    ctx.close()
    lo, hi = info["area"]["min"], info["area"]["max"]
    tk = next(g for g in info["lit"]["gates"] if g["id"] == "bilayer-thickness")["target"]
    tail_mean = sum(tail) / len(tail)
    res = {"seed": SEED, "N": info["n"], "waterPlaced": info["placed"], "startBox": start_box, "areaAfterEquilibration": area_eq,
           "settled": settled, "chunksUsed": chunks_used, "checkpointsChecked": checks,
           "areaPerLipid": tail_mean, "areaTailMin": min(tail), "areaTailMax": max(tail),
           "driftPerChunk": slope, "driftT": t, "acceptedFraction": acc / trials, "trialsTotal": trials,
           "stepsRun": EQUILIBRATE + chunks_used * CHUNK, "stepsPerSec": chunks_used * CHUNK / wall, "events": events,
           **{k: final[k] for k in ("box", "thickness", "clusterFraction", "waterInCore", "totalWater", "headBuriedFraction",
                                     "amphiphileCount", "meanTail")},
           "corridors": {"area": [lo, hi], "thickness": [tk["min"], tk["max"]]},
           "reference": {"areaPerLipid": 1.1734, "areaWindow": [1.1696, 1.1765], "thickness": 4.4689},
           "trajectory": traj}
    res["verdict"] = "passed" if (settled and lo <= min(tail) and max(tail) <= hi and final["thickness"] is not None
                                  and tk["min"] <= final["thickness"] <= tk["max"]) else "FAILED"
    text = json.dumps(res)
    print(text, flush=True)
    if out_path:
        with open(out_path, "w") as fh:
            fh.write(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
