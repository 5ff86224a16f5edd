"""Benchmark: MD steps/s of the 483,268-bead campaign soup on one MI355X.

Workload (the campaign-scale configuration): a 76 sigma periodic box with 483,268 beads
{C: 55988, O: 18659, H: 55988, M: 1452, W: 351181} (seed 19, kT 1.1), charged heads at pH 7.0 and
0.01 M (long-range Debye-Hueckel at rc = 15.2 sigma), bond Monte Carlo every 20 steps, Verlet
rebuild every 10 steps, constant-pH sweep every 1000 steps, seeded jittered-lattice start followed by
the cold-start minimiser. Inputs are resident in HBM before the timed region. The timed region is
exactly --steps integration steps of SoupSystem.step as a caller runs them (step periods replayed
from captured HIP graphs, with the per-1000-step guards and the sweeps that fall inside it) and,
every 1000 steps, the campaign's snapshot analysis (amphiphile recognition, aggregate decomposition,
encapsulated-water flood: detectStage). One constant-pH sweep and one analysis per 1000 steps belong
to the workload: when the window holds none of either, one of each runs explicitly inside it, and
`value` = 1 / (step time + (sweep + analysis time) / 1000), whatever --steps is.

Prints ONE JSON line: steps/s, the roofline of the dominant kernel (k_force_group, the list force,
eight lanes per particle; algorithmic bytes per launch = N * (60 + 4 * Verlet entries read) +
heads * 28) with its average duration measured by HIP events on the library's stream in a separate,
fully bracketed run of --phase-steps steps (which also gives the per-phase breakdown), and the CPU
baseline: the oracle's OpenMP step on the host cores, timed on a bounded sample of the same state.

The line's "batched" object is the replica sweep: --replicas amphiphile soups of uneven size stepped
--replica-steps steps one replica after the other, then as many steps all together (stepReplicas),
both as replica-steps/s.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "protocell-genesis_amd"))
sys.path.insert(0, ROOT)

CAMPAIGN = {
    "box": [76, 76, 76],
    "seed": 19,
    "kT": 1.1,
    "clay": False,
    "start": {"C": 55988, "O": 18659, "H": 55988, "M": 1452, "W": 351181},
    "electrostatics": {"enabled": True, "pH": 7.0, "ionicStrengthMolar": 0.01},
}
HBM_PEAK_GBS = 8000.0
ANALYSIS_EVERY = 1000  # the campaign's aggregate decomposition + encapsulated-water flood cadence
REFERENCE_STEPS_PER_S = 11.5  # 87.2 ms/step at 483,268 beads (reference closeout report)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def replica_sweep(ps, torch, count, steps):
    """The batched replica sweep against the per-replica loop on the same replicas: ``count``
    amphiphile soups of uneven size (protocell_soup.replicas.sweepReplicaParams: edge 14.7-40 sigma
    at density 0.80, pH 5-9, ionic strength 0.001-0.1 M, supply 1.5-4.8 closure thresholds, seed
    1000 + r), each relaxed and warmed up, then ``steps`` steps of every replica one replica after
    the other (SoupSystem.step), then ``steps`` more of all of them together (stepReplicas)."""
    from protocell_soup.replicas import sweepReplicaOpts, sweepReplicaParams

    params = sweepReplicaParams(count)
    t0 = time.perf_counter()
    reps = []
    for p in params:
        s = ps.createSoup(sweepReplicaOpts(p))
        s.relaxColdStart()
        reps.append(s)
    sizes = [s.N for s in reps]
    log(f"replicas: {count} created and relaxed, N {min(sizes)}..{max(sizes)}, total {sum(sizes)} ({time.perf_counter() - t0:.1f} s)")
    for s in reps:
        s.step(40)
    ps.stepReplicas(reps, 40)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k, s in enumerate(reps):
        s.step(steps)
        if k % 16 == 15:
            log(f"per-replica loop: {k + 1}/{count}")
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t
    log(f"per-replica loop: {count * steps / t_loop:.0f} replica-steps/s ({t_loop:.2f} s)")
    t = time.perf_counter()
    ps.stepReplicas(reps, steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t
    log(f"batched: {count * steps / t_batch:.0f} replica-steps/s ({t_batch:.2f} s)")
    for s in reps:
        s.dispose()
    loop_rate, batch_rate = count * steps / t_loop, count * steps / t_batch
    return {"replica_steps_per_s": round(batch_rate, 1), "loop_replica_steps_per_s": round(loop_rate, 1),
            "speedup_vs_loop": round(batch_rate / loop_rate, 3), "replicas": count, "steps_per_replica": steps,
            "beads_total": sum(sizes), "beads_min": min(sizes), "beads_max": max(sizes),
            "bead_steps_per_s": round(sum(sizes) * steps / t_batch, 1),
            "workload": "amphiphile soups at density 0.80: edge U(14.735, 40) sigma, pH U(5, 9), ionic strength "
                        "logU(0.001, 0.1) M, supply U(1.5, 4.8) x 930/76^3 amphiphiles per sigma^3, seed 1000+r, "
                        "charged heads with long-range DH, bond MC/20, rebuild/10, pH sweep/1000"}


def drift_schedule(ps, torch, cadence_sys, steps):
    """The campaign under both Verlet rebuild schedules over the same number of steps (graphs on, the
    constant-pH sweep that falls inside included): the reference's fixed cadence on the benchmark's
    own system, and the displacement-triggered schedule on a second copy of the campaign."""
    import numpy as np

    r0 = cadence_sys.rebuildCount()
    torch.cuda.synchronize()
    t = time.perf_counter()
    cadence_sys.step(steps)
    torch.cuda.synchronize()
    t_cad = time.perf_counter() - t
    rc = cadence_sys.rebuildCount() - r0
    opts = dict(CAMPAIGN, verletOverride={"rebuild": "drift"})
    d = ps.createSoup(opts)
    try:
        d.relaxColdStart()
        d.step(200)
        r1 = d.rebuildCount()
        torch.cuda.synchronize()
        t = time.perf_counter()
        d.step(steps)
        torch.cuda.synchronize()
        t_drift = time.perf_counter() - t
        rd = d.rebuildCount() - r1
        reads = d.readDEBUG("LIST_READS", np.uint64).astype(np.int64)
        cfg = d.verletConfig()
    finally:
        d.dispose()
    log(f"rebuild schedules over {steps} steps: cadence {steps / t_cad:.0f} steps/s with {rc} rebuilds, drift "
        f"{steps / t_drift:.0f} steps/s with {rd} rebuilds")
    return {"steps": steps, "cadence_steps_per_s": round(steps / t_cad, 1), "cadence_rebuilds": rc,
            "drift_steps_per_s": round(steps / t_drift, 1), "drift_rebuilds": rd,
            "drift_check_every": cfg["driftCheckEvery"], "drift_trigger": round(cfg["driftTrigger"], 4),
            "drift_whole_row_force_fraction": round(float(reads[1]) / max(1, int(reads[0] + reads[1])), 3)}


def cpu_baseline(sys_, R, sample_steps):
    """The oracle (C, OpenMP) stepping the same state on the host cores, plus one oracle snapshot
    analysis, both amortised to the campaign's cadence."""
    import numpy as np

    from oracle import analysis as ora  # checker / baseline only
    from oracle import oracle as orc  # checker / baseline only

    # the host's CPU share: OMP_NUM_THREADS where the environment sets it, else every core this process
    # may run on
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    orc.set_threads(threads)
    init = {"positions": sys_.particles(), "velocities": sys_.velocities(), "bondSlots": sys_.bondSlots(),
            "centerLink": sys_.centerLinks(), "centerHeldSteps": sys_.centerHeldSteps(), "frozen": sys_.frozen(),
            "bondRng": sys_.rngState()["bond"], "thermoRng": sys_.rngState()["thermo"]}
    gs = sys_.steps
    start = gs - gs % 20  # the sample starts on a bond step, so it holds rebuilds and bonds at their rates
    o = orc.OracleState(R.cfg, init, sys_.charges(), global_step=start, prng=sys_.protonationRngState())
    o.rebuild()
    o.force_list()
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        o.one_step()
    t_steps = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.protonation_sweep()
    t_sweep = time.perf_counter() - t0
    t0 = time.perf_counter()
    _oracle_snapshot_analysis(ora, sys_, R)
    t_an = time.perf_counter() - t0
    per_step = t_steps / sample_steps + t_sweep / R.sweep_every + t_an / ANALYSIS_EVERY
    return {"value": 1.0 / per_step, "unit": "steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity_cpus": affinity, "cpu_model": _cpu_model(),
            "sample": f"{sample_steps} oracle steps (incl. {sample_steps // 10} rebuilds, {sample_steps // 20} bond MC) "
                      f"+ 1 constant-pH sweep amortised over {R.sweep_every} steps + 1 snapshot analysis amortised over "
                      f"{ANALYSIS_EVERY} steps, same 483,268-bead state; {t_steps:.1f} s + {t_sweep:.1f} s + {t_an:.1f} s"}


def _cpu_model():
    """The host CPU's model name, as lscpu reports it."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _oracle_snapshot_analysis(ora, sys_, R):
    """The oracle's detectStage work on one snapshot: recogniser, clustering, per-aggregate geometry
    and floods, box-wide cavity and head peaks."""
    import numpy as np

    from protocell_soup.analysis import loadStageThresholds, memberCutoff

    t = loadStageThresholds()
    pos = sys_.particles()
    box = list(sys_.box)
    monomers = R.soup["monomers"]
    amph = ora.find_amphiphiles(pos, sys_.bonds(), monomers)
    mem = ora.member_indices(amph)
    water_kind = next(k for k, m in enumerate(monomers) if m["id"] == R.soup["solvent"]["waterId"])
    water = list(np.nonzero(np.rint(pos.reshape(-1, 4)[:, 3]) == water_kind)[0])
    ora.analyze_aggregates(pos, box, monomers, amph, mem, memberCutoff(), t, water)
    if mem:
        ora.enclosed_volume_box(pos, mem, box, t["closureCell"], t["closureRadius"])
    ora.head_peaks(pos, box, monomers, t["headDensityBins"], t["minHeadsPerBin"], t["headPeakSeparationMin"],
                   t["headPeakSeparationMax"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-sample", type=int, default=20)
    ap.add_argument("--phase-steps", type=int, default=1000, help="steps of the separate per-phase profile")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--replicas", type=int, default=64, help="replicas of the batched sweep (0: skip it)")
    ap.add_argument("--replica-steps", type=int, default=2000)
    ap.add_argument("--schedule-steps", type=int, default=1000,
                    help="steps of the cadence-vs-drift rebuild schedule comparison (0: skip it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the FETCH_SIZE/WRITE_SIZE scale reference) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": n, "dispatches": 5}))
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the soup box runs on one MI355X; "
              "multi-GPU sharding is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import protocell_soup as ps

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    R = ps.resolve(CAMPAIGN)
    log(f"resolved N={R.N} dims={tuple(R.plan.dims)} listCap={R.list_cap} esCap={R.cfg.es_cap} ({time.perf_counter() - t0:.1f} s)")
    s = ps.createSoup(CAMPAIGN)
    log(f"created ({time.perf_counter() - t0:.1f} s)")
    relax = s.relaxColdStart()
    log(f"relaxed: max|F| {relax['maxForceBefore']:.3g} -> {relax['maxForceAfter']:.3g} ({time.perf_counter() - t0:.1f} s)")
    done = 0
    while done < args.warmup:
        k = min(1000, args.warmup - done)
        s.step(k)
        done += k
    log(f"warmup {args.warmup} steps done ({time.perf_counter() - t0:.1f} s)")

    # timed region: --steps steps of SoupSystem.step exactly as a caller runs them (step periods
    # replayed from captured HIP graphs, no per-step events), with the campaign's snapshot analysis at
    # every multiple of 1000 steps. Config 3 also runs one constant-pH sweep and one analysis per 1000
    # steps: both are timed inside the region (the natural ones, or one of each run explicitly when
    # the window holds none) and enter `value` at that cadence, whatever --steps is. Only the sweep is
    # bracketed by HIP events, which keeps graph replay on.
    from protocell_soup.analysis import AnalysisContext, detectStage

    actx = AnalysisContext()
    s.profile(True, phases=["sweep"])
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    done = 0
    an_times, stage = [], None
    while done < args.steps:
        k = min(ANALYSIS_EVERY - s.steps % ANALYSIS_EVERY, args.steps - done)
        s.step(k)
        done += k
        if s.steps % ANALYSIS_EVERY == 0:
            ta = time.perf_counter()
            stage, evidence = detectStage(s, actx)
            an_times.append(time.perf_counter() - ta)
        log(f"timed {done}/{args.steps}")
    natural_sweeps = s.profileRead()["sweep"]["count"]
    natural_analyses = len(an_times)
    if natural_sweeps == 0:
        s.protonationSweepDEBUG()
    if natural_analyses == 0:
        ta = time.perf_counter()
        stage, evidence = detectStage(s, actx)
        an_times.append(time.perf_counter() - ta)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    sw = s.profileRead()["sweep"]
    s.profile(False)
    sweep_ms = sw["ms"] / max(1, sw["count"])
    an_ms = 1000.0 * sum(an_times) / len(an_times)
    t_steps = elapsed - sw["ms"] * 1e-3 - sum(an_times)  # the --steps steps alone
    per_step = t_steps / args.steps + (sweep_ms + an_ms) * 1e-3 / ANALYSIS_EVERY
    steps_per_s = 1.0 / per_step
    log(f"timed region {elapsed:.2f} s: {args.steps} steps {t_steps:.3f} s, sweeps {sw['count']} x {sweep_ms:.1f} ms "
        f"({natural_sweeps} in the step schedule), analyses {len(an_times)} x {an_ms:.1f} ms")

    # the list force's average launch, with HIP events on the library's stream, and the phase
    # breakdown: a separate, fully bracketed run of --phase-steps steps (every phase bracketed turns
    # graph replay off)
    reads0 = s.readDEBUG("LIST_READS", np.uint64).astype(np.int64)
    s.profile(True)
    s.step(args.phase_steps)
    phases = s.profileRead()
    s.profile(False)
    # rows read as near parts only and whole; launches in which every row qualified for its near part,
    # and the others (their particles also read their position at the rebuild to decide)
    reads = s.readDEBUG("LIST_READS", np.uint64).astype(np.int64) - reads0
    occ = s.listOccupancyDEBUG()

    f = phases["force"]
    force_ms = f["ms"] / max(1, f["count"])
    lbar = occ["verletMean"]
    lnear = float(s.readDEBUG("VERLET_NEAR").mean())
    frac_whole = float(reads[1]) / max(1, int(reads[0] + reads[1]))  # of the rows a launch reads, those read whole
    frac_test = float(reads[3]) / max(1, int(reads[2] + reads[3]))  # launches that test each particle's drift
    lread = frac_whole * lbar + (1.0 - frac_whole) * lnear  # Verlet entries a force launch reads
    n_heads = R.es_heads
    # algorithmic bytes of one list-force launch: per particle its own position, charge, counts, bond
    # slots, centre link and force (60 B), 4 B per Verlet entry it reads (the near part while its own
    # drift and the measured maximum allow, else the whole row) and, in launches that test each
    # particle's drift, its position at the rebuild (16 B); per charged head its head index and FP64
    # far DH sum (28 B); every neighbour gather is counted once per particle (the table is read, not
    # re-read)
    bytes_per_launch = R.N * (60.0 + 4.0 * lread + 16.0 * frac_test) + n_heads * 28.0
    achieved = bytes_per_launch / (force_ms * 1e-3) / 1e9
    # per step: the force launch, the list build's writes every 10 steps and the bond pass's reads of
    # the near parts every 20 steps, plus the per-particle state every phase streams; per head the
    # refresh (40 B), the far DH kernel (48 B + 4 B per DH entry), the list force's 28 B and the DH
    # build's writes every 10 steps
    bytes_step = R.N * (249.4 + 4.0 * lread + 0.4 * lbar + 0.2 * lnear) + n_heads * (116.0 + 4.4 * occ["esMean"])
    pair_per_step = occ["verletPairs"] + occ["esPairs"]
    log("phase ms/launch: " + ", ".join(f"{k} {v['ms'] / max(1, v['count']):.3f}x{v['count']}" for k, v in phases.items() if v["count"]))
    # measured HBM bytes of a launch (rocprofv3 FETCH_SIZE + WRITE_SIZE of the same command's force
    # dispatches, scaled by a copy calibration)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_force_group_campaign483k_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        traffic = int(tj["hbm_bytes_per_launch"])

    schedules = drift_schedule(ps, torch, s, args.schedule_steps) if args.schedule_steps > 0 else None

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(s, R, args.cpu_sample)
        log(f"cpu baseline {cpu['value']:.3f} steps/s on {cpu['cores']} threads (nproc {cpu['nproc']}, {cpu['cpu_model']})")

    actx.close()
    s.dispose()
    batched = None
    if args.replicas > 0:
        batched = replica_sweep(ps, torch, args.replicas, args.replica_steps)

    out = {
        "metric": "MD steps/s at 483k beads; batched replica-steps/s vs per-replica loop, 1 GPU",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 / steps_per_s, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(steps_per_s / REFERENCE_STEPS_PER_S, 2),
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "campaign soup 76 sigma, 483,268 beads, pH 7.0, 0.01 M, long-range DH, bond MC/20, rebuild/10, "
                        "pH sweep/1000, aggregate decomposition + encapsulated-water flood/1000",
            "N": R.N, "box": 76, "mean_verlet": round(lbar, 2), "mean_verlet_near": round(lnear, 2),
            "force_rows_read_whole": round(frac_whole, 4), "force_launches_testing_drift": round(frac_test, 4), "mean_dh_list": round(occ["esMean"], 2),
            "pair_interactions_per_s": round(pair_per_step * steps_per_s, 1),
            "algorithmic_bytes_per_step": round(bytes_step),
            "step_roofline_frac": round(bytes_step * steps_per_s / (HBM_PEAK_GBS * 1e9), 4),
            "phase_ms_per_step": {k: round(v["ms"] / args.phase_steps, 4) for k, v in phases.items() if v["count"]},
            "steps_only_per_s": round(args.steps / t_steps, 3),
            "sweep_every": R.sweep_every,
            "sweep_ms": round(sweep_ms, 2),
            "sweeps_timed": sw["count"],
            "analysis_every": ANALYSIS_EVERY,
            "analysis_ms_per_sample": round(an_ms, 2),
            "analysis_samples": len(an_times),
            "last_stage": stage,
            "value_basis": "1 / (step time + (sweep + analysis) / 1000): the steps timed with graph replay, one "
                           "constant-pH sweep and one snapshot analysis per 1000 steps, each timed in the region",
        },
        "roofline": {
            "kernel": "k_force_group",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "avg_launch_ms": round(force_ms, 4),
            "bytes_per_launch": round(bytes_per_launch),
        },
        "cpu_baseline": cpu,
        "batched": batched,
        "rebuild_schedules": schedules,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
