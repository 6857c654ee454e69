"""ns/day of the ~24k-atom PME water box on one MI355X.

Workload: 7,900 TIP3P waters (23,700 atoms) in a 6.2 nm cube, seeded
jittered lattice relaxed by 200 steepest-descent steps, PME with rc 0.9 nm,
Ewald tolerance 1e-5, 64^3 grid, order 4, SETTLE waters, NVE velocity
Verlet at 2 fs, mixed precision (f32 pair/PME arithmetic, 64-bit fixed-point
force words, f64 integration and constraints); the pair list is built at
rc + 0.20 nm and rebuilt when an atom moved 0.10 nm.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one velocity-Verlet step of the whole box (neighbor validity
test and conditional rebuild, pair loop, excluded pairs, PME, integration),
replayed from captured HIP graphs.  Prints one JSON line.

`--workload tip3p_333000_waters_pme216` runs the same recipe at 999,000
atoms in a 21.5 nm cube with a 216^3 grid, NPT: Langevin at 300 K and a
Monte Carlo volume move at 1 bar every 100 steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic work per unit, fixed before measuring.
FLOP_PER_COULOMB_PAIR = 50.0     # force-only erfc Ewald pair
FLOP_PER_LJ_COULOMB_PAIR = 60.0  # LJ + erfc Ewald pair
PEAK_FP32_TFLOPS = 157.3         # MI355X FP32 (vector and matrix)
PEAK_HBM_GBS = 8000.0            # MI355X HBM3E
# HBM bytes per launch of the pair kernel on this workload, from two
# rocprofv3 --pmc runs of this script (FETCH_SIZE x 2.000 and WRITE_SIZE x
# 1.000, the factors that make a 256 MiB device copy read and write 256 MiB):
# profiles/nb_pairs_tip3p_7900_pme64.md
NB_PAIRS_TRAFFIC_BYTES = 28.90e6

# The measured workloads: the 23,700-atom box (the default bench line) and
# the 999,000-atom box at full size on one GPU (same recipe, NPT).  The list
# buffer (skin) is each workload's fastest of a sweep: 0.20 nm for the small
# box, where 0.20-0.32 nm ran within noise of each other and 0.15 nm 3 %
# slower (fewer rebuilds outweigh the longer rows); 0.15 nm for the large
# one, where 0.20 and 0.25 nm ran 1 % and 3 % slower (its rows cost more).
WORKLOADS = {
    "tip3p_7900_waters_pme64": {
        "waters": 7900, "edge": 6.2, "grid": 64, "traffic": NB_PAIRS_TRAFFIC_BYTES,
        "ensemble": "nve", "skin": 0.20, "metric": "ns/day on ~24k-atom PME water box (1 GPU)"},
    "tip3p_333000_waters_pme216": {
        "waters": 333000, "edge": 21.5, "grid": 216,
        # the same two --pmc runs with --workload tip3p_333000_waters_pme216:
        # profiles/md_step_tip3p_333000_pme216.md
        "traffic": 1.3870e9, "ensemble": "npt", "skin": 0.15,
        "metric": "ns/day on 999k-atom PME water box (1 GPU)"},
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", choices=sorted(WORKLOADS), default="tip3p_7900_waters_pme64",
                   help="TIP3P box: the 23,700-atom bench line or the 999,000-atom box")
    p.add_argument("--ensemble", choices=("nve", "npt"), default=None,
                   help="nve: velocity Verlet; npt: Langevin 300 K (friction 1/ps) with a "
                        "Monte Carlo volume move at 1 bar every --barostat-every steps. "
                        "Default: the workload's own (nve for the 23,700-atom box, npt for "
                        "the 999,000-atom box)")
    p.add_argument("--barostat-every", type=int, default=100)
    p.add_argument("--steps", type=int, default=5000)
    p.add_argument("--warmup", type=int, default=500)
    p.add_argument("--graph", type=int, default=50, help="steps per captured graph")
    p.add_argument("--skin", type=float, default=None,
                   help="list buffer beyond the cutoff (nm): the list is rebuilt when an "
                        "atom moved half of it.  Default: the workload's (0.20 nm for the "
                        "23,700-atom box, 0.15 nm for the 999,000-atom box)")
    p.add_argument("--prune-skin", type=float, default=0.0,
                   help="buffer of a pruned inner list (nm, below --skin); 0: none")
    p.add_argument("--autotune", type=int, default=7,
                   help="launch-shape trial before the warm-up: bit 0 PME brick shapes, "
                        "bit 1 plane transforms, bit 2 pair-loop segment lengths (0: none)")
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--profile-repeats", type=int, default=50)
    p.add_argument("--measure-steps", type=int, default=1000,
                   help="steps of the run that times the pair kernel in place")
    p.add_argument("--drift-ps", type=float, default=100.0,
                   help="length of the NVE run the energy drift is measured on")
    p.add_argument("--frames", type=int, default=64,
                   help="frames of the force-matching gradient measurement (0: skip)")
    p.add_argument("--frame-repeats", type=int, default=20)
    p.add_argument("--frame-loop-passes", type=int, default=3)
    p.add_argument("--replicas", type=int, default=256,
                   help="also measure the batching workload: this many TIP3P replicas of "
                        "500-8,000 waters, Langevin 300 K, graphs of one chunk (0: skip)")
    p.add_argument("--replica-steps", type=int, default=500)
    p.add_argument("--replica-chunk", type=int, default=10)
    p.add_argument("--replica-launch", type=int, default=1, choices=(0, 1),
                   help="0: one graph of all replicas per chunk; 1: a graph per replica on "
                        "its own stream")
    p.add_argument("--calibrate-copy-mib", type=int, default=0,
                   help="after the bench, run a plain device copy of this many MiB "
                        "5 times (the reference kernel for counter-based traffic)")
    return p.parse_args()


def total_energy(sim):
    _, e = sim.forces()
    return float(e.sum()), e


def pme_bytes(n_atoms, K):
    """Algorithmic HBM bytes per launch of the PME kernels (f32 grid): the
    spread (fused with the z transform) reads N x 36 B (x f64 x3, q f64,
    place index) and writes the half spectrum (8 B per point); the y and x
    passes with the influence product read and write the half spectrum
    once; the backward z pass fused with the gather reads it and N x 36 B
    and writes N x 24 B (64-bit force words)."""
    k3 = K[0] * K[1] * K[2]
    half = K[0] * K[1] * (K[2] // 2 + 1) * 8
    return {
        "pme_spread_rows": n_atoms * 36 + half,
        "pme_yx": 2 * half,
        "pme_gather_rows": half + n_atoms * 36 + n_atoms * 24,
    }


def cpu_baseline(system, x, v, dt, seconds):
    """The CPU baseline: the oracle's C/OpenMP port of the reference's CPU
    path for the same step (velocity Verlet with SETTLE; the neighbor matrix
    at rc + 0.15 nm kept until an atom moved half the buffer; owner-computes
    f32 pair rows, vectorised; 64-bit fixed-point PME spread, f64 line
    FFTs), timed on this host's cores from the GPU run's state: about
    `seconds` of work with all threads and about half that with one."""
    threads = len(os.sched_getaffinity(0))
    from oracle import oracle as O
    o = O.Oracle(system)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass

    def timed(nthreads, budget):
        O.set_threads(nthreads)
        md = o.md(skin=0.15)
        t0 = time.perf_counter()
        x1, v1 = md.run(x, v, dt, 1)          # builds the matrix
        one = time.perf_counter() - t0
        steps = max(2, int(budget / max(one, 1e-3)))
        t0 = time.perf_counter()
        md.run(x1, v1, dt, steps)
        per = (time.perf_counter() - t0) / (steps + 1)  # a run evaluates forces once more
        builds = md.builds
        md.close()
        return per, steps, builds

    per_n, steps_n, builds_n = timed(threads, seconds)
    per_1, steps_1, _ = timed(1, 0.5 * seconds)
    O.set_threads(threads)
    return {"value": round(dt * 1e-3 * 86400.0 / per_n, 4), "unit": "ns/day",
            "cores": threads, "kind": "port",
            "ms_per_step": round(per_n * 1e3, 3),
            "one_thread": {"ms_per_step": round(per_1 * 1e3, 3),
                           "value": round(dt * 1e-3 * 86400.0 / per_1, 4), "steps": steps_1},
            "cpu_model": model,
            "sample": f"{steps_n} velocity-Verlet steps of the same {system.n:,}-atom box from the "
                      f"GPU run's state on {threads} threads ({builds_n} matrix builds) and "
                      f"{steps_1} on one thread: neighbor matrix reused within its 0.15 nm "
                      f"buffer, f32 pair rows, fixed-point PME"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the force path runs on one MI355X; several GPUs would shard it by "
              "slabs of xy columns, which is not built)", file=sys.stderr)
        sys.exit(2)
    from mdir_amd import systems as S
    from mdir_amd.simulation import Simulation

    dt = 0.002
    wl = WORKLOADS[args.workload]
    if args.workload == "tip3p_7900_waters_pme64":
        system = S.tip3p_config2(seed=args.seed)
    else:
        system = S.tip3p_box(wl["waters"], wl["edge"], seed=args.seed, cutoff=0.9,
                             grid=(wl["grid"],) * 3)
    if args.skin is None:
        args.skin = wl["skin"]
    sim = Simulation(system, skin=args.skin, prune_skin=args.prune_skin)
    sim.minimize(200, 0.01)
    x = sim.positions
    relaxed = system.copy()
    relaxed.positions = x
    v = S.initial_velocities(relaxed, 42, 300.0)
    sim.velocities = v
    print(f"setup: {system.n} atoms minimised, velocities at 300 K", file=sys.stderr,
          flush=True)
    ensemble = args.ensemble or wl["ensemble"]
    npt = ensemble == "npt"
    if npt:
        sim.set_langevin(300.0, 1.0, seed=args.seed)

    def advance(steps):
        if npt:
            sim.run_npt(steps, dt, graph=args.graph, every=args.barostat_every, pressure=1.0,
                        temperature=300.0, seed=args.seed)
        else:
            sim.run(steps, dt, graph=args.graph)

    tuned = {}
    if args.autotune:
        tuned = sim.autotune(args.autotune, steps=args.graph, graphs=4, dt=dt)
        best = min(tuned, key=tuned.get)
        print(f"autotune: best bricks of {best[0]} z lines, transform {best[1]}, segments "
              f"of {best[2]} ({tuned[best] * 1e3:.1f} us/step)", file=sys.stderr, flush=True)

    advance(args.warmup)
    # the graphs the timed run replays (its chunks of --graph steps and the
    # remainder), captured before the clock starts; under NPT a kept volume
    # move changes the box the graphs hold, so the run captures again after
    # it, inside the timed region
    sim.prepare(args.barostat_every if npt else args.steps, dt, graph=args.graph)
    builds0 = sim.stats()["builds"]
    mc0 = sim.barostat_stats() if npt else None
    t0 = time.perf_counter()
    advance(args.steps)   # returns after a device sync
    t1 = time.perf_counter()
    builds = sim.stats()["builds"] - builds0
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    ns_day = args.steps * dt * 1e-3 / (elapsed / 86400.0)
    print(f"timed: {args.steps} steps in {elapsed:.3f} s, {ns_day:.1f} ns/day",
          file=sys.stderr, flush=True)

    barostat = None
    if npt:
        mc = sim.barostat_stats()
        box = sim.box
        mass_g = float(system.masses.sum()) / 6.02214076e23
        barostat = {"moves": mc["tried"] - mc0["tried"], "kept": mc["kept"] - mc0["kept"],
                    "every_steps": args.barostat_every, "pressure_bar": 1.0,
                    "box_nm": [round(float(b), 5) for b in box],
                    "density_g_cm3": round(mass_g / (float(np.prod(box)) * 1e-21), 4)}
        print(f"barostat: {barostat['kept']} of {barostat['moves']} volume moves kept, box "
              f"{box[0]:.4f} nm, {barostat['density_g_cm3']} g/cm3", file=sys.stderr,
              flush=True)

    # conserved-energy drift of an NVE run (untimed): end minus start over
    # the run, and the least-squares slope of one sample per picosecond;
    # under NPT there is no conserved energy, and no drift run
    per_ps = int(round(1.0 / dt))
    drift = slope = None
    times = np.array([0.0])
    if not npt:
        samples = max(1, int(round(args.drift_ps)))
        times, energies = [0.0], [total_energy(sim)[0]]
        for k in range(samples):
            sim.run(per_ps, dt, graph=args.graph)
            times.append((k + 1) * per_ps * dt)
            energies.append(total_energy(sim)[0])
        times, energies = np.array(times), np.array(energies)
        drift = (energies[-1] - energies[0]) / times[-1] / system.n  # kJ/mol/ps per atom
        slope = float(np.polyfit(times, energies, 1)[0]) / system.n
        print(f"drift: {samples * per_ps} steps, {drift:.3e} kJ/mol/ps/atom "
              f"(fit {slope:.3e})", file=sys.stderr, flush=True)
    terms = total_energy(sim)[1]

    # The pair kernel's duration inside the replayed steps, on its stream:
    # device-clock readings bracket it in every step of a run of the same
    # kind right after the timed one (events give no elapsed times inside
    # graphs, and the readings cost a few per cent of a step, so they stay
    # out of the timed run).
    pair_ms = sim.run_measured(args.measure_steps, dt, graph=args.graph)
    # per-kernel device time, kernel by kernel on one stream (HIP events)
    prof = sim.profile_step(args.profile_repeats)
    # pairs within rc for the algorithmic FLOP count (f32 test of the
    # reference matrix, exclusions removed)
    pairs = sim.neighbor_pairs(system.cutoff, excluded=True)
    types = system.types
    both_o = (types[pairs[:, 0]] == 0) & (types[pairs[:, 1]] == 0)
    n_lj = int(both_o.sum())
    n_pairs = int(len(pairs))
    flops = FLOP_PER_LJ_COULOMB_PAIR * n_lj + FLOP_PER_COULOMB_PAIR * (n_pairs - n_lj)
    t_pairs = pair_ms * 1e-3
    pair_tflops = flops / t_pairs / 1e12
    kernels_ms = {k: round(v_, 5) for k, v_ in prof.items()}
    pmeb = pme_bytes(system.n, system.pme_grid)
    hbm = {k: round(b / (prof[k] * 1e-3) / 1e9, 1) for k, b in pmeb.items()}

    # the dominant kernel: the pair loop, against the FP32 vector peak
    roofline = {"kernel": "nb_pairs", "bound": "vector ALU", "pipe": "fp32 vector ALU",
                "duration_us": round(pair_ms * 1e3, 2),
                "achieved": round(pair_tflops, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(pair_tflops / PEAK_FP32_TFLOPS, 4),
                "traffic": wl["traffic"],
                "work": {"pairs_within_rc": n_pairs, "lj_pairs": n_lj,
                         "flop_per_pair": [FLOP_PER_COULOMB_PAIR, FLOP_PER_LJ_COULOMB_PAIR],
                         "flop_per_launch": flops}}

    out = {
        "metric": wl["metric"],
        "value": round(ns_day, 2),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "atoms": system.n,
                   "box_nm": wl["edge"], "cutoff_nm": system.cutoff, "pme_grid": list(system.pme_grid),
                   "pme_order": system.pme_order, "ewald_beta": system.beta,
                   "dt_fs": 2.0,
                   "ensemble": ("NPT: Langevin 300 K (1/ps), Monte Carlo volume moves at "
                                f"1 bar every {args.barostat_every} steps") if npt else "NVE",
                   "constraints": "SETTLE",
                   "list_skin_nm": args.skin, "prune_skin_nm": args.prune_skin,
                   "graph_steps": args.graph},
        "roofline": roofline,
        "kernels_ms": kernels_ms,
        "pme_hbm_gbs": hbm,
        "pair_tflops": round(pair_tflops, 3),
        "list_builds_in_timed_steps": int(builds),
        "nve_drift_kj_mol_ps_atom": drift,
        "nve_drift_fit_kj_mol_ps_atom": slope,
        "nve_drift_ps": float(times[-1]),
        "barostat": barostat,
        "energy_terms_kj_mol": [round(float(t), 4) for t in terms],
    }
    if tuned:
        out["autotune_us_per_step"] = {f"seg{sg}_bricks{tl}_transform{tr}": round(v * 1e3, 2)
                                       for (tl, tr, sg), v in sorted(tuned.items())}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(system, sim.positions, sim.velocities, dt,
                                           args.cpu_seconds)
    if args.frames > 0:
        from mdir_amd.frames import force_matching_benchmark
        out["frame_gradients"] = force_matching_benchmark(args.frames, args.frame_repeats,
                                                          args.frame_loop_passes,
                                                          autotune=args.autotune)
    if args.replicas > 0:
        from mdir_amd.replicas import replica_benchmark
        out["replicas"] = replica_benchmark(args.replicas, steps=args.replica_steps,
                                            chunk=args.replica_chunk,
                                            autotune=bool(args.autotune),
                                            launch=args.replica_launch)
        print(f"replicas: {args.replicas}, {out['replicas']['value']} ns/day aggregate, "
              f"{out['replicas']['ms_per_step']} ms per step of all", file=sys.stderr,
              flush=True)
    print(json.dumps(out), flush=True)
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)


def calibrate_copy(mib):
    """A plain device copy of known size through the runtime ABI
    (mgpuMemcpy device to device, the HIP runtime's copy kernel): 5 copies
    of mib MiB, each reading and writing mib MiB, for scaling hardware
    counters of memory traffic."""
    from mdir_amd import runtime as rt
    L = rt.lib()
    size = mib << 20
    s = L.mgpuStreamCreate()
    a = L.mgpuMemAlloc(size, s, False)
    b = L.mgpuMemAlloc(size, s, False)
    L.mgpuMemset32(a, 0x3F800000, size // 4, s)
    for _ in range(5):
        L.mgpuMemcpy(b, a, size, s)
    L.mgpuStreamSynchronize(s)
    L.mgpuMemFree(a, s)
    L.mgpuMemFree(b, s)
    L.mgpuStreamSynchronize(s)
    L.mgpuStreamDestroy(s)
    print(f"copy calibration: 5 x {mib} MiB read + {mib} MiB written", file=sys.stderr,
          flush=True)


if __name__ == "__main__":
    main()
