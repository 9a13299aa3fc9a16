"""Throughput of the engine_step hot path on one MI355X.

Headline workload: the 2,000,000-particle Morse+LJ system (4 types, density
1, cutoff 2.5, tabulated potentials), FP32-mixed path.  One step = binning,
nonbonded pass, bonded pass, forward Euler with migration, all on the GPU
with the particles resident in HBM.  By default the pair pass uses neighbour
lists with a 0.5 skin (rebuilt on the device when a particle has moved more
than half the skin; the pair set of every step is unchanged); the cell pass
is timed beside it.

Prints one JSON line with particle-steps/s (a multi-step call: the captured
step graph replayed), the phase split of the same steps run eagerly with HIP
events on the module's stream, the cell pass and the FP64 GPU path beside
it, the force and energy errors of both precisions against the FP64 CPU
oracle, the roofline of the dominant pair kernel, the CPU baseline (the
oracle restatement of the reference CPU engine, on this host's cores and on
one thread) and, under "legs", the other workloads: the 8,000-particle cell
sorting system in FP64 (1,000 steps), the 250,000-particle DPD droplets
with a Fick flux (500 steps) and the 1,000,000-particle bonded chains (200
steps), each with its own phases, roofline, accuracy and CPU baseline, and
the 512-replica ensemble (neighbour lists, skin 0.3), batched in one launch
sequence against the same replicas stepped one after another.

--leg NAME runs only that workload's GPU timing (for profiler runs);
--calibrate-copy runs only a plain device copy of known size (the reference
for scaling the FETCH_SIZE / WRITE_SIZE counters of a profiler run).
"""
import argparse
import importlib.util
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "tissue-forge_amd")

# HBM3E peak of one MI355X
HBM_PEAK_GBS = 8000.0
PHASES = ("bin_avg", "build_avg", "nonbond_avg", "advance_avg")


def cell_bytes(precision, dpd_flux=False):
    """Tile-streamed bytes per particle-step of the cell pair kernel: every
    particle's position record is read by the 27 cells of its stencil, plus
    its own force written (27*16 + 16 = 448 B in FP32); with DPD and one
    species the stencil also streams velocity and species (27*(16+16+4) + 20
    = 992 B).  FP64 records are twice as wide."""
    if dpd_flux:
        return 992.0 if precision == 0 else 1984.0
    return 448.0 if precision == 0 else 896.0


def list_fixed_bytes(precision):
    """Compulsory bytes per particle of the list pair kernel besides its
    list entries (4 B each): its build-frame position, its type record, its
    entry count and the force written (16 + 16 + 4 + 16 in FP32)."""
    return 52.0 if precision == 0 else 84.0


def profile_path(kernel, workload, kind):
    return os.path.join(ROOT, "profiles", "%s_%s_%s.txt" % (kernel, workload, kind))


def traffic_bytes(kernel, workload):
    """HBM bytes per launch of a kernel on a workload, from the committed
    profiler summary (FETCH_SIZE + WRITE_SIZE, each scaled by the same
    counter's reading for a plain copy of known size); None without it."""
    path = profile_path(kernel, workload, "traffic")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch"):
                return float(line.split("=")[1])
    return None


def counter_metrics(kernel, workload):
    """VALU busy, LDS busy and LDS bank-conflict fractions of a kernel, and
    its L1 (TCP) accesses per CU-cycle where measured, from the committed
    counter summaries (their "derived" lines); {} without them."""
    out, sources = {}, []
    for kind in ("pmc", "l1"):
        path = profile_path(kernel, workload, kind)
        if not os.path.exists(path):
            continue
        with open(path) as f:
            for line in f:
                if line.startswith("derived %s:" % kernel):
                    for item in line.split(":", 1)[1].split():
                        k, v = item.split("=")
                        out[k] = float(v)
                    sources.append(os.path.relpath(path, ROOT))
    if sources:
        out["source"] = ", ".join(sources)
    return out


def load_package():
    spec = importlib.util.spec_from_file_location("tissue_forge_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tissue_forge_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def force_errors(fg, fo):
    """err_max_over_rms = max_i |dF_i| / rms |F|; rel = |dF_i| / max(|F_i|,
    1e-3 rms |F|), its 99th percentile and maximum."""
    d = np.linalg.norm(fg[:, :3] - fo[:, :3], axis=1)
    mag = np.linalg.norm(fo[:, :3], axis=1)
    scale = math.sqrt(float(np.mean(mag ** 2)))
    rel = d / np.maximum(mag, 1e-3 * scale)
    return dict(err_max_over_rms=float(d.max() / scale), rel_p99=float(np.percentile(rel, 99)),
                rel_max=float(rel.max()))


def energy_error(eg, eo):
    return abs(eg["epot"] - eo["epot"]) / max(abs(eo["epot"]), 1e-300)


def host_threads():
    """Host cores the CPU baseline runs on: this process's CPU affinity,
    capped by OMP_NUM_THREADS when set (the host share granted to one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(n, 1)


def host_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cxx = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        cxx = "g++"
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "compiler": cxx,
            "flags": "-O2 -std=c++17 -pthread -ffp-contract=off (oracle/Makefile)"}


def time_gpu(tf, W, scene, pots, precision, steps, warmup, skin=0.0, capacity=0):
    """Timed region: `steps` whole steps replayed from the captured step
    graph, after `warmup` steps.  A second engine prepared the same way (same
    start, same launch shape, so the same trajectory and the same list
    rebuilds) then times the same steps with HIP events around each phase
    (bin, list build, pair pass, advance) on the module's stream.  Launch
    shapes are autotuned for each engine; they never change a result bit.
    skin > 0 steps with neighbour lists (capacity 0: the default)."""
    def prepared():
        eng = W.build_engine(scene, potentials=pots)[0]
        eng.set_verlet(skin, capacity)
        eng.to_device(precision)
        shape = eng.autotune(reps=2)   # result-invariant launch shapes (tfg_autotune)
        return eng, shape

    eng, shape = prepared()
    xq = eng.positions()   # the input positions as the path holds them
    eng.step(1)
    forces = eng.particles()["f"]   # the first step's pair pass, evaluated at xq
    energies = eng.energies()
    eng.step(max(warmup, 1))
    r0 = eng.verlet_state()[1]
    t0 = time.perf_counter()
    eng.step(steps)
    wall_graph = time.perf_counter() - t0
    active, r1, cap = eng.verlet_state()
    lists = None
    if active:
        mean, mx = eng.verlet_entries()
        lists = {"skin": skin, "rebuilds_in_timed_steps": r1 - r0, "capacity": cap, "mean_entries": mean,
                 "max_entries": mx}
    eng.from_device()
    eng.close()

    ev, _ = prepared()
    ev.step(1)
    ev.step(max(warmup, 1))
    q0 = ev.verlet_state()[1]
    ev.set_profiling(True)
    t0 = time.perf_counter()
    ev.step(steps)
    wall = time.perf_counter() - t0
    kt = ev.kernel_times()
    if lists is not None:
        lists["rebuilds_in_evented_steps"] = ev.verlet_state()[1] - q0
    ev.set_profiling(False)
    ev.from_device()
    ev.close()
    return dict(wall_s=wall, wall_graph_s=wall_graph, kernels=kt, forces=forces, energies=energies,
                positions=xq, shape=shape, lists=lists)


def roofline(r, n, precision, workload, dpd_flux=False):
    """Roofline of the pair kernel of a timed run: algorithmic bytes per
    particle x particles / the kernel's average launch time (HIP events on
    the module's stream)."""
    nb_ms = r["kernels"]["nonbond_avg"]
    if r["lists"] is not None:
        kernel = "k_list_force"
        bytes_pp = list_fixed_bytes(precision) + 4.0 * r["lists"]["mean_entries"]
    else:
        kernel = "k_nonbond"
        bytes_pp = cell_bytes(precision, dpd_flux)
    achieved = bytes_pp * n / (nb_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_bytes(kernel, workload),
           "bytes_per_particle": bytes_pp, "avg_launch_ms": nb_ms}
    out.update(counter_metrics(kernel, workload))
    return out


def cpu_baseline(O, scene, pots, threads, warmup, steps, what):
    """The oracle (FP64 restatement of the reference CPU engine: cell tasks,
    Newton-3, sorted sweeps) stepped `warmup` untimed then `steps` timed
    steps on `threads` host threads."""
    orc = O.Oracle(scene, pots)
    if warmup:
        orc.step(warmup, threads)
    t0 = time.perf_counter()
    orc.step(steps, threads)
    s = time.perf_counter() - t0
    n = scene["x"].shape[0]
    return {"value": n * steps / s, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "%d timed steps after %d warm-up steps of %s" % (steps, warmup, what), "seconds": s}


def accuracy(O, scene, pots, r, threads, precision):
    """First-step forces and energy of a timed run against the oracle on the
    positions the path held (FP32-rounded in the mixed path)."""
    sq = dict(scene)
    sq["x"] = r["positions"]
    orq = O.Oracle(sq, pots)
    orq.force(threads)
    out = force_errors(r["forces"], orq.particles()["f"])
    out["energy_rel"] = energy_error(r["energies"], orq.energies())
    return out, orq


LEGS = {
    # name: (builder, precision, gpu steps, (skin, list capacity), dpd_flux, cpu warm-up, cpu steps,
    #        description).  The chains' random walks pack up to ~300 partners within
    #        cutoff + skin, past the default capacity.
    "cell_sorting_8k": (lambda W: W.cell_sorting(), 1, 1000, (0.4, 0), False, 10, 1000,
                        "cell sorting, 8,000 particles, Morse, overdamped, FP64"),
    "dpd_droplets_250k": (lambda W: W.dpd_droplets(), 0, 500, (0.4, 0), True, 10, 20,
                          "DPD droplets, 250,000 particles, Fick flux, FP32-mixed"),
    "bonded_chains_1M": (lambda W: W.bonded_chains(), 0, 200, (0.6, 448), False, 10, 20,
                         "bonded chains, 1,000,000 particles, LJ + bonds + angles, FP32-mixed"),
}


def run_leg(tf, W, O, name, warmup, threads, cpu=True, steps_override=0):
    build, prec, steps, (skin, capacity), dpd_flux, cw, cs, what = LEGS[name]
    steps = steps_override or steps
    scene = build(W)
    pots = [W.make_potential(p) for p in scene["pots"]]
    n = scene["x"].shape[0]
    log("leg %s n=%d precision %d steps %d" % (name, n, prec, steps))
    r = time_gpu(tf, W, scene, pots, prec, steps, warmup, skin, capacity)
    out = {"workload": what, "particles": n, "dtype": "f64" if prec == 1 else "f32", "steps": steps,
           "value": n * steps / r["wall_graph_s"], "unit": "particle-steps/s",
           "ms_per_step": r["wall_graph_s"] / steps * 1e3,
           "phases_ms": {k: r["kernels"][k] for k in PHASES}, "neighbour_lists": r["lists"],
           "roofline": roofline(r, n, prec, name, dpd_flux)}
    log("leg %s: %.4g particle-steps/s, phases %s" % (name, out["value"], out["phases_ms"]))
    if dpd_flux:
        # the same steps with three flux sub-steps per step (two flux-only
        # passes and flux integrations more per step than the line above)
        sub = dict(scene)
        sub["nr_fluxsteps"] = 3
        r3 = time_gpu(tf, W, sub, pots, prec, steps, warmup, skin, capacity)
        out["flux_substeps_3"] = {"ms_per_step": r3["wall_graph_s"] / steps * 1e3,
                                  "ms_per_extra_substep": (r3["wall_graph_s"] - r["wall_graph_s"]) / steps * 1e3 / 2}
        log("leg %s with 3 flux sub-steps: %s" % (name, out["flux_substeps_3"]))
    if cpu:
        out["accuracy"], _ = accuracy(O, scene, pots, r, threads, prec)
        out["cpu_baseline"] = cpu_baseline(O, scene, pots, threads, cw, cs, "the same system")
        log("leg %s: accuracy %s cpu %s" % (name, out["accuracy"], out["cpu_baseline"]))
    return out


def bench_ensemble(tf, W, O, nrep, steps, warmup, threads, cpu=True, skin=0.3):
    """The replica ensemble advanced as one system list (graph replay of the
    batched step) and the same replicas as separate engines, one after
    another, both with neighbour lists of the given skin (0: the cell pass);
    the oracle steps a 16-replica subset one after another."""
    scenes = W.ensemble(nrep=nrep)
    for sc in scenes:
        sc["verlet_skin"] = skin
    total = int(sum(sc["x"].shape[0] for sc in scenes))
    engines = [W.build_engine(sc)[0] for sc in scenes]
    ens = tf.Ensemble(engines, precision=tf.PREC_MIXED)
    shape = ens.autotune(reps=2)
    ens.step(max(warmup, 1), use_graph=True)
    t0 = time.perf_counter()
    ens.step(steps, use_graph=True)
    batched = time.perf_counter() - t0
    ens.close(download=False)
    engines = [W.build_engine(sc)[0] for sc in scenes]
    for e in engines:
        e.to_device(tf.PREC_MIXED)
        e.step(max(warmup, 1))
    t0 = time.perf_counter()
    for e in engines:
        e.step(steps)
    sequential = time.perf_counter() - t0
    for e in engines:
        e.close()
    out = {"replicas": nrep, "particles": total, "steps": steps, "skin": skin,
           "batched_ms_per_step": batched / steps * 1e3, "sequential_ms_per_step": sequential / steps * 1e3,
           "batched_value": total * steps / batched, "sequential_value": total * steps / sequential,
           "speedup": sequential / batched, "autotuned_workgroups": {"pair": shape[0], "stream": shape[1]}}
    if cpu:
        sub = scenes[:16]
        psub = sum(sc["x"].shape[0] for sc in sub)
        cpu_steps = 20
        orcs = []
        for sc in sub:
            pots = [W.make_potential(p) for p in sc["pots"]]
            o = O.Oracle(sc, pots)
            o.step(2, threads)
            orcs.append(o)
        t0 = time.perf_counter()
        for o in orcs:
            o.step(cpu_steps, threads)
        s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": psub * cpu_steps / s, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": "replicas 0-15 (%d particles) stepped one after another, %d timed steps "
                                         "after 2 warm-up steps; the rate applies to the whole ensemble" %
                                         (psub, cpu_steps)}
    return out


def calibrate_copy(nbytes, reps):
    """A plain device-to-device copy of nbytes, reps times (torch's copy
    kernel): reads and writes exactly nbytes each."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copied %d bytes %d times" % (n * 4, reps))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--cpu-steps", type=int, default=10, help="timed CPU steps of the headline workload")
    ap.add_argument("--cpu-warmup", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the host cores granted to this process")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-legs", action="store_true")
    ap.add_argument("--leg", choices=sorted(LEGS), help="time only this workload on the GPU")
    ap.add_argument("--leg-steps", type=int, default=0, help="GPU steps of each leg (0: the workload's own count)")
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--ens-skin", type=float, default=0.3, help="the ensemble's list skin (0: cell pass)")
    ap.add_argument("--ens-steps", type=int, default=50)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    ap.add_argument("--skin", type=float, default=0.5, help="neighbour-list skin (0: cell pass only)")
    ap.add_argument("--no-cell", action="store_true", help="skip the cell-pass comparison run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (--gpus 1): the engine step is measured on a single MI355X; "
              "sharding across GPUs is designed but not implemented", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy, max(args.steps, 1))
        return

    tf = load_package()
    from tissue_forge_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = args.cpu_threads or host_threads()

    if args.leg:
        print(json.dumps({args.leg: run_leg(tf, W, O, args.leg, args.warmup, threads, cpu=False,
                                            steps_override=args.leg_steps)}), flush=True)
        return

    n = args.n
    scene = W.morse_lj(n=n)
    pots = [W.make_potential(p) for p in scene["pots"]]
    log("workload morse_lj n=%d L=%.3f cells=%s" % (n, scene["dim"][0], scene["cells"]))

    mixed = time_gpu(tf, W, scene, pots, tf.PREC_MIXED, args.steps, args.warmup, args.skin)
    log("mixed: wall %.3f s for %d steps (graph replay %.3f s), kernels %s, lists %s" %
        (mixed["wall_s"], args.steps, mixed["wall_graph_s"], mixed["kernels"], mixed["lists"]))
    cell = None
    if args.skin > 0 and not args.no_cell:
        cell = time_gpu(tf, W, scene, pots, tf.PREC_MIXED, args.steps, args.warmup, 0.0)
        log("mixed cell pass: wall %.3f s (graph replay %.3f s), kernels %s" %
            (cell["wall_s"], cell["wall_graph_s"], cell["kernels"]))
    fp64 = None
    if not args.no_fp64:
        fp64 = time_gpu(tf, W, scene, pots, tf.PREC_FP64, args.steps, args.warmup, args.skin)
        log("fp64: wall %.3f s (graph replay %.3f s), kernels %s" % (fp64["wall_s"], fp64["wall_graph_s"],
                                                                    fp64["kernels"]))

    # The timed region is a multi-step call, which replays the captured step
    # graph (the library's path for several steps); the phase split comes
    # from the same number of steps run eagerly with HIP events.
    value = n * args.steps / mixed["wall_graph_s"]
    ms_per_step = mixed["wall_graph_s"] / args.steps * 1e3
    workload = "morse_lj_2M"

    out = {
        "metric": "particle-steps/s at 1 MI355X (FP32-mixed vs FP64, force err ≤1e-5); HBM roofline %",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": "FP32-mixed: FP32 pair and integrator arithmetic, FP64 cutoff-band decisions and reductions",
        "data": "synthetic (seeded uniform positions)",
        "config": {"workload": "morse_lj_2M: %d particles, 4 types, density 1, cutoff 2.5, Morse like / LJ unlike, "
                               "cells %s" % (n, scene["cells"]),
                   "parallelism": "single GPU"},
        "phases_ms": {k: mixed["kernels"][k] for k in PHASES},
        "evented_value": n * args.steps / mixed["wall_s"],
        "neighbour_lists": mixed["lists"],
        "autotuned_workgroups": {"pair": mixed["shape"][0], "stream": mixed["shape"][1]},
        "roofline": roofline(mixed, n, 0, workload),
    }
    if cell is not None:
        out["cell_path"] = {"value": n * args.steps / cell["wall_graph_s"],
                            "ms_per_step": cell["wall_graph_s"] / args.steps * 1e3,
                            "phases_ms": {k: cell["kernels"][k] for k in PHASES},
                            "roofline": roofline(cell, n, 0, workload)}
    if fp64 is not None:
        v64 = n * args.steps / fp64["wall_graph_s"]
        out["fp64_path"] = {"value": v64, "ms_per_step": fp64["wall_graph_s"] / args.steps * 1e3,
                            "phases_ms": {k: fp64["kernels"][k] for k in PHASES},
                            "neighbour_lists": fp64["lists"], "mixed_over_fp64": value / v64}

    if not args.no_cpu:
        t0 = time.perf_counter()
        # The mixed path holds positions in FP32 (cell-centred frame), so its
        # input is the seeded configuration rounded to FP32.  Its arithmetic
        # error is measured against the oracle on exactly those positions;
        # the error against the unrounded inputs, and what that rounding alone
        # does to the FP64 oracle, are reported beside it.
        acc, orq = accuracy(O, scene, pots, mixed, threads, 0)
        orc = O.Oracle(scene, pots)
        orc.force(threads)
        fo = orc.particles()["f"]
        fq = orq.particles()["f"]
        errs = {"mixed": acc,
                "mixed_vs_unrounded_inputs": force_errors(mixed["forces"], fo),
                "fp32_input_rounding_alone": force_errors(fq, fo)}
        errs["mixed_vs_unrounded_inputs"]["energy_rel"] = energy_error(mixed["energies"], orc.energies())
        if fp64 is not None:
            errs["fp64_gpu"] = force_errors(fp64["forces"], fo)
            errs["fp64_gpu"]["energy_rel"] = energy_error(fp64["energies"], orc.energies())
        out["force_error_vs_fp64_oracle"] = errs
        out["accuracy"] = {"rel_max": acc["rel_max"], "rel_p99": acc["rel_p99"],
                           "err_max_over_rms": acc["err_max_over_rms"], "energy_rel": acc["energy_rel"],
                           "target_rel_max": 1e-5, "rel_max_within_target": acc["rel_max"] <= 1e-5,
                           "measure": "rel = |dF_i| / max(|F_i|, 1e-3 rms|F|) against the FP64 oracle on the "
                                      "path's own FP32 positions"}
        log("oracle forces %.1f s" % (time.perf_counter() - t0))
        base = cpu_baseline(O, scene, pots, threads, args.cpu_warmup, args.cpu_steps, "the same %d-particle system" % n)
        one = cpu_baseline(O, scene, pots, 1, 0, 1, "the same %d-particle system, one thread" % n)
        base.update(host_info())
        base["one_thread"] = {"value": one["value"], "sample": one["sample"]}
        out["cpu_baseline"] = base
        log("cpu baseline %s" % base)
    if not args.no_legs:
        out["legs"] = {name: run_leg(tf, W, O, name, args.warmup, threads, cpu=not args.no_cpu,
                                     steps_override=args.leg_steps) for name in LEGS}
    if not args.no_ensemble:
        out["ensemble"] = bench_ensemble(tf, W, O, args.replicas, args.ens_steps, args.warmup, threads,
                                         cpu=not args.no_cpu, skin=args.ens_skin)
        log("ensemble %s" % out["ensemble"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
