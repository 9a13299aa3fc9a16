"""Throughput of the engine_step hot path on one MI355X.

Workload: the 2,000,000-particle Morse+LJ system (4 types, density 1,
cutoff 2.5, tabulated potentials), FP32-mixed path.  One step = binning,
nonbonded pass, bonded pass, forward Euler with migration, all on the GPU
with the particles resident in HBM.  By default the pair pass uses neighbour
lists with a 0.4 skin (rebuilt on the device when a particle has moved more
than half the skin; the pair set of every step is unchanged); the cell pass
is timed beside it.

Prints one JSON line with particle-steps/s (a multi-step call: the captured
step graph replayed), the phase split of the same steps run eagerly with HIP
events on the module's stream, the cell pass and the FP64 GPU path beside
it, the force error of both precisions against the FP64 CPU oracle, the
roofline of the dominant pair kernel, the CPU baseline (the oracle
restatement of the reference CPU engine, timed on a bounded number of steps
on this host's cores) and the 512-replica ensemble, batched in one launch
sequence against the same replicas stepped one after another.

--calibrate-copy runs only a plain device copy of known size (the reference
for scaling the FETCH_SIZE / WRITE_SIZE counters of a profiler run).
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "tissue-forge_amd")

# HBM3E peak of one MI355X; dense FP32 vector peak without packed math
HBM_PEAK_GBS = 8000.0
# Tile-streamed bytes per particle-step of the nonbonded kernel: every
# particle's 16-byte position record is read by the 27 cells of its stencil,
# plus its own force write (27*16 + 16).
NONBOND_BYTES_PER_PARTICLE = 448.0
# Compulsory bytes per particle of the list pair kernel besides its list
# entries (4 B each): its own position and record, its entry count, and the
# force written (16 + 16 + 4 + 16).
LIST_FIXED_BYTES = 52.0
PHASES = ("bin_avg", "build_avg", "nonbond_avg", "advance_avg")


def traffic_bytes(kernel):
    """HBM bytes per launch of a pair kernel on this workload, from the
    committed profiler summary (FETCH_SIZE + WRITE_SIZE, each scaled by the
    same counter's reading for a plain copy of known size); None without it."""
    path = os.path.join(ROOT, "profiles", "%s_morse_lj_2M_traffic.txt" % kernel)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch"):
                return float(line.split("=")[1])
    return None


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
    d = np.linalg.norm(fg[:, :3] - fo[:, :3], axis=1)
    mag = np.linalg.norm(fo[:, :3], axis=1)
    scale = math.sqrt(float(np.mean(mag ** 2)))
    rel = d / np.maximum(mag, 1e-3 * scale)
    return dict(err_rms=float(d.max() / scale), rel_p99=float(np.percentile(rel, 99)), rel_max=float(rel.max()))


def time_gpu(tf, W, scene, pots, precision, steps, warmup, skin=0.0):
    """Timed region: `steps` whole steps replayed from the captured step
    graph, after `warmup` steps.  A second engine prepared the same way (same
    start, same launch shape, so the same trajectory and the same list
    rebuilds) then times the same steps with HIP events around each phase
    (bin, list build, pair pass, advance) on the module's stream.  Launch
    shapes are autotuned for each engine; they never change a result bit.
    skin > 0 steps with neighbour lists."""
    def prepared():
        eng = W.build_engine(scene, potentials=pots)[0]
        eng.set_verlet(skin)
        eng.to_device(precision)
        shape = eng.autotune(reps=2)   # result-invariant launch shapes (tfg_autotune)
        return eng, shape

    eng, shape = prepared()
    xq = eng.positions()   # the input positions as the path holds them
    eng.step(1)
    forces = eng.particles()["f"]   # the first step's pair pass, evaluated at xq
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
    return dict(wall_s=wall, wall_graph_s=wall_graph, kernels=kt, forces=forces, positions=xq, shape=shape,
                lists=lists)


def bench_ensemble(tf, W, nrep, steps, warmup):
    """The replica ensemble advanced as one system list (graph replay of the
    batched step) and the same replicas as separate engines, one after
    another."""
    scenes = W.ensemble(nrep=nrep)
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
    return {"replicas": nrep, "particles": total, "steps": steps,
            "batched_ms_per_step": batched / steps * 1e3, "sequential_ms_per_step": sequential / steps * 1e3,
            "batched_value": total * steps / batched, "sequential_value": total * steps / sequential,
            "speedup": sequential / batched, "autotuned_workgroups": {"pair": shape[0], "stream": shape[1]}}


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
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--ens-steps", type=int, default=50)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    ap.add_argument("--skin", type=float, default=0.4, help="neighbour-list skin (0: cell pass only)")
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
    lists = mixed["lists"]
    if lists is not None:
        # dominant kernel: the list pair pass; its compulsory HBM bytes per
        # particle are its own position, record, count and force plus the
        # list entries (partner positions come from L2 / MALL)
        nb_ms = mixed["kernels"]["nonbond_avg"]
        bytes_pp = LIST_FIXED_BYTES + 4.0 * lists["mean_entries"]
        kernel = "k_list_force"
    else:
        nb_ms = mixed["kernels"]["nonbond_avg"]
        bytes_pp = NONBOND_BYTES_PER_PARTICLE
        kernel = "k_nonbond"
    achieved = bytes_pp * n / (nb_ms * 1e-3) / 1e9

    out = {
        "metric": "particle-steps/s at 1 MI355X (FP32-mixed vs FP64, force err <=1e-5); HBM roofline %",
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
        "neighbour_lists": lists,
        "autotuned_workgroups": {"pair": mixed["shape"][0], "stream": mixed["shape"][1]},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_bytes(kernel),
                     "bytes_per_particle": bytes_pp, "avg_launch_ms": nb_ms},
    }
    if cell is not None:
        cnb = cell["kernels"]["nonbond_avg"]
        ca = NONBOND_BYTES_PER_PARTICLE * n / (cnb * 1e-3) / 1e9
        out["cell_path"] = {"value": n * args.steps / cell["wall_graph_s"],
                            "ms_per_step": cell["wall_graph_s"] / args.steps * 1e3,
                            "phases_ms": {k: cell["kernels"][k] for k in PHASES},
                            "roofline": {"kernel": "k_nonbond", "achieved": ca, "frac": ca / HBM_PEAK_GBS,
                                         "traffic": traffic_bytes("k_nonbond"),
                                         "bytes_per_particle": NONBOND_BYTES_PER_PARTICLE, "avg_launch_ms": cnb}}
    if fp64 is not None:
        v64 = n * args.steps / fp64["wall_graph_s"]
        out["fp64_path"] = {"value": v64, "ms_per_step": fp64["wall_graph_s"] / args.steps * 1e3,
                            "phases_ms": {k: fp64["kernels"][k] for k in PHASES},
                            "neighbour_lists": fp64["lists"], "mixed_over_fp64": value / v64}

    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        orc = O.Oracle(scene, pots)
        orc.force(threads)
        fo = orc.particles()["f"]
        # The mixed path holds positions in FP32 (cell-centred frame), so its
        # input is the seeded configuration rounded to FP32.  Its arithmetic
        # error is measured against the oracle on exactly those positions;
        # the error against the unrounded inputs, and what that rounding alone
        # does to the FP64 oracle, are reported beside it.
        sq = dict(scene)
        sq["x"] = mixed["positions"]
        orq = O.Oracle(sq, pots)
        orq.force(threads)
        fq = orq.particles()["f"]
        out["force_error_vs_fp64_oracle"] = {
            "mixed": force_errors(mixed["forces"], fq),
            "mixed_vs_unrounded_inputs": force_errors(mixed["forces"], fo),
            "fp32_input_rounding_alone": force_errors(fq, fo),
        }
        if fp64 is not None:
            out["force_error_vs_fp64_oracle"]["fp64_gpu"] = force_errors(fp64["forces"], fo)
        log("oracle setup+force %.1f s" % (time.perf_counter() - t0))
        t0 = time.perf_counter()
        orc.step(args.cpu_steps, threads)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * args.cpu_steps / cpu_s, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": "%d full steps of the same %d-particle system (FP64 restatement of the "
                                         "reference CPU engine: cell tasks, Newton-3, sorted sweeps)" %
                                         (args.cpu_steps, n)}
    if not args.no_ensemble:
        out["ensemble"] = bench_ensemble(tf, W, args.replicas, args.ens_steps, args.warmup)
        log("ensemble %s" % out["ensemble"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
