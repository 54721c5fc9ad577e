#!/usr/bin/env python3
"""Throughput of the fused tracking turn on one MI355X.

A step is one turn of the reference's map over the whole bunch
(profile -> induced voltage -> interpolated kick + RF kick -> drift for the
intensity workloads; RF kick -> drift -> profile for c1).  Default workload:
c2 (PSB-like ring, 4e6 macroparticles, 4096 slices, n_fft 16384, 21
resonators), FP64, synthetic seeded Gaussian bunch.

Prints one JSON line with the contract keys plus
  roofline      particle kernel: 32 B per macroparticle-turn (dt, dE read and
                written once) / its average device time (HIP events on the
                engine's stream), against 8 TB/s
  cpu_baseline  the reference's own C++ kernels (oracle/_ref, built with the
                flags of `blond-compile --parallel --optimize`) on the host
                cores, same turn loop with numpy FFTs, bounded sample
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE_TURN = 32  # FP64 dt, dE: 16 B read + 16 B written


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--timing-turns", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--blocks-per-cu", type=int, default=1)
    ap.add_argument("--groups", type=int, default=0, help="count copies (0: automatic)")
    ap.add_argument("--max-peel", type=int, default=0)
    ap.add_argument("--method", choices=["auto", "fft", "direct", "fourstep"], default="auto")
    return ap.parse_args()


def cpu_baseline(snapshot: dict, target_s: float) -> dict:
    """Time the reference's C++ kernels (or, without oracle/_ref, the oracle's
    C restatement) on a bounded sample of the same workload."""
    from oracle import oracle as orc

    lib = orc.ref_lib("opt", "double")
    kind = "reference"
    if lib is None:
        lib, kind = orc.oracle_lib(), "port"
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    # calibrate on one turn, then run enough turns for about target_s seconds
    t0 = time.perf_counter()
    st = orc.track(snapshot, 1, lib)
    one = time.perf_counter() - t0
    turns = max(2, min(int(target_s / max(one, 1e-6)), len(snapshot["acc"]) - 2))
    t0 = time.perf_counter()
    orc.track(st, turns, lib)
    el = time.perf_counter() - t0
    n = len(snapshot["dt"])
    return {"value": n * turns / el, "unit": "macroparticle-turns/s",
            "cores": cores if kind == "reference" else 1, "kind": kind,
            "sample": f"{snapshot['name']}: {n} macroparticles x {turns} turns "
                      f"({el:.1f} s), numpy FFT single-threaded"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (--gpus {a.gpus} "
              "requested); the per-GPU figure is --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    from blond_amd.trackers.fused import FusedTracker
    from blond_amd.utils import bmath as bm
    from blond_amd.workloads import build_workload

    bm.use_gpu(0)
    n_turns = a.warmup + a.repeats * a.steps + a.timing_turns + 2
    w = build_workload(a.workload, n_particles=a.particles or None, n_turns=n_turns)
    n = w.beam.n_macroparticles
    snap = None if a.no_cpu_baseline else w.snapshot()
    ft = FusedTracker(w.tracker, w.profile, w.tiv, use_graph=not a.no_graph, block=a.block,
                      blocks_per_cu=a.blocks_per_cu, groups=a.groups, max_peel=a.max_peel,
                      method=None if a.method == "auto" else a.method)
    ft.track(a.warmup)
    torch.cuda.synchronize()
    times = []
    for _ in range(a.repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ft.track(a.steps)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t_med = statistics.median(times)
    value = n * a.steps / t_med

    # live per-kernel durations (HIP events on the engine's stream)
    ms_ind, ms_turn = ft.time_kernels(a.timing_turns)
    achieved = BYTES_PER_PARTICLE_TURN * n / (ms_turn * 1e-3) / 1e9
    turn_gbs = BYTES_PER_PARTICLE_TURN * n / (t_med / a.steps) / 1e9

    line = {
        "metric": "macroparticle-turns/s + % HBM roofline, plain vs batched vs adjoint, 1 GPU",
        "value": value,
        "unit": "macroparticle-turns/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_med / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian bunch, default_rng(1)/(2); resonators default_rng(42))",
        "config": {"workload": f"{a.workload}: fused intensity turn" if w.intensity else
                   f"{a.workload}: fused plain turn",
                   "macroparticles": n, "n_slices": w.meta.get("n_slices"),
                   "n_fft": w.meta.get("n_fft"), "form": "plain",
                   "graph": not a.no_graph, "repeats": a.repeats,
                   "block": a.block, "blocks_per_cu": a.blocks_per_cu, "groups": a.groups,
                   "max_peel": a.max_peel},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "turn_kernel (interp kick + RF kick + drift + histogram)",
                     "kernel_ms": ms_turn, "induced_ms": ms_ind,
                     "induced_method": ft.method,
                     "whole_turn_gbs": turn_gbs, "whole_turn_frac": turn_gbs / HBM_PEAK_GBS},
    }
    ft.close()
    if snap is not None:
        line["cpu_baseline"] = cpu_baseline(snap, a.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
