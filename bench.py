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
  forms         the other forms of the metric, each on its own workload:
                  large   c3: 1e7 macroparticles, 2 RF systems, 65536 slices,
                          FFT 131072 (four-step), fused intensity turn; with
                          3e7 macroparticles and the FP32 mode as variants
                  batched c4: 256 bunches (1e4..1e6 each), one launch pair per
                          turn for all bunches, autotuned, against 256
                          sequential fused runs of the same bunches
                  adjoint c5: 1e6 macroparticles x 500 turns, forward +
                          reverse pass with checkpoints every 50 turns and
                          recomputation (112 B per macroparticle-turn); as
                          a variant every state kept in HBM (80 B)
                  wake    c2 bunch (4e6, 4096 slices) with a multi-turn
                          wake over 4 revolutions (memory 16384 points,
                          n_fft 16384) plus an inductive Z/n source
                  feedback c2 bunch (4e6, 4096 slices, broadband resonator)
                          under the PSB beam phase loop, one loop workgroup
                          per turn
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
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3"])
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
    ap.add_argument("--forms", default="large,batched,adjoint,wake,feedback",
                    help="comma list of extra forms to measure ('' for none)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the known-size kernel "
                         "that calibrates FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_probe(mib: int):
    """Five device copies of `mib` MiB: each reads and writes mib MiB."""
    import torch

    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read_per_copy": n * 8,
                      "bytes_written_per_copy": n * 8}), flush=True)


def _timed(fn, sync):
    sync()
    t0 = time.perf_counter()
    fn()
    sync()
    return time.perf_counter() - t0


def _large(torch, n_particles, f32, turns):
    from blond_amd.trackers.fused import FusedTracker
    from blond_amd.workloads import build_workload

    w = build_workload("c3", n_particles=n_particles, n_turns=turns + 80)
    if f32:
        # the FP32 mode: float coordinates (the slice axis stays FP64)
        w.beam.dt = w.beam.dt.to(torch.float32).contiguous()
        w.beam.dE = w.beam.dE.to(torch.float32).contiguous()
    n = w.beam.n_macroparticles
    bpp = BYTES_PER_PARTICLE_TURN // 2 if f32 else BYTES_PER_PARTICLE_TURN
    ft = FusedTracker(w.tracker, w.profile, w.tiv)
    ft.track(10)
    el = _timed(lambda: ft.track(turns), torch.cuda.synchronize)
    ms_ind, ms_turn = ft.time_kernels(50)
    out = {"workload": f"c3: {n_particles:.0e} macroparticles, h=35640+142560, 65536 slices, "
                       f"n_fft 131072, {'FP32' if f32 else 'FP64'}",
           "value": n * turns / el, "unit": "macroparticle-turns/s", "turns": turns,
           "ms_per_turn": 1e3 * el / turns, "induced_method": ft.method,
           "bytes_per_macroparticle_turn": bpp,
           "particle_kernel_ms": ms_turn, "induced_ms": ms_ind,
           "particle_kernel_frac": bpp * n / (ms_turn * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "whole_turn_frac": bpp * n / (el / turns) / 1e9 / HBM_PEAK_GBS}
    ft.close()
    del w
    torch.cuda.empty_cache()
    return out


def form_large(torch, turns=200) -> dict:
    """c3: 1e7 and 3e7 macroparticles, 65536 slices, FFT 131072, fused
    intensity turn, FP64 and FP32 (the 1e7 FP64 figures at the top level)."""
    out = _large(torch, 10_000_000, False, turns)
    out["variants"] = {"3e7_f64": _large(torch, 30_000_000, False, turns // 2),
                       "1e7_f32": _large(torch, 10_000_000, True, turns),
                       "3e7_f32": _large(torch, 30_000_000, True, turns // 2)}
    return out


def _form_turns(torch, w, turns, label) -> dict:
    from blond_amd.trackers.fused import FusedTracker

    n = w.beam.n_macroparticles
    ft = FusedTracker(w.tracker, w.profile, w.tiv)
    ft.track(10)
    el = _timed(lambda: ft.track(turns), torch.cuda.synchronize)
    ms_ind, ms_turn = ft.time_kernels(30)
    out = {"workload": label, "value": n * turns / el, "unit": "macroparticle-turns/s",
           "turns": turns, "ms_per_turn": 1e3 * el / turns, "induced_method": ft.method,
           "particle_kernel_ms": ms_turn, "induced_ms": ms_ind,
           "particle_kernel_frac": BYTES_PER_PARTICLE_TURN * n / (ms_turn * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "whole_turn_frac": BYTES_PER_PARTICLE_TURN * n / (el / turns) / 1e9 / HBM_PEAK_GBS}
    ft.close()
    return out


def form_wake(torch, turns=200) -> dict:
    """c2 bunch with a 4-turn multi-turn wake and an inductive source; the
    induced step is the direct convolution over the memory's 16384 values
    plus the wake step."""
    from blond_amd.impedances.impedance import InductiveImpedance
    from blond_amd.workloads import build_wake_workload

    w = build_wake_workload("mtw", n_particles=4_000_000, n_turns=turns + 60, n_slices=4096)
    zon = InductiveImpedance(w.beam, w.profile, [-100.0] * (turns + 61), w.rf)
    zon.to_gpu()
    w.tiv.induced_voltage_list.append(zon)
    src = w.tiv.induced_voltage_list[0]
    out = _form_turns(torch, w, turns, "c2 bunch 4e6, 4096 slices; multi-turn wake over 4 turns "
                      f"(memory {src.n_mtw_memory}, n_fft {src.n_fft}) + inductive Z/n")
    return out


def form_feedback(torch, turns=200) -> dict:
    """c2-ring bunch under the PSB phase and radial loops, with intensity."""
    from blond_amd.workloads import build_feedback_workload

    w = build_feedback_workload("PSB", n_particles=4_000_000, n_turns=turns + 60,
                                n_slices=4096, intensity=True)
    return _form_turns(torch, w, turns, "c2-ring bunch 4e6, 4096 slices, broadband resonator, "
                       "PSB beam phase + radial loop on the device")


def form_batched(torch, turns=100, seq_turns=20) -> dict:
    """c4: 256 bunches in one launch pair per turn vs 256 sequential fused runs."""
    from blond_amd.trackers.fused import FusedTracker
    from blond_amd.workloads import build_ensemble

    E = build_ensemble(n_bunches=256, n_turns=turns + seq_turns + 40)
    n = E.n_particles
    et = E.tracker()
    blk, wpc, _ = et.autotune(turns=3)
    et.track(5)
    el = _timed(lambda: et.track(turns), torch.cuda.synchronize)
    ms_ind, ms_turn = et.time_turns(10)
    et.close()
    # sequential: each bunch alone through the single-bunch fused engine
    # (LDS FFT, the batched kernel's method), seq_turns turns each
    fts = []
    for k in range(len(E.trackers)):
        w = E.bunch(k)
        fts.append(FusedTracker(w.tracker, w.profile, w.tiv, method="fft", graph_chunk=seq_turns))
    for f in fts:
        f.track(2)

    def seq():
        for f in fts:
            f.track(seq_turns)
    el_seq = _timed(seq, torch.cuda.synchronize)
    for f in fts:
        f.close()
    batched = n * turns / el
    sequential = n * seq_turns / el_seq
    return {"workload": "c4: 256 bunches, N_b = 10**U(4,6), 1024 slices, 5 resonators each",
            "macroparticles": n, "value": batched, "unit": "macroparticle-turns/s",
            "turns": turns, "ms_per_turn": 1e3 * el / turns,
            "autotuned": {"block": blk, "wg_per_cu": wpc},
            "particle_kernel_ms": ms_turn, "induced_ms": ms_ind,
            "particle_kernel_frac": BYTES_PER_PARTICLE_TURN * n / (ms_turn * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "sequential_value": sequential, "sequential_turns": seq_turns,
            "speedup_vs_sequential": batched / sequential}


def _adjoint(torch, turns, every) -> dict:
    from blond_amd.trackers.differentiable import DifferentiableTracker
    from blond_amd.workloads import build_adjoint_case

    c = build_adjoint_case(n_particles=1_000_000, n_turns=turns)
    dtr = DifferentiableTracker(c.tracker, c.profile, checkpoint_every=every)
    ts = c.theta_star
    mk = lambda th, g: [torch.tensor(th[k], dtype=torch.float64, device="cuda", requires_grad=g)
                        for k in ("V", "phi", "R", "fr", "Q")]
    with torch.no_grad():
        target = dtr.profile_after(*mk(ts, False)).clone()

    def step():
        ps = mk(c.theta0, True)
        loss = torch.sum((dtr.profile_after(*ps) - target) ** 2)
        loss.backward()
    step()
    el = _timed(step, torch.cuda.synchronize)
    n = c.beam.n_macroparticles
    bpp = 80 if every == 1 else 112  # forward 32 + reverse 48 (+ recompute 32)
    dtr.close()
    how = ("every turn's state kept in HBM, no recompute" if every == 1 else
           f"checkpoints every {every} turns with recomputation")
    return {"workload": f"c5: 1e6 macroparticles x {turns} turns, 256 slices, 1 broadband "
                        f"resonator, {how}",
            "value": n * turns / el, "unit": "macroparticle-turns/s (forward + reverse)",
            "seconds": el, "bytes_per_macroparticle_turn": bpp,
            "whole_frac": bpp * n * turns / el / 1e9 / HBM_PEAK_GBS}


def form_adjoint(torch, turns=500) -> dict:
    """c5 as configured: forward + adjoint with checkpoints every 50 turns
    and recomputation, 1e6 macroparticles x 500 turns; the variant keeps
    every state in HBM (8 GB) and recomputes nothing."""
    out = _adjoint(torch, turns, 50)
    out["variants"] = {"every_state_kept": _adjoint(torch, turns, 1)}
    return out


def measured_traffic(workload: str, n: int):
    """HBM bytes per launch of the particle kernel from the committed
    FETCH_SIZE / WRITE_SIZE profile of this workload (profiles/traffic.json,
    each counter scaled by a 1 GiB copy's reading), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    if not t or t.get("macroparticles") != n:
        return None
    return t["traffic_bytes"]


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
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
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
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(a.workload, n),
                     "algorithmic_bytes": BYTES_PER_PARTICLE_TURN * n,
                     "kernel": "turn_kernel (interp kick + RF kick + drift + histogram)",
                     "kernel_ms": ms_turn, "induced_ms": ms_ind,
                     "induced_method": ft.method,
                     "whole_turn_gbs": turn_gbs, "whole_turn_frac": turn_gbs / HBM_PEAK_GBS},
    }
    ft.close()
    del w
    if snap is not None:
        line["cpu_baseline"] = cpu_baseline(snap, a.cpu_seconds)
    forms = {}
    for name in [f for f in a.forms.split(",") if f]:
        fn = {"large": form_large, "batched": form_batched, "adjoint": form_adjoint,
              "wake": form_wake, "feedback": form_feedback}[name]
        try:
            forms[name] = fn(torch)
        except Exception as exc:  # a form that cannot run is reported, not hidden
            forms[name] = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"form {name} done", file=sys.stderr, flush=True)
    if forms:
        line["forms"] = forms
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
