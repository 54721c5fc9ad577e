#!/usr/bin/env python3
"""Throughput of the fused tracking turn on one MI355X.

A step is one turn of the reference's map over the whole bunch (profile ->
induced voltage -> interpolated kick + RF kick -> drift with intensity
effects; RF kick -> drift -> profile without).  The headline workload is the
large-bunch double-harmonic stress case: 1e7 macroparticles, RF systems at
h = 35640 and 142560, 65536 slices, FFT 131072, 21 resonators, FP64, on a
synthetic seeded Gaussian bunch.

The JSON line carries the contract keys plus
  roofline      the whole fused turn: 32 B per macroparticle-turn (dt and dE
                read and written once) / the turn's average time (HIP events
                on the engine's stream around the timed turns), against
                8 TB/s; `kernel` holds the same figure for the particle
                kernel alone (its own HIP events), which the committed
                rocprof summary confirms, and `traffic` the HBM bytes a
                launch of it moves (FETCH_SIZE + WRITE_SIZE, scaled by a
                copy kernel's reading)
  cpu_baseline  the reference's own C++ kernels (oracle/_ref, built with the
                flags of `blond-compile --parallel --optimize`) on the host
                cores, same turn loop with numpy FFTs, bounded sample
  forms         every form of the metric on its own workload, each with its
                particle working set against the 256 MB last-level cache
                and, where one is defined, its own CPU baseline:
                  plain     c1: 1e6 macroparticles, 100 slices, no intensity,
                            25 turns a launch (and a launch a turn)
                  intensity c2: 4e6, 4096 slices, FFT 16384, 21 resonators
                  large     the headline's variants: 3e7 FP64, 1e7 and 3e7 FP32
                  batched   c4: 256 bunches (1e4..1e6 each), one launch pair
                            per turn, autotuned, against 256 sequential runs
                  adjoint   c5: 1e6 x 500 turns, forward + reverse pass with
                            checkpoints every 50 turns and recomputation
                  wake      c2 bunch with a 4-turn multi-turn wake and Z/n
                  feedback  c2 bunch under the PSB beam phase loop
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
LLC_MB = 256.0  # MI355X last-level (Infinity) cache
BYTES_PER_PARTICLE_TURN = 32  # FP64 dt, dE: 16 B read + 16 B written
METRIC = "macroparticle-turns/s + % HBM roofline, plain vs batched vs adjoint, 1 GPU"
DESCR = {
    "c1": "c1: LHC ring, h=35640, 6 MV, 100 slices, no intensity effects",
    "c2": "c2: PSB-like ring, h=1, 8 kV, 4096 slices, n_fft 16384, 21 resonators",
    "c3": "c3: LHC ring, h=35640+142560, 65536 slices, n_fft 131072, 21 resonators",
}
DEFAULT_N = {"c1": 1_000_000, "c2": 4_000_000, "c3": 10_000_000}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3"])
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--f32", action="store_true", help="the FP32 mode (float coordinates)")
    ap.add_argument("--timing-turns", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--blocks-per-cu", type=int, default=1)
    ap.add_argument("--groups", type=int, default=0, help="count copies (0: automatic)")
    ap.add_argument("--max-peel", type=int, default=0)
    ap.add_argument("--turn-block", type=int, default=25,
                    help="plain turns: at most this many turns a launch (1: a launch a turn)")
    ap.add_argument("--method", choices=["auto", "fft", "direct", "fourstep"], default="auto")
    ap.add_argument("--forms", default="plain,intensity,large,batched,adjoint,wake,feedback",
                    help="comma list of extra forms to measure ('' for none)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the known-size kernel "
                         "that calibrates FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only run the c4 ensemble turns (for counter collection) and exit")
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


def _events_ms(torch, fn):
    """Device time of fn's work on torch's current stream (HIP events)."""
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b)


def working_set(n: int, bytes_per_particle: int = 16) -> dict:
    mb = n * bytes_per_particle / 1e6
    return {"particle_mb": mb, "fits_llc_256mb": mb <= LLC_MB}


def frac(bytes_, seconds) -> float:
    return bytes_ / seconds / 1e9 / HBM_PEAK_GBS


def _workload(torch, name, n, turns, f32=False):
    from blond_amd.workloads import build_workload

    w = build_workload(name, n_particles=n, n_turns=turns)
    if f32:
        # the FP32 mode: float coordinates (the slice axis stays FP64)
        w.beam.dt = w.beam.dt.to(torch.float32).contiguous()
        w.beam.dE = w.beam.dE.to(torch.float32).contiguous()
    return w


def _measure_turns(torch, w, turns, label, f32=False, timing_turns=50, **launch) -> dict:
    """Whole-turn throughput of a fused workload plus its kernel split."""
    from blond_amd.trackers.fused import FusedTracker

    n = w.beam.n_macroparticles
    bpp = BYTES_PER_PARTICLE_TURN // 2 if f32 else BYTES_PER_PARTICLE_TURN
    ft = FusedTracker(w.tracker, w.profile, w.tiv, **launch)
    ft.track(10)
    el = _timed(lambda: ft.track(turns), torch.cuda.synchronize)
    ms_ind, ms_turn = ft.time_kernels(timing_turns)
    out = {"workload": label, "value": n * turns / el, "unit": "macroparticle-turns/s",
           "macroparticles": n, "turns": turns, "ms_per_turn": 1e3 * el / turns,
           "dtype": "f32" if f32 else "f64", "induced_method": ft.method,
           "bytes_per_macroparticle_turn": bpp,
           "whole_turn_frac": frac(bpp * n, el / turns),
           "particle_kernel_ms": ms_turn, "induced_ms": ms_ind,
           "particle_kernel_frac": frac(bpp * n, ms_turn * 1e-3),
           "working_set": working_set(n, bpp // 2)}
    ft.close()
    return out


# ---------------------------------------------------------------------------
# CPU baselines (bounded samples of the same workloads)
# ---------------------------------------------------------------------------
def _cpu_lib():
    from oracle import oracle as orc

    lib = orc.ref_lib("opt", "double")
    if lib is not None:
        return lib, "reference"
    return orc.oracle_lib(), "port"


def _cores() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))


def cpu_baseline(snapshot: dict, target_s: float) -> dict:
    """The reference's C++ kernels (or, without oracle/_ref, the oracle's C
    restatement) on the same workload, for about target_s seconds of turns."""
    from oracle import oracle as orc

    lib, kind = _cpu_lib()
    t0 = time.perf_counter()
    st = orc.track(snapshot, 1, lib)
    one = time.perf_counter() - t0
    turns = max(2, min(int(target_s / max(one, 1e-6)), len(snapshot["acc"]) - 2))
    t0 = time.perf_counter()
    orc.track(st, turns, lib)
    el = time.perf_counter() - t0
    n = len(snapshot["dt"])
    return {"value": n * turns / el, "unit": "macroparticle-turns/s", "cores": _cores(),
            "kind": kind,
            "sample": f"{snapshot['name']}: {n} macroparticles x {turns} turns "
                      f"({el:.1f} s), numpy FFT single-threaded"}


def cpu_baseline_ensemble(E, target_s: float) -> dict:
    """256 sequential CPU runs of the c4 bunches: every bunch tracked for the
    same number of turns, one after the other."""
    from oracle import oracle as orc

    lib, kind = _cpu_lib()
    snaps = [E.bunch(b).snapshot() for b in range(len(E.trackers))]
    t0 = time.perf_counter()
    for s in snaps:
        orc.track(s, 1, lib)
    one = time.perf_counter() - t0
    turns = max(1, int(target_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    for s in snaps:
        orc.track(s, turns, lib)
    el = time.perf_counter() - t0
    n = sum(len(s["dt"]) for s in snaps)
    return {"value": n * turns / el, "unit": "macroparticle-turns/s", "cores": _cores(),
            "kind": kind,
            "sample": f"c4: the 256 bunches one after the other ({n} macroparticles) x "
                      f"{turns} turns ({el:.1f} s), numpy FFT single-threaded"}


def cpu_baseline_adjoint(case, target_s: float) -> dict:
    """The c5 forward map on the CPU (the oracle's float64 restatement, numpy
    vectorised on one core) on a bounded sample of turns."""
    import numpy as np

    from oracle import adjoint as adj

    dm = adj.DiffMap(case.state)
    th = case.theta0
    dt = np.asarray(case.beam.dt.cpu().numpy() if hasattr(case.beam.dt, "cpu") else case.beam.dt)
    dE = np.asarray(case.beam.dE.cpu().numpy() if hasattr(case.beam.dE, "cpu") else case.beam.dE)
    t0 = time.perf_counter()
    dm.forward(dt, dE, th["V"], th["phi"], th["R"], th["fr"], th["Q"], 1)
    one = time.perf_counter() - t0
    turns = max(1, min(int(target_s / max(one, 1e-6)), 400))
    t0 = time.perf_counter()
    dm.forward(dt, dE, th["V"], th["phi"], th["R"], th["fr"], th["Q"], turns)
    el = time.perf_counter() - t0
    n = len(dt)
    return {"value": n * turns / el, "unit": "macroparticle-turns/s (forward only)", "cores": 1,
            "kind": "port",
            "sample": f"c5 forward map: {n} macroparticles x {turns} turns ({el:.1f} s), "
                      "numpy on one core"}


# ---------------------------------------------------------------------------
# forms
# ---------------------------------------------------------------------------
# FP64 operations of one plain c1 macroparticle-turn (1 RF system, simple
# drift), contraction off: the sine's argument (2), Cephes reduction (1 + 6),
# r^2 (1), the octant's degree-6 polynomial (10; the kernel reads the
# branch's coefficients from LDS instead of evaluating both), its final form
# A + B P (5), V sin and the sum (2), charge and acceleration (2), dE += (1),
# drift (2)
C1_FP64_OPS = 32
FP64_PEAK_TFLOPS = 78.6  # published, an FMA counted as two operations


def form_plain(torch, a, turns=1000) -> dict:
    """c1 as configured: 1e6 macroparticles, 100 slices, 1000 turns of
    kick + drift + profile.  The engine runs 25 turns a launch with the
    particles in registers (only the last turn of a launch is
    histogrammed); the variant runs a launch a turn.  With 25 turns a
    launch the bound is FP64 issue, not HBM: the roofline is operations
    per second against the FP64 vector peak."""
    w = _workload(torch, "c1", 1_000_000, 2 * turns + 400)
    snap = None if a.no_cpu_baseline else w.snapshot()
    k = 25
    out = _measure_turns(torch, w, turns, DESCR["c1"] + f", 1e6 macroparticles, FP64, "
                         f"{k} turns a launch", turn_block=k)
    n = out["macroparticles"]
    # dt, dE cross the memory system once a launch
    out["bytes_per_macroparticle_turn"] = BYTES_PER_PARTICLE_TURN / k
    out["whole_turn_frac"] = frac(out["bytes_per_macroparticle_turn"] * n,
                                  out["ms_per_turn"] * 1e-3)
    out["particle_kernel_frac"] = frac(out["bytes_per_macroparticle_turn"] * n,
                                       out["particle_kernel_ms"] * 1e-3)
    tflops = C1_FP64_OPS * n / (out["particle_kernel_ms"] * 1e-3) / 1e12
    out["roofline"] = {"bound": "fp64_valu", "ops_per_macroparticle_turn": C1_FP64_OPS,
                       "achieved_tflops": tflops, "peak_tflops": FP64_PEAK_TFLOPS,
                       "frac": tflops / FP64_PEAK_TFLOPS,
                       "note": "no FMA (bit parity with the reference's build): "
                               "multiplies and adds alone reach at most 0.5"}
    one = _measure_turns(torch, w, turns, DESCR["c1"] + ", 1e6 macroparticles, FP64, "
                         "a launch a turn", turn_block=1)
    out["variants"] = {"a_launch_a_turn": one}
    if snap is not None:
        out["cpu_baseline"] = cpu_baseline(snap, a.cpu_seconds)
    return out


def form_intensity(torch, a, turns=1000) -> dict:
    """c2: 4e6 macroparticles, 4096 slices, direct convolution."""
    w = _workload(torch, "c2", 4_000_000, turns + 200)
    snap = None if a.no_cpu_baseline else w.snapshot()
    out = _measure_turns(torch, w, turns, DESCR["c2"] + ", 4e6 macroparticles, FP64")
    if snap is not None:
        out["cpu_baseline"] = cpu_baseline(snap, a.cpu_seconds)
    return out


def form_large(torch, a, turns=200) -> dict:
    """The headline's variants: 3e7 macroparticles in FP64, 1e7 and 3e7 in
    the FP32 mode."""
    out = {}
    for key, n, f32, t in (("3e7_f64", 30_000_000, False, turns // 2),
                           ("1e7_f32", 10_000_000, True, turns),
                           ("3e7_f32", 30_000_000, True, turns // 2)):
        w = _workload(torch, "c3", n, t + 120, f32)
        out[key] = _measure_turns(torch, w, t, DESCR["c3"] + f", {n:.0e} macroparticles, "
                                  f"{'FP32' if f32 else 'FP64'}", f32=f32)
        out[key]["traffic_bytes"] = measured_traffic(f"c3_{n:.0e}" + ("_f32" if f32 else ""), n)
        del w
        torch.cuda.empty_cache()
    return out


def form_wake(torch, a, turns=200) -> dict:
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
    return _measure_turns(torch, w, turns, "c2 bunch 4e6, 4096 slices; multi-turn wake over 4 "
                          f"turns (memory {src.n_mtw_memory}, n_fft {src.n_fft}) + inductive Z/n",
                          timing_turns=30)


def form_feedback(torch, a, turns=200) -> dict:
    """c2-ring bunch under the PSB phase and radial loops, with intensity."""
    from blond_amd.workloads import build_feedback_workload

    w = build_feedback_workload("PSB", n_particles=4_000_000, n_turns=turns + 60,
                                n_slices=4096, intensity=True)
    return _measure_turns(torch, w, turns, "c2-ring bunch 4e6, 4096 slices, broadband "
                          "resonator, PSB beam phase + radial loop on the device",
                          timing_turns=30)


def form_batched(torch, a, turns=100, seq_turns=20) -> dict:
    """c4: 256 bunches in one launch pair per turn vs 256 sequential fused runs."""
    from blond_amd.trackers.fused import FusedTracker
    from blond_amd.workloads import build_ensemble

    E = build_ensemble(n_bunches=256, n_turns=turns + seq_turns + 40)
    n = E.n_particles
    cpu = None if a.no_cpu_baseline else cpu_baseline_ensemble(E, a.cpu_seconds)
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
    out = {"workload": "c4: 256 bunches, N_b = 10**U(4,6), 1024 slices, 5 resonators each",
           "macroparticles": n, "value": batched, "unit": "macroparticle-turns/s",
           "turns": turns, "ms_per_turn": 1e3 * el / turns,
           "autotuned": {"block": blk, "wg_per_cu": wpc},
           "whole_turn_frac": frac(BYTES_PER_PARTICLE_TURN * n, el / turns),
           "particle_kernel_ms": ms_turn, "induced_ms": ms_ind,
           "particle_kernel_frac": frac(BYTES_PER_PARTICLE_TURN * n, ms_turn * 1e-3),
           "sequential_value": sequential, "sequential_turns": seq_turns,
           "speedup_vs_sequential": batched / sequential,
           "working_set": working_set(n), "traffic_bytes": measured_traffic("c4", n)}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out


def _adjoint(torch, turns, every, case=None, map="reference") -> dict:
    from blond_amd.trackers.differentiable import DifferentiableTracker
    from blond_amd.workloads import build_adjoint_case

    c = case or build_adjoint_case(n_particles=1_000_000, n_turns=turns)
    dtr = DifferentiableTracker(c.tracker, c.profile, checkpoint_every=every, map=map)
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
    what = ("smooth_histogram + linear_interp_kick (the reference's map)" if map == "reference"
            else "quadratic B-spline deposit and kick (smooth variant)")
    return {"workload": f"c5: 1e6 macroparticles x {turns} turns, 256 slices, 1 broadband "
                        f"resonator, {what}, {how}", "map": map,
            "value": n * turns / el, "unit": "macroparticle-turns/s (forward + reverse)",
            "seconds": el, "bytes_per_macroparticle_turn": bpp,
            "whole_frac": frac(bpp * n * turns, el),
            "working_set": working_set(n)}


def form_adjoint(torch, a, turns=500) -> dict:
    """c5 as configured: forward + adjoint of the reference's map
    (smooth_histogram + linear_interp_kick) with checkpoints every 50 turns
    and recomputation, 1e6 macroparticles x 500 turns; the variants keep
    every state in HBM (8 GB) and recompute nothing, or differentiate the
    smooth B-spline map."""
    from blond_amd.workloads import build_adjoint_case

    case = build_adjoint_case(n_particles=1_000_000, n_turns=turns)
    cpu = None if a.no_cpu_baseline else cpu_baseline_adjoint(case, a.cpu_seconds)
    out = _adjoint(torch, turns, 50)
    out["variants"] = {"every_state_kept": _adjoint(torch, turns, 1),
                       "spline_map": _adjoint(torch, turns, 50, map="spline")}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out


FORMS = {"plain": form_plain, "intensity": form_intensity, "large": form_large,
         "batched": form_batched, "adjoint": form_adjoint, "wake": form_wake,
         "feedback": form_feedback}


def measured_traffic(key: str, n: int):
    """HBM bytes per launch of the particle kernel from the committed
    FETCH_SIZE / WRITE_SIZE profile of this workload (profiles/traffic.json,
    each counter scaled by a 1 GiB copy's reading), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    if not t or t.get("macroparticles") != n:
        return None
    return t["traffic_bytes"]


def ensemble_only(torch, turns=20):
    """The c4 ensemble's turns alone (counter collection)."""
    from blond_amd.workloads import build_ensemble

    E = build_ensemble(n_bunches=256, n_turns=turns + 20)
    et = E.tracker()
    et.track(turns)
    torch.cuda.synchronize()
    et.close()
    print(json.dumps({"ensemble_turns": turns, "macroparticles": int(E.n_particles)}), flush=True)


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

    bm.use_gpu(0)
    if a.ensemble_only:
        ensemble_only(torch)
        return
    n_turns = a.warmup + a.repeats * a.steps + a.timing_turns + 2
    n = a.particles or DEFAULT_N[a.workload]
    w = _workload(torch, a.workload, n, n_turns, a.f32)
    bpp = BYTES_PER_PARTICLE_TURN // 2 if a.f32 else BYTES_PER_PARTICLE_TURN
    snap = None if (a.no_cpu_baseline or a.f32) else w.snapshot()
    ft = FusedTracker(w.tracker, w.profile, w.tiv, use_graph=not a.no_graph, block=a.block,
                      blocks_per_cu=a.blocks_per_cu, groups=a.groups, max_peel=a.max_peel,
                      method=None if a.method == "auto" else a.method, turn_block=a.turn_block)
    ft.track(a.warmup)
    torch.cuda.synchronize()
    times, dev_ms = [], []
    for _ in range(a.repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev_ms.append(_events_ms(torch, lambda: ft.track(a.steps)))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t_med = statistics.median(times)
    value = n * a.steps / t_med
    turn_s = statistics.median(dev_ms) * 1e-3 / a.steps  # HIP events, engine's stream

    # live per-kernel durations (HIP events on the engine's stream)
    ms_ind, ms_turn = ft.time_kernels(a.timing_turns)
    # plain turns run turn_block to a launch: the coordinates cross the
    # memory system once a launch
    blocked = not w.intensity and a.turn_block > 1
    alg = bpp * n / (a.turn_block if blocked else 1)
    turn_gbs = alg / turn_s / 1e9
    kern_gbs = alg / (ms_turn * 1e-3) / 1e9
    key = f"{a.workload}_{n:.0e}" + ("_f32" if a.f32 else "")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "macroparticle-turns/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_med / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.f32 else "f64",
        "data": "synthetic (seeded Gaussian bunch, default_rng(1)/(2); resonators default_rng(42))",
        "config": {"workload": DESCR[a.workload] + f", {n:.0e} macroparticles, "
                               f"{'FP32' if a.f32 else 'FP64'}, fused "
                               f"{'intensity' if w.intensity else 'plain'} turn",
                   "macroparticles": n, "n_slices": w.meta.get("n_slices"),
                   "n_fft": w.meta.get("n_fft"),
                   "form": "intensity" if w.intensity else "plain",
                   "turn_block": a.turn_block if blocked else 1,
                   "graph": not a.no_graph, "repeats": a.repeats,
                   "block": a.block, "blocks_per_cu": a.blocks_per_cu, "groups": a.groups,
                   "max_peel": a.max_peel, "working_set": working_set(n, bpp // 2)},
        "roofline": {"bound": "hbm", "achieved": turn_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": turn_gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic(key, n),
                     "algorithmic_bytes": alg,
                     "what": f"whole fused turn (slice axis + particle kernel), {alg / n:g} B "
                             "per macroparticle-turn",
                     "turn_ms": turn_s * 1e3,
                     "kernel": {"name": "turn_kernel (interp kick + RF kick + drift + "
                                        "histogram)",
                                "ms": ms_turn, "achieved": kern_gbs,
                                "frac": kern_gbs / HBM_PEAK_GBS},
                     "slice_axis_ms": ms_ind, "induced_method": ft.method},
    }
    ft.close()
    del w
    torch.cuda.empty_cache()
    if snap is not None:
        line["cpu_baseline"] = cpu_baseline(snap, a.cpu_seconds)
        del snap
    forms = {}
    for name in [f for f in a.forms.split(",") if f]:
        try:
            forms[name] = FORMS[name](torch, a)
        except Exception as exc:  # a form that cannot run is reported, not hidden
            forms[name] = {"error": f"{type(exc).__name__}: {exc}"}
        torch.cuda.empty_cache()
        print(f"form {name} done", file=sys.stderr, flush=True)
    if forms:
        line["forms"] = forms
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
