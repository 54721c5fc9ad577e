"""The multi-GPU decomposition of the turn, rehearsed on the CPU with two
gloo ranks: particles split into contiguous blocks, each rank slices its
own block, one all-reduce of the histogram per turn, every rank computes the
same induced voltage from the summed profile and kicks and drifts its own
particles.  Integer histogram sums make the exchange exact, so each rank's
block must equal the single-process turn loop bit for bit."""
import os
import socket

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import torch.distributed as dist  # noqa: E402
import torch.multiprocessing as mp  # noqa: E402

from oracle import oracle as orc  # noqa: E402

N_TURNS = 6


def _state():
    """A c1 bunch (1e5 particles, 100 slices) with a resonator impedance,
    built on the host (closed-form Z, oracle.resonator_impedance)."""
    from blond_amd.workloads import build_workload

    w = build_workload("c1", n_particles=100_000, n_turns=20, n_slices=100, to_gpu=False)
    st = w.snapshot()
    ns = st["n_slices"]
    bin_size = (st["cut_right"] - st["cut_left"]) / ns
    n_fft = 4 * ns
    freq = np.fft.rfftfreq(n_fft, bin_size)
    Z = orc.resonator_impedance(freq, [2e5, 5e4], [1.2e9, 0.4e9], [1.0, 30.0]) / bin_size
    st.update(intensity=True, Z=Z, n_fft=n_fft, n_induced_voltage=n_fft,
              ratio=float(w.beam.ratio), induced_voltage=np.zeros(ns))
    return st


def _sharded_turns(st, rank, world, n_turns, lib):
    """The per-rank loop: local slicing, all-reduce, redundant induced
    voltage, local kicks and drift (tracker.py:521-615, profile.py:700-719,
    impedance.py:113-146 on a contiguous block of particles)."""
    n = len(st["dt"])
    lo, hi = rank * n // world, (rank + 1) * n // world
    dt = np.array(st["dt"][lo:hi], copy=True)
    dE = np.array(st["dE"][lo:hi], copy=True)
    prof = np.zeros(st["n_slices"])
    for t in range(int(st["turn"]), int(st["turn"]) + n_turns):
        lib.slice_beam(dt, prof, st["cut_left"], st["cut_right"])
        total = torch.from_numpy(prof.copy())
        dist.all_reduce(total, op=dist.ReduceOp.SUM)
        prof = total.numpy().copy()
        volt = orc.induced_voltage_1turn(prof, st["Z"], st["n_fft"], st["charge"], st["ratio"],
                                         st["n_induced_voltage"], st["n_slices"])
        lib.linear_interp_kick(dt, dE, volt, st["bin_centers"], st["charge"], 0.0)
        lib.kick(dt, dE, st["voltage"][:, t], st["omega_rf"][:, t], st["phi_rf"][:, t],
                 st["charge"], st["n_rf"], st["acc"][t])
        k = t + 1
        lib.drift(dt, dE, st["solver"], st["t_rev"][k], st["length_ratio"], st["alpha_order"],
                  st["eta_0"][k], st["eta_1"][k], st["eta_2"][k], st["alpha_0"][k],
                  st["alpha_1"][k], st["alpha_2"][k], st["beta"][k], st["energy"][k])
    return dt, dE, prof, volt


def _worker(rank, world, port, out_dir):
    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist.init_process_group("gloo", rank=rank, world_size=world)
    try:
        st = _state()
        dt, dE, prof, volt = _sharded_turns(st, rank, world, N_TURNS, orc.oracle_lib())
        np.savez(os.path.join(out_dir, f"rank{rank}.npz"), dt=dt, dE=dE, prof=prof, volt=volt)
    finally:
        dist.destroy_process_group()


def _free_port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def test_two_rank_turns_equal_single_process(tmp_path):
    world = 2
    mp.spawn(_worker, args=(world, _free_port(), str(tmp_path)), nprocs=world, join=True)
    ref = orc.track(_state(), N_TURNS)
    parts = [np.load(str(tmp_path / f"rank{r}.npz")) for r in range(world)]
    np.testing.assert_array_equal(np.concatenate([p["dt"] for p in parts]), ref["dt"])
    np.testing.assert_array_equal(np.concatenate([p["dE"] for p in parts]), ref["dE"])
    for p in parts:
        # every rank saw the whole bunch's profile and the same voltage
        np.testing.assert_array_equal(p["volt"], ref["induced_voltage"])
        np.testing.assert_array_equal(p["prof"], parts[0]["prof"])
