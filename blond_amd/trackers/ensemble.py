"""Batched ensemble tracking on the MI355X.

``EnsembleTracker`` takes one (RingAndRFTracker, Profile, TotalInducedVoltage)
triple per bunch -- independent bunches with their own particle counts, RF
settings, cuts, impedances and induced voltages in one ring -- and tracks
them together: every bunch's coordinates live in one pair of device arrays
(each bunch's ``beam.dt``/``beam.dE`` become views of them), and each turn is
two launches for all bunches (induced voltages, one workgroup per bunch;
then one streaming pass over every particle), replayed from a captured HIP
graph.  Per bunch the result equals a FusedTracker run with
``method="fft"`` bit for bit: the arithmetic and the order are the same.
The launch shape is picked by ``autotune`` (timed candidates on the live
state, which is restored).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from .. import _lib
from ..constants import e
from ..utils import bmath as bm
from .fused import (_SOLVER, phase_program, rf_program_array, ring_program_arrays,
                    total_impedance)


def _ptr(t) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


def _host(a):
    return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)


class EnsembleTracker:
    def __init__(self, trackers, profiles, total_induced_voltages=None, use_graph=True,
                 block=1024, wg_per_cu=2, graph_chunk=50, max_peel=0):
        import torch

        if bm.precision.num != 2:
            raise NotImplementedError("the ensemble runs in double precision")
        B = len(trackers)
        if B == 0 or len(profiles) != B:
            raise ValueError("one profile per tracker")
        tivs = list(total_induced_voltages) if total_induced_voltages is not None else None
        if tivs is not None and len(tivs) != B:
            raise ValueError("one total induced voltage per tracker")
        self._torch = torch
        self.trackers, self.profiles, self.tivs = list(trackers), list(profiles), tivs
        self.use_graph = bool(use_graph)
        rp0 = trackers[0].rf_params
        ring0 = ring_program_arrays(rp0)
        for tr, p in zip(trackers, profiles):
            rp = tr.rf_params
            if tr.interpolation:
                raise NotImplementedError("interpolation=True is tracked object by object")
            if (rp.n_turns, rp.n_rf, tr.solver, rp.counter[0]) != (rp0.n_turns, rp0.n_rf,
                                                                   trackers[0].solver,
                                                                   rp0.counter[0]):
                raise ValueError("bunches must share n_turns, n_rf, solver and the turn counter")
            if any(not np.array_equal(a, b) for a, b in zip(ring_program_arrays(rp), ring0)):
                raise ValueError("bunches must share the ring programs")
            if p.n_slices != profiles[0].n_slices or p.beam is not tr.beam:
                raise ValueError("profiles must share n_slices and belong to their bunch")
            if p.smooth:
                raise NotImplementedError("the ensemble bins with the nearest-bin histogram")
            if not np.array_equal(tr.acceleration_kick, trackers[0].acceleration_kick):
                raise ValueError("bunches must share the acceleration kick")
        ns = profiles[0].n_slices
        n_turns, n_rf = rp0.n_turns, rp0.n_rf
        # coordinates: bunch b at offsets[b] (even), one pad slot after an odd bunch
        counts = np.array([tr.beam.n_macroparticles for tr in trackers], dtype=np.int64)
        offsets = np.zeros(B, dtype=np.int64)
        for b in range(1, B):
            offsets[b] = offsets[b - 1] + counts[b - 1] + (counts[b - 1] & 1)
        total = int(offsets[-1] + counts[-1] + (counts[-1] & 1))
        self.counts, self.offsets = counts, offsets
        self.dt = torch.zeros(total, dtype=torch.float64, device="cuda")
        self.dE = torch.zeros(total, dtype=torch.float64, device="cuda")
        for b, tr in enumerate(trackers):
            beam = tr.beam
            o, n = int(offsets[b]), int(counts[b])
            self.dt[o:o + n] = torch.from_numpy(np.ascontiguousarray(_host(beam.dt), dtype=np.float64)).cuda()
            self.dE[o:o + n] = torch.from_numpy(np.ascontiguousarray(_host(beam.dE), dtype=np.float64)).cuda()
            beam.dt = self.dt[o:o + n]
            beam.dE = self.dE[o:o + n]
            beam.id = torch.from_numpy(np.ascontiguousarray(_host(beam.id))).cuda()
            beam._device = "GPU"
        # profiles, bin centres, induced voltages: one [bunch][n_slices] array each
        self.prof = torch.zeros((B, ns), dtype=torch.float64, device="cuda")
        self.bc = torch.from_numpy(np.stack([_host(p.bin_centers) for p in profiles])).cuda()
        for b, p in enumerate(profiles):
            p.to_gpu(recursive=False)
            self.prof[b] = torch.from_numpy(np.asarray(_host(p.n_macroparticles), dtype=np.float64)).cuda()
            p.n_macroparticles = self.prof[b]
            p.bin_centers = self.bc[b]
            p.cut_options.bin_centers = self.bc[b]
        lib = _lib.load()
        h = lib.blmi_ensemble_create(B, n_rf, n_turns, _SOLVER[trackers[0].solver])
        if not h:
            raise _lib.BlmiError(f"blmi_ensemble_create: {_lib.last_error()}")
        self._h = ct.c_void_p(h)
        _lib.call("blmi_ensemble_set_particles", self._h, _ptr(self.dt), _ptr(self.dE),
                  _dptr(offsets), _dptr(counts))
        self._keep = []
        self._upload_programs()
        cl = np.array([float(p.cut_left) for p in profiles])
        cr = np.array([float(p.cut_right) for p in profiles])
        self._keep += [cl, cr]
        _lib.call("blmi_ensemble_set_slices", self._h, ns, _dptr(cl), _dptr(cr), _ptr(self.prof))
        if tivs is not None:
            zs, nffts = zip(*[total_impedance(t) for t in tivs])
            if len(set(nffts)) != 1:
                raise NotImplementedError("bunches must share n_fft")
            Z = np.ascontiguousarray(np.stack(zs))
            scale = np.array([tr.beam.particle.charge * e * tr.beam.ratio for tr in trackers])
            self.volt = torch.zeros((B, ns), dtype=torch.float64, device="cuda")
            self._volt_rows = [self.volt[b] for b in range(B)]
            for b, t in enumerate(tivs):
                if t.profile is not profiles[b]:
                    raise ValueError("each total induced voltage must use its bunch's profile")
                t.to_gpu(recursive=False)
                t.induced_voltage = self._volt_rows[b]
            self._keep += [Z, scale]
            _lib.call("blmi_ensemble_set_induced", self._h, int(nffts[0]), _dptr(Z), _dptr(scale),
                      _ptr(self.volt), _ptr(self.bc))
        self.set_launch(block, wg_per_cu, graph_chunk, max_peel)
        _lib.call("blmi_ensemble_set_turn", self._h, int(rp0.counter[0]))
        self._engine_turn = int(rp0.counter[0])
        self._primed = False

    @property
    def n_particles(self) -> int:
        return int(self.counts.sum())

    def set_launch(self, block=1024, wg_per_cu=2, graph_chunk=50, max_peel=0):
        _lib.call("blmi_ensemble_set_launch", self._h, int(block), int(wg_per_cu),
                  int(graph_chunk), int(max_peel))

    def launch_shape(self) -> tuple[int, int]:
        b, w = ct.c_int(), ct.c_int()
        _lib.call("blmi_ensemble_get_launch", self._h, ct.byref(b), ct.byref(w))
        return b.value, w.value

    def _stream(self):
        return ct.c_void_p(self._torch.cuda.current_stream().cuda_stream)

    def _upload_programs(self):
        """Every bunch's RF programs (its phase offsets accumulated from the
        current turn on) and the shared ring programs to the engine."""
        trackers = self.trackers
        rp0 = trackers[0].rf_params
        ring0 = ring_program_arrays(rp0)
        self._phase = [phase_program(tr.rf_params) for tr in trackers]
        rf = np.ascontiguousarray(np.stack([rf_program_array(tr.rf_params, ph[0])
                                            for tr, ph in zip(trackers, self._phase)]))
        acc = np.zeros(rp0.n_turns + 1, dtype=np.float64)
        acc[: len(trackers[0].acceleration_kick)] = trackers[0].acceleration_kick
        self._programs = [rf, acc, ring0]
        _lib.call("blmi_ensemble_set_programs", self._h, _dptr(rf), _dptr(acc), _dptr(ring0[0]),
                  float(rp0.length_ratio), float(rp0.alpha_order), *[_dptr(a) for a in ring0[1:]],
                  float(rp0.particle.charge))

    def update_programs(self):
        """Re-read the bunches' per-turn RF programs (on the GPU their
        device arrays, which the objects' own turns read) after the caller
        changed them; the next track() uses them from the current turn on."""
        self._torch.cuda.synchronize()  # nothing in flight reads the old tables
        for tr in self.trackers:
            rp = tr.rf_params
            if rp._device == "GPU":
                for name in ("voltage", "omega_rf", "phi_rf"):
                    rp._host[name][...] = getattr(rp, name).cpu().numpy()
        self._upload_programs()

    def _sync(self):
        """Turns tracked by the bunches' own objects between two calls
        (their coordinates are views of the ensemble's arrays, so those
        turns moved them in place): continue from the bunches' common turn
        with the staged histograms rebuilt."""
        turns = {int(tr.rf_params.counter[0]) for tr in self.trackers}
        if len(turns) != 1:
            raise RuntimeError(f"the bunches' RF counters differ ({sorted(turns)}): the "
                               "ensemble tracks all of them through the same turns")
        for tr, o in zip(self.trackers, self.offsets):
            if (tr.beam.dt.data_ptr() != self.dt[int(o):].data_ptr()
                    or tr.beam.dE.data_ptr() != self.dE[int(o):].data_ptr()):
                raise RuntimeError("a bunch's beam.dt or beam.dE is no longer a view of the "
                                   "ensemble's arrays")
        t = turns.pop()
        if t != self._engine_turn:
            _lib.call("blmi_ensemble_set_turn", self._h, t)
            self._engine_turn = t
            self._primed = False

    def _prime(self, stream):
        self._sync()
        if self.tivs is not None and not self._primed:
            _lib.call("blmi_ensemble_prime", self._h, stream)
            self._primed = True

    def autotune(self, turns: int = 5) -> tuple[int, int, float]:
        """Pick the launch shape from timed candidates (state restored);
        returns (block, workgroups per CU, ms per turn)."""
        stream = self._stream()
        self._prime(stream)
        b, w, ms = ct.c_int(), ct.c_int(), ct.c_float()
        _lib.call("blmi_ensemble_autotune", self._h, int(turns), ct.byref(b), ct.byref(w),
                  ct.byref(ms), stream)
        return b.value, w.value, ms.value

    def track(self, n_turns: int = 1):
        n_turns = int(n_turns)
        if n_turns <= 0:
            return
        stream = self._stream()
        self._prime(stream)
        start = self.trackers[0].rf_params.counter[0]
        _lib.call("blmi_ensemble_track", self._h, n_turns, 1 if self.use_graph else 0, stream)
        self._advance(start, start + n_turns)

    def time_turns(self, turns: int) -> tuple[float, float]:
        """Eager turns with each kernel timed on the engine's stream:
        (induced ms, particle kernel ms) per turn; advances like track."""
        stream = self._stream()
        self._prime(stream)
        a, b = ct.c_float(), ct.c_float()
        start = self.trackers[0].rf_params.counter[0]
        _lib.call("blmi_ensemble_time_turns", self._h, int(turns), ct.byref(a), ct.byref(b), stream)
        self._advance(start, start + int(turns))
        return a.value, b.value

    def _advance(self, start: int, end: int):
        self._engine_turn = end
        if self.tivs is not None:
            # the objects' own induced-voltage step (turns tracked between
            # two calls) installs new arrays; the engine writes these rows
            for b, t in enumerate(self.tivs):
                if t.induced_voltage is not self._volt_rows[b]:
                    t.induced_voltage = self._volt_rows[b]
        for tr, (phi, dphi, moves) in zip(self.trackers, self._phase):
            rp, beam = tr.rf_params, tr.beam
            if end > start and moves:
                rp.host_array("phi_rf")[:, start + 1:end + 1] = phi[:, start + 1:end + 1]
                rp.dphi_rf[:] = dphi[:, end]
            rp.counter[0] = end
            beam.beta = rp.beta[end]
            beam.gamma = rp.gamma[end]
            beam.energy = rp.energy[end]
            beam.momentum = rp.momentum[end]

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().blmi_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
