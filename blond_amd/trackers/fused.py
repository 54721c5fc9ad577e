"""Fused tracking turn on the MI355X.

``FusedTracker`` runs the same map a BLonD main script runs object by object,

    with intensity effects: [profile, total_induced_voltage, tracker]
                            (EX_18_robinson_instability.py:226)
    without:                [tracker, profile]  (EX_01_Acceleration.py:152-153)
                            or [tracker]

as at most two kernels per turn (induced voltage, then one streaming pass
doing interpolated kick + RF kick + drift + histogram), replayed from a
captured HIP graph.  Without intensity effects, statistics, a phase loop or
an interpolation table, up to ``turn_block`` turns run in one launch with
the particles held in registers (only the last turn is histogrammed, which
is all a caller can see).  After ``track(k)`` the objects hold exactly what
k passes of the map leave in them: ``beam.dt/dE``, ``profile.n_macroparticles``,
``total_induced_voltage.induced_voltage``, ``rf_station.counter`` and the
beam's synchronous energy.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from .. import _lib
from ..constants import e
from ..impedances.impedance import InducedVoltageFreq, InducedVoltageTime, InductiveImpedance
from ..utils import bmath as bm

_SOLVER = {"simple": 0, "legacy": 1, "exact": 2}
_METHODS = {"fft": 0, "direct": 1, "fourstep": 2}


def _ptr(t) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


def _host_array(a):
    return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)


def phase_program(rp):
    """phi_rf for all turns with the accumulated offsets of tracker.py:556-568
    applied, from the current counter on (the recurrence depends only on the
    programs, so it can be precomputed).  Returns (phi [n_rf, n_turns+1],
    dphi [n_rf, n_turns+1], whether any offset is non-zero)."""
    phi = np.array(rp.host_array("phi_rf"), dtype=np.float64, copy=True)
    omega = rp.host_array("omega_rf")
    dphi = np.array(rp.dphi_rf, dtype=np.float64, copy=True)
    dphi_prog = np.zeros((rp.n_rf, rp.n_turns + 1))
    dphi_prog[:, rp.counter[0]] = dphi
    for t in range(rp.counter[0], rp.n_turns):
        dphi = dphi + (2.0 * np.pi * rp.harmonic[:, t + 1]
                       * (omega[:, t + 1] - rp.omega_rf_d[:, t + 1]) / rp.omega_rf_d[:, t + 1])
        phi[:, t + 1] += dphi
        dphi_prog[:, t + 1] = dphi
    return phi, dphi_prog, bool(np.any(dphi_prog != 0))


def rf_program_array(rp, phi) -> np.ndarray:
    """[n_turns+1][3][n_rf]: voltage, omega_rf and phi_rf per turn."""
    rf = np.empty((rp.n_turns + 1, 3, rp.n_rf), dtype=np.float64)
    rf[:, 0, :] = np.asarray(rp.host_array("voltage"), dtype=np.float64).T
    rf[:, 1, :] = np.asarray(rp.host_array("omega_rf"), dtype=np.float64).T
    rf[:, 2, :] = phi.T
    return np.ascontiguousarray(rf)


def ring_program_arrays(rp) -> list:
    """t_rev, eta_0..2, alpha_0..2, beta, energy per turn (contiguous f64)."""
    return [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in
            (rp.t_rev, rp.eta_0, rp.eta_1, rp.eta_2, rp.alpha_0, rp.alpha_1, rp.alpha_2,
             rp.beta, rp.energy)]


_CONV = (InducedVoltageFreq, InducedVoltageTime)
_DERIV = {"gradient": 0, "diff": 1}


def _conv_sources(tiv) -> list:
    srcs = tiv.induced_voltage_list
    if not srcs or not all(isinstance(s, _CONV + (InductiveImpedance,)) for s in srcs):
        raise NotImplementedError("the fused turn supports InducedVoltageFreq, InducedVoltageTime "
                                  "and InductiveImpedance sources")
    return [s for s in srcs if isinstance(s, _CONV)]


def total_impedance(tiv) -> tuple[np.ndarray | None, int]:
    """Sum of the convolution sources' impedance tables (InducedVoltageFreq:
    Z/bin_size; InducedVoltageTime: rfft of the wake) and their common n_fft;
    (None, 0) when there is none."""
    conv = _conv_sources(tiv)
    if not conv:
        return None, 0
    n_fft = conv[0].n_fft
    if any(s.n_fft != n_fft for s in conv):
        raise NotImplementedError("all convolution sources must share n_fft")
    z = 0
    for s in conv:
        zi = s.total_impedance
        zi = zi.cpu().numpy() if hasattr(zi, "cpu") else np.asarray(zi)
        z = z + zi
    return np.ascontiguousarray(z, dtype=np.complex128), int(n_fft)


def wake_setup(tiv, n_turns: int):
    """The fused turn's wake step for ``tiv``: (mtw source or None, inductive
    coefficient program [n_turns+1] or None, derivative mode).  A multi-turn
    source must be the only convolution source (its own voltage alone enters
    the memory); inductive sources must share a derivative mode (their
    coefficients are summed)."""
    conv = _conv_sources(tiv)
    mtw = [s for s in conv if s.multi_turn_wake]
    if mtw and len(conv) > 1:
        raise NotImplementedError("a multi-turn wake source must be the only convolution source")
    ind = [s for s in tiv.induced_voltage_list if isinstance(s, InductiveImpedance)]
    coef, mode = None, 0
    if ind:
        modes = {s.deriv_mode for s in ind}
        if len(modes) != 1 or next(iter(modes)) not in _DERIV:
            raise NotImplementedError("inductive sources: one derivative mode, 'gradient' or "
                                      "'diff' ('filter1d' is not supported on the GPU)")
        mode = _DERIV[next(iter(modes))]
        coef = np.zeros(n_turns + 1, dtype=np.float64)
        for s in ind:
            for t in range(n_turns + 1):
                try:
                    coef[t] += float(s.coefficient(t))
                except IndexError:
                    # the last entry is never used: turn n_turns is not tracked
                    if t < n_turns:
                        raise ValueError(f"Z_over_n has no entry for turn {t}") from None
                    coef[t] = np.nan
    return (mtw[0] if mtw else None), coef, mode


class FusedTracker:
    def __init__(self, tracker, profile=None, total_induced_voltage=None, use_graph=True,
                 block=1024, blocks_per_cu=1, graph_chunk=50, groups=0, max_peel=0,
                 method=None, sort_every=0, statistics=False, monitor=None, turn_block=25):
        import torch

        self.tracker = tracker
        self.profile = profile
        self.tiv = total_induced_voltage
        self.fb = tracker.beamFB
        if self.fb is not None and profile is None:
            raise ValueError("the beam phase loop needs the profile")
        self.beam = tracker.beam
        self.rf = tracker.rf_params
        self.use_graph = bool(use_graph)
        if self.tiv is not None and self.profile is None:
            raise ValueError("intensity effects need the profile")
        if self.tiv is not None and self.tiv.profile is not self.profile:
            raise ValueError("total_induced_voltage must use the same profile")
        if self.profile is not None and self.profile.smooth:
            raise NotImplementedError("the fused turn bins with the nearest-bin histogram")
        tracker.to_gpu()
        if self.profile is not None:
            self.profile.to_gpu()
        if self.tiv is not None:
            self.tiv.to_gpu()
        self.f32 = self.beam.dt.dtype == torch.float32
        if self.beam.dt.dtype not in (torch.float64, torch.float32):
            raise TypeError("beam coordinates must be float64 or float32")
        self._torch = torch
        rp = self.rf
        n_turns = rp.n_turns
        self.n_turns = n_turns
        n = self.beam.n_macroparticles
        lib = _lib.load()
        h = lib.blmi_engine_create(n, rp.n_rf, n_turns, _SOLVER[tracker.solver])
        if not h:
            raise _lib.BlmiError(f"blmi_engine_create: {_lib.last_error()}")
        self._h = ct.c_void_p(h)
        _lib.call("blmi_engine_set_particles_f32" if self.f32 else "blmi_engine_set_particles",
                  self._h, _ptr(self.beam.dt), _ptr(self.beam.dE), n)
        self._dt, self._dE = self.beam.dt, self.beam.dE
        self._upload_programs()
        if self.profile is not None:
            p = self.profile
            if p.n_macroparticles.dtype != self.beam.dt.dtype:
                p.n_macroparticles = p.n_macroparticles.to(self.beam.dt.dtype).contiguous()
            _lib.call("blmi_engine_set_slices_f32" if self.f32 else "blmi_engine_set_slices",
                      self._h, p.n_slices, float(p.cut_left), float(p.cut_right),
                      _ptr(p.n_macroparticles))
        self._slicing = self._profile_slicing()
        if self.tiv is not None:
            self._setup_induced()
        if self.fb is not None:
            self._setup_feedback()
        if tracker.interpolation:
            if sort_every:
                raise NotImplementedError("interpolation=True with the bin-ordered working set")
            self._setup_interpolation()
        _lib.call("blmi_engine_set_launch", self._h, int(block), int(blocks_per_cu),
                  int(graph_chunk))
        # plain turns: up to turn_block turns a launch (1: a launch a turn)
        _lib.call("blmi_engine_set_turn_block", self._h, int(turn_block))
        _lib.call("blmi_engine_set_hist", self._h, int(groups), int(max_peel))
        _lib.call("blmi_engine_set_order", self._h, int(sort_every))
        if method is not None and self.tiv is not None:
            _lib.call("blmi_engine_set_method", self._h, _METHODS[method])
        self.monitor = monitor
        self._stats = bool(statistics) or monitor is not None
        if self._stats:
            self._setup_statistics()
        _lib.call("blmi_engine_set_turn", self._h, int(rp.counter[0]))
        self._engine_turn = int(rp.counter[0])
        # state that lives on the device between calls and that turns taken
        # outside this tracker would not update: the phase loop's, the
        # multi-turn wake memory's and the bin-ordered particle copies
        self._device_state = (self.fb is not None or bool(sort_every)
                              or getattr(self, "_mtw", None) is not None)
        self._primed = False

    @property
    def method(self) -> str | None:
        """Induced-voltage method in use: 'fft' (one-workgroup LDS FFT),
        'direct' (convolution over all CUs), 'fourstep' (multi-workgroup FFT)
        or None without intensity effects."""
        if self.tiv is None:
            return None
        m = ct.c_int()
        _lib.call("blmi_engine_get_method", self._h, ct.byref(m))
        return {v: k for k, v in _METHODS.items()}.get(m.value)

    def time_kernels(self, turns: int):
        """Track ``turns`` eager turns timing each kernel with HIP events on
        the engine's stream; returns (induced_ms, particle_kernel_ms) averages
        and advances the host state like ``track``."""
        a, b = ct.c_float(), ct.c_float()
        stream = ct.c_void_p(self._torch.cuda.current_stream().cuda_stream)
        self._check_slicing()
        self._sync_particles()
        self._sync_turn()
        if self.tiv is not None and not self._primed:
            _lib.call("blmi_engine_prime", self._h, stream)
            self._primed = True
        start = self.rf.counter[0]
        _lib.call("blmi_engine_time_kernels", self._h, int(turns), ct.byref(a), ct.byref(b), stream)
        self._advance(start, start + int(turns))
        return a.value, b.value

    # ------------------------------------------------------------------
    def _phase_program(self):
        if self.fb is not None:
            # the device loop accumulates the phase offsets turn by turn
            self._dphi_prog, self._phase_moves = None, False
            return np.array(self.rf.host_array("phi_rf"), dtype=np.float64, copy=True)
        phi, self._dphi_prog, self._phase_moves = phase_program(self.rf)
        return phi

    def _upload_programs(self):
        rp = self.rf
        nt = rp.n_turns + 1
        self._phi_prog = self._phase_program()
        rf = rf_program_array(rp, self._phi_prog)
        acc = np.zeros(nt, dtype=np.float64)
        if not rp.empty:
            # tracker.py:585: without voltage the whole kick is skipped; the
            # engine's kick of zero voltage and zero acceleration changes nothing
            acc[: len(self.tracker.acceleration_kick)] = self.tracker.acceleration_kick
        arrs = ring_program_arrays(rp)
        self._keep = (rf, acc, arrs)
        _lib.call("blmi_engine_set_programs", self._h, _dptr(rf), _dptr(acc), _dptr(arrs[0]),
                  float(rp.length_ratio), float(rp.alpha_order), _dptr(arrs[1]), _dptr(arrs[2]),
                  _dptr(arrs[3]), _dptr(arrs[4]), _dptr(arrs[5]), _dptr(arrs[6]),
                  _dptr(arrs[7]), _dptr(arrs[8]), float(rp.particle.charge))

    def _setup_induced(self):
        torch = self._torch
        self._Z, n_fft = total_impedance(self.tiv)
        mtw, coef, mode = wake_setup(self.tiv, self.n_turns)
        if (mtw is not None or coef is not None) and self.f32:
            raise NotImplementedError("multi-turn wakes and inductive sources need FP64 particles")
        scale = self.beam.particle.charge * e * self.beam.ratio
        p = self.profile
        iv = self.tiv.induced_voltage
        dtype = self.beam.dt.dtype
        if not (hasattr(iv, "is_cuda") and iv.is_cuda and iv.numel() == p.n_slices
                and iv.dtype == dtype and iv.is_contiguous()):
            iv = torch.zeros(p.n_slices, dtype=dtype, device="cuda")
            self.tiv.induced_voltage = iv
        self._iv = iv
        bc = p.bin_centers
        if not (hasattr(bc, "is_cuda") and bc.dtype == dtype and bc.is_contiguous()):
            bc = torch.as_tensor(np.asarray(_host_array(bc)), device="cuda").to(dtype).contiguous()
            self._bc = bc
        _lib.call("blmi_engine_set_induced_f32" if self.f32 else "blmi_engine_set_induced",
                  self._h, int(n_fft), _dptr(self._Z) if self._Z is not None else None,
                  float(scale), _ptr(iv), _ptr(bc))
        if mtw is None and coef is None:
            return
        n_mem = n_add = 0
        time_ptr = mem_ptr = None
        if mtw is not None:
            # the engine reads and publishes the source's own memory buffer
            mem = mtw.mtw_memory
            if not (hasattr(mem, "is_cuda") and mem.is_cuda and mem.dtype == torch.float64
                    and mem.is_contiguous()):
                mem = torch.as_tensor(np.asarray(_host_array(mem)), dtype=torch.float64,
                                      device="cuda").contiguous()
                mtw.mtw_memory = mem
            tm = torch.as_tensor(np.asarray(_host_array(mtw.time_mtw)), dtype=torch.float64,
                                 device="cuda").contiguous()
            self._mtw = (mtw, mem, tm)
            n_mem = int(mtw.n_mtw_memory)
            n_iv = int(mtw.n_induced_voltage)
            n_add = slice(n_iv - mtw.front_wake_buffer, None).indices(n_iv)[0]
            time_ptr, mem_ptr = _ptr(tm), _ptr(mem)
            mtw.induced_voltage = mem[:n_iv]
        self._coef = coef
        _lib.call("blmi_engine_set_wake", self._h, n_mem, n_add, time_ptr, mem_ptr,
                  _dptr(coef) if coef is not None else None, int(mode))

    def _setup_interpolation(self):
        """tracker.py:585-601: the kick of every turn is the linear
        interpolation of the RF voltage on the bin centres (plus the induced
        voltage, which is then only summed, not kicked with separately)."""
        torch = self._torch
        if self.profile is None:
            raise ValueError("interpolation=True needs the profile")
        if self.f32:
            raise NotImplementedError("interpolation=True needs FP64 particles")
        if self.tiv is not None and self.tracker.totalInducedVoltage is not self.tiv:
            raise ValueError("with interpolation=True the tracker's total_induced_voltage must "
                             "be the map's induced voltage")
        bc = self.profile.bin_centers
        if not (hasattr(bc, "is_cuda") and bc.dtype == torch.float64 and bc.is_contiguous()):
            bc = torch.as_tensor(np.asarray(_host_array(bc)), dtype=torch.float64,
                                 device="cuda").contiguous()
        self._interp_bc = bc
        _lib.call("blmi_engine_set_interpolation", self._h, _ptr(bc))

    def _setup_feedback(self):
        fb, rp = self.fb, self.rf
        if self.f32:
            raise NotImplementedError("the beam phase loop needs FP64 particles")
        machine, prog, on_time = fb.device_programs(self.n_turns)
        h = np.ascontiguousarray(np.asarray(rp.harmonic, dtype=np.float64).T)
        wd = np.ascontiguousarray(np.asarray(rp.omega_rf_d, dtype=np.float64).T)
        st = fb.device_state(rp.n_rf)
        bc = self.profile.bin_centers
        torch = self._torch
        if not (hasattr(bc, "is_cuda") and bc.dtype == torch.float64 and bc.is_contiguous()):
            bc = torch.as_tensor(np.asarray(_host_array(bc)), dtype=torch.float64,
                                 device="cuda").contiguous()
        self._fb_keep = (h, wd, prog, on_time, st, bc)
        _lib.call("blmi_engine_set_feedback", self._h, int(machine), int(fb.delay),
                  int(fb.first_bin()), float(fb.alpha), float(fb.reference), _ptr(bc), _dptr(h),
                  _dptr(wd), _dptr(prog), _dptr(on_time) if len(on_time) else None,
                  len(on_time), _dptr(st))

    def _setup_statistics(self):
        """Per-turn moments from the particle kernel, shifted by the current
        means (bm.beam_statistics) so that the one-pass variance keeps its
        precision."""
        from ..utils import bmath as bm

        b = self.beam
        if b.n_macroparticles_alive != b.n_macroparticles:
            raise NotImplementedError("fused statistics need every particle alive (the fused "
                                      "turn tracks no losses)")
        _, mt, _, me, _ = bm.beam_statistics(b.dt, b.dE)
        _lib.call("blmi_engine_set_statistics", self._h, 1, float(mt), float(me))

    def statistics(self, turn0: int, turns: int) -> np.ndarray:
        """Rows {n, mean_dt, sigma_dt, mean_dE, sigma_dE} of turns
        [turn0, turn0 + turns) as left by the fused turns (row t + 1 is the
        bunch after turn t; rows not tracked by this engine are zero)."""
        if not self._stats:
            raise RuntimeError("statistics were not enabled (FusedTracker(statistics=True))")
        out = np.zeros((int(turns), 5), dtype=np.float64)
        _lib.call("blmi_engine_get_statistics", self._h, int(turn0), int(turns), _dptr(out))
        return out

    def _fetch_feedback(self, start: int, end: int):
        """The loop's RF program rows start+1..end and its state, back to the
        host objects (tracker.py:551-565 leaves them in RFStation)."""
        rp, n_rf = self.rf, self.rf.n_rf
        rows = np.empty((end - start, 3, n_rf), dtype=np.float64)
        _lib.call("blmi_engine_get_programs", self._h, start + 1, end - start, _dptr(rows))
        for name, k in (("omega_rf", 1), ("phi_rf", 2)):
            host = rp.host_array(name)
            host[:, start + 1:end + 1] = rows[:, k, :].T
            if rp._device == "GPU":
                dev = getattr(rp, name)
                dev[:, start + 1:end + 1] = self._torch.from_numpy(
                    np.ascontiguousarray(host[:, start + 1:end + 1])).to(dev.device)
        st = np.empty(16 + n_rf, dtype=np.float64)
        _lib.call("blmi_engine_get_feedback", self._h, _dptr(st))
        self.fb.set_device_state(st)

    # ------------------------------------------------------------------
    def invalidate(self):
        """Call after changing beam.dt outside the tracker: the staged
        histogram of the current turn is rebuilt on the next track()."""
        self._primed = False

    def update_programs(self):
        """Re-read the RF station's per-turn programs (voltage, omega_rf,
        phi_rf; on the GPU the device arrays, which the objects' own turns
        read too) and the ring's, after the caller changed them; the next
        track() uses them from the current turn on.  With a phase loop the
        device loop owns the phase program: refused."""
        if self.fb is not None:
            raise RuntimeError("with a beam phase loop the phase program lives on the device")
        rp = self.rf
        # nothing still in flight may read the tables being replaced
        self._torch.cuda.synchronize()
        if rp._device == "GPU":
            for name in ("voltage", "omega_rf", "phi_rf"):
                rp._host[name][...] = getattr(rp, name).cpu().numpy()
        self._upload_programs()

    def _profile_slicing(self):
        p = self.profile
        return None if p is None else (int(p.n_slices), float(p.cut_left), float(p.cut_right))

    def _check_slicing(self):
        """The cuts and slice count are the engine's from construction (the
        impedance on the bins and the interpolation table depend on them)."""
        if self._profile_slicing() != self._slicing:
            raise RuntimeError(f"the profile's slicing changed from {self._slicing} to "
                               f"{self._profile_slicing()} since the fused tracker was built: "
                               "build a new FusedTracker for the new slicing")

    def _sync_particles(self):
        """beam.dt or beam.dE replaced by other arrays since the last call
        (an assignment, a copy): track those, with the staged histogram
        rebuilt.  Their precision is the tracker's, fixed at construction."""
        dt, dE = self.beam.dt, self.beam.dE
        if dt is self._dt and dE is self._dE:
            return
        if self._device_state:
            raise RuntimeError("beam.dt or beam.dE was replaced, but the fused tracker's phase "
                               "loop, multi-turn wake or particle order lives on the device")
        if not all(hasattr(a, "is_cuda") and a.is_cuda and a.is_contiguous() for a in (dt, dE)):
            raise TypeError("beam coordinates must be contiguous device tensors")
        if dt.numel() != self.beam.n_macroparticles or dE.numel() != self.beam.n_macroparticles:
            raise ValueError("the number of macroparticles is fixed at construction")
        want = self._torch.float32 if self.f32 else self._torch.float64
        if dt.dtype != want or dE.dtype != want:
            raise TypeError(f"beam coordinates must stay {want} for this tracker")
        _lib.call("blmi_engine_set_particles_f32" if self.f32 else "blmi_engine_set_particles",
                  self._h, _ptr(dt), _ptr(dE), self.beam.n_macroparticles)
        self._dt, self._dE = dt, dE
        self._primed = False

    def _sync_turn(self):
        """Turns tracked outside this tracker (the objects' own track()
        between two calls) moved the RF station's counter: continue from
        that turn, with the staged histogram rebuilt from the coordinates as
        they are.  Refused where the engine keeps state of its own that
        those turns did not update."""
        t = int(self.rf.counter[0])
        if t == self._engine_turn:
            return
        if self._device_state:
            raise RuntimeError(
                f"the RF counter moved from turn {self._engine_turn} to {t} outside the "
                "fused tracker, whose phase loop, multi-turn wake or particle order lives "
                "on the device: track every turn with one of them")
        _lib.call("blmi_engine_set_turn", self._h, t)
        self._engine_turn = t
        self._primed = False

    def track(self, n_turns: int = 1):
        """Advance ``n_turns`` turns (k passes of the map)."""
        n_turns = int(n_turns)
        if n_turns <= 0:
            return
        stream = ct.c_void_p(self._torch.cuda.current_stream().cuda_stream)
        self._check_slicing()
        self._sync_particles()
        self._sync_turn()
        if self.tiv is not None and not self._primed:
            _lib.call("blmi_engine_prime", self._h, stream)
            self._primed = True
        start = self.rf.counter[0]
        _lib.call("blmi_engine_track", self._h, n_turns, 1 if self.use_graph else 0, stream)
        self._advance(start, start + n_turns)

    def _advance(self, start: int, end: int):
        rp = self.rf
        if self.fb is not None and end > start:
            self._fetch_feedback(start, end)
        if self.monitor is not None and end > start:
            self.monitor.record_moments(self.statistics(start + 1, end - start))
        # host-side state the map leaves behind (tracker.py:556-615)
        if end > start and self._phase_moves:
            host_phi = rp.host_array("phi_rf")
            host_phi[:, start + 1:end + 1] = self._phi_prog[:, start + 1:end + 1]
            rp.dphi_rf[:] = self._dphi_prog[:, end]
            if rp._device == "GPU":
                import torch

                rp.phi_rf[:, start + 1:end + 1] = torch.from_numpy(
                    np.ascontiguousarray(self._phi_prog[:, start + 1:end + 1])).to(rp.phi_rf.device)
        rp.counter[0] = end
        self._engine_turn = end
        if self.tiv is not None and self.tiv.induced_voltage is not self._iv:
            # the objects' own induced-voltage step (turns tracked between
            # two calls) installs a new array; the engine writes the one it
            # was given
            self.tiv.induced_voltage = self._iv
        self.beam.beta = rp.beta[end]
        self.beam.gamma = rp.gamma[end]
        self.beam.energy = rp.energy[end]
        self.beam.momentum = rp.momentum[end]

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().blmi_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
