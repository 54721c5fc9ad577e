"""CPU oracle for the BLonD tracking path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module.  The product (``blond_amd``) never does.

Three layers:

* ``Lib``: ctypes access to a BLonD-ABI shared library.  Two kinds load:
  ``oracle/liboracle.so`` (this directory's own C restatement,
  ``blond_oracle.c``) and ``oracle/_ref/libblond_ref_*.so`` (the reference's
  own C++ sources compiled by ``oracle/Makefile``).  Both export the
  reference's symbols (kick.cpp:16, drift.cpp:19, histogram.cpp:22,75,
  linear_interp_kick.cpp:21, kick.cpp:37, fast_resonator.cpp:19).
* numpy restatements of the slice-axis steps that the reference does in
  Python/numpy: the induced voltage of one turn
  (blond/impedances/impedance.py:457-489, 113-131), the multi-turn wake
  memory (impedance.py:405-453, 491-550, mode "time"), the inductive Z/n
  source (impedance.py:986-1006 with profile.py:905-937), the FFT sizing rule
  (next_regular.py:9-59, impedance.py:349-403, 846-874) and the turn loops of
  tracker.py:521-615, profile.py:700-719 and impedance.py:133-146.  numpy's
  interp, gradient, diff and FFTs are the functions the reference itself
  calls there (bm.interp / bm.gradient / bm.diff / bm.rfft / bm.irfft,
  bmath_backends.py:373-449).
* ``CpuAdjoint``: a float64 reverse-mode restatement of the differentiable
  turn, used as the CPU reference gradient.

Physical constants are CODATA 2022 literals (scipy 1.15.3), as the reference
takes them from scipy.constants.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

C_LIGHT = 299792458.0
E_CHARGE = 1.602176634e-19
M_P_EV = 1.67262192595e-27 * C_LIGHT ** 2 / E_CHARGE

_dp = ct.POINTER(ct.c_double)
_fp = ct.POINTER(ct.c_float)


def _ptr(a: np.ndarray, kind=_dp):
    return a.ctypes.data_as(kind)


class Lib:
    """A BLonD-ABI library (``real_t = double`` unless ``single``)."""

    def __init__(self, path: str, single: bool = False):
        self.path = path
        self.lib = ct.CDLL(path)
        self.single = single
        self.real = np.float32 if single else np.float64
        self.c_real = ct.c_float if single else ct.c_double
        self.ptr = _fp if single else _dp

    # kick.cpp:16-35
    def kick(self, dt, dE, voltage, omega_rf, phi_rf, charge, n_rf, acc):
        r = self.real
        v = np.ascontiguousarray(voltage, dtype=r)
        w = np.ascontiguousarray(omega_rf, dtype=r)
        p = np.ascontiguousarray(phi_rf, dtype=r)
        self.lib.kick(_ptr(dt, self.ptr), _ptr(dE, self.ptr), ct.c_int(int(n_rf)),
                      self.c_real(charge), _ptr(v, self.ptr), _ptr(w, self.ptr),
                      _ptr(p, self.ptr), ct.c_int(len(dt)), self.c_real(acc))

    # drift.cpp:19-78
    def drift(self, dt, dE, solver, t_rev, length_ratio, alpha_order, eta_0,
              eta_1, eta_2, alpha_0, alpha_1, alpha_2, beta, energy):
        s = {"simple": 0, "legacy": 1, "exact": 2}[solver] if isinstance(solver, str) else int(solver)
        cr = self.c_real
        self.lib.drift(_ptr(dt, self.ptr), _ptr(dE, self.ptr), ct.c_int(s), cr(t_rev),
                       cr(length_ratio), cr(alpha_order), cr(eta_0), cr(eta_1),
                       cr(eta_2), cr(alpha_0), cr(alpha_1), cr(alpha_2), cr(beta),
                       cr(energy), ct.c_int(len(dt)))

    # histogram.cpp:22-73
    def slice_beam(self, dt, profile, cut_left, cut_right):
        self.lib.histogram(_ptr(dt, self.ptr), _ptr(profile, self.ptr),
                           self.c_real(cut_left), self.c_real(cut_right),
                           ct.c_int(len(profile)), ct.c_int(len(dt)))

    # histogram.cpp:75-132
    def slice_smooth(self, dt, profile, cut_left, cut_right):
        self.lib.smooth_histogram(_ptr(dt, self.ptr), _ptr(profile, self.ptr),
                                  self.c_real(cut_left), self.c_real(cut_right),
                                  ct.c_int(len(profile)), ct.c_int(len(dt)))

    # linear_interp_kick.cpp:21-71
    def linear_interp_kick(self, dt, dE, voltage, bin_centers, charge, acc):
        v = np.ascontiguousarray(voltage, dtype=self.real)
        bc = np.ascontiguousarray(bin_centers, dtype=self.real)
        self.lib.linear_interp_kick(_ptr(dt, self.ptr), _ptr(dE, self.ptr),
                                    _ptr(v, self.ptr), _ptr(bc, self.ptr),
                                    self.c_real(charge), ct.c_int(len(bc)),
                                    ct.c_int(len(dt)), self.c_real(acc))

    # kick.cpp:37-51
    def rf_volt_comp(self, voltage, omega_rf, phi_rf, bin_centers):
        r = self.real
        v = np.ascontiguousarray(voltage, dtype=r)
        w = np.ascontiguousarray(omega_rf, dtype=r)
        p = np.ascontiguousarray(phi_rf, dtype=r)
        bc = np.ascontiguousarray(bin_centers, dtype=r)
        out = np.zeros(len(bc), dtype=r)
        self.lib.rf_volt_comp(_ptr(v, self.ptr), _ptr(w, self.ptr), _ptr(p, self.ptr),
                              _ptr(bc, self.ptr), ct.c_int(len(v)), ct.c_int(len(bc)),
                              _ptr(out, self.ptr))
        return out

    # sparse_histogram.cpp:22-87 (double only).  The reference's own symbol
    # takes no pattern length; the oracle's sparse_histogram_n bounds the
    # bucket lookup by it.
    def sparse_histogram(self, dt, profile, cut_left, cut_right, bunch_indexes, n_slices_bucket):
        cl = np.ascontiguousarray(cut_left, dtype=np.float64)
        cr = np.ascontiguousarray(cut_right, dtype=np.float64)
        bi = np.ascontiguousarray(bunch_indexes, dtype=np.float64)
        if hasattr(self.lib, "sparse_histogram_n"):
            self.lib.sparse_histogram_n(_ptr(dt), _ptr(profile), _ptr(cl), _ptr(cr), _ptr(bi),
                                        ct.c_int(int(n_slices_bucket)), ct.c_int(len(cl)),
                                        ct.c_int(len(dt)), ct.c_int(len(bi)))
        else:
            self.lib.sparse_histogram(_ptr(dt), _ptr(profile), _ptr(cl), _ptr(cr), _ptr(bi),
                                      ct.c_int(int(n_slices_bucket)), ct.c_int(len(cl)),
                                      ct.c_int(len(dt)))

    # beam_phase.cpp:6-39 and 42-65 (double only)
    def beam_phase(self, bin_centers, profile, alpha, omega_rf, phi_rf, bin_size):
        bc = np.ascontiguousarray(bin_centers, dtype=np.float64)
        p = np.ascontiguousarray(profile, dtype=np.float64)
        f = self.lib.beam_phase
        f.restype = ct.c_double
        return f(_ptr(bc), _ptr(p), ct.c_double(alpha), ct.c_double(omega_rf),
                 ct.c_double(phi_rf), ct.c_double(bin_size), ct.c_int(len(p)))

    def beam_phase_fast(self, bin_centers, profile, omega_rf, phi_rf, bin_size):
        bc = np.ascontiguousarray(bin_centers, dtype=np.float64)
        p = np.ascontiguousarray(profile, dtype=np.float64)
        f = self.lib.beam_phase_fast
        f.restype = ct.c_double
        return f(_ptr(bc), _ptr(p), ct.c_double(omega_rf), ct.c_double(phi_rf),
                 ct.c_double(bin_size), ct.c_int(len(p)))

    # fast_resonator.cpp:19-73 (double only)
    def fast_resonator(self, R_S, Q, freq, f_r):
        R = np.ascontiguousarray(R_S, dtype=np.float64)
        Qa = np.ascontiguousarray(Q, dtype=np.float64)
        f = np.ascontiguousarray(freq, dtype=np.float64)
        fr = np.ascontiguousarray(f_r, dtype=np.float64)
        re = np.zeros(len(f))
        im = np.zeros(len(f))
        self.lib.fast_resonator_real_imag(_ptr(re), _ptr(im), _ptr(f), _ptr(R), _ptr(Qa),
                                          _ptr(fr), ct.c_int(len(R)), ct.c_int(len(f)))
        return re + 1j * im


def oracle_lib() -> Lib:
    """The repository's own C restatement (always buildable, no reference needed)."""
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
    return Lib(path)


def ref_lib(kind: str = "strict", precision: str = "double") -> Lib | None:
    """BLonD's own kernels compiled from the reference sources, or None when
    ``oracle/_ref`` was not built (e.g. no reference tree at build time)."""
    path = os.path.join(HERE, "_ref", f"libblond_ref_{kind}_{precision}.so")
    if not os.path.exists(path):
        return None
    return Lib(path, single=(precision == "single"))


# --------------------------------------------------------------------------
# Slice-axis numpy restatements
# --------------------------------------------------------------------------
def next_regular(target: int) -> int:
    """next_regular.py:9-59: smallest 5-smooth number >= target, returned as
    2*regular(ceil(target/2)) (so the result carries a factor 2), except
    that targets <= 6 and powers of two are returned unchanged."""
    if target <= 6:
        return target
    if not (target & (target - 1)):
        return target
    half = -(-target // 2)
    best = None
    p5 = 1
    while True:
        p35 = p5
        while True:
            q = -(-half // p35)
            p2 = 1 << (q - 1).bit_length()
            cand = p2 * p35
            if best is None or cand < best:
                best = cand
            if p35 >= half:
                break
            p35 *= 3
        if p5 >= half:
            break
        p5 *= 5
    return 2 * best


def induced_voltage_1turn(profile, total_impedance, n_fft, charge, ratio,
                          n_induced_voltage, n_slices):
    """impedance.py:457-489 then 113-131: V = -(q e ratio) irfft(Z rfft(p, n_fft))
    cut to n_induced_voltage, then to n_slices."""
    spec = np.fft.rfft(profile, n_fft)
    y = np.fft.irfft(total_impedance * spec)
    v = -(charge * E_CHARGE * ratio * y)
    return np.ascontiguousarray(v[:n_induced_voltage][:n_slices])


def mtw_time_axis(wake_length: float, n_mtw_memory: int) -> np.ndarray:
    """impedance.py:440-447: linspace(0, wake_length, n, endpoint=False)."""
    return np.linspace(0, wake_length, n_mtw_memory, endpoint=False, dtype=np.float64)


def mtw_update(memory, time_mtw, t_rev, voltage, n_induced_voltage, front_wake_buffer):
    """induced_voltage_mtw (impedance.py:491-516) after the one-turn voltage:
    shift_trev_time (533-550), the front-wake tail of this turn's voltage set
    to 0, the voltage added to the first n_induced_voltage points.  Returns
    (new memory, induced voltage = memory[:n_induced_voltage])."""
    v = np.array(voltage, dtype=np.float64, copy=True)
    v[n_induced_voltage - front_wake_buffer:] = 0
    mem = np.interp(time_mtw + t_rev, time_mtw, memory, left=0, right=0)
    mem[:n_induced_voltage] += v
    return mem, mem[:n_induced_voltage]


def profile_derivative(profile, bin_centers, mode: str):
    """Profile.beam_profile_derivative (profile.py:905-937), modes
    'gradient' and 'diff'."""
    dist_centers = bin_centers[1] - bin_centers[0]
    if mode == "gradient":
        return np.gradient(profile, dist_centers)
    if mode == "diff":
        derivative = np.diff(profile) / dist_centers
        diff_centers = bin_centers[0:-1] + dist_centers / 2
        return np.interp(bin_centers, diff_centers, derivative)
    raise RuntimeError("Option for derivative is not recognized.")


def inductive_voltage(profile, bin_centers, Z_over_n, t_rev, turn, charge, ratio, bin_size,
                      mode):
    """InductiveImpedance.induced_voltage_1turn (impedance.py:986-1006)."""
    return -(charge * E_CHARGE / (2 * np.pi) * ratio * Z_over_n[turn] * t_rev[turn] / bin_size
             * profile_derivative(profile, bin_centers, mode))


def induced_voltage_sources(state: dict, profile, turn: int, memories: list):
    """TotalInducedVoltage.induced_voltage_sum (impedance.py:113-131) over
    state["sources"] in list order; multi-turn memories in `memories` are
    replaced by their updated arrays."""
    ns = int(state["n_slices"])
    total = 0
    for k, s in enumerate(state["sources"]):
        if s["kind"] == "inductive":
            v = inductive_voltage(profile, state["bin_centers"], s["Z_over_n"], state["t_rev"],
                                  turn, state["charge"], state["ratio"], state["bin_size"],
                                  s["deriv_mode"])
            v = v[:ns]
        else:
            n_iv = s["n_induced_voltage"]
            spec = np.fft.rfft(profile, s["n_fft"])
            y = np.fft.irfft(s["Z"] * spec)
            v = -(state["charge"] * E_CHARGE * state["ratio"] * y)[:n_iv]
            if s["multi_turn_wake"]:
                tm = mtw_time_axis(s["wake_length"], s["n_mtw_memory"])
                memories[k], v = mtw_update(memories[k], tm, state["t_rev"][turn], v, n_iv,
                                            s["front_wake_buffer"])
        total = total + v[:ns]
    return np.ascontiguousarray(total, dtype=np.float64)


def phase_loop_turn(fb: dict, t: int, omega, phi, profile, bin_centers, bin_size, lib, dE=None):
    """The machine function of BeamFeedback.track at turn t
    (beam_feedback.py:275-367, 414-622) on the loop state ``fb`` (updated in
    place); returns domega_rf.  The beam phase is beam_phase.cpp through
    ``lib``; omega / phi are the [n_rf, n_turns+1] programs as modified so far."""
    m = fb["machine"]
    b0 = fb["first_bin"]
    coeff = lib.beam_phase(bin_centers[b0:], profile[b0:], fb["alpha"], omega[0, t], phi[0, t],
                           bin_size)
    fb["phi_beam"] = np.arctan(coeff) + np.pi
    fb["dphi"] = fb["phi_beam"] - fb["phi_s"][t]
    dphi = fb["dphi"]
    if m == "LHC":
        x = fb["dphi_rf"][0] + fb["reference"]
        fb["domega_rf"] = -fb["gain"] * dphi - fb["gain2"] * (fb["lhc_y"] + fb["lhc_a"][t] * x)
        fb["lhc_y"] = (1 - fb["lhc_t"][t]) * fb["lhc_y"] + (1 - fb["lhc_a"][t]) * fb["lhc_t"][t] * x
    elif m == "LHC_F":
        fb["domega_rf"] = -fb["gain"] * dphi - fb["gain2"] * (
            omega[0, t] - fb["omega_rf_d"][0, t] + fb["reference"])
    elif m == "SPS_RL":
        avg = np.mean(dE[:: fb["sample_dE"]])
        drho = (fb["alpha_0"][t] * fb["ring_radius"] * avg
                / (fb["beta"][t] ** 2.0 * fb["energy"][t]))
        domega_dphi = -fb["gain"] * dphi
        domega_dR = (-np.sign(fb["eta_0"][t]) * fb["gain2"] * (fb["reference"] - drho)
                     / fb["ring_radius"])
        fb["domega_rf"] = domega_dphi + domega_dR
    elif m == "PSB":
        fb["dphi_sum"] += dphi
        pc = fb["PL_counter"]
        on = fb["on_time"]
        if t == on[pc] and t >= fb["delay"]:
            fb["dphi_av"] = fb["dphi_sum"] / (on[pc] - on[pc - 1])
            fb["domega_PL"] = 0.99803799 * fb["domega_PL"] + fb["gain"][t] * (
                0.99901903 * fb["dphi_av"] - 0.99901003 * fb["dphi_av_prev"])
            fb["dphi_av_prev"] = fb["dphi_av"]
            fb["dphi_sum"] = 0.0
            wd = fb["omega_rf_d"][0, t]
            fb["dR_over_R"] = (omega[0, t] - wd) / (
                wd * (1.0 / (fb["alpha_0"][t] * fb["gamma"][t] ** 2) - 1.0))
            fb["domega_RL"] = (fb["domega_RL"] + fb["gain2"][0][t]
                               * (fb["dR_over_R"] - fb["dR_over_R_prev"])
                               + fb["gain2"][1][t] * fb["dR_over_R"])
            fb["dR_over_R_prev"] = fb["dR_over_R"]
            fb["PL_counter"] = pc + 1
        fb["domega_rf"] = -fb["domega_PL"] - fb["domega_RL"]
    else:
        raise ValueError(f"machine {m!r}")
    return fb["domega_rf"]


def phase_update(fb: dict, t: int, omega, phi, domega):
    """beam_feedback.py:246-251 (when the loop ran: domega is not None) then
    tracker.py:551-565: next turn's frequency, accumulated phase offset and
    next turn's phase."""
    h = fb["harmonic"]
    if domega is not None:
        omega[:, t + 1] += domega * h[:, t + 1] / h[0, t + 1]
    wd = fb["omega_rf_d"]
    fb["dphi_rf"] += 2.0 * np.pi * h[:, t + 1] * (omega[:, t + 1] - wd[:, t + 1]) / wd[:, t + 1]
    phi[:, t + 1] += fb["dphi_rf"]


def resonator_impedance(freq, R_S, f_r, Q):
    """Closed form Z(f) = sum_k R_k / (1 + j Q_k (f/f_rk - f_rk/f)), Z(0) = 0."""
    z = np.zeros(len(freq), dtype=np.complex128)
    f = np.asarray(freq[1:], dtype=np.float64)
    for R, fr, q in zip(np.atleast_1d(R_S), np.atleast_1d(f_r), np.atleast_1d(Q)):
        z[1:] += R / (1 + 1j * q * (f / fr - fr / f))
    return z


# --------------------------------------------------------------------------
# Turn loop (tracker.py:521-615, profile.py:700-719, impedance.py:113-146)
# --------------------------------------------------------------------------
def track(state: dict, n_turns: int, lib: Lib | None = None) -> dict:
    """Run ``n_turns`` passes of the reference's map on a snapshot (numpy
    copies of the coordinates and per-turn programs) with a BLonD-ABI
    library: [profile, induced voltage, tracker] with intensity effects,
    [tracker, profile] without.  Returns a new state dict."""
    lib = lib or oracle_lib()
    st = dict(state)
    r = lib.real  # float32 with the reference's single-precision build
    dt = np.array(state["dt"], dtype=r, copy=True)
    dE = np.array(state["dE"], dtype=r, copy=True)
    prof = np.array(state["profile"], dtype=r, copy=True) if "profile" in state else None
    volt_ind = np.array(state["induced_voltage"], copy=True) if state.get("intensity") else None
    t0 = int(state["turn"])
    fb = state.get("feedback")
    if fb is not None:
        # the loop mutates the RF programs and its own state
        fb = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
              for k, v in fb.items()}
        omega_p = np.array(state["omega_rf"], dtype=np.float64, copy=True)
        phi_p = np.array(state["phi_rf"], dtype=np.float64, copy=True)
    else:
        omega_p, phi_p = state["omega_rf"], state["phi_rf"]
    sources = state.get("sources")
    memories = [s.get("memory") for s in sources] if sources else None
    # profile.py:700-719: slice_smooth for a smooth profile, slice_beam otherwise
    slicer = lib.slice_smooth if state.get("smooth") else lib.slice_beam
    interp = bool(state.get("interpolation"))
    for t in range(t0, t0 + n_turns):
        if state.get("intensity"):
            slicer(dt, prof, state["cut_left"], state["cut_right"])
            if sources:
                volt_ind = induced_voltage_sources(state, prof, t, memories)
            else:
                volt_ind = induced_voltage_1turn(prof, state["Z"], state["n_fft"],
                                                 state["charge"], state["ratio"],
                                                 state["n_induced_voltage"], state["n_slices"])
            if not interp:
                # TotalInducedVoltage.track (impedance.py:133-146); with an
                # interpolating tracker the voltage is only summed
                # (induced_voltage_sum) and the tracker kicks with it
                lib.linear_interp_kick(dt, dE, volt_ind, state["bin_centers"], state["charge"],
                                       0.0)
        if fb is not None:
            # RingAndRFTracker.track (tracker.py:551-565) before the kick
            domega = None
            if t >= fb["delay"]:
                domega = phase_loop_turn(fb, t, omega_p, phi_p, prof, state["bin_centers"],
                                         fb["bin_size"], fb.get("lib") or lib, dE)
            phase_update(fb, t, omega_p, phi_p, domega)
        if not state.get("empty", False) and interp:
            # tracker.py:585-601: the RF voltage on the bin centres (plus the
            # induced voltage) interpolated, the acceleration kick inside
            # the table (particles outside it get no kick at all)
            total = lib.rf_volt_comp(state["voltage"][:, t], omega_p[:, t], phi_p[:, t],
                                     state["bin_centers"])
            if state.get("intensity"):
                total = total + volt_ind
            lib.linear_interp_kick(dt, dE, total, state["bin_centers"], state["charge"],
                                   state["acc"][t])
        elif not state.get("empty", False):
            # tracker.py:585: a station without voltage kicks nothing, the
            # acceleration kick included
            lib.kick(dt, dE, state["voltage"][:, t], omega_p[:, t], phi_p[:, t],
                     state["charge"], state["n_rf"], state["acc"][t])
        k = t + 1
        lib.drift(dt, dE, state["solver"], state["t_rev"][k], state["length_ratio"],
                  state["alpha_order"], state["eta_0"][k], state["eta_1"][k], state["eta_2"][k],
                  state["alpha_0"][k], state["alpha_1"][k], state["alpha_2"][k],
                  state["beta"][k], state["energy"][k])
        if not state.get("intensity") and prof is not None:
            slicer(dt, prof, state["cut_left"], state["cut_right"])
    st.update(dt=dt, dE=dE, turn=t0 + n_turns)
    if prof is not None:
        st["profile"] = prof
    if volt_ind is not None:
        st["induced_voltage"] = volt_ind
    if sources:
        st["sources"] = [dict(s, memory=m) if m is not None else dict(s)
                         for s, m in zip(sources, memories)]
    if fb is not None:
        st.update(feedback=fb, omega_rf=omega_p, phi_rf=phi_p)
    return st
