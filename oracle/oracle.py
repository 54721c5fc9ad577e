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
    sources = state.get("sources")
    memories = [s.get("memory") for s in sources] if sources else None
    for t in range(t0, t0 + n_turns):
        if state.get("intensity"):
            lib.slice_beam(dt, prof, state["cut_left"], state["cut_right"])
            if sources:
                volt_ind = induced_voltage_sources(state, prof, t, memories)
            else:
                volt_ind = induced_voltage_1turn(prof, state["Z"], state["n_fft"],
                                                 state["charge"], state["ratio"],
                                                 state["n_induced_voltage"], state["n_slices"])
            lib.linear_interp_kick(dt, dE, volt_ind, state["bin_centers"], state["charge"], 0.0)
        lib.kick(dt, dE, state["voltage"][:, t], state["omega_rf"][:, t], state["phi_rf"][:, t],
                 state["charge"], state["n_rf"], state["acc"][t])
        k = t + 1
        lib.drift(dt, dE, state["solver"], state["t_rev"][k], state["length_ratio"],
                  state["alpha_order"], state["eta_0"][k], state["eta_1"][k], state["eta_2"][k],
                  state["alpha_0"][k], state["alpha_1"][k], state["alpha_2"][k],
                  state["beta"][k], state["energy"][k])
        if not state.get("intensity") and prof is not None:
            lib.slice_beam(dt, prof, state["cut_left"], state["cut_right"])
    st.update(dt=dt, dE=dE, turn=t0 + n_turns)
    if prof is not None:
        st["profile"] = prof
    if volt_ind is not None:
        st["induced_voltage"] = volt_ind
    if sources:
        st["sources"] = [dict(s, memory=m) if m is not None else dict(s)
                         for s, m in zip(sources, memories)]
    return st
