"""bmath-compatible entry points backed by the MI355X library.

Same names, argument order and meaning as BLonD's ``bm`` functions bound by
the C++ backend (blond/utils/butils_wrap_cpp.py:835-1000, 1133-1171) and the
numpy FFTs the backend keeps (blond/utils/bmath_backends.py:447-449):

    kick(dt, dE, voltage, omega_rf, phi_rf, charge, n_rf, acceleration_kick)
    drift(dt, dE, solver, t_rev, length_ratio, alpha_order, eta_0, eta_1, eta_2,
          alpha_0, alpha_1, alpha_2, beta, energy)
    slice_beam(dt, profile, cut_left, cut_right)
    slice_smooth(dt, profile, cut_left, cut_right)
    linear_interp_kick(dt, dE, voltage, bin_centers, charge, acceleration_kick)
    rf_volt_comp(voltages, omega_rf, phi_rf, bin_centers) -> array
    fast_resonator(R_S, Q, frequency_array, frequency_R, impedance=None) -> array
    rfft(a, n=0), irfft(a, n=0), rfftfreq(n, d=1.0)
    interp(x, xp, fp, left=None, right=None), gradient(f, h), diff(a)
    beam_phase(bin_centers, profile, alpha, omega_rf, phi_rf, bin_size)
    beam_phase_fast(bin_centers, profile, omega_rf, phi_rf, bin_size)

Arrays are torch tensors resident on the GPU (the MI355X counterpart of the
reference GPU backend's cupy arrays, bmath_backends.py:717-776).  numpy
arrays are accepted too: they are staged to the device, the HIP kernel runs,
and in/out arrays are written back in place -- the compute is always the HIP
kernel.  As in the reference, element types must match the active precision
(butils_wrap_cpp.py:873-874): ``AssertionError`` otherwise.
"""
from __future__ import annotations

import ctypes as ct
from typing import Any

import numpy as np

from .. import _lib
from ._precision import precision

_SOLVERS = {"simple": 0, "legacy": 1, "exact": 2}


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _lib.LibraryNotAvailable(
            "blond_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


def _stream(torch) -> ct.c_void_p:
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


class _Staged:
    """Device view of one argument; writes numpy in/out arrays back on exit."""

    def __init__(self, torch, arr: Any, writeback: bool, dtype=None, name: str = "array"):
        self.torch = torch
        self.src = arr
        self.writeback = writeback
        dt = dtype if dtype is not None else precision.torch_dtype
        if isinstance(arr, torch.Tensor):
            if not arr.is_cuda:
                raise TypeError(f"{name}: torch tensor must be on the GPU (use to_gpu())")
            assert arr.dtype == dt, f"{name}: dtype {arr.dtype} != {dt}"
            if not arr.is_contiguous():
                if writeback:
                    raise ValueError(f"{name}: in/out tensor must be contiguous")
                arr = arr.contiguous()
            self.dev = arr
            self.writeback = False
        else:
            a = np.asarray(arr)
            if writeback:
                assert a.dtype == np.dtype(precision.real_t), f"{name}: dtype {a.dtype} != {precision.real_t}"
            host = np.ascontiguousarray(a, dtype=precision.real_t if dtype is None else _np_of(dt))
            self.dev = torch.from_numpy(host).to("cuda")

    @property
    def ptr(self) -> ct.c_void_p:
        return ct.c_void_p(self.dev.data_ptr())

    def finish(self):
        if self.writeback:
            self.torch.cuda.current_stream().synchronize()
            np.copyto(self.src, self.dev.cpu().numpy(), casting="no")


def _np_of(tdtype):
    import torch

    return {torch.float64: np.float64, torch.float32: np.float32}[tdtype]


def _suffix() -> str:
    return "" if precision.num == 2 else "_f32"


def _real(x) -> Any:
    return precision.c_real_t(float(x))


def _len(a) -> int:
    return int(a.shape[0]) if hasattr(a, "shape") else len(a)


# ---------------------------------------------------------------------------
# particle kernels
# ---------------------------------------------------------------------------
def kick(dt, dE, voltage, omega_rf, phi_rf, charge, n_rf, acceleration_kick):
    """kick.cpp:16-35 via butils_wrap_cpp.py:863-898."""
    torch = _torch()
    a_dt = _Staged(torch, dt, False, name="dt")
    a_dE = _Staged(torch, dE, True, name="dE")
    v = _Staged(torch, voltage, False, name="voltage")
    w = _Staged(torch, omega_rf, False, name="omega_rf")
    p = _Staged(torch, phi_rf, False, name="phi_rf")
    _lib.call("blmi_kick" + _suffix(), a_dt.ptr, a_dE.ptr, int(n_rf), _real(charge), v.ptr,
              w.ptr, p.ptr, _len(dt), _real(acceleration_kick), _stream(torch))
    a_dE.finish()


def drift(dt, dE, solver, t_rev, length_ratio, alpha_order, eta_0, eta_1, eta_2, alpha_0,
          alpha_1, alpha_2, beta, energy):
    """drift.cpp:19-78 via butils_wrap_cpp.py:901-943."""
    torch = _torch()
    s = _SOLVERS[solver] if isinstance(solver, str) else int(solver)
    a_dt = _Staged(torch, dt, True, name="dt")
    a_dE = _Staged(torch, dE, False, name="dE")
    r = _real
    _lib.call("blmi_drift" + _suffix(), a_dt.ptr, a_dE.ptr, s, r(t_rev), r(length_ratio),
              r(alpha_order), r(eta_0), r(eta_1), r(eta_2), r(alpha_0), r(alpha_1), r(alpha_2),
              r(beta), r(energy), _len(dt), _stream(torch))
    a_dt.finish()


def slice_beam(dt, profile, cut_left, cut_right):
    """histogram.cpp:22-73 via butils_wrap_cpp.py:971-984."""
    torch = _torch()
    a_dt = _Staged(torch, dt, False, name="dt")
    a_p = _Staged(torch, profile, True, name="profile")
    _lib.call("blmi_histogram" + _suffix(), a_dt.ptr, a_p.ptr, _real(cut_left), _real(cut_right),
              _len(profile), _len(dt), _stream(torch))
    a_p.finish()


def slice_smooth(dt, profile, cut_left, cut_right):
    """histogram.cpp:75-132 via butils_wrap_cpp.py:987-1000."""
    torch = _torch()
    a_dt = _Staged(torch, dt, False, name="dt")
    a_p = _Staged(torch, profile, True, name="profile")
    _lib.call("blmi_smooth_histogram" + _suffix(), a_dt.ptr, a_p.ptr, _real(cut_left),
              _real(cut_right), _len(profile), _len(dt), _stream(torch))
    a_p.finish()


def sparse_histogram(dt, profile, cut_left, cut_right, bunch_indexes, n_slices_bucket):
    """sparse_histogram.cpp:22-87 via butils_wrap_cpp.py:1003-1023: profile
    [n_filled, n_slices_bucket] overwritten with the per-bucket counts."""
    torch = _torch()
    a_dt = _Staged(torch, dt, False, name="dt")
    shape = tuple(profile.shape)
    if isinstance(profile, torch.Tensor):
        if not profile.is_contiguous():
            raise ValueError("profile: in/out tensor must be contiguous")
        flat = profile.view(-1)
    else:
        flat = np.asarray(profile).reshape(-1)
        if not np.shares_memory(flat, profile):
            raise ValueError("profile: in/out array must be C-contiguous")
    a_p = _Staged(torch, flat, True, name="profile")
    cl = _Staged(torch, np.asarray(cut_left) if not isinstance(cut_left, torch.Tensor) else cut_left,
                 False, name="cut_left")
    cr = _Staged(torch, np.asarray(cut_right) if not isinstance(cut_right, torch.Tensor) else cut_right,
                 False, name="cut_right")
    bi = _Staged(torch, np.asarray(bunch_indexes, dtype=precision.real_t)
                 if not isinstance(bunch_indexes, torch.Tensor) else bunch_indexes,
                 False, name="bunch_indexes")
    n_filled = shape[0]
    _lib.call("blmi_sparse_histogram" + _suffix(), a_dt.ptr, a_p.ptr, cl.ptr, cr.ptr, bi.ptr,
              int(n_slices_bucket), int(n_filled), _len(bi.dev), _len(dt), _stream(torch))
    a_p.finish()


def linear_interp_kick(dt, dE, voltage, bin_centers, charge, acceleration_kick):
    """linear_interp_kick.cpp:21-71 via butils_wrap_cpp.py:946-968."""
    torch = _torch()
    a_dt = _Staged(torch, dt, False, name="dt")
    a_dE = _Staged(torch, dE, True, name="dE")
    v = _Staged(torch, voltage, False, name="voltage")
    bc = _Staged(torch, bin_centers, False, name="bin_centers")
    _lib.call("blmi_linear_interp_kick" + _suffix(), a_dt.ptr, a_dE.ptr, v.ptr, bc.ptr,
              _real(charge), _len(bin_centers), _len(dt), _real(acceleration_kick),
              _stream(torch))
    a_dE.finish()


def rf_volt_comp(voltages, omega_rf, phi_rf, bin_centers):
    """kick.cpp:37-51 via butils_wrap_cpp.py:835-860; returns a new array of
    the same kind (numpy or device tensor) as ``bin_centers``."""
    torch = _torch()
    v = _Staged(torch, voltages, False, name="voltages")
    w = _Staged(torch, omega_rf, False, name="omega_rf")
    p = _Staged(torch, phi_rf, False, name="phi_rf")
    bc = _Staged(torch, bin_centers, False, name="bin_centers")
    out = torch.zeros(_len(bin_centers), dtype=precision.torch_dtype, device="cuda")
    _lib.call("blmi_rf_volt_comp" + _suffix(), v.ptr, w.ptr, p.ptr, bc.ptr, _len(voltages),
              _len(bin_centers), ct.c_void_p(out.data_ptr()), _stream(torch))
    if isinstance(bin_centers, torch.Tensor):
        return out
    return out.cpu().numpy()


def fast_resonator(R_S, Q, frequency_array, frequency_R, impedance=None):
    """fast_resonator.cpp:19-73 via butils_wrap_cpp.py:1133-1171 (float64).
    Returns complex128 of the same kind as ``frequency_array``."""
    torch = _torch()
    f64 = torch.float64
    R = _Staged(torch, np.asarray(R_S, dtype=np.float64).ravel(), False, f64, "R_S")
    q = _Staged(torch, np.asarray(Q, dtype=np.float64).ravel(), False, f64, "Q")
    fr = _Staged(torch, np.asarray(frequency_R, dtype=np.float64).ravel(), False, f64, "frequency_R")
    if isinstance(frequency_array, torch.Tensor):
        f = _Staged(torch, frequency_array.to(f64), False, f64, "frequency_array")
    else:
        f = _Staged(torch, np.asarray(frequency_array, dtype=np.float64), False, f64,
                    "frequency_array")
    n_f = _len(frequency_array)
    re = torch.zeros(n_f, dtype=f64, device="cuda")
    im = torch.zeros(n_f, dtype=f64, device="cuda")
    _lib.call("blmi_fast_resonator", ct.c_void_p(re.data_ptr()), ct.c_void_p(im.data_ptr()), f.ptr,
              R.ptr, q.ptr, fr.ptr, _len(R.dev), n_f, _stream(torch))
    z = torch.complex(re, im)
    if isinstance(frequency_array, torch.Tensor):
        return z
    out = z.cpu().numpy()
    if impedance is not None and getattr(impedance, "dtype", None) == np.complex128 \
            and len(impedance) == n_f:
        impedance.real = out.real
        impedance.imag = out.imag
        return impedance
    return out


# ---------------------------------------------------------------------------
# FFTs (numpy semantics; bmath_backends.py:447-449)
# ---------------------------------------------------------------------------
def rfft(a, n: int = 0):
    torch = _torch()
    n = int(n) if n else _len(a)
    src = _Staged(torch, a, False, torch.float64, "a")
    out = torch.empty(n // 2 + 1, dtype=torch.complex128, device="cuda")
    _lib.call("blmi_rfft", src.ptr, _len(src.dev), ct.c_void_p(out.data_ptr()), n, _stream(torch))
    return out if isinstance(a, torch.Tensor) else out.cpu().numpy()


def irfft(a, n: int = 0):
    torch = _torch()
    m = _len(a)
    n = int(n) if n else 2 * (m - 1)
    if isinstance(a, torch.Tensor):
        src = a.to(torch.complex128).contiguous()
    else:
        src = torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to("cuda")
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    _lib.call("blmi_irfft", ct.c_void_p(src.data_ptr()), m, ct.c_void_p(out.data_ptr()), n,
              _stream(torch))
    return out if isinstance(a, torch.Tensor) else out.cpu().numpy()


def rfftfreq(n: int, d: float = 1.0):
    """numpy.fft.rfftfreq (host arithmetic: a setup-time frequency axis)."""
    return np.fft.rfftfreq(int(n), d)


# ---------------------------------------------------------------------------
# numpy helpers of the wake and profile-derivative steps (float64 on the device)
# ---------------------------------------------------------------------------
def _f64(torch, a, name):
    """Device float64 contiguous copy/view of a numpy array or tensor."""
    if isinstance(a, torch.Tensor):
        if not a.is_cuda:
            raise TypeError(f"{name}: torch tensor must be on the GPU (use to_gpu())")
        return a.to(torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def _like(torch, out, ref):
    return out if isinstance(ref, torch.Tensor) else out.cpu().numpy()


def interp(x, xp, fp, left=None, right=None):
    """numpy.interp (bm.interp, bmath_backends.py:373) in float64: xp
    increasing; left/right default to fp[0] / fp[-1]."""
    torch = _torch()
    X, XP, FP = _f64(torch, x, "x"), _f64(torch, xp, "xp"), _f64(torch, fp, "fp")
    m = int(XP.numel())
    if m == 0:
        raise ValueError("array of sample points is empty")
    if int(FP.numel()) != m:
        raise ValueError("fp and xp are not of the same length.")
    lv = float(FP[0].item()) if left is None else float(left)
    rv = float(FP[-1].item()) if right is None else float(right)
    out = torch.empty_like(X)
    _lib.call("blmi_interp", ct.c_void_p(X.data_ptr()), int(X.numel()), ct.c_void_p(XP.data_ptr()),
              ct.c_void_p(FP.data_ptr()), m, lv, rv, ct.c_void_p(out.data_ptr()), _stream(torch))
    return _like(torch, out, x)


def gradient(f, h=1.0):
    """numpy.gradient(f, h) for a 1-D array and scalar spacing, edge order 1
    (bm.gradient, bmath_backends.py:429), float64."""
    torch = _torch()
    F = _f64(torch, f, "f")
    n = int(F.numel())
    if n < 2:
        raise ValueError("Shape of array too small to calculate a numerical gradient, "
                         "at least (edge_order + 1) elements are required.")
    out = torch.empty_like(F)
    _lib.call("blmi_gradient", ct.c_void_p(F.data_ptr()), n, float(h), ct.c_void_p(out.data_ptr()),
              _stream(torch))
    return _like(torch, out, f)


def diff(a):
    """numpy.diff(a) of a 1-D array (bm.diff, bmath_backends.py:441), float64."""
    torch = _torch()
    A = _f64(torch, a, "a")
    n = int(A.numel())
    out = torch.empty(max(n - 1, 0), dtype=torch.float64, device="cuda")
    _lib.call("blmi_diff", ct.c_void_p(A.data_ptr()), n, ct.c_void_p(out.data_ptr()),
              _stream(torch))
    return _like(torch, out, a)


def mtw_shift_add(memory, time, t_rev, voltage, n_add=None):
    """One multi-turn-wake update (impedance.py:491-512, 533-550):
    interp(time + t_rev, time, memory, 0, 0) with voltage[:n_add] added to
    its first n_add points.  Returns a new float64 array of the kind of
    ``memory``."""
    torch = _torch()
    M, T = _f64(torch, memory, "memory"), _f64(torch, time, "time")
    n = int(M.numel())
    if int(T.numel()) != n:
        raise ValueError("memory and time differ in length")
    V = _f64(torch, voltage, "voltage") if voltage is not None else None
    k = 0 if V is None else min(int(V.numel()), n if n_add is None else int(n_add))
    out = torch.empty_like(M)
    _lib.call("blmi_mtw_shift_add", ct.c_void_p(M.data_ptr()), ct.c_void_p(out.data_ptr()),
              ct.c_void_p(T.data_ptr()), n, float(t_rev),
              ct.c_void_p(V.data_ptr()) if V is not None else None, k, _stream(torch))
    return _like(torch, out, memory)


def beam_phase(bin_centers, profile, alpha, omega_rf, phi_rf, bin_size) -> float:
    """beam_phase.cpp:6-39 via butils_wrap_cpp.py:785-808: scoeff / ccoeff of
    the exponentially windowed profile (float64, one synchronisation)."""
    torch = _torch()
    bc, p = _f64(torch, bin_centers, "bin_centers"), _f64(torch, profile, "profile")
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.call("blmi_beam_phase", ct.c_void_p(bc.data_ptr()), ct.c_void_p(p.data_ptr()),
              int(p.numel()), float(alpha), float(omega_rf), float(phi_rf), float(bin_size),
              ct.c_void_p(out.data_ptr()), _stream(torch))
    return float(out.item())


def beam_phase_fast(bin_centers, profile, omega_rf, phi_rf, bin_size) -> float:
    """beam_phase.cpp:42-65 via butils_wrap_cpp.py:811-832 (no window)."""
    torch = _torch()
    bc, p = _f64(torch, bin_centers, "bin_centers"), _f64(torch, profile, "profile")
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.call("blmi_beam_phase_fast", ct.c_void_p(bc.data_ptr()), ct.c_void_p(p.data_ptr()),
              int(p.numel()), float(omega_rf), float(phi_rf), float(bin_size),
              ct.c_void_p(out.data_ptr()), _stream(torch))
    return float(out.item())


def strided_mean(x, stride: int = 1) -> float:
    """numpy.mean(x[::stride]) in float64 with a fixed summation order (the
    radial loop's mean energy offset, beam_feedback.py:375-378)."""
    torch = _torch()
    X = _f64(torch, x, "x")
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.call("blmi_strided_mean", ct.c_void_p(X.data_ptr()), int(X.numel()), int(stride),
              ct.c_void_p(out.data_ptr()), _stream(torch))
    return float(out.item())


def beam_statistics(dt, dE, id=None):
    """beam.py:430-458: (n_alive, mean_dt, sigma_dt, mean_dE, sigma_dE) of
    the particles with id != 0 (all when id is None), std with ddof = 0."""
    torch = _torch()
    x = dt if isinstance(dt, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(dt)).cuda()
    e = dE if isinstance(dE, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(dE)).cuda()
    if x.dtype != e.dtype or x.dtype not in (torch.float64, torch.float32):
        raise TypeError("dt and dE must both be float64 or float32")
    x, e = x.contiguous(), e.contiguous()
    ids = None
    if id is not None:
        ids = id if isinstance(id, torch.Tensor) else torch.from_numpy(np.asarray(id)).cuda()
        ids = ids.to(torch.int64).contiguous()
    out = torch.empty(5, dtype=torch.float64, device="cuda")
    name = "blmi_beam_statistics" + ("_f32" if x.dtype == torch.float32 else "")
    _lib.call(name, ct.c_void_p(x.data_ptr()), ct.c_void_p(e.data_ptr()),
              ct.c_void_p(ids.data_ptr()) if ids is not None else None, int(x.numel()),
              ct.c_void_p(out.data_ptr()), _stream(torch))
    r = out.cpu().numpy()
    return int(r[0]), float(r[1]), float(r[2]), float(r[3]), float(r[4])


def induced_voltage_1turn(profile, total_impedance, n_fft: int, scale: float, n_out: int = 0):
    """-(scale * irfft(Z * rfft(profile, n_fft)))[:n_out]: one launch on one
    workgroup, or three (four-step FFT) beyond its LDS (impedance.py:457-489 with scale = charge * e * ratio; n_out defaults to
    len(profile))."""
    torch = _torch()
    p = _Staged(torch, profile, False, torch.float64, "profile")
    if isinstance(total_impedance, torch.Tensor):
        z = total_impedance.to(torch.complex128).contiguous()
    else:
        z = torch.from_numpy(np.ascontiguousarray(total_impedance, dtype=np.complex128)).to("cuda")
    n_out = int(n_out) if n_out else _len(profile)
    out = torch.empty(n_out, dtype=torch.float64, device="cuda")
    _lib.call("blmi_induced_voltage", p.ptr, _len(profile), ct.c_void_p(z.data_ptr()), int(n_fft),
              float(scale), ct.c_void_p(out.data_ptr()), n_out, _stream(torch))
    return out if isinstance(profile, torch.Tensor) else out.cpu().numpy()
