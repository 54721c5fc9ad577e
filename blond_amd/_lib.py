"""ctypes binding of libblond_mi355x.so (include/blond_mi355x.h).

The library is the only compute path of this package: if it is missing or
cannot be loaded, every entry point raises ``LibraryNotAvailable`` -- there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BLOND_MI355X_LIB", os.path.join(_HERE, "libblond_mi355x.so"))

_i = ct.c_int
_i64 = ct.c_int64
_d = ct.c_double
_f = ct.c_float
_p = ct.c_void_p

# name -> argtypes (restype is int unless listed in _RESTYPES)
SIGNATURES = {
    "blmi_last_error": [],
    "blmi_version": [],
    "blmi_init": [_i],
    "blmi_device_info": [ct.POINTER(_i), ct.POINTER(_i)],
    "blmi_kick": [_p, _p, _i, _d, _p, _p, _p, _i64, _d, _p],
    "blmi_kick_f32": [_p, _p, _i, _f, _p, _p, _p, _i64, _f, _p],
    "blmi_drift": [_p, _p, _i, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _d, _i64, _p],
    "blmi_drift_f32": [_p, _p, _i, _f, _f, _f, _f, _f, _f, _f, _f, _f, _f, _f, _i64, _p],
    "blmi_histogram": [_p, _p, _d, _d, _i, _i64, _p],
    "blmi_histogram_f32": [_p, _p, _f, _f, _i, _i64, _p],
    "blmi_sparse_histogram": [_p, _p, _p, _p, _p, _i, _i, _i, _i64, _p],
    "blmi_sparse_histogram_f32": [_p, _p, _p, _p, _p, _i, _i, _i, _i64, _p],
    "blmi_smooth_histogram": [_p, _p, _d, _d, _i, _i64, _p],
    "blmi_smooth_histogram_f32": [_p, _p, _f, _f, _i, _i64, _p],
    "blmi_linear_interp_kick": [_p, _p, _p, _p, _d, _i, _i64, _d, _p],
    "blmi_linear_interp_kick_f32": [_p, _p, _p, _p, _f, _i, _i64, _f, _p],
    "blmi_linear_interp_kick_noalloc": [_p, _p, _p, _p, _d, _i, _i64, _d, _p],
    "blmi_rf_volt_comp": [_p, _p, _p, _p, _i, _i, _p, _p],
    "blmi_rf_volt_comp_f32": [_p, _p, _p, _p, _i, _i, _p, _p],
    "blmi_fast_resonator": [_p, _p, _p, _p, _p, _p, _i, _i, _p],
    "blmi_fft_max_lds_length": [ct.POINTER(_i64)],
    "blmi_rfft": [_p, _i64, _p, _i64, _p],
    "blmi_irfft": [_p, _i64, _p, _i64, _p],
    "blmi_induced_voltage": [_p, _i, _p, _i64, _d, _p, _i, _p],
    "blmi_interp": [_p, _i64, _p, _p, _i64, _d, _d, _p, _p],
    "blmi_gradient": [_p, _i64, _d, _p, _p],
    "blmi_diff": [_p, _i64, _p, _p],
    "blmi_mtw_shift_add": [_p, _p, _p, _i64, _d, _p, _i64, _p],
    "blmi_beam_phase": [_p, _p, _i, _d, _d, _d, _d, _p, _p],
    "blmi_beam_phase_fast": [_p, _p, _i, _d, _d, _d, _p, _p],
    "blmi_strided_mean": [_p, _i64, _i64, _p, _p],
    "blmi_beam_statistics": [_p, _p, _p, _i64, _p, _p],
    "blmi_beam_statistics_f32": [_p, _p, _p, _i64, _p, _p],
    "blmi_engine_create": [_i64, _i, _i, _i],
    "blmi_engine_destroy": [_p],
    "blmi_engine_set_particles": [_p, _p, _p, _i64],
    "blmi_engine_set_programs": [_p, _p, _p, _p, _d, _d, _p, _p, _p, _p, _p, _p, _p, _p, _d],
    "blmi_engine_set_slices": [_p, _i, _d, _d, _p],
    "blmi_engine_set_induced": [_p, _i64, _p, _d, _p, _p],
    "blmi_engine_set_wake": [_p, _i64, _i64, _p, _p, _p, _i],
    "blmi_engine_set_feedback": [_p, _i, _i, _i, _d, _d, _p, _p, _p, _p, _p, _i, _p],
    "blmi_engine_get_feedback": [_p, _p],
    "blmi_engine_get_programs": [_p, _i, _i, _p],
    "blmi_engine_set_statistics": [_p, _i, _d, _d],
    "blmi_engine_get_statistics": [_p, _i, _i, _p],
    "blmi_engine_set_launch": [_p, _i, _i, _i],
    "blmi_engine_set_turn_block": [_p, _i],
    "blmi_engine_set_turn": [_p, _i],
    "blmi_engine_get_turn": [_p, ct.POINTER(_i)],
    "blmi_engine_prime": [_p, _p],
    "blmi_engine_track": [_p, _i, _i, _p],
    "blmi_engine_time_kernels": [_p, _i, ct.POINTER(_f), ct.POINTER(_f), _p],
    "blmi_engine_set_method": [_p, _i],
    "blmi_engine_set_interpolation": [_p, _p],
    "blmi_engine_get_method": [_p, ct.POINTER(_i)],
    "blmi_engine_set_hist": [_p, _i, _i],
    "blmi_engine_set_order": [_p, _i],
    "blmi_engine_set_particles_f32": [_p, _p, _p, _i64],
    "blmi_engine_set_slices_f32": [_p, _i, _d, _d, _p],
    "blmi_engine_set_induced_f32": [_p, _i64, _p, _d, _p, _p],
    "blmi_ensemble_create": [_i, _i, _i, _i],
    "blmi_ensemble_destroy": [_p],
    "blmi_ensemble_set_particles": [_p, _p, _p, _p, _p],
    "blmi_ensemble_set_programs": [_p, _p, _p, _p, _d, _d, _p, _p, _p, _p, _p, _p, _p, _p, _d],
    "blmi_ensemble_set_slices": [_p, _i, _p, _p, _p],
    "blmi_ensemble_set_induced": [_p, _i64, _p, _p, _p, _p],
    "blmi_ensemble_set_launch": [_p, _i, _i, _i, _i],
    "blmi_ensemble_get_launch": [_p, ct.POINTER(_i), ct.POINTER(_i)],
    "blmi_ensemble_set_turn": [_p, _i],
    "blmi_ensemble_prime": [_p, _p],
    "blmi_ensemble_track": [_p, _i, _i, _p],
    "blmi_ensemble_time_turns": [_p, _i, ct.POINTER(_f), ct.POINTER(_f), _p],
    "blmi_ensemble_autotune": [_p, _i, ct.POINTER(_i), ct.POINTER(_i), ct.POINTER(_f), _p],
    "blmi_adjoint_create": [_i64, _i, _i, _i, _i],
    "blmi_adjoint_destroy": [_p],
    "blmi_adjoint_set_particles": [_p, _p, _p],
    "blmi_adjoint_set_programs": [_p, _p, _p, _p, _d, _d, _p, _p, _p, _p, _p, _p, _p, _p, _d],
    "blmi_adjoint_set_slices": [_p, _d, _d],
    "blmi_adjoint_set_map": [_p, _i, _p],
    "blmi_adjoint_set_wake": [_p, _i64, _p, _d],
    "blmi_adjoint_forward": [_p, _i, _p, _p],
    "blmi_adjoint_backward": [_p, _p, _p, _p, _p],
}
_RESTYPES = {
    "blmi_last_error": ct.c_char_p,
    "blmi_engine_create": ct.c_void_p,
    "blmi_ensemble_create": ct.c_void_p,
    "blmi_adjoint_create": ct.c_void_p,
}


class LibraryNotAvailable(RuntimeError):
    """libblond_mi355x.so is not built or cannot be loaded."""


class BlmiError(RuntimeError):
    """A library call returned a non-zero status."""


_LIB = None
_LOAD_ERROR = None


def load(path: str | None = None) -> ct.CDLL:
    """Load (once) and return the library with every signature bound."""
    global _LIB, _LOAD_ERROR
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        _LOAD_ERROR = f"{p} not found: build it with `make -C blond_amd/csrc` (or __graft_entry__.build())"
        raise LibraryNotAvailable(_LOAD_ERROR)
    try:
        lib = ct.CDLL(p)
    except OSError as exc:  # pragma: no cover - depends on the machine
        _LOAD_ERROR = f"cannot load {p}: {exc}"
        raise LibraryNotAvailable(_LOAD_ERROR) from exc
    missing = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            # a library built from older sources: the entry fails when called
            missing.append(name)
            continue
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, ct.c_int)
    lib.blmi_missing = frozenset(missing)
    if path is None:
        _LIB = lib
    return lib


def last_error() -> str:
    msg = load().blmi_last_error()
    return msg.decode() if msg else ""


def call(name: str, *args) -> None:
    """Call ``name`` and raise BlmiError on a non-zero status."""
    lib = load()
    if name in lib.blmi_missing:
        raise LibraryNotAvailable(f"{name} is not in {LIB_PATH}: rebuild the library")
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise BlmiError(f"{name} failed ({rc}): {last_error()}")
