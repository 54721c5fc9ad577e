"""Backend singleton ``bmath`` (blond/utils/__init__.py:119-124).

The reference swaps ``bm.__class__`` between Python/Numba/C++/GPU backends
(blond/utils/bmath_backends.py:230-344).  This package has one backend: the
MI355X library.  ``bmath.use_gpu()`` / ``bmath.use_mi355x()`` select the
device; the hot-path functions are the HIP kernels in every case.
"""
from __future__ import annotations

from . import bmath as _k
from ._precision import PrecisionClass, precision


class MI355XBackend:
    """The ``bm`` object: hot-path functions with the reference signatures."""

    def __init__(self):
        self.precision = precision
        self.device = "GPU"
        self.kick = _k.kick
        self.drift = _k.drift
        self.slice_beam = _k.slice_beam
        self.slice_smooth = _k.slice_smooth
        self.sparse_histogram = _k.sparse_histogram
        self.linear_interp_kick = _k.linear_interp_kick
        self.rf_volt_comp = _k.rf_volt_comp
        self.fast_resonator = _k.fast_resonator
        self.rfft = _k.rfft
        self.irfft = _k.irfft
        self.rfftfreq = _k.rfftfreq
        self.induced_voltage_1turn = _k.induced_voltage_1turn
        self.beam_statistics = _k.beam_statistics
        self.interp = _k.interp
        self.gradient = _k.gradient
        self.diff = _k.diff
        self.mtw_shift_add = _k.mtw_shift_add
        self.beam_phase = _k.beam_phase
        self.beam_phase_fast = _k.beam_phase_fast
        self.strided_mean = _k.strided_mean

    def use_gpu(self, gpu_id: int = 0):
        """Select the GPU (bmath_backends.py:312-344)."""
        import torch

        from .. import _lib

        torch.cuda.set_device(gpu_id)
        _lib.call("blmi_init", int(gpu_id))

    use_mi355x = use_gpu

    def use_cpu(self):
        raise RuntimeError("blond_amd has no CPU backend: the tracking path runs on the MI355X only")

    def use_precision(self, _precision: str = "double"):
        """bmath_backends.py:299-310: 'double' or 'single'."""
        precision.set(_precision)

    def in_mpi(self) -> bool:
        return False

    def report_backend(self):
        print(f"Using the {self.device} (MI355X HIP library)")


bmath = MI355XBackend()
bm = bmath

__all__ = ["bmath", "bm", "precision", "PrecisionClass"]
