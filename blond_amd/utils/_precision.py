"""Floating-point precision switch (blond/utils/__init__.py:18-55)."""
from __future__ import annotations

import ctypes as ct

import numpy as np


class PrecisionClass:
    """Holds the active real type: 'double' (default) or 'single'."""

    def __init__(self, _precision: str = "double"):
        self.set(_precision)

    def set(self, _precision: str = "double"):
        if _precision in ["single", "s", "32", "float32", "float", "f"]:
            self.str = "single"
            self.real_t = np.float32
            self.c_real_t = ct.c_float
            self.complex_t = np.complex64
            self.num = 1
        elif _precision in ["double", "d", "64", "float64"]:
            self.str = "double"
            self.real_t = np.float64
            self.c_real_t = ct.c_double
            self.complex_t = np.complex128
            self.num = 2
        else:
            raise ValueError(f"{_precision=} is not recognized, use 'single' or 'double'")

    @property
    def torch_dtype(self):
        import torch

        return torch.float64 if self.num == 2 else torch.float32


precision = PrecisionClass("double")
