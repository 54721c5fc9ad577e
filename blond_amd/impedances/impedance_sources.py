"""Impedance sources on the tracking path: resonators
(blond/impedances/impedance_sources.py:239-471), in the frequency domain
(imped_calc) and the time domain (wake_calc)."""
from __future__ import annotations

import numpy as np

from ..utils import bmath as bm


class _ImpedanceObject:
    def __init__(self):
        self.time_array = 0
        self.wake = 0
        self.frequency_array = 0
        self.impedance = 0


class Resonators(_ImpedanceObject):
    """Z(f) = sum_k R_k / (1 + j Q_k (f/f_rk - f_rk/f))  (impedance_sources.py:239-414)."""

    def __init__(self, R_S, frequency_R, Q):
        _ImpedanceObject.__init__(self)
        self.R_S = np.array([R_S], dtype=float).flatten()
        self.frequency_R = np.array([frequency_R], dtype=float).flatten()
        self.Q = np.array([Q], dtype=float).flatten()
        if np.count_nonzero(self.Q < 0.5) > 0:
            raise RuntimeError("All quality factors Q must be greater or equal 0.5")
        self.n_resonators = len(self.R_S)

    @property
    def frequency_R(self):
        return self.__frequency_R

    @frequency_R.setter
    def frequency_R(self, frequency_R):
        self.__frequency_R = frequency_R
        self.__omega_R = 2 * np.pi * frequency_R

    @property
    def omega_R(self):
        return self.__omega_R

    @omega_R.setter
    def omega_R(self, omega_R):
        self.__frequency_R = omega_R / 2 / np.pi
        self.__omega_R = omega_R

    def wake_calc(self, time_array):
        """impedance_sources.py:345-380: the resonators' wake in Ohm/s, a
        set-up array formed on the host with numpy as in the reference
        (bm.cos / bm.sin are numpy's there)."""
        self.time_array = time_array
        self.wake = np.zeros(self.time_array.shape, dtype=bm.precision.real_t, order="C")
        for i in range(0, self.n_resonators):
            alpha = self.omega_R[i] / (2 * self.Q[i])
            omega_bar = np.sqrt(self.omega_R[i] ** 2 - alpha ** 2)
            self.wake += ((np.sign(self.time_array) + 1) * self.R_S[i] * alpha
                          * np.exp(-alpha * self.time_array)
                          * (np.cos(omega_bar * self.time_array)
                             - alpha / omega_bar * np.sin(omega_bar * self.time_array)))

    def imped_calc(self, frequency_array):
        """impedance_sources.py:382-414: resonator impedance on the GPU
        (bm.fast_resonator, fast_resonator.cpp:19-73)."""
        if not isinstance(self.impedance, int) and len(self.impedance) == len(frequency_array):
            impedance = self.impedance
        else:
            impedance = None
        self.frequency_array = frequency_array
        self.impedance = bm.fast_resonator(R_S=self.R_S, Q=self.Q,
                                           frequency_array=self.frequency_array,
                                           frequency_R=self.frequency_R, impedance=impedance)
