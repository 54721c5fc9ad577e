"""Induced voltage from impedances, single-turn frequency-domain form
(blond/impedances/impedance.py:48-199, 202-489, 760-940).

Supported: InducedVoltageFreq without multi-turn wake, summed by
TotalInducedVoltage.  Multi-turn wakes, InducedVoltageTime and
InductiveImpedance are not part of this path and raise NotImplementedError.
"""
from __future__ import annotations

import numpy as np

from ..constants import e
from ..toolbox.next_regular import next_regular
from ..utils import bmath as bm


def _is_dev(a) -> bool:
    return hasattr(a, "is_cuda") and a.is_cuda


class TotalInducedVoltage:
    """impedance.py:48-146."""

    def __init__(self, beam, profile, induced_voltage_list):
        self.beam = beam
        self.profile = profile
        self.induced_voltage_list = induced_voltage_list
        self.induced_voltage = np.zeros(int(self.profile.n_slices), dtype=bm.precision.real_t,
                                        order="C")
        self.time_array = self.profile.bin_centers
        self._device = "CPU"

    def reprocess(self):
        for obj in self.induced_voltage_list:
            obj.process()

    def induced_voltage_sum(self):
        """impedance.py:113-131: sum of the sources cut to n_slices."""
        total = 0
        for obj in self.induced_voltage_list:
            obj.induced_voltage_generation()
            total = total + obj.induced_voltage[: self.profile.n_slices]
        if _is_dev(total):
            self.induced_voltage = total.contiguous()
        else:
            self.induced_voltage = np.ascontiguousarray(total, dtype=bm.precision.real_t)

    def track(self):
        """impedance.py:133-146: induced voltage then the interpolated kick."""
        self.induced_voltage_sum()
        bm.linear_interp_kick(dt=self.beam.dt, dE=self.beam.dE, voltage=self.induced_voltage,
                              bin_centers=self.profile.bin_centers,
                              charge=self.beam.particle.charge, acceleration_kick=0.0)

    def to_gpu(self, recursive=True):
        if self._device == "GPU":
            return
        if recursive:
            for obj in self.induced_voltage_list:
                obj.to_gpu()
        import torch

        self.induced_voltage = torch.from_numpy(np.ascontiguousarray(self.induced_voltage)).to(
            device="cuda", dtype=bm.precision.torch_dtype)
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        if recursive:
            for obj in self.induced_voltage_list:
                obj.to_cpu()
        self.induced_voltage = self.induced_voltage.cpu().numpy()
        self._device = "CPU"


class _InducedVoltage:
    """impedance.py:202-489 (single-turn wake)."""

    def __init__(self, beam, profile, frequency_resolution=None, wake_length=None,
                 multi_turn_wake=False, mtw_mode="time", rf_station=None, use_regular_fft=True):
        if multi_turn_wake:
            raise NotImplementedError("multi-turn wakes are not supported")
        self.beam = beam
        self.profile = profile
        self.induced_voltage = np.zeros(int(profile.n_slices), dtype=bm.precision.real_t)
        self.wake_length_input = wake_length
        self.frequency_resolution_input = frequency_resolution
        self.use_regular_fft = use_regular_fft
        self.rf_params = rf_station
        self.multi_turn_wake = multi_turn_wake
        self.mtw_mode = mtw_mode
        self.n_fft = None
        self.wake_length = None
        self.n_induced_voltage = None
        self.total_impedance = None
        self._device = "CPU"
        self.process()

    def process(self):
        """impedance.py:349-403: wake length / frequency resolution -> sizes."""
        p = self.profile
        if self.wake_length_input is not None and self.frequency_resolution_input is None:
            self.n_induced_voltage = int(np.ceil(self.wake_length_input / p.bin_size))
            if self.n_induced_voltage < p.n_slices:
                raise RuntimeError("Error: too short wake length. Increase it above "
                                   "{0:1.2e} s.".format(p.n_slices * p.bin_size))
            self.wake_length = self.n_induced_voltage * p.bin_size
        elif self.frequency_resolution_input is not None and self.wake_length_input is None:
            self.n_induced_voltage = int(np.ceil(1 / (p.bin_size * self.frequency_resolution_input)))
            if self.n_induced_voltage < p.n_slices:
                raise RuntimeError("Error: too large frequency_resolution. Reduce it below "
                                   "{0:1.2e} Hz.".format(1 / (p.cut_right - p.cut_left)))
            self.wake_length = self.n_induced_voltage * p.bin_size
        elif self.wake_length_input is None and self.frequency_resolution_input is None:
            self.wake_length = p.cut_right - p.cut_left
            self.n_induced_voltage = p.n_slices
        else:
            raise RuntimeError("Error: only one of wake_length or frequency_resolution can be "
                               "specified.")
        self.induced_voltage_generation = self.induced_voltage_1turn

    def induced_voltage_1turn(self, beam_spectrum_dict=None):
        """impedance.py:457-489: V = -(q e ratio) irfft(Z * rfft(profile, n_fft)), cut.
        One HIP launch (rfft, impedance product, irfft, scale) when the
        profile is device resident; the numpy-array case runs the same kernel
        through host staging."""
        scale = self.beam.particle.charge * e * self.beam.ratio
        self.induced_voltage = bm.induced_voltage_1turn(
            self.profile.n_macroparticles, self.total_impedance, self.n_fft, scale,
            self.n_induced_voltage)

    def to_gpu(self, recursive=True):
        if self._device == "GPU":
            return
        import torch

        self.induced_voltage = torch.from_numpy(np.ascontiguousarray(self.induced_voltage)).to(
            device="cuda", dtype=bm.precision.torch_dtype)
        self.total_impedance = torch.from_numpy(
            np.ascontiguousarray(self.total_impedance, dtype=np.complex128)).to("cuda")
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        self.induced_voltage = self.induced_voltage.cpu().numpy()
        self.total_impedance = self.total_impedance.cpu().numpy()
        self._device = "CPU"


class InducedVoltageFreq(_InducedVoltage):
    """impedance.py:760-890."""

    def __init__(self, beam, profile, impedance_source_list, frequency_resolution=None,
                 multi_turn_wake=False, front_wake_length=0, rf_station=None, mtw_mode="time",
                 use_regular_fft=True):
        self.impedance_source_list = impedance_source_list
        self.total_impedance = 0
        self.front_wake_length = front_wake_length
        self.freq = None
        self.frequency_resolution = None
        _InducedVoltage.__init__(self, beam, profile, wake_length=None,
                                 frequency_resolution=frequency_resolution,
                                 multi_turn_wake=multi_turn_wake, rf_station=rf_station,
                                 mtw_mode=mtw_mode, use_regular_fft=use_regular_fft)

    def process(self):
        """impedance.py:846-874."""
        _InducedVoltage.process(self)
        if self.use_regular_fft:
            self.n_fft = next_regular(self.n_induced_voltage)
        else:
            self.n_fft = self.n_induced_voltage
        self.profile.beam_spectrum_freq_generation(self.n_fft)
        self.freq = self.profile.beam_spectrum_freq
        self.frequency_resolution = 1 / (self.n_fft * self.profile.bin_size)
        self.sum_impedances(self.freq)

    def sum_impedances(self, freq):
        """impedance.py:876-890: total impedance / bin_size."""
        self.total_impedance = np.zeros(freq.shape, dtype=bm.precision.complex_t, order="C")
        for source in self.impedance_source_list:
            source.imped_calc(freq)
            self.total_impedance += source.impedance
        self.total_impedance /= self.profile.bin_size
