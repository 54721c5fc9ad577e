"""Induced voltage from impedances and wakes (blond/impedances/impedance.py).

TotalInducedVoltage (48-146) sums the sources cut to n_slices; the sources
are InducedVoltageFreq (760-890), InducedVoltageTime (575-757) and
InductiveImpedance (943-1039).  The single-turn and multi-turn ("time" mode,
491-550) wakes run on the MI355X library.  The "freq" multi-turn mode reads
an attribute no code path sets (buffer_extra, impedance.py:414-418): it
raises the same AttributeError here.  Set-up arithmetic (sizes, time axes,
wakes) is host numpy, as in the reference.
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
        self.n_mtw_memory = None
        self.front_wake_buffer = None
        self.buffer_size = None
        self.mtw_memory = None
        self.total_impedance = None
        self.time_mtw = None
        self.shift_trev = None
        self._device = "CPU"
        self.process()

    @property
    def mtw_mode(self):
        return self._mtw_mode

    @mtw_mode.setter
    def mtw_mode(self, mtw_mode):
        """impedance.py:315-326."""
        if mtw_mode not in ("freq", "time"):
            raise ValueError(f"{mtw_mode=} not valid, choose either 'freq' or 'time'")
        self._mtw_mode = mtw_mode

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
        if self.multi_turn_wake:
            # impedance.py:405-453
            self.n_mtw_memory = self.n_induced_voltage
            self.front_wake_buffer = 0
            if self.mtw_mode == "freq":
                self.buffer_size = np.ceil(np.max(self.rf_params.t_rev) / p.bin_size)
                # the reference reads an attribute that nothing sets
                self.buffer_size += np.ceil(np.max(self.buffer_extra) / p.bin_size)
                raise NotImplementedError("multi-turn wake mode 'freq'")
            self.shift_trev = self.shift_trev_time
            self.time_mtw = np.linspace(0, self.wake_length, self.n_mtw_memory, endpoint=False,
                                        dtype=bm.precision.real_t)
            self.mtw_memory = np.zeros(self.n_mtw_memory, dtype=bm.precision.real_t, order="C")
            self.induced_voltage_generation = self.induced_voltage_mtw
        else:
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

    def induced_voltage_mtw(self, beam_spectrum_dict=None):
        """impedance.py:491-516: this turn's voltage (front-wake tail set to
        0) added to the memory shifted by one revolution period; the memory's
        first n_induced_voltage points are the induced voltage.  The shift
        (shift_trev_time, impedance.py:533-550) and the sum are one pass
        (bm.mtw_shift_add) writing the memory in place."""
        self.induced_voltage_1turn(beam_spectrum_dict)
        n_iv = self.n_induced_voltage
        n_add = slice(n_iv - self.front_wake_buffer, None).indices(n_iv)[0]
        t_rev = float(self.rf_params.t_rev[self.rf_params.counter[0]])
        new = bm.mtw_shift_add(self.mtw_memory, self.time_mtw, t_rev, self.induced_voltage, n_add)
        if _is_dev(self.mtw_memory):
            self.mtw_memory.copy_(new)  # in place: a fused tracker may hold this buffer
        else:
            self.mtw_memory = new       # float64, as numpy.interp returns
        self.induced_voltage = self.mtw_memory[:n_iv]

    def shift_trev_time(self):
        """impedance.py:533-550: the memory one revolution period later,
        interp(time + t_rev, time, memory, left=0, right=0)."""
        t_rev = float(self.rf_params.t_rev[self.rf_params.counter[0]])
        new = bm.mtw_shift_add(self.mtw_memory, self.time_mtw, t_rev, None)
        if _is_dev(self.mtw_memory):
            self.mtw_memory.copy_(new)
        else:
            self.mtw_memory = new

    def to_gpu(self, recursive=True):
        if self._device == "GPU":
            return
        import torch

        self.induced_voltage = torch.from_numpy(np.ascontiguousarray(self.induced_voltage)).to(
            device="cuda", dtype=bm.precision.torch_dtype)
        if self.total_impedance is not None:
            self.total_impedance = torch.from_numpy(
                np.ascontiguousarray(self.total_impedance, dtype=np.complex128)).to("cuda")
        if self.mtw_memory is not None:
            # the memory and its time axis are float64 on the device (the
            # reference's numpy.interp returns float64 as well)
            self.mtw_memory = torch.from_numpy(
                np.ascontiguousarray(self.mtw_memory, dtype=np.float64)).to("cuda")
            self.time_mtw = torch.from_numpy(
                np.ascontiguousarray(self.time_mtw, dtype=np.float64)).to("cuda")
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        self.induced_voltage = self.induced_voltage.cpu().numpy()
        if self.total_impedance is not None:
            self.total_impedance = self.total_impedance.cpu().numpy()
        if self.mtw_memory is not None:
            self.mtw_memory = self.mtw_memory.cpu().numpy()
            self.time_mtw = self.time_mtw.cpu().numpy()
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
        if self.front_wake_length:
            # impedance.py:867-871
            self.front_wake_buffer = int(np.ceil(np.max(self.front_wake_length)
                                                 / self.profile.bin_size))
        self.sum_impedances(self.freq)

    def sum_impedances(self, freq):
        """impedance.py:876-890: total impedance / bin_size."""
        self.total_impedance = np.zeros(freq.shape, dtype=bm.precision.complex_t, order="C")
        for source in self.impedance_source_list:
            source.imped_calc(freq)
            self.total_impedance += source.impedance
        self.total_impedance /= self.profile.bin_size


class InducedVoltageTime(_InducedVoltage):
    """impedance.py:575-757: the wake in the time domain, convolved through
    the FFT of the zero-padded total wake (a pseudo-impedance)."""

    def __init__(self, beam, profile, wake_source_list, wake_length=None, multi_turn_wake=False,
                 rf_station=None, mtw_mode="time", use_regular_fft=True):
        self.wake_source_list = wake_source_list
        self.total_wake = 0
        self.frequency_resolution = None
        self.time = None
        _InducedVoltage.__init__(self, beam, profile, frequency_resolution=None,
                                 wake_length=wake_length, multi_turn_wake=multi_turn_wake,
                                 rf_station=rf_station, mtw_mode=mtw_mode,
                                 use_regular_fft=use_regular_fft)

    def process(self):
        """impedance.py:650-683: n_fft covers the linear convolution."""
        _InducedVoltage.process(self)
        n = int(self.n_induced_voltage) + int(self.profile.n_slices) - 1
        self.n_fft = next_regular(n) if self.use_regular_fft else n
        self.frequency_resolution = 1 / (self.n_fft * self.profile.bin_size)
        self.time = np.arange(0, self.wake_length, self.wake_length / self.n_induced_voltage,
                              dtype=bm.precision.real_t)
        self.sum_wakes(self.time)

    def sum_wakes(self, time_array):
        """impedance.py:685-697.  The pseudo-impedance is a set-up transform
        on the host (numpy, as the reference's CPU backends do)."""
        self.total_wake = np.zeros(time_array.shape)
        for wake_object in self.wake_source_list:
            wake_object.wake_calc(time_array)
            self.total_wake += wake_object.wake
        self.total_impedance = np.fft.rfft(self.total_wake, self.n_fft)
        if self._device == "GPU":
            import torch

            self.total_impedance = torch.from_numpy(
                np.ascontiguousarray(self.total_impedance, dtype=np.complex128)).to("cuda")


class InductiveImpedance(_InducedVoltage):
    """impedance.py:943-1039: a constant imaginary Z/n (a per-turn program),
    V = -(q e / 2 pi * ratio * Z/n[turn] * t_rev[turn] / bin_size * dprofile/dt)."""

    def __init__(self, beam, profile, Z_over_n, rf_station, deriv_mode="gradient"):
        self.Z_over_n = Z_over_n
        self.deriv_mode = deriv_mode
        _InducedVoltage.__init__(self, beam, profile, rf_station=rf_station)

    def coefficient(self, index: int) -> float:
        """The scalar of impedance.py:995-1002 for turn ``index``, in the
        reference's left-to-right order."""
        return (self.beam.particle.charge * e / (2 * np.pi) * self.beam.ratio
                * self.Z_over_n[index] * self.rf_params.t_rev[index] / self.profile.bin_size)

    def induced_voltage_1turn(self, beam_spectrum_dict=None):
        """impedance.py:986-1006."""
        index = self.rf_params.counter[0]
        coef = float(self.coefficient(index))
        derivative = self.profile.beam_profile_derivative(self.deriv_mode)[1]
        induced_voltage = -(coef * derivative)
        self.induced_voltage = induced_voltage[: self.n_induced_voltage]

    def to_gpu(self, recursive=True):
        if self._device == "GPU":
            return
        import torch

        self.induced_voltage = torch.from_numpy(np.ascontiguousarray(self.induced_voltage)).to(
            device="cuda", dtype=bm.precision.torch_dtype)
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        self.induced_voltage = self.induced_voltage.cpu().numpy()
        self._device = "CPU"
