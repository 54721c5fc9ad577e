"""Slicing of the beam (blond/beam/profile.py:45-303, 446-991).

The per-turn histogram runs on the MI355X (``bm.slice_beam`` /
``bm.slice_smooth``).  Profile fits and filters (FitOptions / FilterOptions
other than None) are not part of the tracking path and raise
NotImplementedError.
"""
from __future__ import annotations

import numpy as np

from ..utils import bmath as bm


class CutOptions:
    """profile.py:45-303."""

    def __init__(self, cut_left=None, cut_right=None, n_slices: int = 100, n_sigma=None,
                 cuts_unit: str = "s", rf_station=None):
        self.cut_left = float(cut_left) if cut_left is not None else None
        self.cut_right = float(cut_right) if cut_right is not None else None
        self.n_slices = int(n_slices)
        self.n_sigma = float(n_sigma) if n_sigma is not None else None
        self.cuts_unit = cuts_unit
        self.rf_station = rf_station
        if self.cuts_unit == "rad" and self.rf_station is None:
            raise RuntimeError('Argument "rf_station" required convert from radians to seconds')
        if self.cuts_unit not in ["rad", "s"]:
            raise NameError(f'cuts_unit should be "s" or "rad", not {cuts_unit=} !')
        self.edges = np.zeros(n_slices + 1, dtype=bm.precision.real_t, order="C")
        self.bin_centers = np.zeros(n_slices, dtype=bm.precision.real_t, order="C")
        self.bin_size = 0.0
        self._device = "CPU"

    def set_cuts(self, beam=None):
        """profile.py:175-211: explicit cuts, or the beam extent +-5%, or n_sigma."""
        if self.cut_left is None and self.cut_right is None:
            dt = _host(beam.dt)
            if self.n_sigma is None:
                dt_min = dt.min()
                dt_max = dt.max()
                self.cut_left = dt_min - 0.05 * (dt_max - dt_min)
                self.cut_right = dt_max + 0.05 * (dt_max - dt_min)
            else:
                mean_coords = np.mean(dt)
                sigma_coords = np.std(dt)
                self.cut_left = mean_coords - self.n_sigma * sigma_coords / 2
                self.cut_right = mean_coords + self.n_sigma * sigma_coords / 2
        else:
            self.cut_left = float(self.convert_coordinates(self.cut_left, self.cuts_unit))
            self.cut_right = float(self.convert_coordinates(self.cut_right, self.cuts_unit))
        self.edges = np.linspace(self.cut_left, self.cut_right, self.n_slices + 1).astype(
            dtype=bm.precision.real_t, order="C", copy=False)
        self.bin_centers = (self.edges[:-1] + self.edges[1:]) / 2
        self.bin_size = (self.cut_right - self.cut_left) / self.n_slices

    def convert_coordinates(self, value, input_unit_type):
        """profile.py:228-244: 'rad' -> 's' with omega_rf[0, counter]."""
        if input_unit_type == "s":
            return value
        if input_unit_type == "rad":
            return value / float(self.rf_station.host_array("omega_rf")[0, self.rf_station.counter[0]])
        raise NameError(input_unit_type)

    def get_slices_parameters(self):
        return (self.n_slices, self.cut_left, self.cut_right, self.n_sigma, self.edges,
                self.bin_centers, self.bin_size)

    def to_gpu(self, recursive=True):
        if self._device == "GPU":
            return
        self.edges = _dev(self.edges)
        self.bin_centers = _dev(self.bin_centers)
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        self.edges = _host(self.edges)
        self.bin_centers = _host(self.bin_centers)
        self._device = "CPU"


class FitOptions:
    def __init__(self, fit_option=None, fit_extra_options=None):
        self.fit_option = fit_option
        self.fit_extra_options = fit_extra_options


class FilterOptions:
    def __init__(self, filter_method=None, filter_extra_options=None):
        self.filter_method = filter_method
        self.filter_extra_options = filter_extra_options


class OtherSlicesOptions:
    """profile.py:446-480 (direct_slicing defaults to True)."""

    def __init__(self, smooth: bool = False, direct_slicing: bool = True):
        self.smooth = smooth
        self.direct_slicing = direct_slicing


class Profile:
    """profile.py:483-991: the beam line density on n_slices bins."""

    def __init__(self, beam, cut_options=None, fit_options=None, filter_options=None,
                 other_slices_options=None):
        cut_options = cut_options if cut_options is not None else CutOptions()
        fit_options = fit_options if fit_options is not None else FitOptions()
        filter_options = filter_options if filter_options is not None else FilterOptions()
        other_slices_options = (other_slices_options if other_slices_options is not None
                                else OtherSlicesOptions())
        if fit_options.fit_option is not None or filter_options.filter_method is not None:
            raise NotImplementedError("profile fits and filters are outside the tracking path")
        self.cut_options = cut_options
        cut_options.set_cuts(beam)
        self.beam = beam
        (self.n_slices, self.cut_left, self.cut_right, self.n_sigma, self.edges,
         self.bin_centers, self.bin_size) = cut_options.get_slices_parameters()
        self.n_macroparticles = np.zeros(self.n_slices, dtype=bm.precision.real_t, order="C")
        self.beam_spectrum = np.array([], dtype=bm.precision.real_t)
        self.beam_spectrum_freq = np.array([], dtype=bm.precision.real_t)
        self.operations = [self._slice_smooth if other_slices_options.smooth else self._slice]
        self.smooth = bool(other_slices_options.smooth)
        self.fit_option = None
        self._device = "CPU"
        if other_slices_options.direct_slicing and self.beam is not None:
            self.track()

    def track(self):
        """profile.py:700-708."""
        for operation in self.operations:
            operation()

    def _slice(self):
        """profile.py:710-719."""
        bm.slice_beam(self.beam.dt, self.n_macroparticles, self.cut_left, self.cut_right)

    def _slice_smooth(self):
        """profile.py:764-773."""
        bm.slice_smooth(self.beam.dt, self.n_macroparticles, self.cut_left, self.cut_right)

    def beam_spectrum_freq_generation(self, n_sampling_fft: int):
        """profile.py:892-897."""
        self.beam_spectrum_freq = bm.rfftfreq(n_sampling_fft, self.bin_size)

    def beam_spectrum_generation(self, n_sampling_fft: int):
        """profile.py:899-903."""
        self.beam_spectrum = bm.rfft(self.n_macroparticles, n_sampling_fft)

    def beam_profile_derivative(self, mode: str = "gradient"):
        """profile.py:905-937: (bin_centers, d(profile)/dt) with the spacing
        dist = bin_centers[1] - bin_centers[0].  "gradient" is
        numpy.gradient(profile, dist); "diff" interpolates diff(profile)/dist
        from the bin edges' midpoints back onto the bin centres.  "filter1d"
        is CPU-only in the reference (a scipy filter) and raises as its GPU
        backend does."""
        bin_centers = self.bin_centers
        bc2 = _host(bin_centers[:2])
        dist_centers = bc2[1] - bc2[0]
        if mode == "filter1d":
            raise RuntimeError("filter1d mode is not supported in GPU.")
        if mode == "gradient":
            derivative = bm.gradient(self.n_macroparticles, dist_centers)
        elif mode == "diff":
            derivative = bm.diff(self.n_macroparticles) / dist_centers
            diff_centers = bin_centers[0:-1] + dist_centers / 2
            derivative = bm.interp(bin_centers, diff_centers, derivative)
        else:
            raise RuntimeError("Option for derivative is not recognized.")
        return bin_centers, derivative

    def to_gpu(self, recursive=True):
        """profile.py:924-951."""
        if self._device == "GPU":
            return
        if recursive and self.beam is not None:
            self.beam.to_gpu()
        if recursive and self.cut_options is not None:
            self.cut_options.to_gpu()
        self.bin_centers = self.cut_options.bin_centers
        self.edges = self.cut_options.edges
        self.n_macroparticles = _dev(self.n_macroparticles)
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        if recursive and self.beam is not None:
            self.beam.to_cpu()
        if recursive and self.cut_options is not None:
            self.cut_options.to_cpu()
        self.bin_centers = self.cut_options.bin_centers
        self.edges = self.cut_options.edges
        self.n_macroparticles = _host(self.n_macroparticles)
        self._device = "CPU"


def _host(a):
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda",
                                                          dtype=bm.precision.torch_dtype)
