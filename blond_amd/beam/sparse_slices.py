"""Slicing of a sparse bunch train (blond/beam/sparse_slices.py:38-210).

One Profile of ``n_slices_bucket`` slices per filled RF bucket of the
filling pattern, with all profiles filled at once by ``bm.sparse_histogram``
(a HIP kernel; sparse_histogram.cpp:22-87) or one by one.
"""
from __future__ import annotations

import numpy as np

from ..utils import bmath as bm
from .profile import CutOptions, Profile


class SparseSlices:
    """sparse_slices.py:38-210."""

    def __init__(self, rf_station, beam, n_slices_bucket: int, filling_pattern,
                 tracker: str = "C", direct_slicing: bool = False):
        self.beam = beam
        self.rf_station = rf_station
        self.n_slices_bucket = n_slices_bucket
        self.filling_pattern = filling_pattern
        # the row of each RF bucket's profile, -1 where the bucket is empty
        self.bunch_indexes = np.cumsum(filling_pattern) * filling_pattern - 1
        self.n_filled_buckets = int(np.sum(filling_pattern))
        self.cut_left_array = np.zeros(self.n_filled_buckets)
        self.cut_right_array = np.zeros(self.n_filled_buckets)
        self.set_cuts()
        self.profiles_list = []
        self.n_macroparticles_array = np.zeros((self.n_filled_buckets, n_slices_bucket))
        self.bin_centers_array = np.zeros((self.n_filled_buckets, n_slices_bucket))
        self.edges_array = np.zeros((self.n_filled_buckets, n_slices_bucket + 1))
        for i in range(self.n_filled_buckets):
            self.profiles_list.append(Profile(beam, CutOptions(
                cut_left=float(self.cut_left_array[i]), cut_right=float(self.cut_right_array[i]),
                n_slices=n_slices_bucket)))
            self.profiles_list[i].n_macroparticles = self.n_macroparticles_array[i, :]
            self.bin_centers_array[i, :] = self.profiles_list[i].bin_centers
            self.edges_array[i, :] = self.profiles_list[i].edges
            self.profiles_list[i].bin_centers = self.bin_centers_array[i, :]
        if tracker == "C":
            self.track = self._histogram_c
        elif tracker == "onebyone":
            self.track = self._histogram_one_by_one
        else:
            raise NameError(f"{tracker=}")
        if direct_slicing:
            self.track()

    def set_cuts(self):
        """sparse_slices.py:160-177: each filled bucket spans one RF period."""
        t_rf = self.rf_station.t_rf[0, self.rf_station.counter[0]]
        self.cut_left_array = np.zeros(self.n_filled_buckets)
        self.cut_right_array = np.zeros(self.n_filled_buckets)
        for i in range(self.n_filled_buckets):
            bucket_index = np.where(self.filling_pattern)[0][i]
            self.cut_left_array[i] = bucket_index * t_rf
            self.cut_right_array[i] = (bucket_index + 1) * t_rf

    def _histogram_c(self):
        """sparse_slices.py:179-192: all profiles in one launch."""
        bm.sparse_histogram(self.beam.dt, self.n_macroparticles_array, self.cut_left_array,
                            self.cut_right_array, self.bunch_indexes, self.n_slices_bucket)

    def _histogram_one_by_one(self):
        """sparse_slices.py:203-210."""
        for i in range(self.n_filled_buckets):
            self.profiles_list[i].track()

    def to_gpu(self, recursive=True):
        """The profiles as rows of one device array (each Profile keeps a
        row view, as on the host)."""
        import torch

        if isinstance(self.n_macroparticles_array, torch.Tensor):
            return
        if recursive and self.beam is not None:
            self.beam.to_gpu()
        dt = bm.precision.torch_dtype
        self.n_macroparticles_array = torch.from_numpy(
            np.ascontiguousarray(self.n_macroparticles_array)).to(device="cuda", dtype=dt)
        self.bin_centers_array = torch.from_numpy(
            np.ascontiguousarray(self.bin_centers_array)).to(device="cuda", dtype=dt)
        for i, p in enumerate(self.profiles_list):
            p.n_macroparticles = self.n_macroparticles_array[i]
            p.bin_centers = self.bin_centers_array[i]
            p._device = "GPU"
