"""Per-turn bunch statistics writer (blond/monitors/monitors.py:36-467).

``BunchMonitor`` keeps the reference's datasets (group ``Beam``:
n_macroparticles_alive, mean_dt, mean_dE, sigma_dt, sigma_dE, epsn_rms_l and,
with a phase loop, the PL_* / SL_* / RL_* series), its buffering (written
every ``buffer_time`` turns) and its row indexing.  The statistics come from
``Beam.statistics`` (a device reduction) object by object, or from the moments
the fused turn accumulates inside its particle kernel (``FusedTracker(...,
monitor=...)``), with no extra pass over the particles.

Output: ``filename + ".h5"`` through h5py when it is importable (float32
datasets "f", PL_omegaRF float64, gzip 9, as the reference); otherwise the
same datasets, dtypes and rows in ``filename + ".npz"`` under the keys
``"Beam/<name>"``.
"""
from __future__ import annotations

import os

import numpy as np

try:  # pragma: no cover - depends on the environment
    import h5py as _h5py
except ImportError:  # h5py is optional
    _h5py = None

_F32 = np.float32


class _Store:
    """The ``Beam`` group of the output file: HDF5 with h5py, else an npz
    rewritten on every close."""

    def __init__(self, filename: str):
        self.filename = filename
        self.data: dict[str, np.ndarray] = {}
        self.h5 = None

    @property
    def path(self) -> str:
        return self.filename + (".h5" if _h5py is not None else ".npz")

    def create(self, name, dims, dtype):
        if _h5py is not None:
            if self.h5 is None:
                self.h5 = _h5py.File(self.path, "w")
            self.h5.require_group("Beam").create_dataset(name, shape=dims, dtype=dtype,
                                                         compression="gzip", compression_opts=9)
        else:
            self.data[name] = np.zeros(dims, dtype=dtype)

    def open(self):
        if _h5py is not None:
            self.h5 = _h5py.File(self.path, "r+")

    def set(self, name, key, value):
        if _h5py is not None:
            self.h5["Beam"][name][key] = value
        else:
            self.data[name][key] = value

    def close(self):
        if _h5py is not None:
            if self.h5 is not None:
                self.h5.close()
                self.h5 = None
        else:
            np.savez(self.path, **{f"Beam/{k}": v for k, v in self.data.items()})


class BunchMonitor:
    """monitors.py:36-467 (no profile fit, no LHC noise feedback)."""

    _SERIES = (("n_macroparticles_alive", "b_np_alive", _F32),
               ("mean_dt", "b_mean_dt", _F32), ("mean_dE", "b_mean_dE", _F32),
               ("sigma_dt", "b_sigma_dt", _F32), ("sigma_dE", "b_sigma_dE", _F32),
               ("epsn_rms_l", "b_epsn_rms", _F32))
    _PL_SERIES = (("PL_omegaRF", "b_PL_omegaRF", np.float64), ("PL_phiRF", "b_PL_phiRF", _F32),
                  ("PL_bunch_phase", "b_PL_bunch_phase", _F32),
                  ("PL_phase_corr", "b_PL_phase_corr", _F32),
                  ("PL_omegaRF_corr", "b_PL_omegaRF_corr", _F32),
                  ("SL_dphiRF", "b_SL_dphiRF", _F32), ("RL_drho", "b_RL_drho", _F32))

    def __init__(self, ring, rf_parameters, beam, filename, buffer_time=None, profile=None,
                 phase_loop=None, lhc_noise_feedback=None):
        if lhc_noise_feedback is not None:
            raise NotImplementedError("the LHC noise feedback is outside the tracking path")
        self.filename = filename
        self.n_turns = ring.n_turns
        self.i_turn = 0
        self.buffer_time = buffer_time if buffer_time is not None else self.n_turns
        self.rf_params = rf_parameters
        self.beam = beam
        self.profile = profile
        self.fit_option = bool(profile is not None and getattr(profile, "fit_option", None))
        self.phase_loop = phase_loop
        self.lhc_noise_feedback = None
        self.h5file = None
        self._alive = None  # alive count of the fused turn's rows
        self.init_data(self.filename, (self.n_turns + 1,))
        self.track()

    # -- reference flow -------------------------------------------------------
    def track(self):
        """monitors.py:117-129."""
        self.beam.statistics()
        self._record()

    def _record(self):
        self.write_buffer()
        self.i_turn += 1
        if self.i_turn > 0 and (self.i_turn % self.buffer_time) == 0:
            self.open()
            self.write_data(None, (self.n_turns + 1,))
            self.close()
            self.init_buffer()

    def init_data(self, filename, dims):
        """monitors.py:131-292: datasets with the current state in row 0."""
        self.beam.statistics()
        dirname = os.path.dirname(filename)
        if dirname:
            os.makedirs(dirname, exist_ok=True)
        self.h5file = _Store(filename)
        row0 = self._row()
        series = self._SERIES + (self._PL_SERIES if self.phase_loop is not None else ())
        for name, _, dtype in series:
            self.h5file.create(name, dims, dtype)
            self.h5file.set(name, 0, row0[name])
        self.close()
        self.init_buffer()

    def _row(self, turn_for_rf: int = 0) -> dict:
        b = self.beam
        alive = self._alive if self._alive is not None else b.n_macroparticles_alive
        row = {"n_macroparticles_alive": alive, "mean_dt": b.mean_dt,
               "mean_dE": b.mean_dE, "sigma_dt": b.sigma_dt, "sigma_dE": b.sigma_dE,
               "epsn_rms_l": b.epsn_rms_l}
        if self.phase_loop is not None:
            rp, pl = self.rf_params, self.phase_loop
            row.update(PL_omegaRF=rp.host_array("omega_rf")[0, turn_for_rf],
                       PL_phiRF=rp.host_array("phi_rf")[0, turn_for_rf],
                       PL_bunch_phase=pl.phi_beam, PL_phase_corr=pl.dphi,
                       PL_omegaRF_corr=pl.domega_rf, SL_dphiRF=rp.dphi_rf[0],
                       RL_drho=pl.drho)
        return row

    def init_buffer(self):
        """monitors.py:294-323."""
        series = self._SERIES + (self._PL_SERIES if self.phase_loop is not None else ())
        for _, attr, _ in series:
            setattr(self, attr, np.zeros(self.buffer_time))

    def write_buffer(self):
        """monitors.py:325-351: the buffer row of i_turn."""
        i = self.i_turn % self.buffer_time
        row = self._row(self.i_turn)
        series = self._SERIES + (self._PL_SERIES if self.phase_loop is not None else ())
        for name, attr, _ in series:
            getattr(self, attr)[i] = row[name]

    def write_data(self, h5group, dims):
        """monitors.py:353-460: the buffer into rows [i_turn - buffer_time, i_turn)."""
        i1 = self.i_turn - self.buffer_time
        i2 = self.i_turn
        series = self._SERIES + (self._PL_SERIES if self.phase_loop is not None else ())
        for name, attr, _ in series:
            self.h5file.set(name, slice(i1, i2), getattr(self, attr)[:])

    def open(self):
        self.h5file.open()

    def close(self):
        self.h5file.close()

    # -- fused turn ------------------------------------------------------------
    def record_moments(self, rows: np.ndarray):
        """Rows {n, mean_dt, sigma_dt, mean_dE, sigma_dE} of consecutive turns
        from the fused turn's statistics, recorded as track() would after each
        turn (no phase loop: its per-turn series are not kept by the engine)."""
        if self.phase_loop is not None:
            raise NotImplementedError("a monitor with a phase loop records object by object")
        b = self.beam
        try:
            for r in np.asarray(rows, dtype=np.float64):
                self._alive = int(r[0])
                b.mean_dt, b.sigma_dt, b.mean_dE, b.sigma_dE = (float(r[1]), float(r[2]),
                                                                float(r[3]), float(r[4]))
                b.epsn_rms_l = np.pi * b.sigma_dE * b.sigma_dt
                self._record()
        finally:
            self._alive = None
