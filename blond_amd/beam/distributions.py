"""Bunch generation (blond/beam/distributions.py:1194-1279, 1367-1430).

Setup-time host code: draws with numpy's PCG64 ``default_rng`` exactly as the
reference (seed for dt, seed+1 for dE, the standard normal cast to real_t
before scaling), so the same seed gives the same macroparticles.
"""
from __future__ import annotations

import warnings

import numpy as np

from ..utils import bmath as bm


def _get_dE_from_dt(ring, rf_station, dt_amplitude: float) -> float:
    """distributions.py:1367-1430: the dE amplitude matched to a dt
    amplitude in the bucket of the first RF system, from the separatrix
    relation with the synchronous phase; only the first section and the
    first RF system enter."""
    if ring.n_sections > 1:
        warnings.warn("bigaussian(): only the first of several ring sections is used for the "
                      "matched energy spread")
    if rf_station.n_rf > 1:
        warnings.warn("bigaussian(): only the first RF system is used for the matched energy "
                      "spread (higher harmonics ignored)")
    t = rf_station.counter[0]
    voltage = rf_station.charge * rf_station.host_array("voltage")[0, t]
    phi_s = rf_station.phi_s[t]
    phi_b = rf_station.host_array("omega_rf")[0, t] * dt_amplitude + phi_s
    num = voltage * rf_station.energy[t] * rf_station.beta[t] ** 2 * (
        np.cos(phi_b) - np.cos(phi_s) + (phi_b - phi_s) * np.sin(phi_s))
    return np.sqrt(num / (np.pi * rf_station.harmonic[0, t] * np.fabs(rf_station.eta_0[t])))


def bigaussian(ring, rf_station, beam, sigma_dt: float, sigma_dE=None, seed: int = 1234,
               reinsertion: bool = False):
    """distributions.py:1194-1279 (reinsertion=False)."""
    if reinsertion:
        raise NotImplementedError("bigaussian(reinsertion=True) is not supported")
    if sigma_dE is None:
        sigma_dE = _get_dE_from_dt(ring, rf_station, sigma_dt)
    counter = rf_station.counter[0]
    omega_rf = rf_station.host_array("omega_rf")[0, counter]
    phi_s = rf_station.phi_s[counter]
    phi_rf = rf_station.host_array("phi_rf")[0, counter]
    eta0 = rf_station.eta_0[counter]
    if eta0 < 0:
        phi_rf -= np.pi
    beam.sigma_dt = sigma_dt
    beam.sigma_dE = sigma_dE
    rng_dt = np.random.default_rng(seed)
    rng_dE = np.random.default_rng(seed + 1)
    real_t = bm.precision.real_t
    dt = (sigma_dt * rng_dt.normal(size=beam.n_macroparticles).astype(dtype=real_t, order="C",
                                                                       copy=False)
          + (phi_s - phi_rf) / omega_rf)
    dE = sigma_dE * rng_dE.normal(size=beam.n_macroparticles).astype(dtype=real_t, order="C")
    if beam._device == "GPU":
        import torch

        beam.dt = torch.from_numpy(np.ascontiguousarray(dt, dtype=real_t)).to("cuda")
        beam.dE = torch.from_numpy(np.ascontiguousarray(dE, dtype=real_t)).to("cuda")
    else:
        beam.dt = dt
        beam.dE = dE
