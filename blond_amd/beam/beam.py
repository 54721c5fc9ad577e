"""Particles and the macroparticle beam (blond/beam/beam.py:40-458).

``Beam.dt`` / ``Beam.dE`` are structure-of-arrays coordinates: numpy arrays
until ``to_gpu()``, device tensors (HBM) afterwards.
"""
from __future__ import annotations

import numpy as np

from ..constants import c, e, epsilon_0, hbar, m_e, m_mu, m_p
from ..utils import bmath as bm


class Particle:
    """beam.py:40-117: rest mass [eV] and charge [e]."""

    def __init__(self, user_mass: float, user_charge: float, user_decay_rate: float = 0):
        if user_mass > 0.0:
            self.mass = float(user_mass)
            self.charge = float(user_charge)
        else:
            raise RuntimeError("ERROR: Particle mass not recognized!")
        if user_decay_rate >= 0.0:
            self.decay_rate = float(user_decay_rate)
        else:
            raise RuntimeError("ERROR: Invalid particle decay rate!")
        self.radius_cl = 0.25 / (np.pi * epsilon_0) * e ** 2 * self.charge ** 2 / (self.mass * e)
        self.c_gamma = 4 * np.pi / 3 * self.radius_cl / self.mass ** 3
        self.c_q = 55.0 / (32.0 * np.sqrt(3.0)) * hbar * c / (self.mass * e)


class Proton(Particle):
    def __init__(self):
        Particle.__init__(self, m_p * c ** 2 / e, 1)


class Electron(Particle):
    def __init__(self):
        Particle.__init__(self, m_e * c ** 2 / e, -1)


class Positron(Particle):
    def __init__(self):
        Particle.__init__(self, m_e * c ** 2 / e, 1)


class MuPlus(Particle):
    def __init__(self):
        Particle.__init__(self, m_mu * c ** 2 / e, 1, float(1 / 2.1969811e-6))


class MuMinus(Particle):
    def __init__(self):
        Particle.__init__(self, m_mu * c ** 2 / e, -1, float(1 / 2.1969811e-6))


class Beam:
    """beam.py:151-280: coordinates, intensity and statistics."""

    def __init__(self, ring, n_macroparticles: int, intensity: float, dt=None, dE=None):
        self.particle = ring.particle
        self.beta = ring.beta[0][0]
        self.gamma = ring.gamma[0][0]
        self.energy = ring.energy[0][0]
        self.momentum = ring.momentum[0][0]
        if dt is None:
            self.dt = np.zeros([int(n_macroparticles)], dtype=bm.precision.real_t)
        else:
            assert n_macroparticles == len(dt)
            self.dt = np.ascontiguousarray(dt)
        if dE is None:
            self.dE = np.zeros([int(n_macroparticles)], dtype=bm.precision.real_t)
        else:
            assert n_macroparticles == len(dE)
            self.dE = np.ascontiguousarray(dE)
        self.mean_dt = 0.0
        self.mean_dE = 0.0
        self.sigma_dt = 0.0
        self.sigma_dE = 0.0
        self.intensity = float(intensity)
        self.n_macroparticles = int(n_macroparticles)
        self.ratio = self.intensity / self.n_macroparticles
        self.id = np.arange(1, self.n_macroparticles + 1, dtype=int)
        self.epsn_rms_l = 0.0
        self.n_macroparticles_eliminated = 0
        self._device = "CPU"

    @property
    def n_macroparticles_alive(self) -> int:
        if self._device == "GPU":
            return int((self.id != 0).sum().item())
        return int(np.count_nonzero(self.id))

    @property
    def n_macroparticles_not_alive(self) -> int:
        return self.n_macroparticles - self.n_macroparticles_alive

    def statistics(self):
        """beam.py:430-458: mean and r.m.s. (ddof=0) of the alive particles."""
        if self._device == "GPU":
            _, self.mean_dt, self.sigma_dt, self.mean_dE, self.sigma_dE = bm.beam_statistics(
                self.dt, self.dE, self.id)
        else:
            idx = np.nonzero(self.id)[0]
            self.mean_dt = float(np.mean(self.dt[idx]))
            self.sigma_dt = float(np.std(self.dt[idx]))
            self.mean_dE = float(np.mean(self.dE[idx]))
            self.sigma_dE = float(np.std(self.dE[idx]))
        self.epsn_rms_l = np.pi * self.sigma_dE * self.sigma_dt

    def to_gpu(self, recursive=True):
        """beam.py:817-831: move coordinates to HBM as device tensors."""
        if self._device == "GPU":
            return
        import torch

        dt = bm.precision.torch_dtype
        self.dt = torch.from_numpy(np.ascontiguousarray(self.dt)).to(device="cuda", dtype=dt)
        self.dE = torch.from_numpy(np.ascontiguousarray(self.dE)).to(device="cuda", dtype=dt)
        self.id = torch.from_numpy(np.ascontiguousarray(self.id)).to(device="cuda")
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        self.dt = self.dt.cpu().numpy()
        self.dE = self.dE.cpu().numpy()
        self.id = self.id.cpu().numpy()
        self._device = "CPU"
