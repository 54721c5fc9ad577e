"""Ring: per-turn synchronous programs on the host
(blond/input_parameters/ring.py:217-462).

Host-side setup (numpy), evaluated in the same order as the reference so
that every per-turn scalar the kernels receive is bit-identical.  Supported
synchronous data: scalars (constant programs) and arrays of n_turns+1
values per section, as 'momentum', 'total energy' or 'kinetic energy'
(ring_options.py:718-772).  Time-interpolated (tuple) programs are not part
of the tracking path and raise NotImplementedError.
"""
from __future__ import annotations

import warnings

import numpy as np

from ..constants import c


def convert_data(synchronous_data, mass, charge, synchronous_data_type="momentum",
                 bending_radius=None):
    """ring_options.py:718-772: synchronous data -> momentum [eV/c]."""
    if synchronous_data_type == "momentum":
        return synchronous_data
    if synchronous_data_type == "total energy":
        return np.sqrt(synchronous_data ** 2 - mass ** 2)
    if synchronous_data_type == "kinetic energy":
        return np.sqrt((synchronous_data + mass) ** 2 - mass ** 2)
    if synchronous_data_type == "bending field":
        if bending_radius is None:
            raise RuntimeError("ERROR in Ring: bending_radius is not defined and is required "
                               "to compute momentum")
        return synchronous_data * bending_radius * charge * c
    raise RuntimeError("ERROR in Ring: Synchronous data type not recognized!")


def _reshape(input_data, n_turns, n_sections, to_momentum=False, **conv):
    """ring_options.py:196-345 for scalar and per-turn array inputs."""
    if isinstance(input_data, str):
        raise NotImplementedError("loading programs from files is not supported")
    if isinstance(input_data, tuple):
        raise NotImplementedError("time-interpolated (tuple) programs are not supported; "
                                  "pass n_turns+1 values per section")
    if isinstance(input_data, (float, int)):
        v = float(input_data)
        if to_momentum:
            v = convert_data(v, **conv)
        return v * np.ones((n_sections, n_turns + 1))
    data = np.array(input_data, ndmin=2, dtype=float)
    if to_momentum:
        data = convert_data(data, **conv)
    out = np.zeros((n_sections, n_turns + 1), dtype=float)
    if data.size == n_sections:
        data = data.reshape((n_sections, 1))
    if len(data) != n_sections:
        raise RuntimeError("ERROR in Ring: the input data does not match the number of sections")
    for i in range(len(data)):
        if len(data[i]) == 1:
            out[i] = data[i] * np.ones(n_turns + 1)
        elif len(data[i]) == n_turns + 1:
            out[i] = np.array(data[i])
        else:
            raise RuntimeError(f"ERROR in Ring: The 'input_data' should have {n_turns + 1} "
                               f"entries, but shape is {data.shape}")
    return out


class Ring:
    """Ring and synchronous-particle programs (ring.py:217-357)."""

    def __init__(self, ring_length, alpha_0, synchronous_data, particle, n_turns=1,
                 synchronous_data_type="momentum", bending_radius=None, n_sections=1,
                 alpha_1=None, alpha_2=None, ring_options=None):
        if ring_options is not None:
            raise NotImplementedError("RingOptions (ramp interpolation) is not supported")
        self.n_turns = int(n_turns)
        self.n_sections = int(n_sections)
        self.ring_length = np.array(ring_length, ndmin=1, dtype=float)
        self.ring_circumference = np.sum(self.ring_length)
        self.ring_radius = self.ring_circumference / (2 * np.pi)
        self.bending_radius = float(bending_radius) if bending_radius is not None else None
        if self.n_sections != len(self.ring_length):
            raise RuntimeError("ERROR in Ring: Number of sections and ring length size do not match!")
        self.particle = particle
        self.momentum = _reshape(synchronous_data, self.n_turns, self.n_sections, True,
                                 mass=particle.mass, charge=particle.charge,
                                 synchronous_data_type=synchronous_data_type,
                                 bending_radius=self.bending_radius)
        if self.momentum.shape[1] != self.n_turns + 1:
            self.n_turns = self.momentum.shape[1] - 1
            warnings.warn("WARNING in Ring: The number of turns for the simulation was changed "
                          "by passing a momentum program.", stacklevel=2)
        # ring.py:288-306
        self.beta = np.sqrt(1 / (1 + (self.particle.mass / self.momentum) ** 2))
        self.gamma = np.sqrt(1 + (self.momentum / self.particle.mass) ** 2)
        self.energy = np.sqrt(self.momentum ** 2 + self.particle.mass ** 2)
        self.kin_energy = np.sqrt(self.momentum ** 2 + self.particle.mass ** 2) - self.particle.mass
        self.t_rev = np.dot(self.ring_length, 1 / (self.beta * c))
        self.cycle_time = np.cumsum(self.t_rev)
        self.f_rev = 1 / self.t_rev
        self.omega_rev = 2 * np.pi * self.f_rev
        # ring.py:309-317
        if self.n_sections == 1:
            self.delta_E = np.diff(self.energy, axis=1)
        else:
            self.delta_E = np.diff(self.energy.flatten(order="F"))[self.n_sections - 1:].reshape(
                (self.n_sections, self.n_turns))
        self.alpha_0 = _reshape(alpha_0, self.n_turns, self.n_sections)
        self.alpha_order = 0
        if alpha_1 is not None:
            self.alpha_1 = _reshape(alpha_1, self.n_turns, self.n_sections)
            self.alpha_order = 1
        else:
            self.alpha_1 = np.zeros(self.alpha_0.shape)
        if alpha_2 is not None:
            self.alpha_2 = _reshape(alpha_2, self.n_turns, self.n_sections)
            self.alpha_order = 2
        else:
            self.alpha_2 = np.zeros(self.alpha_0.shape)
        self.eta_generation()

    def eta_generation(self):
        """ring.py:403-462: slippage factors eta_0..eta_2."""
        for i in range(self.alpha_order + 1):
            getattr(self, f"_eta{i}")()
        for i in range(self.alpha_order + 1, 3):
            setattr(self, f"eta_{i}", np.zeros([self.n_sections, self.n_turns + 1]))

    def _eta0(self):
        self.eta_0 = np.empty([self.n_sections, self.n_turns + 1])
        for i in range(self.n_sections):
            self.eta_0[i] = self.alpha_0[i] - self.gamma[i] ** (-2.0)

    def _eta1(self):
        self.eta_1 = np.empty([self.n_sections, self.n_turns + 1])
        for i in range(self.n_sections):
            self.eta_1[i] = (3 * self.beta[i] ** 2 / (2 * self.gamma[i] ** 2) + self.alpha_1[i]
                             - self.alpha_0[i] * self.eta_0[i])

    def _eta2(self):
        self.eta_2 = np.empty([self.n_sections, self.n_turns + 1])
        for i in range(self.n_sections):
            self.eta_2[i] = (-(self.beta[i] ** 2) * (5 * self.beta[i] ** 2 - 1)
                             / (2 * self.gamma[i] ** 2) + self.alpha_2[i]
                             - 2 * self.alpha_0[i] * self.alpha_1[i]
                             + self.alpha_1[i] / self.gamma[i] ** 2
                             + self.alpha_0[i] ** 2 * self.eta_0[i]
                             - 3 * self.beta[i] ** 2 * self.alpha_0[i] / (2 * self.gamma[i] ** 2))
