"""RF station programs (blond/input_parameters/rf_parameters.py:239-454).

Per-turn arrays of shape [n_rf, n_turns+1]; ``voltage``, ``phi_rf`` and
``omega_rf`` are Fortran-ordered so that column ``[:, turn]`` is contiguous
(rf_parameters.py:317-328), on the host and -- after ``to_gpu()`` -- on the
device.
"""
from __future__ import annotations

import numpy as np

from ..beam.beam import Proton
from ..constants import c
from ..utils import bmath as bm


def _reshape_rf(input_data, n_turns, n_rf):
    """rf_parameters_options.py:141-229 for scalar / per-turn array inputs."""
    if isinstance(input_data, str):
        raise NotImplementedError("loading RF programs from files is not supported")
    if isinstance(input_data, tuple):
        raise NotImplementedError("time-interpolated (tuple) RF programs are not supported")
    if isinstance(input_data, (float, int)):
        return input_data * np.ones((n_rf, n_turns + 1))
    data = np.array(input_data, ndmin=2, dtype=float)
    if data.size == n_rf:  # one constant value per system
        data = data.reshape((n_rf, 1))
    out = np.zeros((n_rf, n_turns + 1))
    if len(data) != n_rf:
        raise RuntimeError("ERROR in RFStation: the input data does not match the number of "
                           "rf harmonics")
    for i in range(n_rf):
        row = np.atleast_1d(data[i])
        if len(row) == 1:
            out[i] = row[0] * np.ones(n_turns + 1)
        elif len(row) == n_turns + 1:
            out[i] = row
        else:
            raise RuntimeError(f"ERROR in RFStation: expected 1 or {n_turns + 1} values per system")
    return out


class RFStation:
    """RF systems at one station (rf_parameters.py:239-454)."""

    def __init__(self, ring, harmonic, voltage, phi_rf_d, n_rf=1, section_index=1, omega_rf=None,
                 phi_noise=None, phi_modulation=None, rf_station_options=None):
        if phi_noise is not None or phi_modulation is not None or rf_station_options is not None:
            raise NotImplementedError("phase noise, phase modulation and RFStationOptions are "
                                      "outside the tracking path")
        self.counter = [int(0)]
        self.section_index = int(section_index - 1)
        if self.section_index < 0 or self.section_index > ring.n_sections - 1:
            raise RuntimeError("ERROR in RFStation: section_index out of allowed range!")
        self.n_rf = int(n_rf)
        self.particle = ring.particle
        self.n_turns = ring.n_turns
        self.ring_circumference = ring.ring_circumference
        self.section_length = ring.ring_length[self.section_index]
        self.length_ratio = float(self.section_length / self.ring_circumference)
        self.t_rev = ring.t_rev
        self.momentum = ring.momentum[self.section_index]
        self.beta = ring.beta[self.section_index]
        self.gamma = ring.gamma[self.section_index]
        self.energy = ring.energy[self.section_index]
        self.delta_E = ring.delta_E[self.section_index]
        self.alpha_order = ring.alpha_order
        self.charge = self.particle.charge
        for i in range(3):
            setattr(self, f"eta_{i}", getattr(ring, f"eta_{i}")[self.section_index])
            setattr(self, f"alpha_{i}", getattr(ring, f"alpha_{i}")[self.section_index])
        self.sign_eta_0 = np.sign(self.eta_0)
        self.harmonic = _reshape_rf(harmonic, self.n_turns, self.n_rf).astype(
            bm.precision.real_t, order="C", copy=False)
        self.voltage = _reshape_rf(voltage, self.n_turns, self.n_rf).astype(
            bm.precision.real_t, order="F", copy=False)
        self.empty = bool(np.sum(self.voltage) == 0)
        self.phi_rf_d = _reshape_rf(phi_rf_d, self.n_turns, self.n_rf)
        if omega_rf is None:
            # rf_parameters.py:366-373
            self.omega_rf_d = 2.0 * np.pi * self.beta * c * self.harmonic / self.ring_circumference
        else:
            self.omega_rf_d = _reshape_rf(omega_rf, self.n_turns, self.n_rf)
        self.omega_rf_d = self.omega_rf_d.astype(bm.precision.real_t, order="C", copy=False)
        self.phi_noise = None
        self.phi_modulation = None
        self.phi_rf = np.array(self.phi_rf_d).astype(bm.precision.real_t, order="F")
        self.dphi_rf = np.zeros(self.n_rf).astype(bm.precision.real_t, order="F")
        self.omega_rf = np.array(self.omega_rf_d).astype(bm.precision.real_t, order="F")
        self.t_rf = 2 * np.pi / self.omega_rf
        if not self.empty:
            self.phi_s = calculate_phi_s(self, self.particle)
            self.Q_s = calculate_Q_s(self, self.particle)
            self.omega_s0 = self.Q_s * ring.omega_rev
        self._device = "CPU"
        # host mirrors used by the per-turn phase bookkeeping after to_gpu()
        self._host = None

    # -- device residency (rf_parameters.py to_gpu) -----------------------------
    def to_gpu(self, recursive=True):
        if self._device == "GPU":
            return
        import torch

        self._host = {"voltage": self.voltage, "phi_rf": self.phi_rf, "omega_rf": self.omega_rf}
        dt = bm.precision.torch_dtype
        for name in ("voltage", "phi_rf", "omega_rf"):
            arr = np.asfortranarray(getattr(self, name))
            # keep Fortran strides: column [:, turn] stays contiguous on the device
            t = torch.from_numpy(arr).to(device="cuda", dtype=dt)
            setattr(self, name, t)
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        for name in ("voltage", "phi_rf", "omega_rf"):
            setattr(self, name, np.asfortranarray(getattr(self, name).cpu().numpy()))
        self._host = None
        self._device = "CPU"

    def host_array(self, name):
        """Host copy of voltage / phi_rf / omega_rf whichever the residency."""
        if self._device == "GPU":
            return self._host[name]
        return getattr(self, name)


def calculate_Q_s(rf_station, particle=None):
    """rf_parameters.py:610-636 (single harmonic, no intensity effects)."""
    if particle is None:
        particle = Proton()
    return np.sqrt(rf_station.harmonic[0] * np.abs(particle.charge) * rf_station.voltage[0]
                   * np.abs(rf_station.eta_0 * np.cos(rf_station.phi_s))
                   / (2 * np.pi * rf_station.beta ** 2 * rf_station.energy))


def calculate_phi_s(rf_station, particle=None, accelerating_systems="as_single"):
    """rf_parameters.py:640-721, the default 'as_single' branch: analytic
    synchronous phase from the energy gain per turn and the main voltage,
    shifted by pi according to the sign of eta_0 and of the charge."""
    if particle is None:
        particle = Proton()
    if accelerating_systems != "as_single":
        raise NotImplementedError("only accelerating_systems='as_single' is supported")
    eta0 = rf_station.eta_0
    denergy = np.append(rf_station.delta_E, rf_station.delta_E[-1])
    acceleration_ratio = denergy / (particle.charge * rf_station.voltage[0, :])
    acceleration_test = ((acceleration_ratio > -1) & (acceleration_ratio < 1)) == 0
    if np.count_nonzero(acceleration_test) > 0:
        print("WARNING in calculate_phi_s(): acceleration is not possible (momentum increment "
              "is too big or voltage too low) at index " + str(acceleration_test.nonzero()[0]))
    phi_s = np.arcsin(acceleration_ratio)
    eta0_middle_points = (eta0[1:] + eta0[:-1]) / 2
    eta0_middle_points = np.append(eta0_middle_points, eta0[-1])
    index = np.where(eta0_middle_points > 0)[0]
    index_below = np.where(eta0_middle_points < 0)[0]
    phi_s[index] = (np.heaviside(np.sign(particle.charge), 0) * np.pi - phi_s[index]) % (2 * np.pi)
    phi_s[index_below] = (np.heaviside(np.sign(particle.charge), 0) * np.pi
                          + phi_s[index_below]) % (2 * np.pi)
    return phi_s
