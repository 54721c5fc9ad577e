"""RF kick + drift per turn (blond/trackers/tracker.py:256-615).

``RingAndRFTracker.track()`` calls ``bm.kick`` (or ``bm.linear_interp_kick``
with interpolation) then ``bm.drift`` exactly as the reference; both are
HIP kernels.  The per-turn scalars come from the host programs (no device
synchronisation per turn).  A beam phase loop (BeamFeedback) runs before the
phase update as in the reference; phase noise/modulation, cavity and noise
feedbacks, periodicity and xsuite coupling are outside the tracking path and
raise NotImplementedError.
"""
from __future__ import annotations

import numpy as np

from ..utils import bmath as bm


class RingAndRFTracker:
    """tracker.py:256-397 (constructor checks) and 521-615 (track)."""

    def __init__(self, rf_station, beam, solver="simple", beam_feedback=None, noise_feedback=None,
                 cavity_feedback=None, periodicity=False, interpolation=False, with_xsuite=False,
                 profile=None, total_induced_voltage=None):
        if noise_feedback is not None or cavity_feedback is not None:
            raise NotImplementedError("noise and cavity feedback loops are outside the tracking path")
        if periodicity:
            raise NotImplementedError("periodicity=True is not supported")
        if with_xsuite:
            raise NotImplementedError("xsuite coupling is not supported")
        self.rf_params = rf_station
        self.counter = rf_station.counter
        self.acceleration_kick = -rf_station.delta_E
        self.beam = beam
        self.solver = str(solver)
        if self.solver not in ["simple", "exact", "legacy"]:
            raise RuntimeError("ERROR in RingAndRFTracker: Choice of longitudinal solver not "
                               "recognised!")
        if self.rf_params.alpha_order > 1:
            self.solver = "exact"
        self.beamFB = beam_feedback
        self.noiseFB = None
        self.cavityFB = None
        self.periodicity = False
        self.interpolation = bool(interpolation)
        self.profile = profile
        self.totalInducedVoltage = total_induced_voltage
        if self.interpolation and self.profile is None:
            raise RuntimeError("ERROR in RingAndRFTracker: Please specify a Profile object to "
                               "use the interpolation option")
        self._device = "CPU"

    def kick(self, beam_dt, beam_dE, index: int):
        """tracker.py:399-422."""
        bm.kick(beam_dt, beam_dE, self.rf_params.voltage[:, index],
                self.rf_params.omega_rf[:, index], self.rf_params.phi_rf[:, index],
                self.rf_params.particle.charge, self.rf_params.n_rf,
                self.acceleration_kick[index])

    def drift(self, beam_dt, beam_dE, index: int):
        """tracker.py:424-468."""
        rp = self.rf_params
        bm.drift(beam_dt, beam_dE, self.solver, rp.t_rev[index], rp.length_ratio, rp.alpha_order,
                 rp.eta_0[index], rp.eta_1[index], rp.eta_2[index], rp.alpha_0[index],
                 rp.alpha_1[index], rp.alpha_2[index], rp.beta[index], rp.energy[index])

    def rf_voltage_calculation(self):
        """tracker.py:470-519 (no cavity feedback)."""
        t = self.counter[0]
        rp = self.rf_params
        self.rf_voltage = bm.rf_volt_comp(rp.voltage[:, t], rp.omega_rf[:, t], rp.phi_rf[:, t],
                                          self.profile.bin_centers)

    def _update_phase(self, turn: int):
        """tracker.py:556-568: accumulated phase offset from frequency offsets
        (zero unless omega_rf differs from the design omega_rf_d)."""
        rp = self.rf_params
        omega = rp.host_array("omega_rf")
        dphi = (2.0 * np.pi * rp.harmonic[:, turn + 1]
                * (omega[:, turn + 1] - rp.omega_rf_d[:, turn + 1]) / rp.omega_rf_d[:, turn + 1])
        rp.dphi_rf += dphi
        if np.any(rp.dphi_rf != 0):
            host_phi = rp.host_array("phi_rf")
            host_phi[:, turn + 1] += rp.dphi_rf
            if rp._device == "GPU":
                import torch

                rp.phi_rf[:, turn + 1] = torch.from_numpy(
                    np.ascontiguousarray(host_phi[:, turn + 1])).to(rp.phi_rf.device)

    def track(self):
        """tracker.py:521-615."""
        turn = self.counter[0]
        if self.beamFB is not None and turn >= self.beamFB.delay:
            self.beamFB.track()
        self._update_phase(turn)
        if self.rf_params.empty is False:
            if self.interpolation:
                self.rf_voltage_calculation()
                if self.totalInducedVoltage is not None:
                    total = self.rf_voltage + self.totalInducedVoltage.induced_voltage
                else:
                    total = self.rf_voltage
                bm.linear_interp_kick(dt=self.beam.dt, dE=self.beam.dE, voltage=total,
                                      bin_centers=self.profile.bin_centers,
                                      charge=self.beam.particle.charge,
                                      acceleration_kick=self.acceleration_kick[turn])
            else:
                self.kick(self.beam.dt, self.beam.dE, turn)
        self.drift(self.beam.dt, self.beam.dE, turn + 1)
        rp = self.rf_params
        self.beam.beta = rp.beta[turn + 1]
        self.beam.gamma = rp.gamma[turn + 1]
        self.beam.energy = rp.energy[turn + 1]
        self.beam.momentum = rp.momentum[turn + 1]
        self.counter[0] += 1

    def to_gpu(self, recursive=True):
        """tracker.py:732-757."""
        if self._device == "GPU":
            return
        if recursive:
            if self.profile is not None:
                self.profile.to_gpu()
            if self.totalInducedVoltage is not None:
                self.totalInducedVoltage.to_gpu()
            if self.beam is not None:
                self.beam.to_gpu()
            self.rf_params.to_gpu()
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        if self._device == "CPU":
            return
        if recursive:
            if self.profile is not None:
                self.profile.to_cpu()
            if self.totalInducedVoltage is not None:
                self.totalInducedVoltage.to_cpu()
            if self.beam is not None:
                self.beam.to_cpu()
            self.rf_params.to_cpu()
        self._device = "CPU"
