"""Beam phase loop (blond/llrf/beam_feedback.py:35-622).

``BeamFeedback`` keeps the reference's configuration, state and per-machine
transfer functions.  Object by object, the beam phase of the profile is a
HIP reduction (``bm.beam_phase``, beam_phase.cpp:6-39) and the loop's scalar
recurrences run on the host as in the reference; in the fused turn
(FusedTracker) the whole loop of the LHC, LHC_F and PSB machines runs on
the device, one workgroup per turn.  SPS_F needs ``bm.trapz``, which the
reference's backends do not define (bmath_backends.py:87, 421), so it raises
the same AttributeError; SPS_RL runs object by object.
"""
from __future__ import annotations

import numpy as np

from ..utils import bmath as bm


class BeamFeedback:
    """beam_feedback.py:35-233."""

    def __init__(self, ring, rf_station, profile, configuration, PhaseNoise=None,
                 LHCNoiseFB=None, delay: int = 0):
        if PhaseNoise is not None or LHCNoiseFB is not None:
            raise NotImplementedError("RF phase noise is outside the tracking path")
        self.ring = ring
        self.rf_station = rf_station
        self.profile = profile
        self.config = configuration
        self.delay = delay
        self.machine = self.config.get("machine", "LHC")
        self.alpha = self.config.get("window_coefficient", 0.0)
        self.time_offset = self.config.get("time_offset", None)
        try:
            self.gain = self.config["PL_gain"]
        except Exception:
            raise RuntimeError("You need to specify the Phase Loop gain! Aborting")
        if self.machine == "LHC":
            self.gain2 = self.config.get("SL_gain", 0.0)
            self.lhc_y = 0
            if self.gain2 != 0:
                self.lhc_a = 5.25 - self.rf_station.omega_s0 / (np.pi * 40.0)
                self.lhc_t = ((2 * np.pi * self.rf_station.Q_s * np.sqrt(self.lhc_a))
                              / np.sqrt(1 + self.gain / self.gain2
                                        * np.sqrt((1 + 1 / self.lhc_a) / (1 + self.lhc_a))))
            else:
                self.lhc_a = np.zeros(self.rf_station.n_turns + 1)
                self.lhc_t = np.zeros(self.rf_station.n_turns + 1)
        elif self.machine == "LHC_F":
            self.gain2 = self.config.get("FL_gain", 0.0)
        elif self.machine == "SPS_RL":
            self.gain2 = self.config.get("RL_gain", 0.0)
            self.sample_dE = self.config.get("sample_dE", 1)
        elif self.machine == "SPS_F":
            self.gain2 = self.config.get("FL_gain", 0.0)
        elif self.machine == "PSB":
            self.gain = self.gain * np.ones(ring.n_turns + 1)
            self.gain2 = self.config.get("RL_gain", [0.0, 0.0])
            self.gain2[0] = self.gain2[0] * np.ones(ring.n_turns + 1)
            self.gain2[1] = self.gain2[1] * np.ones(ring.n_turns + 1)
            self.dt = self.config.get("period", 10.0e-6)
            self.PL_counter = 0
            self.on_time = np.array([])
            self.precalculate_time(ring)
            self.coefficients = self.config.get(
                "coefficients", [0.999019, -0.999019, 0.0, 1.0, -0.998038, 0.0])
            self.dphi_sum = 0.0
            self.dphi_av = 0.0
            self.dphi_av_prev = 0.0
            self.dR_over_R_prev = 0.0
            self.domega_PL = 0.0
            self.domega_RL = 0.0
            self.dR_over_R = 0
        self.drho = 0.0
        self.domega_rf = 0.0
        self.phi_beam = 0.0
        self.dphi = 0.0
        self.reference = 0.0
        self.RFnoise = None
        self.noiseFB = None
        self._device = "CPU"

    # ------------------------------------------------------------------
    def track(self):
        """beam_feedback.py:236-251: the machine's correction, then the RF
        frequency of all systems for the next turn."""
        getattr(self, self.machine)()
        rp = self.rf_station
        counter = rp.counter[0] + 1
        omega = rp.host_array("omega_rf")
        omega[:, counter] += self.domega_rf * rp.harmonic[:, counter] / rp.harmonic[0, counter]
        if rp._device == "GPU":
            import torch

            rp.omega_rf[:, counter] = torch.from_numpy(
                np.ascontiguousarray(omega[:, counter])).to(rp.omega_rf.device)

    def precalculate_time(self, ring):
        """beam_feedback.py:253-273: the turns on which the PSB loop acts."""
        if self.dt > 0:
            n = self.delay + 1
            while n < ring.t_rev.size:
                summa = 0
                while summa < self.dt:
                    try:
                        summa += ring.t_rev[n]
                        n += 1
                    except Exception:
                        self.on_time = np.append(self.on_time, 0)
                        return
                self.on_time = np.append(self.on_time, n - 1)
        else:
            self.on_time = np.arange(ring.t_rev.size)

    def _profile_arrays(self):
        p = self.profile
        bc, prof = p.bin_centers, p.n_macroparticles
        if self.time_offset is not None:
            first = self.first_bin()
            bc, prof = bc[first:], prof[first:]
        return bc, prof

    def first_bin(self) -> int:
        """First bin with bin_centers >= time_offset (beam_feedback.py:296-298)."""
        if self.time_offset is None:
            return 0
        bc = self.profile.bin_centers
        bc = bc.detach().cpu().numpy() if hasattr(bc, "detach") else np.asarray(bc)
        return int(np.searchsorted(bc, self.time_offset, side="left"))

    def beam_phase(self):
        """beam_feedback.py:275-307: phi_beam = arctan(coeff) + pi."""
        turn = self.rf_station.counter[0]
        omega_rf = self.rf_station.host_array("omega_rf")[0, turn]
        phi_rf = self.rf_station.host_array("phi_rf")[0, turn]
        bc, prof = self._profile_arrays()
        coeff = bm.beam_phase(bc, prof, self.alpha, omega_rf, phi_rf, self.profile.bin_size)
        self.phi_beam = np.arctan(coeff) + np.pi

    def beam_phase_sharpWindow(self):
        """beam_feedback.py:309-347.  The reference calls bm.trapz and
        bm.logical_and, which its backends do not define: AttributeError."""
        turn = self.rf_station.counter[0]
        omega_rf = self.rf_station.host_array("omega_rf")[0, turn]
        if self.alpha != 0.0:
            bm.logical_and((self.time_offset - np.pi / omega_rf) <= self.profile.bin_centers,
                           self.profile.bin_centers
                           <= (-1 / self.alpha + self.time_offset - 2 * np.pi / omega_rf))
        bm.trapz(None)

    def phase_difference(self):
        """beam_feedback.py:349-367 (no RF phase noise)."""
        counter = self.rf_station.counter[0]
        self.dphi = self.phi_beam - self.rf_station.phi_s[counter]

    def radial_difference(self):
        """beam_feedback.py:369-388: relative radial displacement from the
        mean energy offset of every sample_dE-th particle."""
        counter = self.rf_station.counter[0]
        self.average_dE = bm.strided_mean(self.profile.beam.dE, self.sample_dE)
        self.drho = (self.ring.alpha_0[0, counter] * self.ring.ring_radius * self.average_dE
                     / (self.ring.beta[0, counter] ** 2.0 * self.ring.energy[0, counter]))

    def radial_steering_from_freq(self):
        """beam_feedback.py:390-412.  The reference adds to
        rf_station.dphi_rf_steering, which nothing defines: AttributeError."""
        counter = self.rf_station.counter[0]
        rp = self.rf_station
        self.radial_steering_domega_rf = (-rp.omega_rf_d[0, counter] * rp.eta_0[counter]
                                          / self.ring.alpha_0[0, counter] * self.reference
                                          / self.ring.ring_radius)
        omega = rp.host_array("omega_rf")
        omega[:, counter] += (self.radial_steering_domega_rf * rp.harmonic[:, counter]
                              / rp.harmonic[0, counter])
        rp.dphi_rf_steering += 0  # raises as in the reference

    def LHC_F(self):
        """beam_feedback.py:414-441."""
        counter = self.rf_station.counter[0]
        self.beam_phase()
        self.phase_difference()
        omega = self.rf_station.host_array("omega_rf")
        self.domega_rf = -self.gain * self.dphi - self.gain2 * (
            omega[0, counter] - self.rf_station.omega_rf_d[0, counter] + self.reference)

    def SPS_F(self):
        """beam_feedback.py:443-462."""
        counter = self.rf_station.counter[0]
        self.beam_phase_sharpWindow()
        self.phase_difference()
        omega = self.rf_station.host_array("omega_rf")
        self.domega_dphi = -self.gain * self.dphi
        self.domega_df = -self.gain2 * (omega[0, counter] - self.rf_station.omega_rf_d[0, counter])
        self.domega_rf = self.domega_dphi + self.domega_df

    def SPS_RL(self):
        """beam_feedback.py:464-512."""
        counter = self.rf_station.counter[0]
        if self.reference != 0:
            self.radial_steering_from_freq()
        self.beam_phase()
        self.phase_difference()
        self.radial_difference()
        self.domega_dphi = -self.gain * self.dphi
        self.domega_dR = (-np.sign(self.rf_station.eta_0[counter]) * self.gain2
                          * (self.reference - self.drho) / self.ring.ring_radius)
        self.domega_rf = self.domega_dphi + self.domega_dR

    def LHC(self):
        """beam_feedback.py:514-561."""
        counter = self.rf_station.counter[0]
        dphi_rf = self.rf_station.dphi_rf[0]
        self.beam_phase()
        self.phase_difference()
        self.domega_rf = -self.gain * self.dphi - self.gain2 * (
            self.lhc_y + self.lhc_a[counter] * (dphi_rf + self.reference))
        self.lhc_y = (1 - self.lhc_t[counter]) * self.lhc_y + (
            1 - self.lhc_a[counter]) * self.lhc_t[counter] * (dphi_rf + self.reference)

    def PSB(self):
        """beam_feedback.py:563-622."""
        counter = self.rf_station.counter[0]
        self.beam_phase()
        self.phase_difference()
        self.dphi_sum += self.dphi
        if counter == self.on_time[self.PL_counter] and counter >= self.delay:
            self.dphi_av = self.dphi_sum / (self.on_time[self.PL_counter]
                                            - self.on_time[self.PL_counter - 1])
            self.domega_PL = 0.99803799 * self.domega_PL + self.gain[counter] * (
                0.99901903 * self.dphi_av - 0.99901003 * self.dphi_av_prev)
            self.dphi_av_prev = self.dphi_av
            self.dphi_sum = 0.0
            omega = self.rf_station.host_array("omega_rf")
            self.dR_over_R = (omega[0, counter] - self.rf_station.omega_rf_d[0, counter]) / (
                self.rf_station.omega_rf_d[0, counter]
                * (1.0 / (self.ring.alpha_0[0, counter] * self.rf_station.gamma[counter] ** 2)
                   - 1.0))
            self.domega_RL = (self.domega_RL + self.gain2[0][counter]
                              * (self.dR_over_R - self.dR_over_R_prev)
                              + self.gain2[1][counter] * self.dR_over_R)
            self.dR_over_R_prev = self.dR_over_R
            self.PL_counter += 1
        self.domega_rf = -self.domega_PL - self.domega_RL

    # ------------------------------------------------------------------
    # device loop state (FusedTracker)
    _STATE = ("lhc_y", "dphi_sum", "dphi_av", "dphi_av_prev", "domega_PL", "domega_RL",
              "dR_over_R", "dR_over_R_prev", "PL_counter", "domega_rf", "phi_beam", "dphi")

    def device_programs(self, n_turns: int):
        """(machine id, per-turn program [n_turns+1][8], on_time) for the
        device loop: gain, gain2 (P), gain2 (I), lhc_a, lhc_t, phi_s,
        alpha_0, gamma per turn."""
        ids = {"LHC": 0, "LHC_F": 1, "PSB": 2}
        if self.machine not in ids:
            raise NotImplementedError(f"the fused turn runs the LHC, LHC_F and PSB loops, "
                                      f"not {self.machine}")
        nt = n_turns + 1
        prog = np.zeros((nt, 8), dtype=np.float64)
        rp = self.rf_station
        if self.machine == "PSB":
            prog[:, 0] = self.gain[:nt]
            prog[:, 1] = self.gain2[0][:nt]
            prog[:, 2] = self.gain2[1][:nt]
        else:
            prog[:, 0] = self.gain
            prog[:, 1] = self.gain2
        if self.machine == "LHC":
            prog[:, 3] = np.asarray(self.lhc_a, dtype=np.float64)[:nt]
            prog[:, 4] = np.asarray(self.lhc_t, dtype=np.float64)[:nt]
        prog[:, 5] = np.asarray(rp.phi_s, dtype=np.float64)[:nt]
        prog[:, 6] = np.asarray(self.ring.alpha_0[0], dtype=np.float64)[:nt]
        prog[:, 7] = np.asarray(rp.gamma, dtype=np.float64)[:nt]
        on_time = np.asarray(getattr(self, "on_time", np.zeros(0)), dtype=np.float64)
        return ids[self.machine], np.ascontiguousarray(prog), np.ascontiguousarray(on_time)

    def device_state(self, n_rf: int) -> np.ndarray:
        st = np.zeros(16 + n_rf, dtype=np.float64)
        for k, name in enumerate(self._STATE):
            st[k] = float(getattr(self, name, 0.0))
        st[16:] = np.asarray(self.rf_station.dphi_rf, dtype=np.float64)
        return st

    def set_device_state(self, st: np.ndarray):
        for k, name in enumerate(self._STATE):
            if hasattr(self, name) or name in ("domega_rf", "phi_beam", "dphi"):
                v = float(st[k])
                setattr(self, name, int(v) if name == "PL_counter" else v)
        self.rf_station.dphi_rf[:] = st[16:]

    def to_gpu(self, recursive=True):
        self._device = "GPU"

    def to_cpu(self, recursive=True):
        self._device = "CPU"
