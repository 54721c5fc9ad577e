"""Synthetic benchmark workloads (seeded, no external data).

  c1  LHC-like ring (26658.883 m, 1/55.759505^2, 450 GeV/c), h=35640, 6 MV,
      Gaussian bunch 4 sigma = 1 ns, 100 slices over [0, t_rf]; turn =
      tracker, profile.
  c2  PSB-like ring (157.08 m, gamma_t = 4.4, 310 MeV/c), h=1, 8 kV, phi=pi,
      Gaussian bunch 4 sigma = 400 ns (a quarter of the 1.67 us bucket),
      intensity 1e13, 4096 slices over [0, t_rf],
      n_fft = 16384; broadband resonator (5 kOhm, 1 GHz, Q=1) + 20 narrowband
      resonators drawn from default_rng(42); turn = profile, induced
      voltage, tracker.
  c3  c1 ring with a second RF system at 4h, V2/V1 = 0.25 in
      bunch-lengthening phase (phi2 = 0: the two slopes cancel at the bunch
      centre, phi = pi), Gaussian bunch as c1, 65536 slices, n_fft = 131072,
      resonators as c2.
  c4  ensemble of independent bunches in the c2 ring: particle counts
      round(10**U(4,6)), V = U(4,8) kV, phi = pi + U(-0.1, 0.1), each with
      its own 1024-slice profile over [0, t_rf], intensity 1e12, and 5
      narrowband resonators drawn as in c2 (n_fft = 1024).  Seeds: counts,
      V, phi from default_rng(seed); bunch b from bigaussian(seed*1e5 + 2b),
      its resonators from default_rng(seed*1e5 + 5e4 + b).
  feedback workloads (build_feedback_workload): a displaced bunch under a
      beam phase loop; see that function.
  wake workloads (build_wake_workload): a c2-ring bunch (intensity 1e12,
      256 slices over [0, t_rf]) with a multi-turn wake or an inductive
      source; see that function.

Bunch draws follow bigaussian(seed=1): dt from default_rng(1), dE from
default_rng(2).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .beam.beam import Beam, Proton
from .beam.distributions import bigaussian
from .beam.profile import CutOptions, OtherSlicesOptions, Profile
from .constants import e
from .impedances.impedance import (InducedVoltageFreq, InducedVoltageTime, InductiveImpedance,
                                   TotalInducedVoltage)
from .impedances.impedance_sources import Resonators
from .input_parameters.rf_parameters import RFStation
from .llrf.beam_feedback import BeamFeedback
from .input_parameters.ring import Ring
from .trackers.tracker import RingAndRFTracker


C2_SIGMA_DT = 100e-9


def resonator_set(n_narrow: int = 20, seed: int = 42, broadband=True):
    """Broadband (5 kOhm, 1 GHz, Q=1) plus narrowband resonators with
    R = 10**U(3,6), f_r = U(1e6, 2e8), Q = 10**U(2,4), drawn in that order."""
    rng = np.random.default_rng(seed)
    R = 10 ** rng.uniform(3, 6, n_narrow)
    fr = rng.uniform(1e6, 2e8, n_narrow)
    Q = 10 ** rng.uniform(2, 4, n_narrow)
    if broadband:
        R = np.concatenate([[5e3], R])
        fr = np.concatenate([[1e9], fr])
        Q = np.concatenate([[1.0], Q])
    return R, fr, Q


@dataclass
class Workload:
    name: str
    ring: Ring
    rf: RFStation
    beam: Beam
    tracker: RingAndRFTracker
    profile: Profile | None = None
    tiv: TotalInducedVoltage | None = None
    intensity: bool = False
    meta: dict = field(default_factory=dict)
    _fused: object = None

    def map(self):
        """The per-turn map as a BLonD main script runs it.  With an
        interpolating tracker the induced voltage is only summed
        (induced_voltage_sum) and the tracker kicks with it
        (unittests/impedances/test_impedance.py:174-176)."""
        if self.intensity:
            if self.tracker.interpolation:
                return [self.profile, _InducedSum(self.tiv), self.tracker]
            return [self.profile, self.tiv, self.tracker]
        return [self.tracker] + ([self.profile] if self.profile is not None else [])

    def to_gpu(self):
        self.tracker.to_gpu()
        if self.profile is not None:
            self.profile.to_gpu()
        if self.tiv is not None:
            self.tiv.to_gpu()

    def fused(self, n_turns: int = 1, **launch):
        """Track n_turns with the fused engine (created on first use)."""
        from .trackers.fused import FusedTracker

        if self._fused is None:
            self._fused = FusedTracker(self.tracker, self.profile, self.tiv, **launch)
        self._fused.track(n_turns)
        return self._fused

    def snapshot(self) -> dict:
        """Host copies of everything a CPU restatement of the turn needs."""
        rp = self.rf
        st = {
            "name": self.name,
            "intensity": self.intensity,
            "interpolation": bool(self.tracker.interpolation),
            "turn": int(rp.counter[0]),
            "dt": _h(self.beam.dt).copy(),
            "dE": _h(self.beam.dE).copy(),
            "n_rf": rp.n_rf,
            "empty": bool(rp.empty),
            "charge": float(rp.particle.charge),
            "voltage": np.array(rp.host_array("voltage"), dtype=np.float64),
            "omega_rf": np.array(rp.host_array("omega_rf"), dtype=np.float64),
            "phi_rf": np.array(rp.host_array("phi_rf"), dtype=np.float64),
            "acc": np.array(self.tracker.acceleration_kick, dtype=np.float64),
            "solver": self.tracker.solver,
            "t_rev": np.array(rp.t_rev), "length_ratio": float(rp.length_ratio),
            "alpha_order": float(rp.alpha_order),
            "eta_0": np.array(rp.eta_0), "eta_1": np.array(rp.eta_1), "eta_2": np.array(rp.eta_2),
            "alpha_0": np.array(rp.alpha_0), "alpha_1": np.array(rp.alpha_1),
            "alpha_2": np.array(rp.alpha_2), "beta": np.array(rp.beta),
            "energy": np.array(rp.energy),
        }
        if self.profile is not None:
            p = self.profile
            st.update(n_slices=p.n_slices, cut_left=float(p.cut_left), cut_right=float(p.cut_right),
                      bin_centers=_h(p.bin_centers).copy(),
                      profile=_h(p.n_macroparticles).copy())
        fb = self.tracker.beamFB
        if fb is not None:
            st["feedback"] = _feedback_state(fb, rp, self.profile)
        if self.tiv is not None:
            srcs = self.tiv.induced_voltage_list
            conv = [s for s in srcs if isinstance(s, (InducedVoltageFreq, InducedVoltageTime))]
            st.update(ratio=float(self.beam.ratio),
                      induced_voltage=_h(self.tiv.induced_voltage).copy())
            if conv:
                z = 0
                for s in conv:
                    z = z + _h(s.total_impedance)
                st.update(Z=np.asarray(z, dtype=np.complex128), n_fft=int(conv[0].n_fft),
                          n_induced_voltage=int(conv[0].n_induced_voltage))
            if len(conv) != len(srcs) or any(s.multi_turn_wake for s in conv):
                # per-source form: multi-turn memories and inductive sources
                st["bin_size"] = float(self.profile.bin_size)
                st["sources"] = [_source_state(s) for s in srcs]
        return st

    def oracle_copy(self) -> dict:
        return self.snapshot()


class _InducedSum:
    """A map element that sums the induced voltage without kicking."""

    def __init__(self, tiv):
        self.tiv = tiv

    def track(self):
        self.tiv.induced_voltage_sum()


def _feedback_state(fb, rp, profile) -> dict:
    """Host copy of a BeamFeedback's configuration and state for a CPU
    restatement of the loop."""
    ring = fb.ring
    d = dict(machine=fb.machine, alpha=float(fb.alpha), first_bin=fb.first_bin(),
             delay=int(fb.delay), reference=float(fb.reference),
             bin_size=float(profile.bin_size),
             phi_s=np.array(rp.phi_s, dtype=np.float64),
             harmonic=np.array(rp.harmonic, dtype=np.float64),
             omega_rf_d=np.array(rp.omega_rf_d, dtype=np.float64),
             dphi_rf=np.array(rp.dphi_rf, dtype=np.float64),
             phi_beam=float(fb.phi_beam), dphi=float(fb.dphi), domega_rf=float(fb.domega_rf),
             alpha_0=np.array(ring.alpha_0[0], dtype=np.float64),
             gamma=np.array(rp.gamma, dtype=np.float64))
    if fb.machine == "PSB":
        d.update(gain=np.array(fb.gain, dtype=np.float64),
                 gain2=[np.array(fb.gain2[0], dtype=np.float64),
                        np.array(fb.gain2[1], dtype=np.float64)],
                 on_time=np.array(fb.on_time, dtype=np.float64), PL_counter=int(fb.PL_counter),
                 dphi_sum=float(fb.dphi_sum), dphi_av=float(fb.dphi_av),
                 dphi_av_prev=float(fb.dphi_av_prev), domega_PL=float(fb.domega_PL),
                 domega_RL=float(fb.domega_RL), dR_over_R=float(fb.dR_over_R),
                 dR_over_R_prev=float(fb.dR_over_R_prev))
    else:
        d.update(gain=float(fb.gain), gain2=float(fb.gain2))
    if fb.machine == "LHC":
        d.update(lhc_a=np.array(fb.lhc_a, dtype=np.float64),
                 lhc_t=np.array(fb.lhc_t, dtype=np.float64), lhc_y=float(fb.lhc_y))
    if fb.machine == "SPS_RL":
        d.update(sample_dE=int(fb.sample_dE), ring_radius=float(ring.ring_radius),
                 beta=np.array(ring.beta[0], dtype=np.float64),
                 energy=np.array(ring.energy[0], dtype=np.float64),
                 eta_0=np.array(rp.eta_0, dtype=np.float64))
    return d


def _source_state(s) -> dict:
    """Host copy of one induced-voltage source for a CPU restatement."""
    if isinstance(s, InductiveImpedance):
        return {"kind": "inductive", "Z_over_n": np.asarray(s.Z_over_n, dtype=np.float64),
                "deriv_mode": s.deriv_mode}
    d = {"kind": "convolution", "Z": np.asarray(_h(s.total_impedance), dtype=np.complex128),
         "n_fft": int(s.n_fft), "n_induced_voltage": int(s.n_induced_voltage),
         "multi_turn_wake": bool(s.multi_turn_wake)}
    if s.multi_turn_wake:
        d.update(memory=np.array(_h(s.mtw_memory), dtype=np.float64),
                 n_mtw_memory=int(s.n_mtw_memory), wake_length=float(s.wake_length),
                 front_wake_buffer=int(s.front_wake_buffer))
    return d


def _h(a):
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _profile(beam, cut_left, cut_right, n_slices):
    return Profile(beam, CutOptions(cut_left=cut_left, cut_right=cut_right, n_slices=n_slices),
                   other_slices_options=OtherSlicesOptions(direct_slicing=False))


def build_workload(name: str, n_particles: int | None = None, n_turns: int | None = None,
                   n_slices: int | None = None, n_fft: int | None = None, seed: int = 1,
                   n_resonators: int = 20, to_gpu: bool = True,
                   intensity: float | None = None, interpolation: bool = False) -> Workload:
    """Seeded synthetic workload ``c1``, ``c2`` or ``c3`` (see module doc).
    ``intensity`` replaces the workload's bunch intensity (c2, c3);
    ``interpolation`` gives the tracker interpolation=True (tracker.py:585-601:
    the RF voltage on the bin centres, plus the induced voltage, kicks by
    linear interpolation).  Device residency (``to_gpu``) needs a GPU; with
    ``to_gpu=False`` only the host setup runs."""
    if name == "c1":
        n_particles = n_particles or 1_000_000
        n_turns = n_turns or 1000
        n_slices = n_slices or 100
        ring = Ring(26658.883, 1 / 55.759505 ** 2, 450e9, Proton(), n_turns)
        rf = RFStation(ring, [35640], [6e6], [0.0])
        beam = Beam(ring, n_particles, 1e9)
        bigaussian(ring, rf, beam, sigma_dt=1.0e-9 / 4, seed=seed)
        prof = _profile(beam, 0.0, float(rf.t_rf[0, 0]), n_slices)
        tr = RingAndRFTracker(rf, beam, interpolation=interpolation, profile=prof)
        w = Workload("c1", ring, rf, beam, tr, prof, None, False)
    elif name in ("c2", "c3"):
        if name == "c2":
            n_particles = n_particles or 4_000_000
            n_turns = n_turns or 2000
            n_slices = n_slices or 4096
            n_fft = n_fft or 16384
            ring = Ring(157.08, 1 / 4.4 ** 2, 310e6, Proton(), n_turns)
            rf = RFStation(ring, [1], [8e3], [np.pi])
            intensity = intensity or 1e13
        else:
            n_particles = n_particles or 10_000_000
            n_turns = n_turns or 500
            n_slices = n_slices or 65536
            n_fft = n_fft or 131072
            ring = Ring(26658.883, 1 / 55.759505 ** 2, 450e9, Proton(), n_turns)
            rf = RFStation(ring, [35640, 142560], [6e6, 1.5e6], [0.0, 0.0], n_rf=2)
            intensity = intensity or 1e11
        beam = Beam(ring, n_particles, intensity)
        bigaussian(ring, rf, beam, sigma_dt=C2_SIGMA_DT if name == "c2" else 1.0e-9 / 4,
                   seed=seed)
        prof = _profile(beam, 0.0, float(rf.t_rf[0, 0]), n_slices)
        R, fr, Q = resonator_set(n_resonators)
        res = Resonators(R, fr, Q)
        # frequency resolution chosen so that n_induced_voltage = n_fft exactly
        f_res = (1 + 1e-9) / (n_fft * prof.bin_size)
        ivf = InducedVoltageFreq(beam, prof, [res], frequency_resolution=f_res)
        tiv = TotalInducedVoltage(beam, prof, [ivf])
        tr = RingAndRFTracker(rf, beam, interpolation=interpolation, profile=prof,
                              total_induced_voltage=tiv if interpolation else None)
        w = Workload(name, ring, rf, beam, tr, prof, tiv, True)
        w.meta["n_fft"] = ivf.n_fft
    else:
        raise ValueError(f"unknown workload {name!r}")
    w.meta.update(n_particles=n_particles, n_turns=n_turns, n_slices=n_slices)
    if to_gpu:
        w.to_gpu()
    return w


WAKE_KINDS = ("mtw", "mtw_front", "mtw_time", "zovern", "zovern_diff", "mixed")


def build_wake_workload(kind: str, n_particles: int = 50_000, n_turns: int = 40,
                        n_slices: int = 256, memory_turns: int = 4, seed: int = 1,
                        to_gpu: bool = True) -> Workload:
    """A c2-ring bunch (h = 1, 8 kV, Gaussian 4 sigma = 400 ns, intensity
    1e12, n_slices over [0, t_rf] = one revolution) whose induced voltage
    comes from:

      mtw          InducedVoltageFreq with a multi-turn wake (mode 'time',
                   impedance.py:491-550) over `memory_turns` revolutions
                   (frequency resolution 1/(memory_turns * t_rev)), of a
                   broadband resonator (5 kOhm, 1 GHz, Q = 1) and a
                   long-range one (20 kOhm at 1.5 f_rev, Q = 200)
      mtw_front    the same with a front-wake cut of 8 bins (impedance.py:867-871)
      mtw_time     InducedVoltageTime with a multi-turn wake of the same
                   resonators, wake length memory_turns * t_rf
      zovern       InductiveImpedance alone, Z/n = -100 Ohm every turn,
                   derivative 'gradient'
      zovern_diff  the same with derivative 'diff'
      mixed        InducedVoltageFreq (single turn, the resonators above)
                   plus InductiveImpedance (Z/n = 50 Ohm, 'diff')
    Map: profile, induced voltage, tracker."""
    if kind not in WAKE_KINDS:
        raise ValueError(f"unknown wake workload {kind!r}")
    ring = Ring(157.08, 1 / 4.4 ** 2, 310e6, Proton(), n_turns)
    rf = RFStation(ring, [1], [8e3], [np.pi])
    beam = Beam(ring, n_particles, 1e12)
    bigaussian(ring, rf, beam, sigma_dt=C2_SIGMA_DT, seed=seed)
    prof = _profile(beam, 0.0, float(rf.t_rf[0, 0]), n_slices)
    f_rev = float(ring.f_rev[0])
    res = Resonators([5e3, 2e4], [1e9, 1.5 * f_rev], [1.0, 200.0])
    f_res = (1 + 1e-9) / (memory_turns * n_slices * prof.bin_size)
    srcs = []
    if kind in ("mtw", "mtw_front"):
        srcs.append(InducedVoltageFreq(beam, prof, [res], frequency_resolution=f_res,
                                       multi_turn_wake=True, rf_station=rf, mtw_mode="time",
                                       front_wake_length=8 * prof.bin_size if kind == "mtw_front"
                                       else 0))
    elif kind == "mtw_time":
        srcs.append(InducedVoltageTime(beam, prof, [res],
                                       wake_length=memory_turns * float(rf.t_rf[0, 0]),
                                       multi_turn_wake=True, rf_station=rf))
    elif kind in ("zovern", "zovern_diff"):
        srcs.append(InductiveImpedance(beam, prof, [-100.0] * (n_turns + 1), rf,
                                       deriv_mode="gradient" if kind == "zovern" else "diff"))
    else:
        srcs.append(InducedVoltageFreq(beam, prof, [res], frequency_resolution=f_res))
        srcs.append(InductiveImpedance(beam, prof, [50.0] * (n_turns + 1), rf, deriv_mode="diff"))
    tiv = TotalInducedVoltage(beam, prof, srcs)
    w = Workload(kind, ring, rf, beam, RingAndRFTracker(rf, beam), prof, tiv, True)
    w.meta.update(n_particles=n_particles, n_turns=n_turns, n_slices=n_slices)
    if to_gpu:
        w.to_gpu()
    return w


FEEDBACK_KINDS = ("LHC", "LHC_SL", "LHC_F", "PSB", "SPS_RL")


def build_feedback_workload(kind: str, n_particles: int = 50_000, n_turns: int = 60,
                            n_slices: int = 1024, intensity: bool = False, seed: int = 1234,
                            to_gpu: bool = True) -> Workload:
    """A bunch displaced from the synchronous phase under a beam phase loop.

      LHC, LHC_F, SPS_RL: the SPS-like case of the reference's loop test
          (unittests/llrf/test_beam_feedback.py:24-60): 6911.5038 m,
          gamma_t 17.95142852, 25.92 GeV/c, h = 4620, 4.5 MV, Gaussian
          sigma = 0.05 ns moved by 0.1 ns, n_slices over [0, t_rf];
          PL_gain 1000/s; LHC_F adds a frequency loop (FL_gain 0.05 per turn),
          SPS_RL a radial loop (RL_gain 1e4, every 7th particle).
      LHC_SL: the c1 LHC ring (h = 35640, 6 MV), bunch 4 sigma = 1 ns moved
          by 0.1 ns, phase and synchro loops as EX_28_phase_loop_lhc.py
          (PL_gain 1/(5 t_rev), SL_gain PL_gain/10).
      PSB: the c2 ring (157.08 m, 310 MeV/c, h = 1, 8 kV, phi = pi) with
          the loop of EX_08_Phase_Loop.py:78 (PL_gain 1/25 us, period
          10 us) and a radial loop (RL_gain [1e-3, 1e2]), bunch of 100 ns
          rms moved by 50 ns.
    Map without intensity: tracker (with the loop), profile.  With
    ``intensity`` a broadband resonator adds induced voltage and the map is
    profile, induced voltage, tracker."""
    if kind not in FEEDBACK_KINDS:
        raise ValueError(f"unknown feedback workload {kind!r}")
    if kind == "PSB":
        ring = Ring(157.08, 1 / 4.4 ** 2, 310e6, Proton(), n_turns)
        rf = RFStation(ring, [1], [8e3], [np.pi])
        beam = Beam(ring, n_particles, 1e12)
        bigaussian(ring, rf, beam, sigma_dt=C2_SIGMA_DT, seed=seed)
        offset = 50e-9
        config = {"machine": "PSB", "PL_gain": 1.0 / 25.0e-6, "period": 10.0e-6,
                  "RL_gain": [1e-3, 1e2]}
    elif kind == "LHC_SL":
        ring = Ring(26658.883, 1 / 55.759505 ** 2, 450e9, Proton(), n_turns)
        rf = RFStation(ring, [35640], [6e6], [0.0])
        beam = Beam(ring, n_particles, 1.15e11 if intensity else 1e9)
        bigaussian(ring, rf, beam, sigma_dt=1.0e-9 / 4, seed=seed)
        offset = 0.1e-9
        pl = 1 / (5 * float(ring.t_rev[0]))
        config = {"machine": "LHC", "PL_gain": pl, "SL_gain": pl / 10}
    else:
        ring = Ring(6911.5038, 1.0 / 17.95142852 ** 2, 25.92e9, Proton(), n_turns)
        rf = RFStation(ring, [4620], [4.5e6], [0.0])
        beam = Beam(ring, n_particles, 1.2e11 if intensity else 1.2e6)
        bigaussian(ring, rf, beam, sigma_dt=0.05e-9, seed=seed)
        offset = 0.1e-9
        config = {"machine": kind, "PL_gain": 1000.0}
        if kind == "LHC_F":
            config["FL_gain"] = 0.05
        elif kind == "SPS_RL":
            config.update(RL_gain=1e4, sample_dE=7)
    beam.dt += offset
    prof = _profile(beam, 0.0, float(rf.t_rf[0, 0]), n_slices)
    fb = BeamFeedback(ring, rf, prof, config)
    tiv = None
    if intensity:
        res = Resonators([5e5], [1e9], [1.0])
        ivf = InducedVoltageFreq(beam, prof, [res],
                                 frequency_resolution=(1 + 1e-9) / (4 * n_slices * prof.bin_size))
        tiv = TotalInducedVoltage(beam, prof, [ivf])
    tr = RingAndRFTracker(rf, beam, beam_feedback=fb, profile=prof)
    w = Workload(f"fb_{kind}", ring, rf, beam, tr, prof, tiv, intensity)
    w.meta.update(n_particles=n_particles, n_turns=n_turns, n_slices=n_slices)
    if to_gpu:
        w.to_gpu()
        prof.track()  # the loop reads the profile of the current coordinates
    return w


def induced_scale(beam) -> float:
    """charge * e * ratio (impedance.py:475-485)."""
    return beam.particle.charge * e * beam.ratio


@dataclass
class Ensemble:
    """Per-bunch reference-style objects of the c4 workload."""
    ring: Ring
    trackers: list
    profiles: list
    tivs: list
    counts: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def n_particles(self) -> int:
        return int(self.counts.sum())

    def tracker(self, **kw):
        from .trackers.ensemble import EnsembleTracker

        return EnsembleTracker(self.trackers, self.profiles, self.tivs, **kw)

    def bunch(self, b: int) -> Workload:
        """Bunch b as a single-bunch workload (the sequential form)."""
        tr = self.trackers[b]
        return Workload(f"c4[{b}]", self.ring, tr.rf_params, tr.beam, tr, self.profiles[b],
                        self.tivs[b], True)


def build_ensemble(n_bunches: int = 256, n_turns: int = 1000, n_slices: int = 1024,
                   n_min: float = 1e4, n_max: float = 1e6, seed: int = 7,
                   n_resonators: int = 5, intensity: float = 1e12) -> Ensemble:
    """The c4 ensemble on the host (coordinates as numpy arrays); the
    ensemble tracker moves them to HBM."""
    rng = np.random.default_rng(seed)
    counts = np.rint(10 ** rng.uniform(np.log10(n_min), np.log10(n_max), n_bunches)).astype(np.int64)
    volts = rng.uniform(4e3, 8e3, n_bunches)
    phis = np.pi + rng.uniform(-0.1, 0.1, n_bunches)
    ring = Ring(157.08, 1 / 4.4 ** 2, 310e6, Proton(), n_turns)
    trackers, profiles, tivs = [], [], []
    for b in range(n_bunches):
        rf = RFStation(ring, [1], [volts[b]], [phis[b]])
        beam = Beam(ring, int(counts[b]), intensity)
        bigaussian(ring, rf, beam, sigma_dt=C2_SIGMA_DT, seed=seed * 100000 + 2 * b)
        prof = _profile(beam, 0.0, float(rf.t_rf[0, 0]), n_slices)
        R, fr, Q = resonator_set(n_resonators, seed=seed * 100000 + 50000 + b, broadband=False)
        ivf = InducedVoltageFreq(beam, prof, [Resonators(R, fr, Q)])
        tivs.append(TotalInducedVoltage(beam, prof, [ivf]))
        trackers.append(RingAndRFTracker(rf, beam))
        profiles.append(prof)
    return Ensemble(ring, trackers, profiles, tivs, counts,
                    meta=dict(n_bunches=n_bunches, n_turns=n_turns, n_slices=n_slices,
                              n_fft=int(tivs[0].induced_voltage_list[0].n_fft), seed=seed))


@dataclass
class AdjointCase:
    """c5: the c1 ring with one broadband resonator, a continuous profile
    on the bucket and a target profile made under hidden parameters."""
    ring: Ring
    rf: RFStation
    beam: Beam
    tracker: RingAndRFTracker
    profile: Profile
    n_fft: int
    theta0: dict
    theta_star: dict
    state: dict


def build_adjoint_case(n_particles: int = 1_000_000, n_turns: int = 500, n_slices: int = 256,
                       seed: int = 11, intensity: float = 1e9) -> AdjointCase:
    """Nominal parameters theta0 = (V 6 MV, phi 0, broadband R 1 MOhm,
    f_r 2 GHz, Q 1); hidden theta* drawn from default_rng(seed) within a few
    per cent of them (phi within +-0.02 rad).  Host-only (no GPU needed)."""
    ring = Ring(26658.883, 1 / 55.759505 ** 2, 450e9, Proton(), n_turns)
    rf = RFStation(ring, [35640], [6e6], [0.0])
    beam = Beam(ring, n_particles, intensity)
    bigaussian(ring, rf, beam, sigma_dt=1.0e-9 / 4, seed=seed)
    prof = Profile(beam, CutOptions(cut_left=0.0, cut_right=float(rf.t_rf[0, 0]),
                                    n_slices=n_slices),
                   other_slices_options=OtherSlicesOptions(direct_slicing=False))
    tr = RingAndRFTracker(rf, beam)
    from .toolbox.next_regular import next_regular

    n_fft = next_regular(n_slices)
    theta0 = dict(V=np.array([6e6]), phi=np.array([0.0]), R=np.array([1e6]),
                  fr=np.array([2e9]), Q=np.array([1.0]))
    rng = np.random.default_rng(seed + 1000)
    theta_star = dict(V=theta0["V"] * (1 + rng.uniform(-0.03, 0.03)),
                      phi=np.array([rng.uniform(-0.02, 0.02)]),
                      R=theta0["R"] * (1 + rng.uniform(-0.1, 0.1)),
                      fr=theta0["fr"] * (1 + rng.uniform(-0.05, 0.05)),
                      Q=theta0["Q"] * (1 + rng.uniform(-0.1, 0.1)))
    rp = rf
    coeff = (np.asarray(rp.t_rev) * rp.length_ratio * np.asarray(rp.eta_0[0])
             / (np.asarray(rp.beta) ** 2 * np.asarray(rp.energy)))
    state = dict(omega=np.array(rp.host_array("omega_rf"), dtype=np.float64),
                 coeff=coeff, acc=np.zeros(n_turns + 1),
                 charge=float(beam.particle.charge), cut_left=float(prof.cut_left),
                 cut_right=float(prof.cut_right), n_slices=n_slices,
                 bin_centers=np.asarray(prof.bin_centers, dtype=np.float64).copy(),
                 n_fft=n_fft, bin_size=float(prof.bin_size),
                 scale=float(beam.particle.charge * e * beam.ratio))
    return AdjointCase(ring, rf, beam, tr, prof, n_fft, theta0, theta_star, state)
