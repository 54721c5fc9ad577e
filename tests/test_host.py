"""Host-side setup of the MI355X package (CPU only): ring and RF programs,
cuts, FFT sizing, bunch generation -- the values the kernels consume."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

from blond_amd.beam.beam import Beam, Proton
from blond_amd.beam.distributions import bigaussian
from blond_amd.beam.profile import CutOptions, OtherSlicesOptions, Profile
from blond_amd.impedances.impedance import InducedVoltageFreq
from blond_amd.input_parameters.rf_parameters import RFStation
from blond_amd.input_parameters.ring import Ring
from blond_amd.toolbox.next_regular import next_regular
from blond_amd.trackers.tracker import RingAndRFTracker
from blond_amd.workloads import build_workload, resonator_set


def lhc_ring(n_turns=10):
    return Ring(26658.883, 1 / 55.759505 ** 2, 450e9, Proton(), n_turns)


def test_proton_mass_codata_2022():
    assert Proton().mass == pytest.approx(938272089.4282573, rel=0, abs=1e-6)


def test_ring_closed_forms_lhc():
    """Closed forms of ring.py:288-306, 427-432 at 450 GeV/c."""
    r = lhc_ring()
    m = Proton().mass
    beta = 1 / math.sqrt(1 + (m / 450e9) ** 2)
    gamma = math.sqrt(1 + (450e9 / m) ** 2)
    assert r.beta[0, 0] == pytest.approx(beta, rel=1e-15)
    assert r.gamma[0, 0] == pytest.approx(gamma, rel=1e-15)
    assert r.t_rev[0] == pytest.approx(26658.883 / (beta * 299792458.0), rel=1e-15)
    assert r.eta_0[0, 0] == pytest.approx(1 / 55.759505 ** 2 - 1 / gamma ** 2, rel=1e-13)
    assert np.all(r.delta_E == 0)
    # the values quoted for this ring
    assert r.beta[0, 0] == 0.9999978262922387
    assert r.t_rev[0] == 8.892465516509709e-05


def test_rf_station_programs():
    r = lhc_ring()
    rf = RFStation(r, [35640], [6e6], [0.0])
    assert rf.omega_rf[0, 0] == 2518229887.2245054
    assert rf.t_rf[0, 0] == 2.495080111254127e-09
    assert rf.voltage.flags.f_contiguous and rf.voltage[:, 3].flags.c_contiguous
    # above transition, no acceleration: phi_s = pi (rf_parameters.py:690-705)
    assert np.allclose(rf.phi_s, np.pi)
    psb = Ring(157.08, 1 / 4.4 ** 2, 310e6, Proton(), 5)
    rf2 = RFStation(psb, [1], [8e3], [np.pi])
    assert psb.eta_0[0, 0] < 0 and np.allclose(rf2.phi_s, np.pi)


def test_kinetic_energy_program():
    r = Ring(2 * np.pi * 25.0, 1 / 4.1 ** 2, 160e6, Proton(), 3,
             synchronous_data_type="kinetic energy")
    m = Proton().mass
    assert r.momentum[0, 0] == pytest.approx(math.sqrt((160e6 + m) ** 2 - m ** 2), rel=1e-15)


def test_cut_options_rad_and_default():
    r = Ring(125, 0.001, 1e9, Proton(), 1)
    rf = RFStation(r, [1], [7e6], [0.0])
    c = CutOptions(cut_left=0, cut_right=2 * np.pi, n_slices=200, cuts_unit="rad", rf_station=rf)
    c.set_cuts()
    assert c.cut_right == pytest.approx(2 * np.pi / rf.omega_rf[0, 0], rel=1e-15)
    assert len(c.bin_centers) == 200 and c.bin_size == pytest.approx(c.cut_right / 200)
    beam = Beam(r, 1000, 1e10)
    beam.dt = np.linspace(1e-9, 2e-9, 1000)
    p = Profile(beam, other_slices_options=OtherSlicesOptions(direct_slicing=False))
    assert p.cut_left == pytest.approx(1e-9 - 0.05e-9) and p.cut_right == pytest.approx(2.05e-9)
    with pytest.raises(RuntimeError):
        CutOptions(cut_left=0, cut_right=1, cuts_unit="rad")


def test_next_regular_matches_oracle():
    for t in list(range(1, 500)) + [1067, 16384, 16385, 131072, 99991]:
        assert next_regular(t) == orc.next_regular(t)


def _profile16():
    return Profile(None, cut_options=CutOptions(cut_left=0, cut_right=5e-9, n_slices=16))


def test_induced_voltage_freq_sizing():
    """unittests/impedances/test_impedance.py:38-88 (sizes only, no source)."""
    iv = InducedVoltageFreq(None, _profile16(), [])
    assert iv.n_fft == 16
    assert iv.wake_length == pytest.approx(_profile16().bin_size * 16)
    assert iv.frequency_resolution == pytest.approx(1 / (_profile16().bin_size * 16))
    iv = InducedVoltageFreq(None, _profile16(), [], frequency_resolution=3e6)
    assert iv.n_induced_voltage == 1067 and iv.n_fft == 1080
    assert iv.frequency_resolution == pytest.approx(iv.freq[1] - iv.freq[0])
    with pytest.raises(RuntimeError):
        InducedVoltageFreq(None, _profile16(), [], frequency_resolution=1e12)


def test_bigaussian_draws_follow_default_rng():
    """distributions.py:1243-1255: default_rng(seed) for dt, seed+1 for dE."""
    r = lhc_ring()
    rf = RFStation(r, [35640], [6e6], [0.0])
    beam = Beam(r, 5000, 1e9)
    bigaussian(r, rf, beam, sigma_dt=0.25e-9, seed=1)
    z1 = np.random.default_rng(1).normal(size=5000)
    z2 = np.random.default_rng(2).normal(size=5000)
    np.testing.assert_array_equal(beam.dt, 0.25e-9 * z1 + (np.pi - 0.0) / rf.omega_rf[0, 0])
    # matched dE amplitude from the single-harmonic Hamiltonian
    phi_b = rf.omega_rf[0, 0] * 0.25e-9 + np.pi
    sdE = math.sqrt(6e6 * r.energy[0, 0] * r.beta[0, 0] ** 2
                    * (math.cos(phi_b) - math.cos(np.pi) + (phi_b - np.pi) * math.sin(np.pi))
                    / (np.pi * 35640 * abs(r.eta_0[0, 0])))
    assert beam.sigma_dE == pytest.approx(sdE, rel=1e-14)
    np.testing.assert_allclose(beam.dE, beam.sigma_dE * z2, rtol=1e-15)


def test_tracker_setup_checks():
    r = lhc_ring()
    rf = RFStation(r, [35640], [6e6], [0.0])
    beam = Beam(r, 10, 1e9)
    t = RingAndRFTracker(rf, beam)
    assert t.solver == "simple" and np.all(t.acceleration_kick == -rf.delta_E)
    with pytest.raises(RuntimeError):
        RingAndRFTracker(rf, beam, solver="bogus")
    r2 = Ring(100.0, 1e-3, 1e9, Proton(), 4, alpha_1=1e-4, alpha_2=1e-5)
    t2 = RingAndRFTracker(RFStation(r2, [1], [1e3], [0.0]), Beam(r2, 4, 1e9))
    assert t2.solver == "exact"  # tracker.py:343-346


def test_workload_c1_host_setup():
    w = build_workload("c1", n_particles=1000, n_turns=5, to_gpu=False)
    assert w.profile.n_slices == 100
    assert w.profile.cut_right == w.rf.t_rf[0, 0]
    s = w.snapshot()
    assert s["voltage"].shape == (1, 6) and s["acc"].shape == (5,)
    R, fr, Q = resonator_set()
    assert len(R) == 21 and R[0] == 5e3 and fr[0] == 1e9 and Q[0] == 1.0
    assert np.all((R[1:] >= 1e3) & (R[1:] <= 1e6)) and np.all((Q[1:] >= 100) & (Q[1:] <= 1e4))


def test_oracle_turn_loop_runs_and_conserves_particles():
    w = build_workload("c1", n_particles=4000, n_turns=6, to_gpu=False)
    st = orc.track(w.snapshot(), 5)
    assert st["turn"] == 5 and st["profile"].sum() <= 4000
    assert np.all(np.isfinite(st["dt"])) and np.all(np.isfinite(st["dE"]))


def test_bunch_monitor_buffering_rows(tmp_path):
    """monitors.py:117-129, 325-460: row 0 from the constructor, then the
    buffer of turns [k*b, (k+1)*b) written when i_turn reaches (k+1)*b;
    float32 datasets; written as npz when h5py is absent."""
    from blond_amd.monitors.monitors import BunchMonitor, _h5py
    from blond_amd.workloads import build_workload

    w = build_workload("c1", n_particles=4000, n_turns=12, n_slices=100, to_gpu=False)
    mon = BunchMonitor(w.ring, w.rf, w.beam, str(tmp_path / "mon"), buffer_time=4)
    rows = np.zeros((8, 5))
    rows[:, 0] = 4000
    rows[:, 1] = 1e-9 + np.arange(8) * 1e-12
    rows[:, 2] = 2e-10
    rows[:, 3] = np.arange(8) * 10.0
    rows[:, 4] = 5e5
    mon.record_moments(rows)
    if _h5py is not None:  # pragma: no cover - h5py not installed here
        return
    d = np.load(str(tmp_path / "mon.npz"))
    mean_dt = d["Beam/mean_dt"]
    assert mean_dt.dtype == np.float32 and mean_dt.shape == (13,)
    w.beam.statistics()
    assert mean_dt[0] == np.float32(w.beam.mean_dt)
    # buffer row i holds turn i; rows 1..7 are the recorded turns 1..7 (turn 8
    # sits in the buffer until i_turn reaches 12)
    np.testing.assert_array_equal(mean_dt[1:8], rows[:7, 1].astype(np.float32))
    assert np.all(mean_dt[8:] == 0)
    np.testing.assert_array_equal(d["Beam/epsn_rms_l"][1:8],
                                  (np.pi * rows[:7, 4] * rows[:7, 2]).astype(np.float32))
    assert mon.i_turn == 9
