"""The oracle's turn loop (oracle.track) against closed forms.

Every GPU turn test compares the engine with ``oracle.track``; these tests pin
that loop itself: its map order and its turn indexing (tracker.py:521-615,
profile.py:700-719, impedance.py:113-146), independently of its kernels.  The
cases are in tests/turn_cases.py."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import turn_cases as tc

LIB = orc.oracle_lib()
REF = orc.ref_lib("strict")


@pytest.mark.parametrize("solver,a1,a2,form", tc.DRIFT_CASES)
def test_drift_closed_forms_2000_turns(solver, a1, a2, form):
    """unittests/trackers/test_drift.py:105-492 (rtol 1e-9 there and here)."""
    w = tc.drift_case(solver, a1, a2)
    st = w.snapshot()
    assert st["empty"]
    out = orc.track(st, 2000, LIB)
    expect = tc.drift_closed_form(w.ring, st["dE"], form, 2000)
    np.testing.assert_allclose(out["dt"], expect, rtol=1e-9, atol=0)
    np.testing.assert_array_equal(out["dE"], st["dE"])  # no voltage: no kick at all
    if REF is not None:
        ref = orc.track(st, 2000, REF)
        np.testing.assert_array_equal(out["dt"], ref["dt"])


def test_empty_station_skips_the_acceleration_kick():
    """tracker.py:585: without voltage even the acceleration kick is skipped."""
    w = tc.eta_step_case(n_particles=11, n_turns=20, k=10)
    st = w.snapshot()
    st["voltage"] = np.zeros_like(st["voltage"])
    st["empty"] = True
    out = orc.track(st, 20, LIB)
    np.testing.assert_array_equal(out["dE"], st["dE"])


@pytest.mark.parametrize("n", [120, 121, 200])
def test_ramp_indexing_closed_form(n):
    """Quadratic ramp (delta_E varies every turn) with eta_0 non-zero on turn
    k = 120 only.  The acceleration kick at [t] (not [t + 1]), the drift with
    the programs at [t + 1] (not [t]) and kick-before-drift each change the
    result by far more than the tolerance (checked below on the oracle with
    its indices shifted)."""
    w = tc.eta_step_case()
    st = w.snapshot()
    out = orc.track(st, n, LIB)
    dt, dE = tc.eta_step_closed_form(w, st["dE"], n)
    scale = np.abs(dE).max()
    np.testing.assert_allclose(out["dt"], dt, rtol=1e-12, atol=0)
    np.testing.assert_allclose(out["dE"], dE, rtol=0, atol=1e-12 * scale)


def test_ramp_indexing_detects_shifted_indices():
    """The same closed form fails for a loop whose acceleration program is
    read one turn late, and for one whose drift programs are read one turn
    early: the test above would catch either mistake."""
    w = tc.eta_step_case()
    st = w.snapshot()
    n = 200
    dt, dE = tc.eta_step_closed_form(w, st["dE"], n)
    scale = np.abs(dE).max()
    late_acc = dict(st, acc=np.append(st["acc"][1:], st["acc"][-1]))
    bad = orc.track(late_acc, n, LIB)
    assert np.abs(bad["dE"] - dE).max() > 1e3 * 1e-12 * scale
    early = dict(st)
    for key in ("t_rev", "eta_0", "eta_1", "eta_2", "alpha_0", "alpha_1", "alpha_2", "beta",
                "energy"):
        early[key] = np.concatenate([[st[key][0]], st[key][:-1]])
    bad = orc.track(early, n, LIB)
    assert np.abs(bad["dt"] - dt).max() > 1e3 * 1e-12 * np.abs(dt).max()


def test_synchronous_particle_stays_on_a_ramp():
    """V[t] = delta_E[t] / (q r): the particle at phi_s with dE = 0 stays
    there over 1000 accelerating turns (to rounding), while a kick that reads
    the acceleration program one turn late lets it wander."""
    w = tc.synchronous_case(n_turns=1000, spread=False)
    st = w.snapshot()
    out = orc.track(st, 1000, LIB)
    dEs = np.abs(w.rf.delta_E).max()
    assert abs(out["dE"][0]) < 1e-10 * dEs, out["dE"][0]
    assert abs(out["dt"][0] - w.meta["dt_s"]) < 1e-12 * w.meta["dt_s"]
    late = dict(st, acc=np.append(st["acc"][1:], st["acc"][-1]))
    bad = orc.track(late, 1000, LIB)
    assert abs(bad["dE"][0]) > 1e-3 * dEs


def test_intensity_map_order_one_turn():
    """One intensity turn of oracle.track from known coordinates: the profile
    it leaves is the histogram of the coordinates *before* the turn (profile
    first), the kick is evaluated at the old dt (numpy.interp of the induced
    voltage on the bin centres plus the RF sine), and the drift uses the
    kicked dE.  Closed forms with numpy primitives, not the oracle's kernels."""
    rng = np.random.default_rng(5)
    n, ns = 20001, 64
    t_rf = 2.495080111254127e-09
    st = tc.eta_step_case(n_particles=n, n_turns=4, k=2).snapshot()
    st["eta_0"] = np.full_like(st["eta_0"], 3.172867586042721e-04)
    st["dt"] = rng.uniform(0.1, 0.9, n) * t_rf
    st["dE"] = rng.normal(0, 1e8, n)
    st["acc"] = np.zeros_like(st["acc"])
    edges = np.linspace(0.0, t_rf, ns + 1)
    bc = (edges[:-1] + edges[1:]) / 2
    Z = orc.resonator_impedance(np.fft.rfftfreq(2 * ns, t_rf / ns), [1e6], [2e9], [1.0]) * ns / t_rf
    st.update(intensity=True, n_slices=ns, cut_left=0.0, cut_right=t_rf, bin_centers=bc,
              profile=np.zeros(ns), ratio=1e11 / n, induced_voltage=np.zeros(ns), Z=Z,
              n_fft=2 * ns, n_induced_voltage=2 * ns)
    out = orc.track(st, 1, LIB)
    counts = np.histogram(st["dt"], bins=edges)[0]
    np.testing.assert_array_equal(out["profile"], counts)
    v_ind = -(orc.E_CHARGE * st["ratio"]) * np.fft.irfft(Z * np.fft.rfft(counts, 2 * ns))[:ns]
    np.testing.assert_allclose(out["induced_voltage"], v_ind, rtol=0,
                               atol=1e-12 * np.abs(v_ind).max())
    inside = (st["dt"] >= bc[0]) & (st["dt"] < bc[-1])
    kick = np.where(inside, np.interp(st["dt"], bc, v_ind), 0.0)
    kick = kick + st["voltage"][0, 0] * np.sin(st["omega_rf"][0, 0] * st["dt"] + st["phi_rf"][0, 0])
    dE1 = st["dE"] + kick
    np.testing.assert_allclose(out["dE"], dE1, rtol=0, atol=1e-9 * np.abs(dE1).max())
    coeff = st["t_rev"][1] * st["eta_0"][1] / (st["beta"][1] ** 2 * st["energy"][1])
    np.testing.assert_allclose(out["dt"], st["dt"] + coeff * dE1, rtol=1e-13, atol=0)


@pytest.mark.parametrize("intensity", [False, True])
def test_interpolated_turn_closed_form(intensity):
    """tracker.py:585-601 with interpolation=True: one turn kicks by linear
    interpolation of the RF voltage on the bin centres (plus the summed
    induced voltage, which does not kick by itself), the acceleration kick
    only inside the table; then the drift.  Closed form with numpy.sin and
    numpy.interp on a ramp (acceleration kick non-zero)."""
    rng = np.random.default_rng(9)
    n, ns = 20001, 64
    w = tc.eta_step_case(n_particles=n, n_turns=4, k=2)
    st = w.snapshot()
    t_rf = 2 * np.pi / st["omega_rf"][0, 0]
    st["eta_0"] = np.full_like(st["eta_0"], 3.172867586042721e-04)
    st["dt"] = rng.uniform(-0.05, 1.05, n) * t_rf
    st["dE"] = rng.normal(0, 1e8, n)
    edges = np.linspace(0.0, t_rf, ns + 1)
    bc = (edges[:-1] + edges[1:]) / 2
    st.update(interpolation=True, n_slices=ns, cut_left=0.0, cut_right=t_rf, bin_centers=bc,
              profile=np.zeros(ns))
    v_ind = np.zeros(ns)
    if intensity:
        Z = orc.resonator_impedance(np.fft.rfftfreq(2 * ns, t_rf / ns), [1e6], [2e9],
                                    [1.0]) * ns / t_rf
        st.update(intensity=True, ratio=1e11 / n, induced_voltage=np.zeros(ns), Z=Z,
                  n_fft=2 * ns, n_induced_voltage=2 * ns)
        counts = np.histogram(st["dt"], bins=edges)[0]
        v_ind = -(orc.E_CHARGE * st["ratio"]) * np.fft.irfft(Z * np.fft.rfft(counts, 2 * ns))[:ns]
    out = orc.track(st, 1, LIB)
    acc = st["acc"][0]
    assert acc != 0
    total = st["voltage"][0, 0] * np.sin(st["omega_rf"][0, 0] * bc + st["phi_rf"][0, 0]) + v_ind
    inside = (st["dt"] >= bc[0]) & (st["dt"] < bc[-1])
    dE1 = st["dE"] + np.where(inside, np.interp(st["dt"], bc, total) + acc, 0.0)
    np.testing.assert_allclose(out["dE"], dE1, rtol=0, atol=1e-9 * np.abs(dE1).max())
    np.testing.assert_array_equal(out["dE"][~inside], st["dE"][~inside])
    coeff = st["t_rev"][1] * st["eta_0"][1] / (st["beta"][1] ** 2 * st["energy"][1])
    np.testing.assert_allclose(out["dt"], st["dt"] + coeff * dE1, rtol=1e-13, atol=0)
