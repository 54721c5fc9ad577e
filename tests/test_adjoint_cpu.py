"""The float64 CPU reference of the differentiable map (oracle/adjoint.py)
pinned on the CPU:

* its resonator impedance against the oracle's restatement of
  fast_resonator.cpp, and its derivatives against finite differences;
* its forward pass, both maps, against the oracle's turn loop
  (oracle.track): with no impedance the particles are the same bits as the
  reference's kick and drift, and the reference map's profile is the
  reference's smooth_histogram of them; its induced voltage matches the
  FFT route (impedance.py:457-489) to 1e-12 of its maximum; with impedance,
  whole turns of the reference map follow oracle.track's smooth-profile
  intensity map;
* its adjoint gradient against central finite differences of its own
  forward map: the spline map at any size, the reference map (whose loss
  jumps when a particle crosses a bin centre or edge) on cases where no
  particle crosses one between the two finite-difference runs -- the
  crossing count is computed and required to be zero."""
import numpy as np
import pytest

from oracle import adjoint as adj
from oracle import oracle as orc

from blond_amd.workloads import Workload, build_adjoint_case

KEYS = ("V", "phi", "R", "fr", "Q")
LIB = orc.oracle_lib()


def test_resonator_matches_oracle_fast_resonator():
    f = np.linspace(0, 5e9, 301)
    R, fr, Q = [1e6, 3e3], [2e9, 4e8], [1.0, 40.0]
    np.testing.assert_allclose(adj.resonator_Z(f, R, fr, Q), orc.resonator_impedance(f, R, fr, Q),
                               rtol=1e-13, atol=0)


def test_resonator_derivatives_vs_finite_differences():
    f = np.linspace(0, 5e9, 101)
    R, fr, Q = np.array([1e6]), np.array([2e9]), np.array([1.3])
    dR, dfr, dQ = adj.resonator_dZ(f, R, fr, Q)
    for d, (a, i) in ((dR, ("R", 0)), (dfr, ("fr", 0)), (dQ, ("Q", 0))):
        th = dict(R=R.copy(), fr=fr.copy(), Q=Q.copy())
        h = 1e-6 * th[a][i]
        tp = {k: v.copy() for k, v in th.items()}
        tm = {k: v.copy() for k, v in th.items()}
        tp[a][i] += h
        tm[a][i] -= h
        fd = (adj.resonator_Z(f, **tp) - adj.resonator_Z(f, **tm)) / (2 * h)
        np.testing.assert_allclose(d[0], fd, rtol=1e-7, atol=1e-9 * np.abs(fd).max())


# ---------------------------------------------------------------------------
# the forward pass against the oracle's turn loop
# ---------------------------------------------------------------------------
def _oracle_state(c, th, intensity):
    """The adjoint case as a state of oracle.track: the same ring, programs
    and bunch; with ``intensity`` the smooth profile and the resonator of
    ``th`` through the FFT route."""
    w = Workload("c5", c.ring, c.rf, c.beam, c.tracker, c.profile, None, False)
    st = w.snapshot()
    st["voltage"] = np.full_like(st["voltage"], th["V"][0])
    st["phi_rf"] = np.full_like(st["phi_rf"], th["phi"][0])
    st["smooth"] = True
    if intensity:
        s = c.state
        Z = adj.resonator_Z(np.fft.rfftfreq(s["n_fft"], s["bin_size"]), th["R"], th["fr"],
                            th["Q"]) / s["bin_size"]
        st.update(intensity=True, ratio=s["scale"] / (s["charge"] * orc.E_CHARGE), Z=Z,
                  n_fft=s["n_fft"], n_induced_voltage=s["n_fft"],
                  induced_voltage=np.zeros(s["n_slices"]))
    return st


@pytest.mark.parametrize("mode", adj.MODES)
def test_forward_without_impedance_is_the_reference_kick_and_drift(mode):
    """R = 0: no induced voltage, so each turn is the reference's kick and
    drift: the same bits as oracle.track; the reference map's final profile
    is smooth_histogram of the final coordinates."""
    c = build_adjoint_case(n_particles=20001, n_turns=40, n_slices=256, seed=11)
    th = dict(c.theta0, R=np.array([0.0]))
    m = adj.DiffMap(c.state, mode=mode)
    p, x, e = m.forward(c.beam.dt, c.beam.dE, th["V"], th["phi"], th["R"], th["fr"], th["Q"], 30,
                        coords=True)
    ref = orc.track(_oracle_state(c, th, False), 30, LIB)
    np.testing.assert_array_equal(x, ref["dt"])
    np.testing.assert_array_equal(e, ref["dE"])
    if mode == "reference":
        prof = np.zeros(256)
        LIB.slice_smooth(ref["dt"], prof, c.state["cut_left"], c.state["cut_right"])
        np.testing.assert_array_equal(p, prof)


def test_induced_voltage_matches_fft_route():
    c = build_adjoint_case(n_particles=20001, n_turns=4, n_slices=256, seed=11, intensity=1e11)
    th = c.theta0
    m = adj.DiffMap(c.state)
    s = c.state
    p = m.deposit(np.ascontiguousarray(c.beam.dt))
    v = m.induced(p, m.wake(th["R"], th["fr"], th["Q"]))
    Z = adj.resonator_Z(np.fft.rfftfreq(s["n_fft"], s["bin_size"]), th["R"], th["fr"],
                        th["Q"]) / s["bin_size"]
    ref = orc.induced_voltage_1turn(p, Z, s["n_fft"], s["charge"],
                                    s["scale"] / (s["charge"] * orc.E_CHARGE), s["n_fft"],
                                    s["n_slices"])
    assert np.abs(v).max() > 0
    np.testing.assert_allclose(v, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("turns", [1, 12])
def test_reference_map_turns_follow_the_oracle_intensity_map(turns):
    """With impedance: the reference map's turns are oracle.track's
    smooth-profile intensity map (profile, induced voltage, interpolated
    kick, RF kick, drift) up to the induced voltage's rounding (direct
    circular product here, FFT there)."""
    c = build_adjoint_case(n_particles=20001, n_turns=turns + 2, n_slices=256, seed=11,
                           intensity=1e11)
    th = c.theta0
    m = adj.DiffMap(c.state, mode="reference")
    p, x, e = m.forward(c.beam.dt, c.beam.dE, th["V"], th["phi"], th["R"], th["fr"], th["Q"],
                        turns, coords=True)
    ref = orc.track(_oracle_state(c, th, True), turns, LIB)
    assert np.abs(ref["induced_voltage"]).max() > 1e-3 * 6e6
    np.testing.assert_allclose(x, ref["dt"], rtol=0, atol=1e-12 * np.abs(ref["dt"]).max())
    np.testing.assert_allclose(e, ref["dE"], rtol=0, atol=1e-12 * np.abs(ref["dE"]).max())


# ---------------------------------------------------------------------------
# the adjoint against central finite differences
# ---------------------------------------------------------------------------
def _fd_vs_adjoint(c, m, turns, h_rel=1e-6):
    """Adjoint gradient, central finite differences (step h_rel relative, or
    absolute for a zero parameter) and, per parameter, the number of
    (particle, turn) pairs whose piece of the map differs between the two
    finite-difference runs."""
    th, ts = c.theta0, c.theta_star
    dt0, dE0 = c.beam.dt.copy(), c.beam.dE.copy()
    target = m.forward(dt0, dE0, ts["V"], ts["phi"], ts["R"], ts["fr"], ts["Q"], turns)

    def run(t):
        p, hist, _ = m.forward(dt0, dE0, t["V"], t["phi"], t["R"], t["fr"], t["Q"], turns,
                               keep=True)
        _, xT, _ = m.forward(dt0, dE0, t["V"], t["phi"], t["R"], t["fr"], t["Q"], turns,
                             coords=True)
        cells = [m.cells(h[0]) for h in hist] + [m.cells(xT)]
        return np.sum((p - target) ** 2), p, cells

    L, p, _ = run(th)
    assert L > 0
    g = m.gradient(dt0, dE0, th["V"], th["phi"], th["R"], th["fr"], th["Q"], turns,
                   2 * (p - target))
    np.testing.assert_array_equal(g["profile"], p)
    out = {}
    for k in KEYS:
        x0 = th[k][0]
        h = h_rel * abs(x0) if x0 != 0 else h_rel
        tp, tm = dict(th), dict(th)
        tp[k] = np.array([x0 + h])
        tm[k] = np.array([x0 - h])
        Lp, _, cp = run(tp)
        Lm, _, cm = run(tm)
        crossings = sum(int(np.count_nonzero(a != b)) for a, b in zip(cp, cm))
        out[k] = (g[k][0], (Lp - Lm) / (2 * h), crossings)
    return out


@pytest.mark.parametrize("n_slices,turns", [(32, 10), (64, 25)])
def test_spline_adjoint_vs_central_differences(n_slices, turns):
    c = build_adjoint_case(n_particles=600, n_turns=turns + 2, n_slices=n_slices, seed=11)
    m = adj.DiffMap(c.state, mode="spline")
    for k, (g, fd, _) in _fd_vs_adjoint(c, m, turns).items():
        assert abs(g - fd) <= 1e-5 * abs(fd), (k, g, fd)


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_reference_adjoint_vs_central_differences_without_crossings(seed):
    """Step 1e-6 relative (1e-6 rad for phi = 0), 400 particles x 30 turns,
    64 slices, intensity 1e11 (the induced voltage at a few per cent of the
    RF): no particle crosses a bin centre or edge between the two runs of
    any parameter, and the adjoint agrees to 1e-6 (measured: <= 6e-9)."""
    c = build_adjoint_case(n_particles=400, n_turns=32, n_slices=64, seed=seed, intensity=1e11)
    m = adj.DiffMap(c.state, mode="reference")
    for k, (g, fd, crossings) in _fd_vs_adjoint(c, m, 30).items():
        print(f"seed {seed} {k}: adjoint {g:+.9e} fd {fd:+.9e} crossings {crossings}")
        assert crossings == 0, (k, crossings)
        assert abs(g - fd) <= 1e-6 * abs(fd), (k, g, fd)


def test_reference_crossings_break_finite_differences():
    """The same map where particles do cross (1000 particles x 100 turns,
    256 slices): the crossing count is positive and the finite difference of
    V departs from the adjoint by far more than 1e-6 -- the loss jumps, the
    adjoint differentiates the pieces."""
    c = build_adjoint_case(n_particles=1000, n_turns=102, n_slices=256, seed=11)
    m = adj.DiffMap(c.state, mode="reference")
    g, fd, crossings = _fd_vs_adjoint(c, m, 100)["V"]
    assert crossings > 0
    assert abs(g - fd) > 1e-3 * abs(fd)


def test_spline_deposit_is_smooth_and_conserves_weight():
    c = build_adjoint_case(n_particles=10, n_turns=2, n_slices=16)
    m = adj.DiffMap(c.state, mode="spline")
    bw = (c.state["cut_right"] - c.state["cut_left"]) / 16
    # a particle sweeping across a bin centre or edge: the profile and its
    # derivative move continuously
    for x0 in (5.5, 6.0):
        xs = c.state["cut_left"] + bw * (x0 + np.linspace(-1e-6, 1e-6, 3))
        ps = [m.deposit(np.array([x])) for x in xs]
        assert np.abs(ps[0] - ps[2]).max() < 3e-6
        np.testing.assert_allclose(ps[2] - ps[1], ps[1] - ps[0], atol=1e-11)
    inner = c.state["cut_left"] + bw * np.array([1.7, 3.2, 14.3])
    assert m.deposit(inner).sum() == pytest.approx(3.0, abs=1e-14)


def test_reference_deposit_jumps_at_centres_and_edges():
    """smooth_histogram's weights (histogram.cpp:100-113): a particle moving
    across the centre of bin k moves its whole weight from bin k-1 to bin
    k+1; across the edge between bins k-1 and k, weights (-0.5, 1.5) become
    (0.5, 0.5); the weight sums to one."""
    c = build_adjoint_case(n_particles=10, n_turns=2, n_slices=16)
    m = adj.DiffMap(c.state, mode="reference")
    cl = c.state["cut_left"]
    bw = (c.state["cut_right"] - cl) / 16
    below, above = [m.deposit(np.array([cl + bw * (6.5 + s)])) for s in (-1e-9, 1e-9)]
    np.testing.assert_allclose(below[[5, 6, 7]], [1.0, 0.0, 0.0], atol=1e-7)
    np.testing.assert_allclose(above[[5, 6, 7]], [0.0, 0.0, 1.0], atol=1e-7)
    below, above = [m.deposit(np.array([cl + bw * (6.0 + s)])) for s in (-1e-9, 1e-9)]
    np.testing.assert_allclose(below[[5, 6]], [-0.5, 1.5], atol=1e-7)
    np.testing.assert_allclose(above[[5, 6]], [0.5, 0.5], atol=1e-7)
    assert m.deposit(cl + bw * np.array([1.7, 3.2, 14.3])).sum() == pytest.approx(3.0, abs=1e-14)
