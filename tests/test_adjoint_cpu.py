"""The float64 CPU reference of the differentiable map (oracle/adjoint.py)
pinned on the CPU: its resonator impedance against the oracle's restatement
of fast_resonator.cpp, its derivatives against finite differences of the
impedance, and its adjoint gradient against central finite differences of
its own forward map."""
import numpy as np
import pytest

from oracle import adjoint as adj
from oracle import oracle as orc

from blond_amd.workloads import build_adjoint_case


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


@pytest.mark.parametrize("n_slices,turns", [(32, 10), (64, 25)])
def test_adjoint_gradient_vs_central_differences(n_slices, turns):
    c = build_adjoint_case(n_particles=600, n_turns=turns + 2, n_slices=n_slices, seed=11)
    m = adj.DiffMap(c.state)
    th, ts = c.theta0, c.theta_star
    dt0, dE0 = c.beam.dt.copy(), c.beam.dE.copy()
    target = m.forward(dt0, dE0, ts["V"], ts["phi"], ts["R"], ts["fr"], ts["Q"], turns)

    def loss(t):
        p = m.forward(dt0, dE0, t["V"], t["phi"], t["R"], t["fr"], t["Q"], turns)
        return np.sum((p - target) ** 2), p

    L, p = loss(th)
    assert L > 0
    g = m.gradient(dt0, dE0, th["V"], th["phi"], th["R"], th["fr"], th["Q"], turns,
                   2 * (p - target))
    np.testing.assert_array_equal(g["profile"], p)
    for k in ("V", "phi", "R", "fr", "Q"):
        x0 = th[k][0]
        h = 1e-6 * abs(x0) if x0 != 0 else 1e-6
        tp, tm = dict(th), dict(th)
        tp[k] = np.array([x0 + h])
        tm[k] = np.array([x0 - h])
        fd = (loss(tp)[0] - loss(tm)[0]) / (2 * h)
        assert abs(g[k][0] - fd) <= 1e-5 * abs(fd), (k, g[k][0], fd)


def test_deposit_is_smooth_and_conserves_weight():
    c = build_adjoint_case(n_particles=10, n_turns=2, n_slices=16)
    m = adj.DiffMap(c.state)
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
