"""The CPU oracle pinned: against the reference build's golden vectors, the
reference's own profile fixture, and closed forms / brute force chosen so
that a dropped term, a wrong sign or an off-by-one fails.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "kernels_golden.npz"))
ORC = orc.oracle_lib()
W1 = 2.5182298872245054e9


# ---------------------------------------------------------------- golden ----
@pytest.mark.parametrize("n_rf", [1, 2])
def test_kick_golden(n_rf):
    e = G["dE"].copy()
    ORC.kick(G["dt"], e, G[f"kick_rf{n_rf}_V"], G[f"kick_rf{n_rf}_w"], G[f"kick_rf{n_rf}_phi"],
             1.0, n_rf, -1234.5)
    np.testing.assert_array_equal(e, G[f"kick_rf{n_rf}"])


@pytest.mark.parametrize("solver", ["simple", "legacy", "exact"])
@pytest.mark.parametrize("ao", [0, 1, 2])
def test_drift_golden(solver, ao):
    a = G["drift_args"]
    t = G["dt"].copy()
    ORC.drift(t, G["dE"], solver, a[0], a[1], ao, a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9])
    np.testing.assert_array_equal(t, G[f"drift_{solver}_ao{ao}"])


@pytest.mark.parametrize("ns", [1, 100, 4096])
def test_histogram_golden(ns):
    h = np.zeros(ns)
    ORC.slice_beam(G[f"hist_{ns}_dt"], h, 0.0, 2.495080111254127e-09)
    np.testing.assert_array_equal(h, G[f"hist_{ns}"])


@pytest.mark.parametrize("ns", [8, 150])
def test_smooth_golden(ns):
    h = np.zeros(ns)
    ORC.slice_smooth(G["dt"], h, 0.0, 2.495080111254127e-09)
    np.testing.assert_array_equal(h, G[f"smooth_{ns}"])


@pytest.mark.parametrize("ns", [2, 100])
def test_interp_golden(ns):
    e = G["dE"].copy()
    ORC.linear_interp_kick(G["dt"], e, G[f"interp_{ns}_V"], G[f"interp_{ns}_bc"], 1.0, 12.5)
    np.testing.assert_array_equal(e, G[f"interp_{ns}"])


def test_rf_volt_comp_and_resonator_golden():
    got = ORC.rf_volt_comp(np.array([6e6, 1.5e6]), np.array([W1, 4 * W1]), np.array([0.0, 0.2]),
                           G["rfv_bc"])
    np.testing.assert_array_equal(got, G["rfv"])
    z = ORC.fast_resonator(G["res_R"], G["res_Q"], G["res_f"], G["res_fr"])
    np.testing.assert_array_equal(z, G["res_Z"])


# ------------------------------------------- reference profile fixture ----
def _profile_fixture():
    exp = json.load(open(os.path.join(HERE, "golden", "profile_expected.json")))
    dt = np.load(os.path.join(HERE, "golden", "dt_coordinates.npz"))["arr_0"]
    return exp, dt


def test_reference_profiles_nearest_bin():
    """test_beam_profile_object.py:169-204: default cuts (beam extent +-5%,
    100 slices) and 'rad' cuts [0, 2 pi] / omega_rf (200 slices), exact."""
    exp, dt = _profile_fixture()
    lo, hi = dt.min(), dt.max()
    cl, cr = lo - 0.05 * (hi - lo), hi + 0.05 * (hi - lo)
    h = np.zeros(100)
    ORC.slice_beam(dt, h, cl, cr)
    assert h.tolist() == exp["profile1"]
    # ring of the test: C = 125 m, momentum 1 GeV/c, h = 1
    mp = orc.M_P_EV
    beta = math.sqrt(1 / (1 + (mp / 1e9) ** 2))
    omega_rf = 2.0 * np.pi * beta * orc.C_LIGHT * 1.0 / 125.0
    assert abs(125.0 / (beta * orc.C_LIGHT) - exp["t_rev0"]) < 1e-14
    h = np.zeros(200)
    ORC.slice_beam(dt, h, 0.0 / omega_rf, 2 * np.pi / omega_rf)
    assert h.tolist() == exp["profile2"]


def test_reference_profile_smooth():
    """test_beam_profile_object.py:206-230: smooth profile on [0, t_rev],
    150 slices, rtol 1e-6 -- including the negative bin of the weight rule."""
    exp, dt = _profile_fixture()
    mp = orc.M_P_EV
    beta = math.sqrt(1 / (1 + (mp / 1e9) ** 2))
    t_rev = 125.0 / (beta * orc.C_LIGHT)
    h = np.zeros(150)
    ORC.slice_smooth(dt, h, 0.0, t_rev)
    np.testing.assert_allclose(h, exp["profile3"], rtol=1e-6, atol=0)
    assert min(exp["profile3"]) < 0 and h.min() < 0


# ---------------------------------------------- closed forms / brute force ----
def test_cephes_sine_against_libm():
    """The oracle's sine (the Cephes method the reference's vdt::fast_sin
    uses) is within a few ulp of libm over the kick's argument range."""
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-50, 50, 20000), np.linspace(-7, 7, 2001)])
    dE = np.zeros_like(x)
    ORC.kick(x, dE, np.array([1.0]), np.array([1.0]), np.array([0.0]), 1.0, 1, 0.0)
    assert np.max(np.abs(dE - np.sin(x))) <= 4e-16 * 4
    # octant symmetry: sin(-x) = -sin(x) exactly
    e2 = np.zeros_like(x)
    ORC.kick(-x, e2, np.array([1.0]), np.array([1.0]), np.array([0.0]), 1.0, 1, 0.0)
    np.testing.assert_array_equal(e2, -dE)


def test_kick_closed_form_two_harmonics():
    dt, dE = G["dt"], G["dE"]
    V, w, ph = np.array([6e6, 1.5e6]), np.array([W1, 4 * W1]), np.array([0.1, 0.3])
    e = dE.copy()
    ORC.kick(dt, e, V, w, ph, -1.0, 2, 7.0)
    ref = dE + (-1.0 * (V[0] * np.sin(w[0] * dt + ph[0]) + V[1] * np.sin(w[1] * dt + ph[1])) + 7.0)
    np.testing.assert_allclose(e, ref, rtol=0, atol=1e-15 * np.abs(dE).max() + 1e-8)


def test_drift_closed_forms():
    """unittests/trackers/test_drift.py:40-96: linear, legacy and exact drifts
    as functions of dE (textbook slippage / momentum relations)."""
    dE = np.linspace(-0.1, 0.1, 2001) * 160e6
    T, beta, E = 1e-6, 0.5, 1.1e9
    a0, a1, a2 = 0.06, 0.6, 6.0
    eta0, eta1, eta2 = 0.05, 0.2, -0.3
    d = dE / (beta ** 2 * E)
    cases = {
        ("simple", 0): T * eta0 * d,
        ("legacy", 2): T * (1 / (1 - eta0 * d - eta1 * d ** 2 - eta2 * d ** 3) - 1),
    }
    delta = np.sqrt(1 + (1 / beta ** 2) * ((dE / E) ** 2 + 2 * dE / E)) - 1
    cases[("exact", 2)] = T * ((1 + a0 * delta + a1 * delta ** 2 + a2 * delta ** 3)
                                * (1 + dE / E) / (1 + delta) - 1)
    for (solver, ao), ref in cases.items():
        t = np.zeros_like(dE)
        ORC.drift(t, dE, solver, T, 1.0, ao, eta0, eta1, eta2, a0, a1, a2, beta, E)
        np.testing.assert_allclose(t, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_histogram_brute_force():
    rng = np.random.default_rng(3)
    for ns in (1, 7, 100):
        x = rng.uniform(-0.2, 1.2, 5000)
        x[:5] = [0.0, 1.0, 0.5, np.nan, -0.0]
        h = np.zeros(ns)
        ORC.slice_beam(x, h, 0.0, 1.0)
        ref = np.zeros(ns)
        for v in x:
            if np.isfinite(v):
                b = math.floor((v - 0.0) * (ns / 1.0))
                if 0 <= b < ns:
                    ref[b] += 1
        np.testing.assert_array_equal(h, ref)


def test_smooth_histogram_weight_rule():
    """histogram.cpp:99-116 on single particles: d <= 0.5 gives 0.5-d to the
    own bin and 0.5+d to the bin below; d > 0.5 gives 0.5-d (negative) to the
    own bin and 0.5+d to the bin above; outer half-bins dropped."""
    ns, cl, cr = 10, 0.0, 10.0
    for f, own, other in [(3.25, 0.25, 0.75), (3.75, -0.25, 1.25), (3.5, 0.0, 1.0)]:
        h = np.zeros(ns)
        ORC.slice_smooth(np.array([f]), h, cl, cr)
        k = int(f)
        d = f - k
        k2 = k + 1 if d > 0.5 else k - 1
        exp = np.zeros(ns)
        exp[k] += own
        exp[k2] += other
        np.testing.assert_allclose(h, exp, atol=1e-15)
    for f in (0.2, 9.8):
        h = np.zeros(ns)
        ORC.slice_smooth(np.array([f]), h, cl, cr)
        assert not h.any()


def test_linear_interp_kick_closed_form():
    ns = 64
    bc = np.linspace(0, 1, ns)
    V = np.cos(3 * bc) * 1e3
    x = np.random.default_rng(9).uniform(-0.1, 1.1, 5000)
    e = np.zeros_like(x)
    ORC.linear_interp_kick(x, e, V, bc, -2.0, 5.0)
    inside = (x >= bc[0]) & (x < bc[-1])
    ref = np.where(inside, -2.0 * np.interp(x, bc, V) + 5.0, 0.0)
    np.testing.assert_allclose(e, ref, rtol=0, atol=1e-9)


def test_resonator_closed_form():
    f = np.linspace(0, 3e9, 301)
    R, fr, Q = np.array([5e3, 2e4]), np.array([1e9, 3e8]), np.array([1.0, 50.0])
    z = ORC.fast_resonator(R, Q, f, fr)
    ref = np.zeros_like(f, dtype=complex)
    for r, a, q in zip(R, fr, Q):
        ref[1:] += r / (1 + 1j * q * (f[1:] / a - a / f[1:]))
    np.testing.assert_allclose(z, ref, rtol=1e-13, atol=0)
    assert z[0] == 0
    # at resonance the impedance is real and equal to R (plus the other term)
    k = np.argmin(np.abs(f - 1e9))
    assert abs(z[k].real - (5e3 + (2e4 / (1 + 1j * 50 * (1e9 / 3e8 - 3e8 / 1e9))).real)) < 1e-9


def test_next_regular_brute_force_and_reference_values():
    def smooth5(m):
        while True:
            r = m
            for p in (2, 3, 5):
                while r % p == 0:
                    r //= p
            if r == 1:
                return m
            m += 1
    for t in range(1, 3000):
        exp = t if (t <= 6 or not (t & (t - 1))) else 2 * smooth5(-(-t // 2))
        assert orc.next_regular(t) == exp
    # test_impedance.py:57, 71, 78
    assert orc.next_regular(16) == 16 and orc.next_regular(1067) == 1080


def test_induced_voltage_numpy_vs_direct_dft():
    """The oracle's FFT convolution equals the circular convolution of the
    profile with irfft(Z), computed here by a direct O(n^2) DFT."""
    n_fft, ns = 64, 24
    rng = np.random.default_rng(4)
    p = rng.integers(0, 50, ns).astype(float)
    f = np.fft.rfftfreq(n_fft, 1e-9)
    Z = orc.resonator_impedance(f, [1e3], [1e8], [3.0]) / 1e-9
    v = orc.induced_voltage_1turn(p, Z, n_fft, 1.0, 2.5, n_fft, ns)
    k = np.arange(n_fft)
    Zfull = np.concatenate([Z, np.conj(Z[-2:0:-1])])
    Zfull[n_fft // 2] = Z[n_fft // 2].real
    h = np.real(np.exp(2j * np.pi * np.outer(k, k) / n_fft) @ Zfull) / n_fft
    ref = np.array([sum(p[j] * h[(i - j) % n_fft] for j in range(ns)) for i in range(ns)])
    ref = -(orc.E_CHARGE * 2.5) * ref
    np.testing.assert_allclose(v, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


# -------------------------------------------- oracle vs reference builds ----
def test_oracle_matches_reference_builds_when_available():
    strict = orc.ref_lib("strict")
    opt = orc.ref_lib("opt")
    if strict is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(12)
    dt = rng.normal(1.25e-9, 0.3e-9, 50001)
    dE = rng.normal(0, 1e8, 50001)
    for lib in (strict, opt):
        e1, e2 = dE.copy(), dE.copy()
        ORC.kick(dt, e1, np.array([6e6, 1e6]), np.array([W1, 4 * W1]), np.array([0.0, 1.0]), 1.0,
                 2, 3.0)
        lib.kick(dt, e2, np.array([6e6, 1e6]), np.array([W1, 4 * W1]), np.array([0.0, 1.0]), 1.0,
                 2, 3.0)
        if lib is strict:
            np.testing.assert_array_equal(e1, e2)
        else:  # fast-math + FMA build
            np.testing.assert_allclose(e1, e2, rtol=0, atol=1e-12 * np.abs(e2).max())
        h1, h2 = np.zeros(4096), np.zeros(4096)
        ORC.slice_beam(dt, h1, 0.0, 2.5e-9)
        lib.slice_beam(dt, h2, 0.0, 2.5e-9)
        np.testing.assert_array_equal(h1, h2)


def test_oracle_header_states_its_role():
    src = open(os.path.join(os.path.dirname(HERE), "oracle", "blond_oracle.c")).read()
    assert "TEST INFRASTRUCTURE ONLY" in src
