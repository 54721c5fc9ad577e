"""The oracle's multi-turn wake and inductive-source steps pinned on the CPU:
exact brute force on grids where every interpolation point is a sample
(so a shift in the wrong direction, a dropped front-wake cut or a wrong
summation order fails bit for bit), closed forms of the profile derivative,
and the host-side set-up of the sources.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as orc


def test_mtw_shift_is_toward_later_times_exact():
    """time = 0, 1, 2, ... and t_rev = 3: the memory after a turn holds at i
    what was at i + 3 (the wake seen one revolution later), zeros past the
    end, plus this turn's voltage."""
    n = 12
    tm = orc.mtw_time_axis(12.0, n)
    np.testing.assert_array_equal(tm, np.arange(n, dtype=np.float64))
    mem = np.arange(1.0, n + 1.0)
    v = np.full(n, 0.5)
    new, volt = orc.mtw_update(mem, tm, 3.0, v, n, 0)
    expect = np.concatenate([mem[3:], np.zeros(3)]) + 0.5
    np.testing.assert_array_equal(new, expect)
    np.testing.assert_array_equal(volt, expect)


def test_mtw_memory_is_sum_of_shifted_turns_exact():
    """With a fixed one-turn voltage v and t_rev = k samples, after N turns
    the memory is sum_j v[i + j k] over the turns still inside the window,
    accumulated oldest first (brute force)."""
    rng = np.random.default_rng(3)
    n, k, turns = 40, 7, 9
    tm = orc.mtw_time_axis(float(n), n)
    v = rng.standard_normal(n)
    mem = np.zeros(n)
    for _ in range(turns):
        mem, _ = orc.mtw_update(mem, tm, float(k), v, n, 0)
    expect = np.zeros(n)
    for i in range(n):
        acc = 0.0
        for j in range(turns - 1, -1, -1):
            if i + j * k < n:
                acc = acc + v[i + j * k]
        expect[i] = acc
    np.testing.assert_array_equal(mem, expect)


def test_mtw_front_wake_cut():
    """front_wake_buffer = b: the last b points of this turn's voltage do not
    enter the memory (impedance.py:506-508)."""
    n = 16
    tm = orc.mtw_time_axis(16.0, n)
    mem0 = np.zeros(n)
    v = np.ones(n)
    new, _ = orc.mtw_update(mem0, tm, 0.0, v, n, 5)
    np.testing.assert_array_equal(new, np.r_[np.ones(n - 5), np.zeros(5)])
    new0, _ = orc.mtw_update(mem0, tm, 0.0, v, n, 0)
    np.testing.assert_array_equal(new0, np.ones(n))


def test_mtw_fractional_shift_is_linear_interpolation():
    """A memory linear in time stays linear under a fractional shift (up to
    rounding) and is 0 past the end of the window."""
    n = 64
    wl = 64 * 2.5e-9
    tm = orc.mtw_time_axis(wl, n)
    mem = 3.0 + 1e9 * tm
    t_rev = 10.37 * 2.5e-9
    new, _ = orc.mtw_update(mem, tm, t_rev, np.zeros(n), n, 0)
    inside = tm + t_rev <= tm[-1]
    np.testing.assert_allclose(new[inside], 3.0 + 1e9 * (tm[inside] + t_rev), rtol=1e-14)
    np.testing.assert_array_equal(new[~inside], 0.0)
    assert inside.sum() == n - 11


def test_profile_derivative_closed_forms():
    bc = (np.arange(50) + 0.5) * 2e-9
    h = bc[1] - bc[0]
    lin = 7.0 * np.arange(50.0)
    g = orc.profile_derivative(lin, bc, "gradient")
    np.testing.assert_allclose(g, 7.0 / h, rtol=1e-12)
    d = orc.profile_derivative(lin, bc, "diff")
    np.testing.assert_allclose(d, 7.0 / h, rtol=1e-12)
    # quadratic: central differences are exact inside, one-sided at the ends
    q = np.arange(50.0) ** 2
    g = orc.profile_derivative(q, bc, "gradient")
    np.testing.assert_allclose(g[1:-1], 2 * np.arange(1, 49.0) / h, rtol=1e-12)
    assert g[0] == pytest.approx(1.0 / h) and g[-1] == pytest.approx(97.0 / h)
    # diff mode: midpoint slopes interpolated back onto the centres
    dq = orc.profile_derivative(q, bc, "diff")
    mids = (2 * np.arange(49.0) + 1) / h
    np.testing.assert_allclose(dq[1:-1], (mids[:-1] + mids[1:]) / 2, rtol=1e-12)
    assert dq[0] == pytest.approx(mids[0]) and dq[-1] == pytest.approx(mids[-1])
    with pytest.raises(RuntimeError):
        orc.profile_derivative(q, bc, "spline")


def test_inductive_voltage_of_a_ramp_is_constant():
    """V = -(q e/(2 pi) ratio Z/n t_rev / bin_size * dp/dt): a linear profile
    gives a constant voltage of the opposite sign to Z/n for q > 0."""
    bc = (np.arange(32) + 0.5) * 1e-9
    p = 3.0 * np.arange(32.0)
    zon = np.full(5, 100.0)
    trev = np.full(5, 1e-6)
    v = orc.inductive_voltage(p, bc, zon, trev, 2, 1.0, 1e9, 1e-9, "gradient")
    expect = -(orc.E_CHARGE / (2 * np.pi) * 1e9 * 100.0 * 1e-6 / 1e-9 * (3.0 / 1e-9))
    np.testing.assert_allclose(v, expect, rtol=1e-12)


def test_sources_sum_single_convolution_matches_one_turn():
    """The per-source loop with one single-turn convolution source is the
    one-turn formula of impedance.py:457-489."""
    rng = np.random.default_rng(5)
    ns, n_fft = 64, 256
    prof = rng.integers(0, 50, ns).astype(np.float64)
    Z = rng.standard_normal(n_fft // 2 + 1) + 1j * rng.standard_normal(n_fft // 2 + 1)
    st = {"n_slices": ns, "charge": 1.0, "ratio": 1e9, "bin_centers": np.arange(ns) + 0.5,
          "t_rev": np.ones(4), "bin_size": 1.0,
          "sources": [{"kind": "convolution", "Z": Z, "n_fft": n_fft,
                       "n_induced_voltage": n_fft, "multi_turn_wake": False}]}
    v = orc.induced_voltage_sources(st, prof, 0, [None])
    ref = orc.induced_voltage_1turn(prof, Z, n_fft, 1.0, 1e9, n_fft, ns)
    np.testing.assert_array_equal(v, ref)


def test_inductive_workload_setup_on_host():
    """InductiveImpedance set-up needs no GPU: sizes and the per-turn
    coefficient in the reference's left-to-right order."""
    from blond_amd.constants import e
    from blond_amd.workloads import build_wake_workload

    w = build_wake_workload("zovern", n_particles=2000, n_turns=6, to_gpu=False)
    src = w.tiv.induced_voltage_list[0]
    assert src.n_induced_voltage == w.profile.n_slices
    c = (w.beam.particle.charge * e / (2 * np.pi) * w.beam.ratio * -100.0
         * w.rf.t_rev[3] / w.profile.bin_size)
    assert src.coefficient(3) == c
    st = w.snapshot()
    assert st["sources"][0]["kind"] == "inductive"
    r = orc.track(st, 3)
    assert np.isfinite(r["dE"]).all() and np.abs(r["induced_voltage"]).max() > 0


def test_mtw_freq_mode_raises_as_reference():
    """The reference's 'freq' multi-turn mode reads an attribute nothing
    sets (impedance.py:414-418): AttributeError."""
    from blond_amd.impedances.impedance import _InducedVoltage
    from blond_amd.workloads import build_wake_workload

    w = build_wake_workload("zovern", n_particles=2000, n_turns=6, to_gpu=False)
    with pytest.raises(AttributeError, match="buffer_extra"):
        _InducedVoltage(w.beam, w.profile, multi_turn_wake=True, mtw_mode="freq",
                        rf_station=w.rf)
    with pytest.raises(ValueError):
        _InducedVoltage(w.beam, w.profile, multi_turn_wake=True, mtw_mode="fft",
                        rf_station=w.rf)
