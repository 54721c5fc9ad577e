"""GPU parity: every HIP entry point against the CPU oracle on seeded inputs.

Bars: bit-identical results where the arithmetic is the same IEEE sequence
(kick, drift simple/legacy, histogram counts, interp kick, rf_volt_comp);
1e-12 relative where a correctly rounded sqrt or a different summation order
enters (exact drift, smooth histogram); FFT-based results within 1e-10 of
their maximum (the induced-voltage tolerance of the tracking path).
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.utils import bmath as bm  # noqa: E402

ORC = orc.oracle_lib()
W1 = 2.5182298872245054e9  # omega_rf of the LHC-like ring, h = 35640


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    return t.detach().cpu().numpy()


def bunch(n, seed=3, center=1.25e-9, sdt=0.3e-9, sdE=1e8):
    rng = np.random.default_rng(seed)
    return rng.normal(center, sdt, n), rng.normal(0.0, sdE, n)


@pytest.mark.parametrize("n", [0, 1, 7, 1000, 100001])
@pytest.mark.parametrize("n_rf", [1, 2, 3])
def test_kick_bit_exact(n, n_rf):
    dt, dE = bunch(n)
    V = np.array([6e6, 1.5e6, 3e5])[:n_rf]
    w = np.array([W1, 4 * W1, 2 * W1])[:n_rf]
    ph = np.array([0.0, 0.3, -2.0])[:n_rf]
    ref = dE.copy()
    ORC.kick(dt, ref, V, w, ph, 1.0, n_rf, -1234.5)
    g_dE = dev(dE)
    bm.kick(dev(dt), g_dE, dev(V), dev(w), dev(ph), 1.0, n_rf, -1234.5)
    np.testing.assert_array_equal(host(g_dE), ref)


def test_kick_unaligned_views_and_numpy_staging():
    dt, dE = bunch(10001)
    V, w, ph = np.array([6e6]), np.array([W1]), np.array([0.1])
    ref = dE[1:].copy()
    ORC.kick(dt[1:].copy(), ref, V, w, ph, -1.0, 1, 0.0)
    g_dt, g_dE = dev(dt), dev(dE)
    bm.kick(g_dt[1:], g_dE[1:], dev(V), dev(w), dev(ph), -1.0, 1, 0.0)  # 8-byte offset views
    np.testing.assert_array_equal(host(g_dE)[1:], ref)
    h = dE[1:].copy()
    bm.kick(dt[1:].copy(), h, V, w, ph, -1.0, 1, 0.0)  # numpy in/out, staged through HBM
    np.testing.assert_array_equal(h, ref)


def test_kick_large_phase_arguments():
    """Arguments far outside [-pi, pi] exercise the octant reduction."""
    rng = np.random.default_rng(11)
    dt = rng.uniform(-2e-6, 2e-6, 50001)
    dE = np.zeros_like(dt)
    V, w, ph = np.array([1.0, 2.0]), np.array([W1, 4 * W1]), np.array([0.0, 1.0])
    ref = dE.copy()
    ORC.kick(dt, ref, V, w, ph, 1.0, 2, 0.0)
    g = dev(dE)
    bm.kick(dev(dt), g, dev(V), dev(w), dev(ph), 1.0, 2, 0.0)
    np.testing.assert_array_equal(host(g), ref)
    # and the sine itself is within a few ulp of libm
    x = w[0] * dt
    np.testing.assert_allclose(ref, np.sin(x) + 2 * np.sin(w[1] * dt + 1.0), rtol=0, atol=5e-12)


DRIFT_ARGS = dict(t_rev=8.892465516509709e-05, length_ratio=1.0, eta_0=3.172867586042721e-04,
                  eta_1=2e-6, eta_2=-3e-7, alpha_0=3.2e-4, alpha_1=1e-5, alpha_2=1e-6,
                  beta=0.9999978262922387, energy=450.0009780e9)


@pytest.mark.parametrize("solver", ["simple", "legacy", "exact"])
@pytest.mark.parametrize("alpha_order", [0, 1, 2])
def test_drift(solver, alpha_order):
    dt, dE = bunch(100003)
    a = DRIFT_ARGS
    args = (solver, a["t_rev"], a["length_ratio"], alpha_order, a["eta_0"], a["eta_1"],
            a["eta_2"], a["alpha_0"], a["alpha_1"], a["alpha_2"], a["beta"], a["energy"])
    ref = dt.copy()
    ORC.drift(ref, dE, *args)
    g = dev(dt)
    bm.drift(g, dev(dE), *args)
    if solver == "exact":
        # sqrt enters: allow a few ulp of the drift increment
        np.testing.assert_allclose(host(g) - dt, ref - dt, rtol=1e-12, atol=1e-12 * np.abs(ref - dt).max())
    else:
        np.testing.assert_array_equal(host(g), ref)


@pytest.mark.parametrize("n_slices", [1, 100, 4096, 20000, 65536])
def test_histogram_counts_identical(n_slices):
    dt, _ = bunch(300001)
    cl, cr = 0.0, 2.495080111254127e-09
    # edge cases: exact edges, outside, NaN, +-inf
    bw = (cr - cl) / n_slices
    extra = np.array([cl, cr, cl - 1e-20, cr - 1e-25, cl + bw, np.nan, np.inf, -np.inf, 1e30,
                      -1e30, cl + 0.5 * bw])
    dt = np.concatenate([dt, extra])
    ref = np.zeros(n_slices)
    ORC.slice_beam(dt, ref, cl, cr)
    g = torch.zeros(n_slices, dtype=torch.float64, device="cuda")
    bm.slice_beam(dev(dt), g, cl, cr)
    np.testing.assert_array_equal(host(g), ref)


def test_histogram_empty_and_repeat_determinism():
    g = torch.full((100,), 7.0, dtype=torch.float64, device="cuda")
    bm.slice_beam(torch.zeros(0, dtype=torch.float64, device="cuda"), g, 0.0, 1.0)
    assert host(g).sum() == 0.0  # output overwritten, not accumulated
    dt, _ = bunch(200000)
    outs = []
    for _ in range(3):
        bm.slice_beam(dev(dt), g, 0.0, 2.5e-9)
        outs.append(host(g).copy())
    assert all(np.array_equal(outs[0], o) for o in outs)


@pytest.mark.parametrize("n_slices", [8, 150, 256, 9000])
def test_smooth_histogram(n_slices):
    dt, _ = bunch(200001)
    cl, cr = 0.0, 2.5e-9
    ref = np.zeros(n_slices)
    ORC.slice_smooth(dt, ref, cl, cr)
    g = torch.zeros(n_slices, dtype=torch.float64, device="cuda")
    bm.slice_smooth(dev(dt), g, cl, cr)
    got = host(g)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))
    assert abs(got.sum() - ref.sum()) < 1e-9


@pytest.mark.parametrize("n_slices", [2, 100, 4096])
def test_linear_interp_kick_bit_exact(n_slices):
    dt, dE = bunch(100001)
    bc = np.linspace(0, 2.5e-9, n_slices + 1)[:-1] + 2.5e-9 / (2 * n_slices)
    rng = np.random.default_rng(5)
    V = rng.normal(0, 1e3, n_slices)
    dt = np.concatenate([dt, [bc[0], bc[-1], bc[0] - 1e-18, np.nan, -1.0]])
    dE = np.concatenate([dE, np.zeros(5)])
    ref = dE.copy()
    ORC.linear_interp_kick(dt, ref, V, bc, 1.0, 12.5)
    g = dev(dE)
    bm.linear_interp_kick(dev(dt), g, dev(V), dev(bc), 1.0, 12.5)
    np.testing.assert_array_equal(host(g)[:-2], ref[:-2])
    assert np.isnan(host(g)[-2]) == np.isnan(ref[-2])


def test_linear_interp_kick_on_two_streams():
    """The interpolation table is library workspace shared by every call on
    a device: calls on two streams, queued without a synchronisation between
    them (each stream's first kick behind a long kernel), must still each
    kick by their own table."""
    cases = []
    for k, n_slices in enumerate((4096, 100)):
        dt, dE = bunch(2_000_001 if k == 0 else 100_001)
        bc = np.linspace(0, 2.5e-9, n_slices + 1)[:-1] + 2.5e-9 / (2 * n_slices)
        V = np.random.default_rng(11 + k).normal(0, 1e3, n_slices)
        ref = dE.copy()
        ORC.linear_interp_kick(dt, ref, V, bc, 1.0, 0.0)
        for _ in range(3):
            ORC.linear_interp_kick(dt, ref, V, bc, 1.0, 0.0)
        cases.append((dev(dt), dev(dE), dev(V), dev(bc), ref))
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    for _ in range(4):
        for (gdt, gdE, gV, gbc, _), st in zip(cases, streams):
            with torch.cuda.stream(st):
                bm.linear_interp_kick(gdt, gdE, gV, gbc, 1.0, 0.0)
    torch.cuda.synchronize()
    for gdt, gdE, gV, gbc, ref in cases:
        np.testing.assert_array_equal(host(gdE), ref)


def test_rf_volt_comp_bit_exact():
    bc = np.linspace(0, 2.5e-9, 257)[:-1]
    V, w, ph = np.array([6e6, 1.5e6]), np.array([W1, 4 * W1]), np.array([0.0, 0.2])
    ref = ORC.rf_volt_comp(V, w, ph, bc)
    got = host(bm.rf_volt_comp(dev(V), dev(w), dev(ph), dev(bc)))
    np.testing.assert_array_equal(got, ref)


def test_fast_resonator():
    f = np.fft.rfftfreq(16384, 2.5e-9 / 4096)
    R = np.array([5e3, 1e4, 3e5]); fr = np.array([1e9, 2e8, 5e7]); Q = np.array([1.0, 100.0, 5e3])
    ref = ORC.fast_resonator(R, Q, f, fr)
    got = bm.fast_resonator(R, Q, f, fr)
    np.testing.assert_allclose(got.real, ref.real, rtol=1e-15, atol=0)
    np.testing.assert_allclose(got.imag, ref.imag, rtol=1e-15, atol=1e-300)
    assert got[0] == 0


@pytest.mark.parametrize("n", [8, 16, 1080, 1024, 16384, 2000, 20480])
def test_rfft_irfft_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n)
    X = bm.rfft(dev(x), n)
    Xr = np.fft.rfft(x)
    assert np.max(np.abs(host(X) - Xr)) <= 1e-13 * np.max(np.abs(Xr))
    Y = Xr * (1 + 0.5j)
    y = host(bm.irfft(dev(Y), n))
    yr = np.fft.irfft(Y, n)
    assert np.max(np.abs(y - yr)) <= 1e-13 * np.max(np.abs(yr))
    # zero padding / truncation semantics of numpy's n argument
    np.testing.assert_allclose(host(bm.rfft(dev(x[: n // 3]), n)), np.fft.rfft(x[: n // 3], n),
                               rtol=0, atol=1e-12 * np.abs(Xr).max())


@pytest.mark.parametrize("n_slices,n_fft", [(40, 64), (1000, 1024), (1500, 2048), (1000, 1080),
                                            (4096, 16384), (65536, 131072), (20000, 48000),
                                            (1000, 262144), (100000, 2097152)])
def test_induced_voltage_within_1e10_of_max(n_slices, n_fft):
    """One-workgroup LDS transform up to its limit (one wavefront a
    transform for half-lengths 32, 512 and 1024, workgroup-wide stages for
    540 and 8192), the four-step transform beyond it (131072 = the
    2^16-slice workload, 256 x 256; 48000 = 2^7 3 5^3, workgroup-wide
    stages; 262144 with a short profile, 512 x 256; 2^21, 1024 x 1024)."""
    dt, _ = bunch(400000)
    prof = np.zeros(n_slices)
    ORC.slice_beam(dt, prof, 0.0, 2.5e-9)
    bin_size = 2.5e-9 / n_slices
    f = np.fft.rfftfreq(n_fft, bin_size)
    Z = orc.resonator_impedance(f, [5e3, 1e5], [1e9, 2e8], [1.0, 300.0]) / bin_size
    ref = orc.induced_voltage_1turn(prof, Z, n_fft, 1.0, 1e13 / 4e5, n_fft, n_slices)
    scale = 1.0 * orc.E_CHARGE * (1e13 / 4e5)
    got = host(bm.induced_voltage_1turn(dev(prof), dev(Z), n_fft, scale))
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"\ninduced voltage n_fft={n_fft}: max error / max |V| = {err:.2e}")
    assert err <= 1e-10
    assert err <= 1e-13  # what the kernels achieve


def test_induced_voltage_fourstep_full_length_output():
    """n_out = n_fft (the whole irfft) through the four-step transform."""
    n_slices, n_fft = 30000, 65536 * 2
    rng = np.random.default_rng(5)
    prof = rng.integers(0, 50, n_slices).astype(np.float64)
    Z = rng.normal(size=n_fft // 2 + 1) + 1j * rng.normal(size=n_fft // 2 + 1)
    Z[0] = 0
    ref = -(2.5 * np.fft.irfft(Z * np.fft.rfft(prof, n_fft), n_fft))
    got = host(bm.induced_voltage_1turn(dev(prof), dev(Z), n_fft, 2.5, n_out=n_fft))
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


# ---------------------------------------------------------------------------
# FP32 mode (the reference's libblond_single.so: real_t = float, fast_sinf)
# ---------------------------------------------------------------------------
REF32 = orc.ref_lib("strict", "single")
needs_ref32 = pytest.mark.skipif(REF32 is None, reason="oracle/_ref single build not present")


@pytest.fixture
def single():
    bm.use_precision("single")
    yield
    bm.use_precision("double")


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


@needs_ref32
@pytest.mark.parametrize("n_rf", [1, 2])
def test_f32_kick_drift_bit_exact_vs_reference_single(single, n_rf):
    dt, dE = (f32(a) for a in bunch(100001))
    V, w, phi = f32([6e6, 1.5e6][:n_rf]), f32([2.5182e9, 1.00728e10][:n_rf]), f32([0.0, 0.3][:n_rf])
    rdt, rdE = dt.copy(), dE.copy()
    REF32.kick(rdt, rdE, V, w, phi, 1.0, n_rf, -12.5)
    REF32.drift(rdt, rdE, "simple", 8.892465516509709e-05, 1.0, 0.0, 3.17e-4, 0.0, 0.0, 3.2e-4,
                0.0, 0.0, 0.99999782, 450e9)
    gdt, gdE = dev(dt), dev(dE)
    bm.kick(gdt, gdE, dev(V), dev(w), dev(phi), 1.0, n_rf, -12.5)
    bm.drift(gdt, gdE, "simple", 8.892465516509709e-05, 1.0, 0.0, 3.17e-4, 0.0, 0.0, 3.2e-4,
             0.0, 0.0, 0.99999782, 450e9)
    np.testing.assert_array_equal(host(gdE), rdE)
    np.testing.assert_array_equal(host(gdt), rdt)


@needs_ref32
def test_f32_histogram_and_interp_kick_vs_reference_single(single):
    dt, dE = (f32(a) for a in bunch(200001))
    ref = np.zeros(1000, np.float32)
    REF32.slice_beam(dt, ref, 0.0, 2.5e-9)
    got = dev(np.zeros(1000, np.float32))
    bm.slice_beam(dev(dt), got, 0.0, 2.5e-9)
    np.testing.assert_array_equal(host(got), ref)
    bc = f32((np.arange(1000) + 0.5) * 2.5e-12)
    V = f32(np.sin(np.arange(1000) / 50.0) * 1e4)
    rdE = dE.copy()
    REF32.linear_interp_kick(dt, rdE, V, bc, 1.0, 0.0)
    gdE = dev(dE)
    bm.linear_interp_kick(dev(dt), gdE, dev(V), dev(bc), 1.0, 0.0)
    np.testing.assert_allclose(host(gdE), rdE, rtol=0, atol=2e-7 * np.abs(rdE).max())


# ---------------------------------------------------------------------------
# bunch moments (beam.py:430-458)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 1000, 1000003])
def test_beam_statistics_vs_numpy(n):
    dt, dE = bunch(n, seed=9)
    ids = np.arange(1, n + 1, dtype=np.int64)
    ids[::7] = 0  # lost particles are excluded
    got = bm.beam_statistics(dev(dt), dev(dE), dev(ids))
    a = ids != 0
    want = (int(a.sum()), np.mean(dt[a]), np.std(dt[a]), np.mean(dE[a]), np.std(dE[a]))
    assert got[0] == want[0]
    if want[0]:
        np.testing.assert_allclose(got[1:], want[1:], rtol=1e-13, atol=0)
    # without ids every particle counts; repeated calls are identical
    g1 = bm.beam_statistics(dev(dt), dev(dE))
    g2 = bm.beam_statistics(dev(dt), dev(dE))
    assert g1 == g2 and g1[0] == n
    np.testing.assert_allclose(g1[1:], (np.mean(dt), np.std(dt), np.mean(dE), np.std(dE)),
                               rtol=1e-13, atol=0)
