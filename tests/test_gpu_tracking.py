"""GPU parity of whole turns: the object-by-object map (RingAndRFTracker,
Profile, TotalInducedVoltage on the MI355X) and the fused engine against the
oracle's turn loop (the reference's map restated on the CPU)."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd import _lib  # noqa: E402

ORC_LIB = orc.oracle_lib()
from blond_amd.workloads import build_workload  # noqa: E402


def host(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def test_device_info_printed():
    import ctypes as ct

    cus, lds = ct.c_int(), ct.c_int()
    _lib.call("blmi_device_info", ct.byref(cus), ct.byref(lds))
    n = ct.c_int64()
    _lib.call("blmi_fft_max_lds_length", ct.byref(n))
    print(f"\nCUs={cus.value} max LDS/workgroup={lds.value} max LDS FFT length={n.value}")
    assert cus.value > 0 and lds.value >= 65536


@pytest.mark.parametrize("n_particles", [20000, 20001])
def test_plain_turns_fused_bit_exact(n_particles):
    w = build_workload("c1", n_particles=n_particles, n_turns=40, n_slices=100)
    ref = orc.track(w.snapshot(), 25)
    w.fused(25)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    assert w.rf.counter[0] == 25


def test_plain_turns_object_by_object_bit_exact():
    w = build_workload("c1", n_particles=30001, n_turns=20, n_slices=100)
    ref = orc.track(w.snapshot(), 10)
    for _ in range(10):
        for m in w.map():
            m.track()
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])


def test_plain_turns_graph_vs_eager_identical():
    a = build_workload("c1", n_particles=50000, n_turns=130, n_slices=100)
    b = build_workload("c1", n_particles=50000, n_turns=130, n_slices=100)
    a.fused(0)
    fa = a.fused(120)  # 2 graph replays of 50 + 20 eager turns
    from blond_amd.trackers.fused import FusedTracker

    fb = FusedTracker(b.tracker, b.profile, None, use_graph=False)
    fb.track(120)
    np.testing.assert_array_equal(host(a.beam.dt), host(b.beam.dt))
    np.testing.assert_array_equal(host(a.beam.dE), host(b.beam.dE))
    np.testing.assert_array_equal(host(a.profile.n_macroparticles), host(b.profile.n_macroparticles))
    del fa


def _check_intensity(w, ref, rtol):
    """With intensity effects the induced voltage of the GPU (direct
    convolution or LDS FFT) and of numpy's pocketfft differ by ~1e-16 of its
    maximum; fed back through dE and dt, that moves the odd particle across a
    histogram edge, and one count moved changes the wake locally.  Multi-turn
    parity is therefore held statistically: moments (mean, rms of dt and dE)
    to `rtol` = 1e-9 relative, at least 99.9% of particles to `rtol`, every
    particle to 1e-6, profiles to a few counts."""
    dt, dE = host(w.beam.dt), host(w.beam.dE)
    for a, b in ((dt, ref["dt"]), (dE, ref["dE"])):
        sc = np.abs(b).max()
        err = np.abs(a - b) / sc
        assert np.mean(err > rtol) <= 1e-3, f"{np.mean(err > rtol):.2e} of particles above {rtol}"
        assert err.max() <= 1e-6
        assert abs(a.mean() - b.mean()) <= rtol * max(abs(b.mean()), b.std())
        assert abs(a.std() - b.std()) <= rtol * b.std()
    dp = np.abs(host(w.profile.n_macroparticles) - ref["profile"]).sum()
    assert dp <= 8, f"profiles differ by {dp} counts"
    v = host(w.tiv.induced_voltage)
    assert np.max(np.abs(v - ref["induced_voltage"])) <= 1e-10 * np.abs(ref["induced_voltage"]).max()


@pytest.mark.parametrize("n_slices,n_fft,method", [(256, 1024, None), (4096, 16384, None),
                                                   (4096, 16384, "fft"),
                                                   (4096, 16384, "fourstep"),
                                                   (256, 1024, "fourstep")])
def test_intensity_turns_fused(n_slices, n_fft, method):
    """Fused intensity turns against the oracle's; the four-step transform
    at n_fft 16384 (64 x 128, 4 columns a workgroup) and 1024 (32 x 16, 8
    columns a workgroup: pass C's 9 transforms on 8 wavefronts)."""
    w = build_workload("c2", n_particles=40001, n_turns=30, n_slices=n_slices, n_fft=n_fft)
    ref = orc.track(w.snapshot(), 20)
    ft = w.fused(20, method=method)
    assert ft.method == (method or "direct")
    _check_intensity(w, ref, 1e-9)


def test_intensity_turns_fused_2e16_slices():
    """The large-bunch shape: 65536 slices (u16-pair LDS histogram), FFT of
    131072 (four-step transform), interpolation evaluated from V itself."""
    w = build_workload("c3", n_particles=200001, n_turns=12, n_resonators=20)
    ref = orc.track(w.snapshot(), 8)
    ft = w.fused(8)
    assert ft.method == "fourstep"
    _check_intensity(w, ref, 1e-9)


@pytest.mark.parametrize("n_particles", [300001, 17_000_001])
def test_intensity_first_turn_exact_2e16_slices(n_particles):
    """The first fused c3 turn against the oracle's; at 17e6 macroparticles
    (272 MB of coordinates, above the last-level cache) the particle kernel
    reads and writes them with non-temporal loads and stores."""
    w = build_workload("c3", n_particles=n_particles, n_turns=4)
    ref = orc.track(w.snapshot(), 1)
    w.fused(1)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    v = host(w.tiv.induced_voltage)
    assert np.max(np.abs(v - ref["induced_voltage"])) <= 1e-10 * np.abs(ref["induced_voltage"]).max()
    sc = np.abs(ref["dE"]).max()
    np.testing.assert_allclose(host(w.beam.dE), ref["dE"], rtol=0, atol=1e-12 * sc)


@pytest.mark.parametrize("n_particles", [100001, 17_000_001])
def test_plain_turns_2e16_slices_bit_exact(n_particles):
    """65536 slices count in u16 pairs in LDS; above 65535 particles per
    workgroup (17e6 on 256 CUs) the grid grows so no counter can overflow."""
    w = build_workload("c1", n_particles=n_particles, n_turns=4, n_slices=65536)
    ref = orc.track(w.snapshot(), 2)
    w.fused(2)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])



@pytest.mark.parametrize("n_slices", [16384, 16385, 24576, 32768])
def test_plain_turns_histogram_mode_boundary_bit_exact(n_slices):
    """u32 LDS counts up to 16384 slices, u16 pairs with partial rows above
    (where u32 would still fit in LDS): both sides of the switch."""
    w = build_workload("c1", n_particles=300001, n_turns=4, n_slices=n_slices)
    ref = orc.track(w.snapshot(), 3)
    w.fused(3)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])

def test_intensity_turns_object_by_object():
    w = build_workload("c2", n_particles=40001, n_turns=12, n_slices=512, n_fft=2048)
    ref = orc.track(w.snapshot(), 8)
    for _ in range(8):
        for m in w.map():
            m.track()
    _check_intensity(w, ref, 1e-9)


def test_intensity_first_turn_exact_profile_and_voltage():
    """One turn from identical coordinates: the profile is bit-identical and
    the induced voltage within 1e-10 of its maximum (the per-kernel bars)."""
    w = build_workload("c2", n_particles=100001, n_turns=5, n_slices=4096, n_fft=16384)
    ref = orc.track(w.snapshot(), 1)
    w.fused(1)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    v = host(w.tiv.induced_voltage)
    assert np.max(np.abs(v - ref["induced_voltage"])) <= 1e-10 * np.abs(ref["induced_voltage"]).max()
    sc = np.abs(ref["dE"]).max()
    np.testing.assert_allclose(host(w.beam.dE), ref["dE"], rtol=0, atol=1e-12 * sc)


# ---------------------------------------------------------------------------
# FP32 mode of the fused engine against the reference's single build
# ---------------------------------------------------------------------------
REF32 = orc.ref_lib("strict", "single")
needs_ref32 = pytest.mark.skipif(REF32 is None, reason="oracle/_ref single build not present")


def _to_f32(w):
    w.beam.dt = w.beam.dt.to(torch.float32).contiguous()
    w.beam.dE = w.beam.dE.to(torch.float32).contiguous()


@needs_ref32
@pytest.mark.parametrize("n_slices", [100, 65536])
def test_plain_turns_fused_f32_bit_exact(n_slices):
    """Float particles, float programs and cuts: bit-identical to the
    reference's single-precision kernels (u32 and u16-pair histograms)."""
    w = build_workload("c1", n_particles=100001, n_turns=20, n_slices=n_slices)
    _to_f32(w)
    ref = orc.track(w.snapshot(), 12, REF32)
    ft = w.fused(12)
    assert ft.f32
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    assert w.profile.n_macroparticles.dtype == torch.float32


@needs_ref32
def test_intensity_turns_fused_f32_large_shape():
    """2 RF systems, 65536 slices, FFT 131072, float particles.  The engine
    forms the induced voltage in FP64 (the reference's single mode in
    float32), so the comparison is at float resolution: the first turn's
    profile exactly, moments after 8 turns to 1e-5."""
    w = build_workload("c3", n_particles=200001, n_turns=12)
    _to_f32(w)
    snap = w.snapshot()
    ft = w.fused(1)
    ref1 = orc.track(snap, 1, REF32)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref1["profile"])
    ft.track(7)
    ref = orc.track(snap, 8, REF32)
    for a, b in ((host(w.beam.dt), ref["dt"]), (host(w.beam.dE), ref["dE"])):
        a, b = a.astype(np.float64), b.astype(np.float64)
        assert abs(a.mean() - b.mean()) <= 1e-5 * max(abs(b.mean()), b.std())
        assert abs(a.std() - b.std()) <= 1e-5 * b.std()
    v, vr = host(w.tiv.induced_voltage), ref["induced_voltage"]
    assert np.max(np.abs(v - vr)) <= 1e-4 * np.abs(vr).max()


# ---------------------------------------------------------------------------
# bin-ordered working set (slice counts above 4096)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("sort_every", [0, 1, 3])
def test_ordered_plain_turns_bit_exact(sort_every):
    """Grouping the particles by coarse bin changes nothing but the order the
    engine visits them in: bit-identical to the oracle, caller's order kept."""
    w = build_workload("c1", n_particles=200001, n_turns=20, n_slices=16384)
    ref = orc.track(w.snapshot(), 11, ORC_LIB)
    w.fused(11, sort_every=sort_every)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])


def test_ordered_graph_vs_eager_identical():
    """Captured chunks hold an even number of sorts (chunk 50 -> 48 with a
    sort every 8 turns) and eager remainders continue the pattern."""
    a = build_workload("c3", n_particles=100001, n_turns=130)
    b = build_workload("c3", n_particles=100001, n_turns=130)
    from blond_amd.trackers.fused import FusedTracker

    fa = FusedTracker(a.tracker, a.profile, a.tiv, use_graph=True)
    fa.track(7)
    fa.track(110)
    fb = FusedTracker(b.tracker, b.profile, b.tiv, use_graph=False)
    fb.track(117)
    np.testing.assert_array_equal(host(a.beam.dt), host(b.beam.dt))
    np.testing.assert_array_equal(host(a.beam.dE), host(b.beam.dE))
    np.testing.assert_array_equal(host(a.profile.n_macroparticles), host(b.profile.n_macroparticles))
    fa.close()
    fb.close()


def test_ordered_intensity_matches_unordered():
    """Intensity turns at 65536 slices: ordered and unordered engines agree
    bit for bit (same per-particle arithmetic, integer counts)."""
    a = build_workload("c3", n_particles=150001, n_turns=30)
    b = build_workload("c3", n_particles=150001, n_turns=30)
    a.fused(20, sort_every=4)
    b.fused(20, sort_every=0)
    np.testing.assert_array_equal(host(a.beam.dt), host(b.beam.dt))
    np.testing.assert_array_equal(host(a.beam.dE), host(b.beam.dE))
    np.testing.assert_array_equal(host(a.tiv.induced_voltage), host(b.tiv.induced_voltage))


# ---------------------------------------------------------------------------
# interpolation=True (tracker.py:585-601)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("n_slices", [100, 65536])
def test_interpolation_plain_turns_bit_exact(fused, n_slices):
    """The RF voltage on the bin centres interpolated (rf_volt_comp +
    linear_interp_kick with the acceleration kick), no intensity: the same
    bits as the oracle's turns, fused (table in LDS / in L2) and object by
    object."""
    w = build_workload("c1", n_particles=100001, n_turns=30, n_slices=n_slices,
                       interpolation=True)
    snap = w.snapshot()
    assert snap["interpolation"]
    ref = orc.track(snap, 20, ORC_LIB)
    if fused:
        w.fused(20)
    else:
        for _ in range(20):
            for m in w.map():
                m.track()
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("n_slices,n_fft", [(4096, 16384), (65536, 131072)])
def test_interpolation_intensity_turns(fused, n_slices, n_fft):
    """interpolation=True with intensity effects: the induced voltage is
    summed (not kicked with by itself) and added to the RF voltage on the
    bin centres (the reference's pattern, unittests/impedances/
    test_impedance.py:174-176).  First turn: profile bit-identical, voltage
    within 1e-10 of its maximum, dE within 1e-12; after 20 turns the
    statistical bars of the other intensity tests."""
    name = "c2" if n_slices == 4096 else "c3"
    w = build_workload(name, n_particles=60001, n_turns=30, n_slices=n_slices, n_fft=n_fft,
                       intensity=1e11, interpolation=True)
    snap = w.snapshot()
    ref1 = orc.track(snap, 1, ORC_LIB)

    def turns(k):
        if fused:
            w.fused(k)
        else:
            for _ in range(k):
                for m in w.map():
                    m.track()
    turns(1)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref1["profile"])
    v = host(w.tiv.induced_voltage)
    assert np.max(np.abs(v - ref1["induced_voltage"])) <= 1e-10 * np.abs(ref1["induced_voltage"]).max()
    sc = np.abs(ref1["dE"]).max()
    np.testing.assert_allclose(host(w.beam.dE), ref1["dE"], rtol=0, atol=1e-12 * sc)
    turns(19)
    _check_intensity(w, orc.track(snap, 20, ORC_LIB), 1e-9)


@needs_ref32
@pytest.mark.parametrize("name,n_slices,n_fft", [("c2", 4096, 16384), ("c2", 16384, 32768),
                                                 ("c3", 65536, 131072)])
def test_intensity_turn_f32_bit_exact_given_the_voltage(name, n_slices, n_fft):
    """FP32 mode, one intensity turn.  The reference's single build holds
    the induced voltage as a float array and forms the interpolation table
    in float (linear_interp_kick.cpp:33-48 with real_t = float); the engine
    does the same from its own voltage.  Replaying the turn on the CPU with
    the reference's single-precision kernels from the engine's published
    (float) voltage gives the engine's coordinates bit for bit, and the
    profile is the single build's histogram.  (The voltage itself differs:
    numpy transforms the float profile in single precision, the engine in
    double.)  The float table sits in LDS at 4096 slices, is gathered from
    global memory at 16384 (u32 histogram, table too large for LDS) and
    partly from an LDS window at 65536."""
    w = build_workload(name, n_particles=100001, n_turns=4, n_slices=n_slices, n_fft=n_fft)
    _to_f32(w)
    snap = w.snapshot()
    w.fused(1)
    v32 = host(w.tiv.induced_voltage)
    assert v32.dtype == np.float32
    dt = np.array(snap["dt"], dtype=np.float32)
    dE = np.array(snap["dE"], dtype=np.float32)
    prof = np.zeros(n_slices, dtype=np.float32)
    REF32.slice_beam(dt, prof, snap["cut_left"], snap["cut_right"])
    REF32.linear_interp_kick(dt, dE, v32, snap["bin_centers"], snap["charge"], 0.0)
    REF32.kick(dt, dE, snap["voltage"][:, 0], snap["omega_rf"][:, 0], snap["phi_rf"][:, 0],
               snap["charge"], snap["n_rf"], snap["acc"][0])
    REF32.drift(dt, dE, snap["solver"], snap["t_rev"][1], snap["length_ratio"],
                snap["alpha_order"], snap["eta_0"][1], snap["eta_1"][1], snap["eta_2"][1],
                snap["alpha_0"][1], snap["alpha_1"][1], snap["alpha_2"][1], snap["beta"][1],
                snap["energy"][1])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), prof)
    np.testing.assert_array_equal(host(w.beam.dE), dE)
    np.testing.assert_array_equal(host(w.beam.dt), dt)
