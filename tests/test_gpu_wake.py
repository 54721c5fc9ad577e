"""GPU parity of the wake steps beyond the one-turn convolution: numpy's
interp / gradient / diff on the device (bit-exact), the multi-turn wake
memory update (bit-exact against numpy), and whole turns with a multi-turn
wake (InducedVoltageFreq and InducedVoltageTime, mode 'time'), an inductive
Z/n source ('gradient' and 'diff') and a mixture, object by object and
fused, against the oracle's turn loop."""
import ctypes as ct

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd import _lib  # noqa: E402
from blond_amd.utils import bmath as bm  # noqa: E402
from blond_amd.workloads import WAKE_KINDS, build_wake_workload  # noqa: E402
from tests.test_gpu_tracking import _check_intensity, host  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


# ------------------------------------------------------------ numpy helpers --
@pytest.mark.parametrize("m", [1, 2, 5, 1000, 70001])
def test_interp_bit_exact_random(m):
    rng = np.random.default_rng(m)
    xp = np.sort(rng.uniform(-3, 5, m))
    fp = rng.standard_normal(m)
    x = np.concatenate([rng.uniform(-4, 6, 5000), xp[::7], [xp[0], xp[-1], np.nan, -np.inf,
                                                            np.inf]])
    for left, right in ((None, None), (0.0, 0.0), (-2.5, 7.0)):
        out = bm.interp(dev(x), dev(xp), dev(fp), left, right)
        ref = np.interp(x, xp, fp, left=left, right=right)
        np.testing.assert_array_equal(host(out), ref)


def test_interp_uniform_grid_and_numpy_inputs():
    tm = np.linspace(0, 3.2e-6, 4096, endpoint=False)
    fp = np.sin(tm * 1e7)
    x = tm + 1.2345e-7
    out = bm.interp(x, tm, fp, left=0, right=0)  # numpy in, numpy out
    assert isinstance(out, np.ndarray)
    np.testing.assert_array_equal(out, np.interp(x, tm, fp, left=0, right=0))


@pytest.mark.parametrize("n", [2, 3, 1000, 65536])
def test_gradient_and_diff_bit_exact(n):
    rng = np.random.default_rng(n)
    f = rng.integers(0, 1000, n).astype(np.float64)
    h = 6.53e-9
    np.testing.assert_array_equal(host(bm.gradient(dev(f), h)), np.gradient(f, h))
    np.testing.assert_array_equal(host(bm.diff(dev(f))), np.diff(f))


def test_gradient_rejects_one_point():
    with pytest.raises(ValueError):
        bm.gradient(dev(np.ones(1)), 1.0)


@pytest.mark.parametrize("n_add_frac", [1.0, 0.75])
def test_mtw_shift_add_bit_exact(n_add_frac):
    """interp(time + t_rev, time, memory, 0, 0) then += v[:n_add]: the
    reference's two numpy steps, bit for bit."""
    rng = np.random.default_rng(11)
    n = 16384
    wl = n * 6.5e-9
    tm = np.linspace(0, wl, n, endpoint=False)
    mem = rng.standard_normal(n)
    v = rng.standard_normal(n)
    n_add = int(n * n_add_frac)
    for t_rev in (1.6702e-6, 0.0, 257 * 6.5e-9, wl * 2):
        out = bm.mtw_shift_add(dev(mem), dev(tm), t_rev, dev(v), n_add)
        ref = np.interp(tm + t_rev, tm, mem, left=0, right=0)
        ref[:n_add] += v[:n_add]
        np.testing.assert_array_equal(host(out), ref)


# ---------------------------------------------------------------- turns ----
def _check_sources(w, ref):
    """Multi-turn memories of the sources within 1e-10 of their maximum."""
    for s, r in zip(w.tiv.induced_voltage_list, ref.get("sources", [])):
        if r.get("multi_turn_wake"):
            m = host(s.mtw_memory)
            sc = np.abs(r["memory"]).max()
            assert sc > 0
            assert np.max(np.abs(m - r["memory"])) <= 1e-10 * sc


@pytest.mark.parametrize("kind", WAKE_KINDS)
def test_wake_turns_object_by_object(kind):
    w = build_wake_workload(kind, n_particles=40001, n_turns=12)
    ref = orc.track(w.snapshot(), 8)
    for _ in range(8):
        for m in w.map():
            m.track()
    _check_intensity(w, ref, 1e-9)
    _check_sources(w, ref)


@pytest.mark.parametrize("kind", WAKE_KINDS)
def test_wake_first_turn_fused_exact_profile(kind):
    w = build_wake_workload(kind, n_particles=100001, n_turns=4)
    ref = orc.track(w.snapshot(), 1)
    w.fused(1)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    v = host(w.tiv.induced_voltage)
    assert np.max(np.abs(v - ref["induced_voltage"])) <= 1e-10 * np.abs(ref["induced_voltage"]).max()
    sc = np.abs(ref["dE"]).max()
    np.testing.assert_allclose(host(w.beam.dE), ref["dE"], rtol=0, atol=1e-12 * sc)
    _check_sources(w, ref)


CONV_KINDS = ("mtw", "mtw_front", "mtw_time", "mixed")


@pytest.mark.parametrize("kind", WAKE_KINDS)
@pytest.mark.parametrize("method", [None, "fft"])
def test_wake_turns_fused(kind, method):
    """The automatic method (the direct convolution forming the whole
    memory at these sizes) and the LDS FFT."""
    if method is not None and kind not in CONV_KINDS:
        pytest.skip("no impedance convolution in this workload")
    w = build_wake_workload(kind, n_particles=40001, n_turns=30)
    ref = orc.track(w.snapshot(), 20)
    ft = w.fused(20, method=method)
    if kind in CONV_KINDS:
        assert ft.method == (method or "direct")
    _check_intensity(w, ref, 1e-9)
    _check_sources(w, ref)


def test_mtw_fused_graph_vs_eager_and_split_calls():
    """Graph replays (chunks of 10 turns) and eager turns give identical
    coordinates and memories, also when the memory is handed back to the
    caller between track() calls."""
    a = build_wake_workload("mtw", n_particles=30001, n_turns=60)
    b = build_wake_workload("mtw", n_particles=30001, n_turns=60)
    from blond_amd.trackers.fused import FusedTracker

    fa = FusedTracker(a.tracker, a.profile, a.tiv, graph_chunk=10)
    fa.track(25)
    fa.track(20)
    fb = FusedTracker(b.tracker, b.profile, b.tiv, use_graph=False)
    for _ in range(45):
        fb.track(1)
    np.testing.assert_array_equal(host(a.beam.dt), host(b.beam.dt))
    np.testing.assert_array_equal(host(a.beam.dE), host(b.beam.dE))
    np.testing.assert_array_equal(host(a.tiv.induced_voltage_list[0].mtw_memory),
                                  host(b.tiv.induced_voltage_list[0].mtw_memory))
    assert a.rf.counter[0] == 45


@pytest.mark.parametrize("method", [None, "fourstep", "fft"])
def test_mtw_fused_large_memory(method):
    """4096 slices with a 4-turn memory (n_fft 16384, 16384 memory values):
    the engine's own choice is the direct convolution over all 16384
    outputs; the four-step and LDS transforms give the same turns within
    the induced-voltage bar."""
    w = build_wake_workload("mtw", n_particles=200001, n_turns=8, n_slices=4096)
    ref = orc.track(w.snapshot(), 4)
    ft = w.fused(0, method=method)
    ft.track(4)
    assert ft.method == (method or "direct")
    _check_intensity(w, ref, 1e-9)
    _check_sources(w, ref)


def test_mtw_direct_beyond_work_limit_takes_fft():
    """A memory of 8 x 8192 values over 8192 slices (2^29 products a turn)
    is past the direct method's work limit: the engine takes a transform."""
    w = build_wake_workload("mtw", n_particles=20001, n_turns=4, n_slices=8192,
                            memory_turns=8)
    ft = w.fused(1)
    assert ft.method in ("fft", "fourstep")


def test_engine_wake_arguments_rejected():
    w = build_wake_workload("zovern", n_particles=10001, n_turns=6)
    ft = w.fused(0)
    tm = torch.zeros(256, dtype=torch.float64, device="cuda")
    mem = torch.zeros(256, dtype=torch.float64, device="cuda")
    p = (ct.c_void_p(tm.data_ptr()), ct.c_void_p(mem.data_ptr()))
    with pytest.raises(_lib.BlmiError, match="convolution"):
        # a memory needs a convolution; this engine has none
        _lib.call("blmi_engine_set_wake", ft._h, 256, 256, *p, None, 0)
    with pytest.raises(_lib.BlmiError, match="derivative"):
        _lib.call("blmi_engine_set_wake", ft._h, 0, 0, None, None, None, 7)
    with pytest.raises(_lib.BlmiError, match="n_add"):
        _lib.call("blmi_engine_set_wake", ft._h, 0, 5, None, None, None, 0)
