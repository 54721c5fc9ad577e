"""Small and odd bunch sizes through the fused engine and the batched
ensemble: one macroparticle, a few, either side of a wavefront (64) and of
a workgroup's pairs (1024 x 2).  The particle kernels' odd-tail particle,
their idle lanes (every lane of a wavefront runs the first lane's iteration
count; past its own pairs a lane's results go to the sink), the grid
sizing of a nearly empty launch and the histogram windows placed around a
handful of particles are taken only at such sizes.  Plain turns are
bit-identical to the oracle; an intensity turn's profile is exact and its
coordinates within the first-turn bar of the large tests."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.trackers.ensemble import EnsembleTracker  # noqa: E402
from blond_amd.workloads import build_ensemble, build_workload  # noqa: E402

from tests.test_gpu_tracking import host  # noqa: E402

ORC = orc.oracle_lib()
SIZES = [1, 2, 3, 63, 65, 2049, 4097]


@pytest.mark.parametrize("turn_block", [25, 1])
@pytest.mark.parametrize("n", SIZES)
def test_plain_turns_small_bunches_bit_exact(n, turn_block):
    """c1 (100 slices): 13 turns as 7 + 6, in multi-turn launches and one
    launch a turn."""
    w = build_workload("c1", n_particles=n, n_turns=40)
    snap = w.snapshot()
    w.fused(7, turn_block=turn_block)
    w.fused(6)
    ref = orc.track(snap, 13, ORC)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])


def _first_intensity_turn(name, n, **kw):
    w = build_workload(name, n_particles=n, n_turns=4, **kw)
    ref = orc.track(w.snapshot(), 1, ORC)
    w.fused(1)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    vr = ref["induced_voltage"]
    assert np.max(np.abs(host(w.tiv.induced_voltage) - vr)) <= 1e-10 * max(np.abs(vr).max(), 1e-300)
    sc = max(np.abs(ref["dE"]).max(), 1.0)
    np.testing.assert_allclose(host(w.beam.dE), ref["dE"], rtol=0, atol=1e-12 * sc)
    np.testing.assert_allclose(host(w.beam.dt), ref["dt"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("n", [1, 3, 65, 4097])
def test_intensity_turn_small_bunches_c2(n):
    """c2 (4096 slices, direct convolution, table and u32 histogram in
    LDS)."""
    _first_intensity_turn("c2", n)


@pytest.mark.parametrize("n", [1, 3, 65, 4097])
def test_intensity_turn_small_bunches_c3(n):
    """c3 (65536 slices, four-step transform, u16 partial rows over a
    window of the bins and a window of the table in LDS, both placed on
    the few particles there are)."""
    _first_intensity_turn("c3", n)


def test_plain_ensemble_tiny_bunches_bit_exact():
    """Bunches of 1 to ~70 macroparticles (odd ones own a pad slot) in one
    batched launch: each equals the oracle's bunch bit for bit."""
    ens = build_ensemble(n_bunches=12, n_turns=20, n_slices=100, n_min=1, n_max=70, seed=5)
    assert ens.counts.min() <= 3 and np.any(ens.counts % 2 == 1)
    snaps = [ens.bunch(k).snapshot() for k in range(12)]
    for s in snaps:
        s["intensity"] = False
    et = EnsembleTracker(ens.trackers, ens.profiles, None, graph_chunk=5)
    et.track(9)
    for k in range(12):
        ref = orc.track(snaps[k], 9)
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dt), ref["dt"])
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dE), ref["dE"])
        np.testing.assert_array_equal(host(ens.profiles[k].n_macroparticles), ref["profile"])
    et.close()


def test_intensity_ensemble_tiny_bunches_first_turn():
    """The same tiny bunches with their induced voltages (one workgroup a
    bunch, one wavefront a 512-point transform): each bunch's first turn
    against the oracle's (exact profile, induced voltage within 1e-10 of its
    maximum, dE within 1e-12 of its maximum)."""
    ens = build_ensemble(n_bunches=12, n_turns=20, n_slices=100, n_min=1, n_max=70, seed=5)
    snaps = [ens.bunch(k).snapshot() for k in range(12)]
    et = ens.tracker()
    et.track(1)
    for k in range(12):
        ref = orc.track(snaps[k], 1)
        np.testing.assert_array_equal(host(ens.profiles[k].n_macroparticles), ref["profile"])
        vr = ref["induced_voltage"]
        v = host(ens.tivs[k].induced_voltage)
        assert np.max(np.abs(v - vr)) <= 1e-10 * max(np.abs(vr).max(), 1e-300)
        sc = max(np.abs(ref["dE"]).max(), 1.0)
        np.testing.assert_allclose(host(ens.trackers[k].beam.dE), ref["dE"], rtol=0, atol=1e-12 * sc)
    et.close()
