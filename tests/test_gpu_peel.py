"""The histogram's wavefront-merged count (max_peel > 0: up to max_peel
times the lanes of a wavefront holding the first pending lane's bin are
counted with one LDS atomic, the rest one each) against the oracle.  It is
off by default, so every other test takes the unmerged count; here it runs
in each histogram layout: u32 LDS counts (c1 plain turns, one launch a
turn and 25 turns a launch; c2 intensity), u16 pairs over a window of the
bins with out-of-window device atomics (c3), and the batched ensemble.
Counts are exact; plain coordinates bit-identical; an intensity turn's
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


@pytest.mark.parametrize("max_peel", [1, 4, 64])
@pytest.mark.parametrize("turn_block", [25, 1])
def test_plain_turns_merged_count_bit_exact(max_peel, turn_block):
    """c1 (100 slices: a wavefront's 64 particles share a few bins, the
    case merging is for), 13 turns as 7 + 6."""
    w = build_workload("c1", n_particles=50_001, n_turns=40)
    snap = w.snapshot()
    w.fused(7, turn_block=turn_block, max_peel=max_peel)
    w.fused(6)
    ref = orc.track(snap, 13, ORC)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])


@pytest.mark.parametrize("name,n", [("c2", 100_001), ("c3", 100_001)])
def test_intensity_turn_merged_count(name, n):
    """c2: u32 counts in LDS; c3: u16 pairs over the bin window, the bins
    outside it by device atomics (merged lanes there count one each)."""
    w = build_workload(name, n_particles=n, n_turns=4)
    ref = orc.track(w.snapshot(), 1, ORC)
    w.fused(1, max_peel=4)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    vr = ref["induced_voltage"]
    assert np.max(np.abs(host(w.tiv.induced_voltage) - vr)) <= 1e-10 * np.abs(vr).max()
    sc = np.abs(ref["dE"]).max()
    np.testing.assert_allclose(host(w.beam.dE), ref["dE"], rtol=0, atol=1e-12 * sc)
    np.testing.assert_allclose(host(w.beam.dt), ref["dt"], rtol=1e-12, atol=0)


def test_plain_ensemble_merged_count_bit_exact():
    """Batched bunches of 20 to 3000 macroparticles, merged counts."""
    ens = build_ensemble(n_bunches=10, n_turns=20, n_slices=100, n_min=20, n_max=3000, seed=9)
    snaps = [ens.bunch(k).snapshot() for k in range(10)]
    for s in snaps:
        s["intensity"] = False
    et = EnsembleTracker(ens.trackers, ens.profiles, None, graph_chunk=5, max_peel=4)
    et.track(7)
    for k in range(10):
        ref = orc.track(snaps[k], 7)
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dt), ref["dt"])
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dE), ref["dE"])
        np.testing.assert_array_equal(host(ens.profiles[k].n_macroparticles), ref["profile"])
    et.close()
