"""GPU parity of the batched ensemble: every bunch of a batched run equals
the same bunch tracked alone by the fused engine (same arithmetic, same
order: bit-exact), plain ensembles equal the oracle bit for bit, and
intensity ensembles follow the oracle to the statistical bar of the
single-bunch tests."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.trackers.fused import FusedTracker  # noqa: E402
from blond_amd.workloads import build_ensemble  # noqa: E402

from tests.test_gpu_tracking import _check_intensity, host  # noqa: E402


def small(n_bunches=6, n_turns=40, n_slices=256, seed=3):
    return build_ensemble(n_bunches=n_bunches, n_turns=n_turns, n_slices=n_slices, n_min=2e3,
                          n_max=3e4, seed=seed)


def test_counts_and_offsets_layout():
    ens = small()
    et = ens.tracker()
    assert et.n_particles == int(ens.counts.sum())
    assert np.all(et.offsets % 2 == 0)
    for b, tr in enumerate(ens.trackers):
        assert tr.beam.dt.data_ptr() == et.dt.data_ptr() + 8 * int(et.offsets[b])
        assert tr.beam.dt.numel() == ens.counts[b]
    et.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batched_equals_sequential_bit_exact(use_graph):
    a, b = small(), small()
    et = a.tracker(use_graph=use_graph, graph_chunk=10)
    et.track(25)
    for k in range(len(b.trackers)):
        w = b.bunch(k)
        w.to_gpu()
        ft = FusedTracker(w.tracker, w.profile, w.tiv, method="fft", graph_chunk=10)
        ft.track(25)
        np.testing.assert_array_equal(host(a.trackers[k].beam.dt), host(w.beam.dt))
        np.testing.assert_array_equal(host(a.trackers[k].beam.dE), host(w.beam.dE))
        np.testing.assert_array_equal(host(a.profiles[k].n_macroparticles),
                                      host(w.profile.n_macroparticles))
        np.testing.assert_array_equal(host(a.tivs[k].induced_voltage), host(w.tiv.induced_voltage))
        assert a.trackers[k].rf_params.counter[0] == 25
        ft.close()
    et.close()


def test_intensity_ensemble_vs_oracle():
    ens = small(n_bunches=4, n_turns=20)
    snaps = [ens.bunch(k).snapshot() for k in range(4)]
    et = ens.tracker()
    et.track(12)
    for k in range(4):
        _check_intensity(ens.bunch(k), orc.track(snaps[k], 12), 1e-9)
    et.close()


def test_plain_ensemble_bit_exact_vs_oracle():
    ens = small(n_bunches=5, n_turns=20, n_slices=100)
    snaps = [ens.bunch(k).snapshot() for k in range(5)]
    for s in snaps:
        s["intensity"] = False
    from blond_amd.trackers.ensemble import EnsembleTracker

    et = EnsembleTracker(ens.trackers, ens.profiles, None, graph_chunk=5)
    et.track(12)
    for k in range(5):
        ref = orc.track(snaps[k], 12)
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dt), ref["dt"])
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dE), ref["dE"])
        np.testing.assert_array_equal(host(ens.profiles[k].n_macroparticles), ref["profile"])
    et.close()


def test_autotune_restores_state_and_keeps_results():
    a, b = small(n_turns=30), small(n_turns=30)
    ea, eb = a.tracker(), b.tracker()
    blk, wpc, ms = ea.autotune(turns=3)
    print(f"\nautotune: block={blk} wg_per_cu={wpc} {ms:.4f} ms/turn")
    assert (blk, wpc) == ea.launch_shape() and ms > 0
    assert a.trackers[0].rf_params.counter[0] == 0
    ea.track(10)
    eb.track(10)
    np.testing.assert_array_equal(host(ea.dt), host(eb.dt))
    np.testing.assert_array_equal(host(ea.dE), host(eb.dE))
    np.testing.assert_array_equal(host(ea.prof), host(eb.prof))
    ea.close()
    eb.close()
