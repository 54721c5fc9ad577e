"""Fused turns interleaved with the objects' own turns (the map a BLonD
main script runs, element by element) on one workload: the fused tracker
continues from the RF station's counter as the objects left it, with the
histogram of the current turn rebuilt from the coordinates as they are.
Plain turns are bit-identical to the oracle's uninterrupted run; intensity
turns held as the other multi-turn intensity tests hold them.  Where the
engine keeps state of its own on the device (here the bin-ordered particle
copies), a turn taken outside it is refused with an error instead.  New
arrays assigned to beam.dt and beam.dE between two calls are followed (in
the tracker's precision; another is refused).  The batched ensemble the
same: its bunches' own turns in between are picked up, bunches whose
counters diverged are refused."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.workloads import build_workload  # noqa: E402

from tests.test_gpu_tracking import _check_intensity, host  # noqa: E402

ORC = orc.oracle_lib()


def _objects(w, turns):
    for _ in range(turns):
        for m in w.map():
            m.track()


@pytest.mark.parametrize("turn_block", [25, 1])
def test_plain_fused_and_object_turns_bit_exact(turn_block):
    """c1: 5 fused, 3 by the objects, 4 fused, 1 by the objects, 2 fused."""
    w = build_workload("c1", n_particles=40_001, n_turns=40)
    ref = orc.track(w.snapshot(), 15, ORC)
    w.fused(5, turn_block=turn_block)
    _objects(w, 3)
    w.fused(4)
    _objects(w, 1)
    w.fused(2)
    assert w.rf.counter[0] == 15
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_intensity_fused_and_object_turns(name):
    """2 fused, 1 by the objects, 3 fused, 2 by the objects, 2 fused."""
    w = build_workload(name, n_particles=60_001, n_turns=20)
    ref = orc.track(w.snapshot(), 10, ORC)
    w.fused(2)
    _objects(w, 1)
    w.fused(3)
    _objects(w, 2)
    w.fused(2)
    assert w.rf.counter[0] == 10
    _check_intensity(w, ref, 1e-9)


def test_turns_outside_an_ordered_engine_refused():
    w = build_workload("c1", n_particles=20_001, n_turns=20, n_slices=16384)
    w.fused(2, sort_every=2)
    _objects(w, 1)
    with pytest.raises(RuntimeError, match="outside the fused tracker"):
        w.fused(1)


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_replaced_coordinate_arrays_followed(name):
    """beam.dt and beam.dE assigned new arrays (copies) between two calls:
    the tracker tracks those, the old ones are left alone."""
    w = build_workload(name, n_particles=40_001, n_turns=20)
    ref = orc.track(w.snapshot(), 9, ORC)
    w.fused(4)
    old_dt = w.beam.dt
    kept = host(old_dt).copy()
    w.beam.dt = w.beam.dt.clone()
    w.beam.dE = w.beam.dE.clone()
    w.fused(5)
    np.testing.assert_array_equal(host(old_dt), kept)
    if name == "c1":
        np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
        np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
        np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    else:
        _check_intensity(w, ref, 1e-9)


def test_replaced_coordinates_of_another_precision_refused():
    w = build_workload("c1", n_particles=20_001, n_turns=20)
    w.fused(2)
    w.beam.dt = w.beam.dt.to(torch.float32).contiguous()
    w.beam.dE = w.beam.dE.to(torch.float32).contiguous()
    with pytest.raises(TypeError, match="must stay"):
        w.fused(1)


def test_ensemble_and_object_turns_plain_bit_exact():
    """Batched turns, then each bunch's own tracker and profile for 2
    turns (their coordinates are views of the ensemble's arrays), then
    batched turns again: every bunch equals the oracle's 9 turns."""
    from blond_amd.trackers.ensemble import EnsembleTracker
    from blond_amd.workloads import build_ensemble

    ens = build_ensemble(n_bunches=6, n_turns=20, n_slices=100, n_min=50, n_max=3000, seed=3)
    snaps = [ens.bunch(k).snapshot() for k in range(6)]
    for s in snaps:
        s["intensity"] = False
    et = EnsembleTracker(ens.trackers, ens.profiles, None, graph_chunk=5)
    et.track(3)
    for k in range(6):
        for _ in range(2):
            ens.trackers[k].track()
            ens.profiles[k].track()
    et.track(4)
    for k in range(6):
        ref = orc.track(snaps[k], 9)
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dt), ref["dt"])
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dE), ref["dE"])
        np.testing.assert_array_equal(host(ens.profiles[k].n_macroparticles), ref["profile"])
    et.close()


def test_ensemble_and_object_turns_intensity():
    """The same with every bunch's induced voltage: 2 batched, 1 turn of
    each bunch's map [profile, induced voltage, tracker], 2 batched."""
    from blond_amd.workloads import build_ensemble

    ens = build_ensemble(n_bunches=4, n_turns=20, n_slices=256, n_min=20_000, n_max=40_000,
                         seed=4)
    snaps = [ens.bunch(k).snapshot() for k in range(4)]
    et = ens.tracker()
    et.track(2)
    for k in range(4):
        for m in ens.bunch(k).map():
            m.track()
    et.track(2)
    for k in range(4):
        _check_intensity(ens.bunch(k), orc.track(snaps[k], 5), 1e-9)
    et.close()


def test_ensemble_with_diverged_counters_refused():
    from blond_amd.trackers.ensemble import EnsembleTracker
    from blond_amd.workloads import build_ensemble

    ens = build_ensemble(n_bunches=3, n_turns=20, n_slices=100, n_min=50, n_max=500, seed=2)
    et = EnsembleTracker(ens.trackers, ens.profiles, None, graph_chunk=5)
    et.track(1)
    ens.trackers[0].track()
    with pytest.raises(RuntimeError, match="counters differ"):
        et.track(1)
    et.close()


def test_resliced_profile_refused():
    """Slicing is fixed at construction (the impedance on the bins and the
    interpolation table depend on it): new cuts are refused, not ignored."""
    w = build_workload("c1", n_particles=20_001, n_turns=20)
    w.fused(1)
    w.profile.cut_left = w.profile.cut_left - 1e-10
    with pytest.raises(RuntimeError, match="slicing changed"):
        w.fused(1)


def test_updated_rf_programs_followed():
    """The RF voltage of every turn from the current one on raised by 10 %
    (on the device arrays, which the objects' own turns read) between two
    calls: after update_programs() the fused turns equal the oracle's run
    of the new program."""
    w = build_workload("c1", n_particles=30_001, n_turns=30)
    ft = w.fused(3)
    w.rf.voltage[:, 3:] *= 1.1
    ft.update_programs()
    snap = w.snapshot()
    assert snap["voltage"][0, 3] == pytest.approx(1.1 * snap["voltage"][0, 2])
    ref = orc.track(snap, 6, ORC)
    w.fused(6)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])


def test_ensemble_updated_rf_programs_followed():
    from blond_amd.trackers.ensemble import EnsembleTracker
    from blond_amd.workloads import build_ensemble

    ens = build_ensemble(n_bunches=5, n_turns=20, n_slices=100, n_min=50, n_max=2000, seed=6)
    et = EnsembleTracker(ens.trackers, ens.profiles, None, graph_chunk=5)
    et.track(3)
    for tr in ens.trackers:
        tr.rf_params.voltage[:, 3:] *= 0.9
    et.update_programs()
    snaps = [ens.bunch(k).snapshot() for k in range(5)]
    for s in snaps:
        s["intensity"] = False
    et.track(4)
    for k in range(5):
        ref = orc.track(snaps[k], 4)
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dt), ref["dt"])
        np.testing.assert_array_equal(host(ens.trackers[k].beam.dE), ref["dE"])
    et.close()
