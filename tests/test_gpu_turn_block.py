"""Several plain turns a launch (FusedTracker(turn_block=K)): the particles
stay in registers across up to K turns and only the last turn is
histogrammed.  Coordinates and profile must be bit-identical to one launch a
turn (turn_block=1) after every track() call -- whatever the split into
launches (odd and even counts, graph chunks and eager remainders) -- and to
the oracle's turns; for 1, 2 and 3 RF systems, the three drift solvers, an
accelerating ring, FP32 coordinates and an odd particle count.  A workload
with intensity effects or statistics keeps one launch a turn (nothing to
compare there but that it still runs: covered by the other GPU tests)."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import turn_cases as tc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.beam.beam import Beam, Proton  # noqa: E402
from blond_amd.beam.distributions import bigaussian  # noqa: E402
from blond_amd.beam.profile import CutOptions, Profile  # noqa: E402
from blond_amd.input_parameters.rf_parameters import RFStation  # noqa: E402
from blond_amd.input_parameters.ring import Ring  # noqa: E402
from blond_amd.trackers.fused import FusedTracker  # noqa: E402
from blond_amd.trackers.tracker import RingAndRFTracker  # noqa: E402
from blond_amd.workloads import Workload, build_workload  # noqa: E402

ORC = orc.oracle_lib()
SPLITS = (1, 2, 24, 26, 57, 3)  # 113 turns: single turns, even and odd, a graph chunk + rest


def host(t):
    return t.detach().cpu().numpy()


def _ring_case(n_rf, solver, n_particles=20001, n_turns=140, ramp=False, seed=3):
    """LHC-like ring with 1-3 RF systems (harmonics h, 2h, 3h; phases 0),
    100 slices over the first bucket."""
    p = np.linspace(450e9, 450.3e9, n_turns + 1) if ramp else 450e9
    ring = Ring(26658.883, 1 / 55.759505 ** 2, p, Proton(), n_turns)
    h = [35640, 71280, 106920][:n_rf]
    V = [6e6, 1e6, 0.5e6][:n_rf]
    rf = RFStation(ring, h, V, [0.0] * n_rf, n_rf=n_rf)
    beam = Beam(ring, n_particles, 1e9)
    bigaussian(ring, rf, beam, sigma_dt=1.0e-9 / 4, seed=seed)
    prof = Profile(beam, CutOptions(cut_left=0.0, cut_right=float(rf.t_rf[0, 0]), n_slices=100))
    tr = RingAndRFTracker(rf, beam, solver=solver, profile=prof)
    w = Workload(f"ring{n_rf}_{solver}", ring, rf, beam, tr, prof, None, False,
                 meta=dict(n_particles=n_particles, n_turns=n_turns))
    w.to_gpu()
    return w


def _compare_splits(build, f32=False):
    wa, wb = build(), build()
    if f32:
        for w in (wa, wb):
            w.beam.dt = w.beam.dt.to(torch.float32).contiguous()
            w.beam.dE = w.beam.dE.to(torch.float32).contiguous()
    snap = wa.snapshot()
    fa = FusedTracker(wa.tracker, wa.profile, turn_block=25)
    fb = FusedTracker(wb.tracker, wb.profile, turn_block=1)
    done = 0
    for k in SPLITS:
        fa.track(k)
        fb.track(k)
        done += k
        np.testing.assert_array_equal(host(wa.beam.dt), host(wb.beam.dt), err_msg=f"dt, {done}")
        np.testing.assert_array_equal(host(wa.beam.dE), host(wb.beam.dE), err_msg=f"dE, {done}")
        np.testing.assert_array_equal(host(wa.profile.n_macroparticles),
                                      host(wb.profile.n_macroparticles), err_msg=f"profile, {done}")
    assert wa.rf.counter[0] == done
    return wa, snap, done


@pytest.mark.parametrize("n_rf", [1, 2, 3])
@pytest.mark.parametrize("solver", ["simple", "legacy", "exact"])
def test_turn_block_equals_a_launch_a_turn(n_rf, solver):
    _compare_splits(lambda: _ring_case(n_rf, solver))


def test_turn_block_matches_the_oracle_on_a_ramp():
    wa, snap, done = _compare_splits(lambda: _ring_case(2, "simple", ramp=True))
    ref = orc.track(snap, done, ORC)
    np.testing.assert_array_equal(host(wa.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(wa.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(wa.profile.n_macroparticles), ref["profile"])


@pytest.mark.parametrize("solver", ["simple", "exact"])
def test_turn_block_three_rf_matches_the_oracle(solver):
    """Three RF systems take the run-time-count sine loops of both kernels
    (the multi-turn one's and the one-turn kernel's table sine): the
    oracle's coordinates and profile after 113 turns."""
    wa, snap, done = _compare_splits(lambda: _ring_case(3, solver))
    ref = orc.track(snap, done, ORC)
    np.testing.assert_array_equal(host(wa.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(wa.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(wa.profile.n_macroparticles), ref["profile"])


def test_turn_block_c1_odd_count_matches_the_oracle():
    """c1 (100 slices) at an odd particle count, 113 turns in the splits."""
    wa, snap, done = _compare_splits(lambda: build_workload("c1", n_particles=100_001, n_turns=200))
    ref = orc.track(snap, done, ORC)
    np.testing.assert_array_equal(host(wa.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(wa.profile.n_macroparticles), ref["profile"])


def test_turn_block_f32():
    _compare_splits(lambda: build_workload("c1", n_particles=50_000, n_turns=200), f32=True)


def test_turn_block_ex01_2000_turns_closed_loop():
    """The ex01 acceleration case tracked 2000 turns in blocks: the same
    coordinates as the oracle's turns."""
    w = tc.ex01_case(n_particles=20000, n_turns=2000)
    w.to_gpu()
    snap = w.snapshot()
    ft = FusedTracker(w.tracker, w.profile)
    ft.track(2000)
    ref = orc.track(snap, 2000, ORC)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])


def test_turn_block_time_kernels_per_turn():
    w = build_workload("c1", n_particles=200_000, n_turns=200)
    ft = FusedTracker(w.tracker, w.profile)
    ft.track(3)
    _, ms = ft.time_kernels(50)
    assert 0.0 < ms < 1.0
    assert w.rf.counter[0] == 53
