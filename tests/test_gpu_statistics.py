"""GPU parity of the per-turn bunch moments the fused turn accumulates in
its particle kernel (Beam.statistics, beam.py:430-458) and of BunchMonitor
(monitors.py:36-467) fed from them, against numpy moments of the oracle's
coordinates and against the object-by-object monitor."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.workloads import build_workload  # noqa: E402
from tests.test_gpu_tracking import host  # noqa: E402

ORC = orc.oracle_lib()


@pytest.mark.parametrize("n_particles", [100001, 3_000_001])
def test_fused_moments_match_numpy_of_oracle_turns(n_particles):
    """Plain turns are bit-exact to the oracle, so the per-turn moments can
    be checked against numpy's two-pass mean / std of the oracle's own
    coordinates, turn by turn."""
    w = build_workload("c1", n_particles=n_particles, n_turns=30, n_slices=100)
    st = w.snapshot()
    from blond_amd.trackers.fused import FusedTracker

    ft = FusedTracker(w.tracker, w.profile, None, statistics=True, graph_chunk=8)
    ft.track(20)
    rows = ft.statistics(1, 20)
    for t in range(20):
        st = orc.track(st, 1, ORC)
        dt, dE = st["dt"], st["dE"]
        assert rows[t, 0] == n_particles
        # means to 1e-12 of the spread, spreads to 1e-10 relative
        assert abs(rows[t, 1] - dt.mean()) <= 1e-12 * dt.std(), (t, rows[t, 1], dt.mean())
        assert abs(rows[t, 3] - dE.mean()) <= 1e-12 * dE.std(), (t, rows[t, 3], dE.mean())
        assert rows[t, 2] == pytest.approx(dt.std(), rel=1e-10)
        assert rows[t, 4] == pytest.approx(dE.std(), rel=1e-10)
    np.testing.assert_array_equal(host(w.beam.dt), st["dt"])


def test_statistics_do_not_change_the_turns():
    a = build_workload("c2", n_particles=200001, n_turns=30, n_slices=1024, n_fft=4096)
    b = build_workload("c2", n_particles=200001, n_turns=30, n_slices=1024, n_fft=4096)
    from blond_amd.trackers.fused import FusedTracker

    fa = FusedTracker(a.tracker, a.profile, a.tiv, statistics=True)
    fb = FusedTracker(b.tracker, b.profile, b.tiv)
    fa.track(20)
    fb.track(20)
    np.testing.assert_array_equal(host(a.beam.dt), host(b.beam.dt))
    np.testing.assert_array_equal(host(a.beam.dE), host(b.beam.dE))
    r = fa.statistics(20, 1)[0]
    b.beam.statistics()
    assert r[1] == pytest.approx(b.beam.mean_dt, rel=1e-12)
    assert r[2] == pytest.approx(b.beam.sigma_dt, rel=1e-10)
    assert r[4] == pytest.approx(b.beam.sigma_dE, rel=1e-10)


def test_bunch_monitor_fused_vs_object_by_object(tmp_path):
    from blond_amd.monitors.monitors import BunchMonitor, _h5py
    from blond_amd.trackers.fused import FusedTracker

    a = build_workload("c1", n_particles=50001, n_turns=24, n_slices=100)
    b = build_workload("c1", n_particles=50001, n_turns=24, n_slices=100)
    ma = BunchMonitor(a.ring, a.rf, a.beam, str(tmp_path / "a"), buffer_time=6)
    mb = BunchMonitor(b.ring, b.rf, b.beam, str(tmp_path / "b"), buffer_time=6)
    for _ in range(18):
        for m in a.map() + [ma]:
            m.track()
    fb = FusedTracker(b.tracker, b.profile, None, monitor=mb, graph_chunk=5)
    fb.track(11)
    fb.track(7)
    np.testing.assert_array_equal(host(a.beam.dt), host(b.beam.dt))
    if _h5py is not None:  # pragma: no cover - h5py not installed here
        return
    da, db = np.load(str(tmp_path / "a.npz")), np.load(str(tmp_path / "b.npz"))
    assert sorted(da.files) == sorted(db.files)
    for k in da.files:
        np.testing.assert_allclose(db[k], da[k], rtol=2e-7, atol=0, err_msg=k)
    assert ma.i_turn == mb.i_turn == 19
