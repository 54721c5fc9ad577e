"""Non-finite and out-of-range coordinates through the fused engine, as the
reference treats them: a NaN or infinite dt is never counted (every bin
comparison is false, histogram.cpp:40-60), takes no interpolated kick
(linear_interp_kick.cpp:56-64) and stays non-finite through the RF kick and
drift; a finite dt outside the cuts is tracked and not counted.  A few such
particles are planted in c1 and c2 bunches; coordinates (NaN where the
oracle has NaN) and profiles must match the oracle's, plain turns bit for
bit and an intensity turn to the first-turn bar of the large tests."""
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


def _plant(w):
    """NaN, +-inf and far-outside dt; a NaN dE; spread over the bunch so that
    they land in different workgroups and lanes (one is the odd tail)."""
    dt, dE = host(w.beam.dt).copy(), host(w.beam.dE).copy()
    n = dt.size
    idx = [0, 1, 63, 64, n // 3, n // 2, n - 2, n - 1]
    dt[idx[0]] = np.nan
    dt[idx[1]] = np.inf
    dt[idx[2]] = -np.inf
    dt[idx[3]] = -1e-6       # far left of the cuts
    dt[idx[4]] = 1e-6        # far right
    dE[idx[5]] = np.nan
    dt[idx[6]] = dt[idx[6]] + 1e-7
    dt[idx[7]] = np.nan      # the odd tail
    w.beam.dt.copy_(torch.from_numpy(dt))
    w.beam.dE.copy_(torch.from_numpy(dE))


@pytest.mark.parametrize("turn_block", [25, 1])
def test_plain_turns_with_nonfinite_particles_bit_exact(turn_block):
    w = build_workload("c1", n_particles=20001, n_turns=30)
    _plant(w)
    snap = w.snapshot()
    w.fused(11, turn_block=turn_block)
    ref = orc.track(snap, 11, ORC)
    np.testing.assert_array_equal(host(w.beam.dt), ref["dt"])
    np.testing.assert_array_equal(host(w.beam.dE), ref["dE"])
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    assert np.isnan(ref["dt"][0]) and np.isnan(ref["dE"][20001 // 2])


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_intensity_turn_with_nonfinite_particles(name):
    w = build_workload(name, n_particles=40001, n_turns=4)
    _plant(w)
    snap = w.snapshot()
    ref = orc.track(snap, 1, ORC)
    w.fused(1)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    vr = ref["induced_voltage"]
    assert np.max(np.abs(host(w.tiv.induced_voltage) - vr)) <= 1e-10 * np.abs(vr).max()
    dE, dt = host(w.beam.dE), host(w.beam.dt)
    np.testing.assert_array_equal(np.isnan(dE), np.isnan(ref["dE"]))
    np.testing.assert_array_equal(np.isnan(dt), np.isnan(ref["dt"]))
    ok = np.isfinite(ref["dE"])
    sc = np.abs(ref["dE"][ok]).max()
    np.testing.assert_allclose(dE[ok], ref["dE"][ok], rtol=0, atol=1e-12 * sc)
