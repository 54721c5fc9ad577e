"""Parity of the fused engine at the measured workloads' full sizes: c2
(4e6 macroparticles, 4096 slices, n_fft 16384, direct convolution), c3
(1e7 macroparticles, two RF systems, 65536 slices, n_fft 131072, four-step
FFT, u16-pair histograms over 512 workgroups) and the c4 ensemble (256
bunches, 5.5e7 macroparticles).

From identical coordinates, one turn against the oracle's turn: the profile
bit-identical, the induced voltage within 1e-10 of its maximum, dE within
1e-12 of its maximum.  After further turns, a size-independent property:
the profile the engine forms in the next turn is the exact histogram of the
coordinates it left (the oracle's histogram of the same array), and it
counts every macroparticle inside the cuts once."""
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


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_full_size_first_turn_and_histogram_of_own_coordinates(name):
    w = build_workload(name, n_turns=40)
    st = w.snapshot()
    print(f"\n{name}: {len(st['dt'])} macroparticles, {st['n_slices']} slices", flush=True)
    ref = orc.track(st, 1)
    ft = w.fused(1)
    np.testing.assert_array_equal(host(w.profile.n_macroparticles), ref["profile"])
    v = host(w.tiv.induced_voltage)
    assert np.max(np.abs(v - ref["induced_voltage"])) <= 1e-10 * np.abs(ref["induced_voltage"]).max()
    sc = np.abs(ref["dE"]).max()
    np.testing.assert_allclose(host(w.beam.dE), ref["dE"], rtol=0, atol=1e-12 * sc)
    ft.track(12)
    dt = host(w.beam.dt).copy()
    ft.track(1)  # its first step slices dt
    prof = host(w.profile.n_macroparticles)
    want = np.zeros(st["n_slices"])
    ORC.slice_beam(dt, want, st["cut_left"], st["cut_right"])
    np.testing.assert_array_equal(prof, want)
    # every macroparticle whose bin (histogram.cpp:40-56) is in range, once
    f = np.floor((dt - st["cut_left"]) * (st["n_slices"] / (st["cut_right"] - st["cut_left"])))
    assert prof.sum() == np.count_nonzero((f >= 0) & (f < st["n_slices"]))


def test_full_size_batched_ensemble_first_turn():
    """c4 at its size: 256 bunches of 1e4..1e6 macroparticles (5.5e7 in
    all), 1024 slices each, one launch pair per turn; every bunch's first
    turn against the oracle's turn of that bunch alone."""
    from blond_amd.workloads import build_ensemble

    ens = build_ensemble(n_bunches=256, n_turns=10)
    snaps = [ens.bunch(k).snapshot() for k in range(256)]
    print(f"\nc4: {int(ens.counts.sum())} macroparticles in 256 bunches", flush=True)
    et = ens.tracker()
    et.track(1)
    for k in range(256):
        ref = orc.track(snaps[k], 1)
        np.testing.assert_array_equal(host(ens.profiles[k].n_macroparticles), ref["profile"])
        v = host(ens.tivs[k].induced_voltage)
        vmax = np.abs(ref["induced_voltage"]).max()
        assert np.max(np.abs(v - ref["induced_voltage"])) <= 1e-10 * vmax, f"bunch {k}"
        sc = np.abs(ref["dE"]).max()
        np.testing.assert_allclose(host(ens.trackers[k].beam.dE), ref["dE"], rtol=0,
                                   atol=1e-12 * sc)
    et.close()
