"""GPU parity of the sparse histogram (per-bucket profiles of a filling
pattern): counts identical to the oracle (and, in single precision, to the
reference's single build), SparseSlices 'C' and 'onebyone' against each
other and against one uniform profile over the train."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.utils import bmath as bm  # noqa: E402
from tests.test_gpu_tracking import host  # noqa: E402
from tests.test_sparse_cpu import T_RF, _cuts, _train  # noqa: E402

ORC = orc.oracle_lib()
REF32 = orc.ref_lib("strict", "single")


@pytest.mark.parametrize("pattern,nsb,n_per", [
    ([1, 0, 0, 0, 0, 1], 32, 20000),
    ([1, 1, 1, 0, 1, 0, 0, 1, 0], 64, 50000),
    ([0, 1, 0, 1], 16, 10000),
    ([1] * 72 + [0] * 8 + [1] * 48, 256, 4000),     # 30720 bins: global counts
])
def test_sparse_histogram_bit_exact(pattern, nsb, n_per):
    pattern = np.array(pattern, dtype=np.float64)
    rng = np.random.default_rng(nsb)
    cl, cr, bidx = _cuts(pattern)
    dt = np.concatenate([_train(pattern, n_per, rng), cl, cr[:-1], cr[-1:],
                         [cl[0] - 1e-12, cr[-1] + 1e-12, np.nan]])
    ref = np.zeros((len(cl), nsb))
    ORC.sparse_histogram(dt, ref, cl, cr, bidx, nsb)
    out = torch.zeros((len(cl), nsb), dtype=torch.float64, device="cuda")
    bm.sparse_histogram(torch.from_numpy(dt).cuda(), out, cl, cr, bidx, nsb)
    np.testing.assert_array_equal(host(out), ref)
    # numpy arrays in and out (staged through the device)
    out2 = np.zeros((len(cl), nsb))
    bm.sparse_histogram(dt, out2, cl, cr, bidx, nsb)
    np.testing.assert_array_equal(out2, ref)


@pytest.mark.skipif(REF32 is None, reason="oracle/_ref single build not present")
def test_sparse_histogram_f32_matches_reference_single():
    from blond_amd.utils._precision import precision

    pattern = np.zeros(11, dtype=np.float32)
    pattern[::2] = 1
    rng = np.random.default_rng(5)
    cl, cr, bidx = (x.astype(np.float32) for x in _cuts(pattern))
    dt = _train(pattern, 30000, rng).astype(np.float32)
    ref = np.zeros((len(cl), 48), dtype=np.float32)
    f = REF32.lib.sparse_histogram
    f(*[x.ctypes.data_as(orc._fp) for x in (dt, ref, cl, cr, bidx)], orc.ct.c_int(48),
      orc.ct.c_int(len(cl)), orc.ct.c_int(len(dt)))
    precision.set("single")
    try:
        out = np.zeros((len(cl), 48), dtype=np.float32)
        bm.sparse_histogram(dt, out, cl, cr, bidx, 48)
    finally:
        precision.set("double")
    np.testing.assert_array_equal(out, ref)


def test_sparse_slices_c_onebyone_and_uniform():
    """unittests/beam_profile/test_sparse_profile.py: two bunches five
    buckets apart, 32 slices per bucket."""
    from blond_amd.beam.beam import Beam, Proton
    from blond_amd.beam.profile import CutOptions, Profile
    from blond_amd.beam.sparse_slices import SparseSlices
    from blond_amd.input_parameters.rf_parameters import RFStation
    from blond_amd.input_parameters.ring import Ring

    ring = Ring(6911.5038, 1.0 / 17.95142852 ** 2, 25.92e9, Proton(), n_turns=1)
    rf = RFStation(ring, 4620, 3.5e6, 0, n_rf=1)
    t_rf = rf.t_rf[0, 0]
    pattern = np.zeros(6)
    pattern[::5] = 1
    beam = Beam(ring, 20000, 2e11)
    rng = np.random.default_rng(1984)
    beam.dt[:] = np.concatenate([rng.normal(0.5 * t_rf, 0.2e-9, 10000),
                                 rng.normal(5.5 * t_rf, 0.2e-9, 10000)])
    beam.dE[:] = 0.0
    uni = Profile(beam, CutOptions(cut_left=0.0, cut_right=6 * t_rf, n_slices=6 * 32))
    c = SparseSlices(rf, beam, 32, pattern, tracker="C", direct_slicing=True)
    one = SparseSlices(rf, beam, 32, pattern, tracker="onebyone", direct_slicing=True)
    np.testing.assert_array_equal(c.n_macroparticles_array, one.n_macroparticles_array)
    for b, lo in ((0, 0), (1, 5 * 32)):
        np.testing.assert_allclose(c.bin_centers_array[b], uni.bin_centers[lo:lo + 32], rtol=1e-6)
        np.testing.assert_array_equal(c.n_macroparticles_array[b],
                                      uni.n_macroparticles[lo:lo + 32])
        np.testing.assert_array_equal(c.profiles_list[b].n_macroparticles,
                                      c.n_macroparticles_array[b])
    # device-resident rows: the same counts, each profile a row view
    beam.to_gpu()
    c.to_gpu()
    c.track()
    np.testing.assert_array_equal(host(c.n_macroparticles_array), one.n_macroparticles_array)
    assert c.profiles_list[1].n_macroparticles.data_ptr() == \
        c.n_macroparticles_array[1].data_ptr()
