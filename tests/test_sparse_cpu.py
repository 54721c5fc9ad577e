"""The oracle's sparse histogram pinned on the CPU: the reference's own
sparse_histogram.cpp (bit for bit, on inputs where the reference stays in
bounds), a brute-force loop, and the invariant of the reference's
unittests/beam_profile/test_sparse_profile.py: the per-bucket profiles are
the matching slices of one uniform histogram over the whole train."""
import numpy as np
import pytest

from oracle import oracle as orc

ORC = orc.oracle_lib()
REF = orc.ref_lib("strict", "double")
T_RF = 2.5e-9


def _train(pattern, n_per_bucket, rng, sigma=0.3e-9):
    dt = []
    for b in np.where(pattern)[0]:
        dt.append(rng.normal((b + 0.5) * T_RF, sigma, n_per_bucket))
    return np.concatenate(dt)


def _cuts(pattern):
    idx = np.where(pattern)[0]
    return idx * T_RF, (idx + 1) * T_RF, np.cumsum(pattern) * pattern - 1


def _brute(dt, cl, cr, bidx, nsb):
    out = np.zeros((len(cl), nsb))
    inv_b = 1.0 / (cr[0] - cl[0])
    inv_w = inv_b * nsb
    for a in dt:
        if a < cl[0] or a > cr[-1] or np.isnan(a):
            continue
        fb = int((a - cl[0]) * inv_b)
        if fb >= len(bidx):
            continue
        ib = int(bidx[fb])
        if ib == -1:
            continue
        b = ib * nsb + int((a - cl[ib]) * inv_w)
        if 0 <= b < out.size:
            out.flat[b] += 1
    return out


@pytest.mark.parametrize("pattern", [[1, 0, 0, 0, 0, 1], [1, 1, 1, 0, 1, 0, 0, 1, 0],
                                     [0, 1, 0, 1]])
def test_sparse_oracle_brute_force_and_reference(pattern):
    pattern = np.array(pattern, dtype=np.float64)
    rng = np.random.default_rng(int(pattern.sum()))
    dt = _train(pattern, 3000, rng)
    cl, cr, bidx = _cuts(pattern)
    # edges and points outside: left cut, inside right cuts, beyond the train
    dt = np.concatenate([dt, cl, cr[:-1], [cl[0] - 1e-12, cr[-1] + 1e-12]])
    out = np.zeros((len(cl), 32))
    ORC.sparse_histogram(dt, out, cl, cr, bidx, 32)
    np.testing.assert_array_equal(out, _brute(dt, cl, cr, bidx, 32))
    if REF is not None:
        ref = np.zeros((len(cl), 32))
        ORC_ref = orc.Lib(REF.path)
        ORC_ref.lib.sparse_histogram(*[x.ctypes.data_as(orc._dp) for x in
                                       (dt, ref, cl, cr, bidx.astype(np.float64))],
                                     orc.ct.c_int(32), orc.ct.c_int(len(cl)),
                                     orc.ct.c_int(len(dt)))
        np.testing.assert_array_equal(out, ref)
    if pattern[0] == 0:
        # the reference looks buckets up from the first filled bucket's cut
        # but indexes bunch_indexes by absolute bucket: the bunches of this
        # pattern land on empty entries (only edge points are counted)
        bulk = np.zeros((len(cl), 32))
        ORC.sparse_histogram(_train(pattern, 3000, rng), bulk, cl, cr, bidx, 32)
        assert bulk.sum() == 0


def test_sparse_rows_are_slices_of_the_uniform_histogram():
    """test_sparse_profile.py:95-120: bins of a 32-slice-per-bucket uniform
    histogram over the train equal the sparse rows."""
    pattern = np.zeros(6)
    pattern[::5] = 1
    rng = np.random.default_rng(1984)
    dt = _train(pattern, 10000, rng, sigma=0.2e-9)
    cl, cr, bidx = _cuts(pattern)
    out = np.zeros((2, 32))
    ORC.sparse_histogram(dt, out, cl, cr, bidx, 32)
    uni = np.zeros(6 * 32)
    ORC.slice_beam(dt, uni, 0.0, 6 * T_RF)
    np.testing.assert_array_equal(out[0], uni[:32])
    np.testing.assert_array_equal(out[1], uni[5 * 32:])
    assert out.sum() == len(dt) - np.sum((dt < 0) | (dt >= 6 * T_RF)
                                         | ((dt >= T_RF) & (dt < 5 * T_RF)))
