"""blmi_linear_interp_kick_noalloc (no device allocation: the table lives in
a fixed per-device array) against blmi_linear_interp_kick through the C ABI:
bit-identical kicks on two calls in a row, the second with fewer slices and
particles; sizes beyond the fixed table are refused with an error; a call
leaves the device's free memory unchanged."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd import _lib  # noqa: E402

_ARGS = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_int,
         ct.c_int64, ct.c_double, ct.c_void_p]
MAX_SLICES = 65536


def _fn(lib, name):
    f = getattr(lib, name)
    f.argtypes = _ARGS
    f.restype = ct.c_int
    return f


def _case(n, ns, seed):
    """Bin centres over [1, 2] ns, particles over [0.9, 2.1] ns (some outside
    the table, which get no kick), voltages of a few kV."""
    rng = np.random.default_rng(seed)
    bc = np.linspace(1e-9, 2e-9, ns)
    v = rng.normal(0.0, 1e3, ns)
    dt = rng.uniform(0.9e-9, 2.1e-9, n)
    dE = rng.normal(0.0, 1e6, n)
    return [torch.from_numpy(a).to("cuda") for a in (dt, dE, v, bc)]


def _call(f, dt, dE, v, bc, stream):
    return f(dt.data_ptr(), dE.data_ptr(), v.data_ptr(), bc.data_ptr(), 1.0, len(bc), len(dt),
             -3.5, stream)


def test_noalloc_interp_kick_matches_allocating_entry_point():
    lib = _lib.load()
    ref = _fn(lib, "blmi_linear_interp_kick")
    noalloc = _fn(lib, "blmi_linear_interp_kick_noalloc")
    stream = torch.cuda.current_stream().cuda_stream
    for n, ns, seed in ((1_000_001, MAX_SLICES, 1), (50_001, 1000, 2)):
        print(f"\nno-alloc interp kick: n={n} slices={ns} seed={seed}", flush=True)
        dt, dE, v, bc = _case(n, ns, seed)
        a, b = dE.clone(), dE.clone()
        assert _call(ref, dt, a, v, bc, stream) == 0, _lib.last_error()
        assert _call(noalloc, dt, b, v, bc, stream) == 0, _lib.last_error()
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        assert not torch.equal(b, dE)  # the kick did something
        untouched = (dt < bc[0]) | (dt >= bc[-1])
        assert torch.equal(b[untouched], dE[untouched])


def test_noalloc_interp_kick_refuses_oversize_and_allocates_nothing():
    lib = _lib.load()
    noalloc = _fn(lib, "blmi_linear_interp_kick_noalloc")
    stream = torch.cuda.current_stream().cuda_stream
    dt, dE, v, bc = _case(1001, MAX_SLICES + 1, 3)
    assert _call(noalloc, dt, dE, v, bc, stream) != 0
    assert "exceed" in _lib.last_error()
    dt, dE, v, bc = _case(200_001, 4096, 4)
    assert _call(noalloc, dt, dE, v, bc, stream) == 0, _lib.last_error()
    dt, dE, v, bc = _case(100_001, MAX_SLICES, 5)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    assert _call(noalloc, dt, dE, v, bc, stream) == 0, _lib.last_error()
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
