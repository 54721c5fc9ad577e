"""Several fused engines in one process: their turns interleaved give
exactly what each gives alone.  Every engine owns its particle, count,
partial-row, window and transform buffers; only read-only FFT twiddle
tables are shared between engines of the same transform length."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.workloads import build_workload  # noqa: E402

from tests.test_gpu_tracking import host  # noqa: E402


def _state(w):
    return (host(w.beam.dt).copy(), host(w.beam.dE).copy(),
            host(w.profile.n_macroparticles).copy())


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("names", [("c3", "c3"), ("c2", "c3"), ("c1", "c2")])
def test_interleaved_engines_equal_engines_alone(names):
    """Two engines (the c3 pair shares the 2^17 four-step plan's shape but
    not its buffers; seeds differ so the bunches differ) stepped turn by
    turn in alternation, against the same two tracked one after the
    other."""
    def make(k, name):
        return build_workload(name, n_particles=100_001, n_turns=30, seed=11 + k)

    alone = []
    for k, name in enumerate(names):
        w = make(k, name)
        w.fused(9)
        alone.append(_state(w))
    ws = [make(k, name) for k, name in enumerate(names)]
    for _ in range(3):
        for w in ws:
            w.fused(3)
    for w, ref in zip(ws, alone):
        _same(_state(w), ref)
