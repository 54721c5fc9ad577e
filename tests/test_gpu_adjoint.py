"""GPU parity of the differentiable tracker: the forward profile against
the float64 CPU reference map, the hand-written adjoint against the CPU
reference gradient and against central finite differences of the GPU
forward (step 1e-6 relative, 1e-6 relative agreement)."""
import numpy as np
import pytest

from oracle import adjoint as adj

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from blond_amd.trackers.differentiable import DifferentiableTracker  # noqa: E402
from blond_amd.workloads import build_adjoint_case  # noqa: E402

KEYS = ("V", "phi", "R", "fr", "Q")


def params(th):
    return [torch.tensor(th[k], dtype=torch.float64, device="cuda", requires_grad=True)
            for k in KEYS]


def setup(n, turns, n_slices, every, seed=11):
    c = build_adjoint_case(n_particles=n, n_turns=turns, n_slices=n_slices, seed=seed)
    dtr = DifferentiableTracker(c.tracker, c.profile, checkpoint_every=every)
    return c, dtr


def test_forward_profile_matches_cpu_reference():
    c, dtr = setup(20000, 40, 64, 16)
    m = adj.DiffMap(c.state)
    th = c.theta0
    p_gpu = dtr.profile_after(*params(th), turns=30).detach().cpu().numpy()
    p_cpu = m.forward(c.beam.dt, c.beam.dE, th["V"], th["phi"], th["R"], th["fr"], th["Q"], 30)
    assert p_gpu.sum() == pytest.approx(p_cpu.sum(), rel=1e-12)
    np.testing.assert_allclose(p_gpu, p_cpu, rtol=0, atol=1e-8 * p_cpu.max())
    dtr.close()


def test_adjoint_matches_cpu_reference_gradient():
    c, dtr = setup(20000, 40, 64, 16)
    m = adj.DiffMap(c.state)
    th, ts = c.theta0, c.theta_star
    target = m.forward(c.beam.dt, c.beam.dE, ts["V"], ts["phi"], ts["R"], ts["fr"], ts["Q"], 35)
    ps = params(th)
    p = dtr.profile_after(*ps, turns=35)
    loss = torch.sum((p - torch.from_numpy(target).cuda()) ** 2)
    loss.backward()
    ref = m.gradient(c.beam.dt, c.beam.dE, th["V"], th["phi"], th["R"], th["fr"], th["Q"], 35,
                     2 * (p.detach().cpu().numpy() - target))
    for k, t in zip(KEYS, ps):
        got = t.grad.cpu().numpy()[0]
        want = ref[k][0]
        print(f"\n{k}: gpu {got:.12e} cpu {want:.12e} rel {abs(got - want) / abs(want):.2e}")
        assert abs(got - want) <= 1e-6 * abs(want)
    dtr.close()


def test_adjoint_vs_central_differences_with_checkpoints():
    """100 turns with checkpoints every 30 (segments of 30, 30, 30, 10)."""
    c, dtr = setup(100000, 100, 256, 30)
    th, ts = c.theta0, c.theta_star
    with torch.no_grad():
        target = dtr.profile_after(*params(ts), turns=100).clone()

    def loss_at(t):
        with torch.no_grad():
            p = dtr.profile_after(*params(t), turns=100)
            return float(torch.sum((p - target) ** 2))

    ps = params(th)
    p = dtr.profile_after(*ps, turns=100)
    loss = torch.sum((p - target) ** 2)
    loss.backward()
    for k, t in zip(KEYS, ps):
        x0 = th[k][0]
        h = 1e-6 * abs(x0) if x0 != 0 else 1e-6
        tp, tm = dict(th), dict(th)
        tp[k] = np.array([x0 + h])
        tm[k] = np.array([x0 - h])
        fd = (loss_at(tp) - loss_at(tm)) / (2 * h)
        got = t.grad.cpu().numpy()[0]
        print(f"\n{k}: adjoint {got:.12e} fd {fd:.12e} rel {abs(got - fd) / abs(fd):.2e}")
        assert abs(got - fd) <= 1e-6 * abs(fd)
    dtr.close()


def test_backward_after_another_forward_recomputes():
    c, dtr = setup(5000, 20, 32, 8)
    th, ts = c.theta0, c.theta_star
    p1 = dtr.profile_after(*(a := params(th)), turns=15)
    l1 = torch.sum(p1 ** 2)
    with torch.no_grad():
        dtr.profile_after(*params(ts), turns=15)  # overwrites the checkpoints
    l1.backward()
    g1 = [t.grad.clone() for t in a]
    b = params(th)
    torch.sum(dtr.profile_after(*b, turns=15) ** 2).backward()
    for x, y in zip(g1, b):
        assert torch.allclose(x, y.grad, rtol=1e-10, atol=0)
    dtr.close()
