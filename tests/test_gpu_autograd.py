"""The autograd wrapper: backward is the matched adjoint kernel. Checked by the dot-product
test, by FP64 central finite differences (torch.autograd.gradcheck) and against the oracle's
adjoint."""
import numpy as np
import pytest

import oracle
from backends import geom_to_descriptors

pytestmark = pytest.mark.gpu


def make(dim, dtype):
    if dim == 2:
        g = oracle.circle_geometry(2, (24, 20), 9, 360, 2400.0, 120.0, det=(34,))
    else:
        g = oracle.circle_geometry(3, (12, 10, 14), 7, 360, 1200.0, 60.0, det=(20, 18))
    import elsa_amd as E

    vd, dd = geom_to_descriptors(g)
    return g, E.JosephsMethodCUDA(vd, dd, dtype=dtype)


@pytest.mark.parametrize("dim", [2, 3])
def test_gradcheck_fp64(dim):
    """FP64 central finite differences (torch.autograd.gradcheck, eps 1e-6) agree with the
    backward of project() and of backproject(): the projection is linear, so the check is
    exact up to rounding. The FP64 back projection sums in 64-bit fixed point, so repeated
    backward passes agree bit for bit (nondet_tol 0)."""
    import torch

    from elsa_amd.autograd import backproject, project

    g, op = make(dim, np.float64)
    x = torch.rand(g.vol_shape, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: project(op, v), (x,), eps=1e-6, atol=1e-8, rtol=1e-6,
                                    fast_mode=True, nondet_tol=0.0)
    y = torch.rand(g.sino_shape, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: backproject(op, v), (y,), eps=1e-6, atol=1e-8, rtol=1e-6,
                                    fast_mode=True, nondet_tol=0.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_backward_is_the_matched_adjoint(dim):
    """grad of <w, A x> is A^T w: equal to the operator's applyAdjoint bit for bit, within
    FP32 tolerance of the oracle's adjoint, and <A x, w> = <x, A^T w> (dot-product test)."""
    import torch

    from elsa_amd.autograd import project

    g, op = make(dim, np.float32)
    rng = np.random.default_rng(3)
    x = torch.tensor(rng.random(g.vol_shape, dtype=np.float32), device="cuda", requires_grad=True)
    w = torch.tensor(rng.standard_normal(g.sino_shape).astype(np.float32), device="cuda")
    loss = (project(op, x) * w).sum()
    loss.backward()
    grad = x.grad.detach()
    np.testing.assert_array_equal(grad.cpu().numpy(), op.applyAdjoint(w).cpu().numpy())
    ref = oracle.joseph(g, w.cpu().numpy(), adjoint=True, mode=oracle.MODE_DIRECT)
    assert np.linalg.norm(grad.cpu().numpy() - ref) <= 1e-6 * np.linalg.norm(ref)
    lhs = float((op.apply(x.detach()).double() * w.double()).sum())
    rhs = float((x.detach().double() * grad.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_batched_project_and_training_step():
    """A leading batch dimension runs as one fused batch of the operator and its clones (one
    launch per pass), each element bit for bit the operator's own forward and backward; a
    least-squares step with torch's optimiser lowers the loss."""
    import torch

    from elsa_amd.autograd import _batch_of, backproject, project

    g, op = make(3, np.float32)
    xt = torch.rand((3,) + g.vol_shape, device="cuda")
    b = project(op, xt)
    assert b.shape == (3,) + g.sino_shape
    assert _batch_of(op, 3).fused
    for i in range(3):
        np.testing.assert_array_equal(b[i].cpu().numpy(), op.apply(xt[i]).cpu().numpy())
    back = backproject(op, b)
    for i in range(3):
        np.testing.assert_array_equal(back[i].cpu().numpy(), op.applyAdjoint(b[i]).cpu().numpy())
    xt = xt[:2].contiguous()
    b = b[:2].contiguous()
    x = torch.zeros_like(xt, requires_grad=True)
    opt = torch.optim.SGD([x], lr=1e-3)
    losses = []
    for _ in range(5):
        opt.zero_grad()
        loss = 0.5 * ((project(op, x) - b) ** 2).sum()
        loss.backward()
        opt.step()
        losses.append(float(loss))
    assert losses[-1] < losses[0]
