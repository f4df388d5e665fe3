"""The checkpointed rollout's bookkeeping (luna_hydro.autograd) on the CPU with a stand-in solver whose step is a fixed
linear map: the gradient through the rollout must equal plain autograd through the same maps, for every checkpoint
spacing; and the single-step Function. (The GPU tests run it over the library's adjoint kernels.)"""
import numpy as np
import pytest
import torch

from luna_hydro import autograd


class LinearSolver:
    """step: q <- A_k q with A_k chosen by dt (so each step can differ); adjoint: A_k^T qbar."""

    def __init__(self, n, seed=0):
        g = torch.Generator().manual_seed(seed)
        self.mats = {}
        self.n, self.g = n, g
        self.q = None
        self.steps = 0

    def A(self, dt):
        if dt not in self.mats:
            self.mats[dt] = torch.eye(self.n, dtype=torch.float64) + 0.1 * torch.randn(self.n, self.n, generator=self.g, dtype=torch.float64)
        return self.mats[dt]

    def set_state(self, q):
        self.q = q.detach().clone()

    def state(self):
        return self.q.clone()

    def step(self, dt):
        self.q = self.A(dt) @ self.q
        self.steps += 1

    def adjoint(self, dt, qbar):
        return self.A(dt).T @ qbar


@pytest.mark.parametrize("every", [1, 3, 7, 10, 25])
def test_rollout_gradient_matches_autograd(every):
    n, steps = 6, 10
    solver = LinearSolver(n)
    dts = [0.1 * (k % 4 + 1) for k in range(steps)]
    q0 = torch.randn(n, 2, dtype=torch.float64, generator=torch.Generator().manual_seed(1), requires_grad=True)
    w = torch.randn(n, 2, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    (w * autograd.rollout(q0, solver, dts, every)).sum().backward()
    ref = q0.detach().clone().requires_grad_(True)
    x = ref
    for dt in dts:
        x = solver.A(dt) @ x
    (w * x).sum().backward()
    assert torch.allclose(q0.grad, ref.grad, rtol=1e-13, atol=1e-13)
    # recomputation: forward steps + the backward's (steps - segments) recomputed steps
    segments = len(autograd.checkpoints(steps, every))
    assert solver.steps == steps + (steps - segments)


def test_single_step_function():
    n = 5
    solver = LinearSolver(n, seed=3)
    q0 = torch.randn(n, 1, dtype=torch.float64, generator=torch.Generator().manual_seed(4), requires_grad=True)
    out = autograd.step(q0, solver, 0.5)
    assert torch.allclose(out, solver.A(0.5) @ q0.detach())
    out.sum().backward()
    assert torch.allclose(q0.grad, solver.A(0.5).T @ torch.ones(n, 1, dtype=torch.float64))


def test_checkpoint_schedule():
    assert autograd.checkpoints(200, 20) == list(range(0, 200, 20))
    assert autograd.checkpoints(5, 20) == [0]
    with pytest.raises(ValueError):
        autograd.checkpoints(5, 0)


class PairSolver:
    """A stand-in moving solver: the state is (q, x); step: q <- A q + B x, x <- x + C q (matrices by dt); adjoint: the
    transpose."""

    def __init__(self, n, seed=3):
        self.g = torch.Generator().manual_seed(seed)
        self.n, self.mats, self.state_ = n, {}, None

    def M(self, dt):
        if dt not in self.mats:
            n = self.n
            r = lambda: 0.1 * torch.randn(n, n, generator=self.g, dtype=torch.float64)  # noqa: E731
            self.mats[dt] = (torch.eye(n, dtype=torch.float64) + r(), r(), r())
        return self.mats[dt]

    def set_state(self, state):
        self.state_ = (state[0].detach().clone(), state[1].detach().clone())

    def state(self):
        return (self.state_[0].clone(), self.state_[1].clone())

    def step(self, dt):
        A, B, C = self.M(dt)
        q, x = self.state_
        self.state_ = (A @ q + B @ x, x + C @ q)

    def adjoint(self, dt, bars):
        A, B, C = self.M(dt)
        qb, xb = bars
        return (A.T @ qb + C.T @ xb, B.T @ qb + xb)


@pytest.mark.parametrize("every", [1, 2, 5, 9])
def test_moving_rollout_gradient_matches_autograd(every):
    """moving_rollout's checkpointed backward over a pair state equals plain autograd through the same maps."""
    n, steps = 5, 9
    solver = PairSolver(n)
    dts = [0.1 * (k % 3 + 1) for k in range(steps)]
    gen = torch.Generator().manual_seed(4)
    q0 = torch.randn(n, 2, dtype=torch.float64, generator=gen, requires_grad=True)
    x0 = torch.randn(n, 2, dtype=torch.float64, generator=gen, requires_grad=True)
    wq, wx = torch.randn(n, 2, dtype=torch.float64, generator=gen), torch.randn(n, 2, dtype=torch.float64, generator=gen)
    qT, xT = autograd.moving_rollout(q0, x0, solver, dts, every)
    ((wq * qT).sum() + (wx * xT).sum()).backward()
    rq, rx = q0.detach().clone().requires_grad_(True), x0.detach().clone().requires_grad_(True)
    q, x = rq, rx
    for dt in dts:
        A, B, C = solver.M(dt)
        q, x = A @ q + B @ x, x + C @ q
    ((wq * q).sum() + (wx * x).sum()).backward()
    assert torch.allclose(q0.grad, rq.grad, rtol=1e-13, atol=1e-13) and torch.allclose(x0.grad, rx.grad, rtol=1e-13, atol=1e-13)
