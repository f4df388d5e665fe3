"""The static step and a checkpointed rollout of steps as torch.autograd Functions over libluna_hydro.so's hand-written
adjoint kernels (FiniteVolume::step_adjoint, finite_volume.hpp:2201-2231).

    from luna_hydro import autograd, hydro
    fv = hydro.FiniteVolume(mesh, 5 / 3)                  # static, order 1, HLLC
    q0 = torch.tensor(Q0, device="cuda", requires_grad=True)
    qT = autograd.rollout(q0, autograd.Solver(fv), [dt] * 200, every=20)
    loss(qT).backward()                                   # q0.grad = J^T dL/dqT through the 200 steps

Q is the extensive conserved state, N x (dim + 2) float64 on the GPU; dt is fixed per step (the derivative does not go
through the time-step rule, as the reference's). The rollout keeps Q every `every` steps and recomputes each segment's
states in the backward pass: memory is (steps / every + every) states instead of steps.

Any object with set_state(tensor), step(dt) -> None, state() -> tensor and adjoint(dt, qbar) -> tensor (J^T qbar at the
state last set) can be the solver; Solver adapts a hydro.FiniteVolume.

On a moving mesh the state is (Q, x) (moving_rollout, MovingSolver): each step rebuilds the mesh of the new positions with
the project's Voronoi producer, and the backward pass runs FiniteVolume::step_moving_adjoint (luna_fv_moving_adjoint) with both
meshes' geometry derivatives; checkpoints keep (Q, x):

    solver = autograd.MovingSolver(hydro.FiniteVolume(mesh, 5 / 3, motion="moving"))
    qT, xT = autograd.moving_rollout(q0, x0, solver, [dt] * 200, every=20)
    autograd.density_misfit(qT, xT, target).backward()      # q0.grad, x0.grad
"""
import ctypes

import numpy as np
import torch


class _DeviceRows:
    """A device pointer as torch sees one (__cuda_array_interface__), to copy into the solver's own Q."""

    def __init__(self, ptr, shape):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": "<f8", "data": (ptr, False), "version": 3, "strides": None}


class Solver:
    """Adapts hydro.FiniteVolume (a static mesh) to the rollout: Q lives in the library's device rows."""

    def __init__(self, fv):
        from . import hydro

        self.fv = fv
        self._hydro = hydro
        self.shape = (fv.N, fv.V)

    def _q(self):
        ptr = ctypes.POINTER(ctypes.c_double)()
        self._hydro._check(self._hydro.lib().luna_fv_state_device(self.fv._h, ctypes.byref(ptr)), "state_device")
        addr = ctypes.cast(ptr, ctypes.c_void_p).value
        return torch.as_tensor(_DeviceRows(addr, self.shape), device="cuda")

    def set_state(self, q):
        self._q().copy_(q)
        torch.cuda.synchronize()  # the library runs on its own stream

    def state(self):
        out = self._q().clone()
        torch.cuda.synchronize()
        return out

    def step(self, dt):
        self.fv.step(dt)  # the step sweeps Q first (the cached sweep went with the new state)
        self._hydro.synchronize()

    def adjoint(self, dt, qbar):
        out = qbar.detach().clone().contiguous()
        torch.cuda.synchronize()
        self.fv.step_adjoint(dt, out)
        return out


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, solver, dt):
        q0 = q.detach().contiguous()
        solver.set_state(q0)
        solver.step(dt)
        ctx.save_for_backward(q0)
        ctx.solver, ctx.dt = solver, dt
        return solver.state()

    @staticmethod
    def backward(ctx, qbar):
        (q0,) = ctx.saved_tensors
        ctx.solver.set_state(q0)
        return ctx.solver.adjoint(ctx.dt, qbar), None, None


def step(q, solver, dt):
    """One static step Q -> Q' at fixed dt, differentiable in Q."""
    return _Step.apply(q, solver, dt)


def checkpoints(steps, every):
    """The step indices whose input states the rollout keeps: 0, every, 2 every, ... (< steps)."""
    if every < 1:
        raise ValueError("every >= 1")
    return list(range(0, steps, every))


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, solver, dts, every):
        kept = {}
        state = q.detach().contiguous()
        solver.set_state(state)
        for k, dt in enumerate(dts):
            if k % every == 0:
                kept[k] = solver.state()
            solver.step(dt)
        ctx.solver, ctx.dts, ctx.every, ctx.kept = solver, list(dts), every, kept
        return solver.state()

    @staticmethod
    def backward(ctx, qbar):
        solver, dts, every = ctx.solver, ctx.dts, ctx.every
        bar = qbar.detach().contiguous()
        for start in reversed(checkpoints(len(dts), every)):
            end = min(start + every, len(dts))
            states = [ctx.kept[start]]  # recompute the segment's step inputs from its checkpoint
            solver.set_state(states[0])
            for k in range(start, end - 1):
                solver.step(dts[k])
                states.append(solver.state())
            for k in reversed(range(start, end)):
                solver.set_state(states[k - start])
                bar = solver.adjoint(dts[k], bar)
        return bar, None, None, None


def rollout(q, solver, dts, every=20):
    """len(dts) static steps from Q at the given dts, differentiable in Q, keeping Q every `every` steps and recomputing the
    rest in the backward pass."""
    return _Rollout.apply(q, solver, list(dts), int(every))


class MovingSolver:
    """Adapts hydro.FiniteVolume on a moving mesh to moving_rollout: the state is (Q, x), host float64 tensors (N x V, N x D);
    set_state builds the mesh of x; step runs moving_begin, the producer, moving_finish; adjoint(dt, (qbar, xbar)) at the
    state last set runs that step's forward to reach the mesh of t^n+1, then luna_fv_moving_adjoint. steps_fixed: hold
    the whole grid steps (derivative 0) instead of taking the quantisation as the identity."""

    def __init__(self, fv, side=1.0, steps_fixed=False):
        from . import mesh

        self.fv, self.side, self.steps_fixed = fv, side, steps_fixed
        self._mesh = mesh
        self.mesh = None

    def set_state(self, state):
        q, x = state
        self.mesh = self._mesh.build(x.detach().cpu().numpy(), self.side)
        self.fv.set_geometry(self.mesh)
        self.fv.set_state(q.detach().cpu().numpy())

    def state(self):
        return (torch.tensor(self.fv.state()), torch.tensor(self.mesh["positions"]))

    def step(self, dt):
        x1 = self.fv.moving_begin(dt)
        self.mesh = self._mesh.build(x1, self.side)
        self.fv.moving_finish(self.mesh, dt)

    def adjoint(self, dt, bars):
        m0, q0 = self.mesh, self.fv.state()
        self.step(dt)
        m1 = self.mesh
        self.fv.set_state(q0)
        qb, xb = self.fv.moving_adjoint(dt, m0, m1, bars[0].detach().cpu().numpy(), bars[1].detach().cpu().numpy(), self.steps_fixed)
        self.mesh = m0
        return (torch.tensor(qb), torch.tensor(xb))


class _MovingRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, x, solver, dts, every):
        kept = {}
        solver.set_state((q.detach(), x.detach()))
        for k, dt in enumerate(dts):
            if k % every == 0:
                kept[k] = solver.state()
            solver.step(dt)
        ctx.solver, ctx.dts, ctx.every, ctx.kept = solver, list(dts), every, kept
        return solver.state()

    @staticmethod
    def backward(ctx, qbar, xbar):
        solver, dts, every = ctx.solver, ctx.dts, ctx.every
        bars = (qbar.detach().contiguous(), xbar.detach().contiguous())
        for start in reversed(checkpoints(len(dts), every)):
            end = min(start + every, len(dts))
            states = [ctx.kept[start]]
            solver.set_state(states[0])
            for k in range(start, end - 1):
                solver.step(dts[k])
                states.append(solver.state())
            for k in reversed(range(start, end)):
                solver.set_state(states[k - start])
                bars = solver.adjoint(dts[k], bars)
        return bars[0], bars[1], None, None, None


def moving_rollout(q, x, solver, dts, every=20):
    """len(dts) moving steps from (Q, x) at the given dts, differentiable in both, keeping (Q, x) every `every` steps and
    recomputing the rest (and each step's mesh) in the backward pass. Returns (Q, x) after the last step."""
    return _MovingRollout.apply(q, x, solver, list(dts), int(every))


class _DensityMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, x, target, side):
        from . import mesh

        m = mesh.build(x.detach().cpu().numpy(), side)
        vol = torch.tensor(m["volume"])
        rho = q.detach()[:, 0] / vol
        ctx.save_for_backward(rho, vol, target)
        ctx.mesh = m
        return 0.5 * torch.sum(vol * (rho - target) ** 2)

    @staticmethod
    def backward(ctx, gL):
        from . import mesh

        rho, vol, target = ctx.saved_tensors
        m = ctx.mesh
        N, D = m["cells"], m["dim"]
        qbar = torch.zeros(N, D + 2, dtype=torch.float64)
        qbar[:, 0] = gL * (rho - target)  # L = 1/2 sum (q_0^2 / V - 2 q_0 rho* + rho*^2 V)
        Vbar = (gL * 0.5 * (target ** 2 - rho ** 2)).numpy()
        g = mesh.GeometryDerivative(m)
        xbar = g.adjoint(Vbar, np.zeros((N, D)), np.zeros((m["faces"], D)), np.zeros((m["faces"], D)))
        g.close()
        return qbar, torch.tensor(xbar), None, None


def density_misfit(q, x, target, side=1.0):
    """L = 1/2 sum_i V_i (rho_i - rho*_i)^2 of the state (Q, x) on the mesh of x (rho = Q_0 / V), differentiable in Q and x
    (V through the mesh library's geometry derivative)."""
    return _DensityMisfit.apply(q, x, target, side)
