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
"""
import ctypes

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
