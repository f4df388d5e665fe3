"""The oracle's chain of moving-step adjoints over many steps of the smooth acoustic wave (rho = 1 + 1e-3 sin(2 pi (x + y)),
v = 0, p = 3/5, gamma 5/3; the density misfit after the last step), against central differences of the oracle's forward
with the whole grid steps held.

A smooth state is where such a check is hardest: the face-centroid limiter clips at smooth extrema and HLLC switches at
S* = w on faces that move with the flow, so the forward is only piecewise smooth, and a perturbation of size h crosses
a number of switches proportional to h. Along directions of the initial density, the difference between central
differences and the adjoint therefore shrinks with h (pinned: below 1e-4 at h = 1e-9, and at least 10x below its
value at h = 1e-6). Directions of the initial velocity start exactly on the limiter's switch (v = 0 and p uniform: the
neighbours' range of v collapses to a point, and the limited gradient is a positively homogeneous, non-linear function
of the perturbation), so there the two differ at any h; the reference reports its own multi-step HLLC check for the same
reasons (hydro_derivatives_test.cpp:1641-1643)."""
import numpy as np
import pytest

import oracle
from luna_hydro import mesh, problems

GAMMA = 5.0 / 3.0
PRESSURE = 3.0 / 5.0


def conserved(rho, vel, vol):
    energy = PRESSURE / (GAMMA - 1) + 0.5 * rho * np.sum(vel ** 2, axis=1)
    return np.column_stack([rho, rho[:, None] * vel, energy]) * vol[:, None]


class Chain:
    def __init__(self, n, steps, riemann):
        self.m0 = mesh.build(problems.jittered_lattice(n, 3, 0.25, 1), 1.0)
        self.N, self.steps, self.riemann = self.m0["cells"], steps, riemann
        q0 = problems.acoustic_state(self.m0, GAMMA)
        self.rho0, self.vel0 = q0[:, 0] / self.m0["volume"], np.zeros((self.N, 3))
        fv = oracle.FiniteVolume(self.m0, GAMMA, moving=True, riemann=riemann)
        fv.set_state(q0)
        self.dt = fv.time_step(0.3)
        self.meshes = None
        noise = np.random.Generator(np.random.MT19937(7)).normal(size=self.N)
        qT, ms, _, _ = self.run(self.rho0 * (1 + 1e-4 * noise), self.vel0)
        self.target = qT[:, 0] / ms[-1]["volume"]
        self.qT, self.meshes, self.qs, self.held = self.run(self.rho0, self.vel0)

    def run(self, rho, vel, held=None):
        """The forward from (rho, v) on the first mesh; with `held` the whole grid steps of each step are given (the positions,
        and so the meshes, are those of the unperturbed run)."""
        fv = oracle.FiniteVolume(self.m0, GAMMA, moving=True, riemann=self.riemann)
        fv.set_state(conserved(rho, vel, self.m0["volume"]))
        qs, meshes, steps = [], [self.m0], []
        for k in range(self.steps):
            qs.append(fv.q.copy())
            fv.time_step(0.3)
            if held is not None:
                fv.c.hold_steps = 1
                fv.steps[...] = held[k]
            x1 = fv.moving_begin(self.dt)
            steps.append(fv.steps.copy())
            m1 = self.meshes[k + 1] if held is not None else mesh.build(x1, 1.0)
            fv.moving_finish(m1, self.dt)
            meshes.append(m1)
        return fv.q.copy(), meshes, qs, steps

    def loss(self, q, m):
        return 0.5 * np.sum(m["volume"] * (q[:, 0] / m["volume"] - self.target) ** 2)

    def gradient(self):
        """dL/d(rho0, v0) by the chain of the oracle's step adjoints with the grid steps held, then Q0(rho0, v0)'s."""
        mT, qT, N = self.meshes[-1], self.qT, self.N
        rho = qT[:, 0] / mT["volume"]
        qbar = np.zeros_like(qT)
        qbar[:, 0] = rho - self.target
        gT = mesh.GeometryDerivative(mT)
        xbar = gT.adjoint(0.5 * (self.target ** 2 - rho ** 2), np.zeros((N, 3)), np.zeros((mT["faces"], 3)), np.zeros((mT["faces"], 3)))
        for k in reversed(range(self.steps)):
            fv = oracle.FiniteVolume(self.meshes[k], GAMMA, moving=True, riemann=self.riemann)
            fv.set_state(self.qs[k])
            fv.c.hold_steps = 1
            fv.steps[...] = self.held[k]
            g0, g1 = mesh.GeometryDerivative(self.meshes[k]), mesh.GeometryDerivative(self.meshes[k + 1])
            qbar, xbar = fv.moving_adjoint(self.dt, qbar, xbar, self.meshes[k + 1], lambda *a: g0.adjoint(*a),
                                           lambda *a: g1.adjoint(*a), steps_fixed=True)
        vol = self.m0["volume"]
        # Q0 = V (rho, rho v, p / (gamma - 1) + rho |v|^2 / 2) at v = 0: dQ_0 / drho = V, dQ_1+d / dv_d = V rho
        return qbar[:, 0] * vol, qbar[:, 1:4] * (vol * self.rho0)[:, None]

    def central(self, u_rho, u_vel, h):
        L = []
        for sign in (1, -1):
            q, ms, _, _ = self.run(self.rho0 + sign * h * u_rho, self.vel0 + sign * h * u_vel, self.held)
            L.append(self.loss(q, ms[-1]))
        return (L[0] - L[1]) / (2 * h)


@pytest.fixture(scope="module")
def chain():
    return Chain(6, 20, "hllc")


def test_acoustic_rollout_density_directions_converge(chain):
    g_rho, g_vel = chain.gradient()
    N = chain.N
    for j in range(3):
        u = np.random.Generator(np.random.MT19937(1000 + j)).normal(size=N)
        ad = float(g_rho @ u)
        errs = {h: abs(chain.central(u, np.zeros((N, 3)), h) - ad) / abs(ad) for h in (1e-6, 1e-9)}
        print("direction %d: |central - adjoint| / |adjoint| at h 1e-6 %.2e, 1e-9 %.2e" % (j, errs[1e-6], errs[1e-9]))
        assert errs[1e-9] < 1e-4, errs
        assert errs[1e-9] < 0.1 * errs[1e-6] or errs[1e-6] < 1e-6, errs


def test_acoustic_rollout_velocity_directions_start_on_the_limiter_switch(chain):
    """At v = 0 a velocity perturbation's limited gradients scale with its size (positively homogeneous), so central
    differences stay off the adjoint as h shrinks: reported, with the density part's agreement as the control."""
    g_rho, g_vel = chain.gradient()
    u = np.random.Generator(np.random.MT19937(2000)).normal(size=(chain.N, 3))
    ad = float(np.sum(g_vel * u))
    errs = [abs(chain.central(np.zeros(chain.N), u, h) - ad) / abs(ad) for h in (1e-6, 1e-9)]
    print("velocity direction: |central - adjoint| / |adjoint| at h 1e-6 %.2e, 1e-9 %.2e" % tuple(errs))
    assert min(errs) > 1e-6  # not a smooth point: the two do not meet
