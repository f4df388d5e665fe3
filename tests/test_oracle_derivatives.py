"""Pins the oracle's derivatives of a static step (finite_volume_1d.hpp:710-935, euler.hpp:115-832): the 1D step's tangent
and adjoint through 5 fixed steps against LUNA's JAX jvp / vjp (hydro_derivatives_test.cpp, luna_derivative_values.hpp,
tolerance 1e-8), and each pointwise derivative (primitive with its floors, the half step, HLLC in every wave case) against
central differences and its own transpose (the dot test <ybar, J dx> = <J^T ybar, dx>)."""
import numpy as np
import pytest

import oracle
from golden_data import load
from luna_hydro import problems

LIMITER = {0: "minmod", 1: "face", 2: "vanleer"}
# Sod with MC starts exactly flat on either side of the jump, where MC has no derivative: LUNA's jvp takes the zero branch
# of its where(a b > 0, ...), the reference (and this oracle) the unlimited central slope. The reference reports that case
# and does not gate it (hydro_derivatives_test.cpp:1077-1081); the noisy Sod with no flat cell is gated.
UNGATED = {"luna_d_sod_64_l1"}
GATED = sorted(set(load("luna_1d_forward")) - UNGATED)


def golden_line(c):
    pb = problems.problem_1d(c["problem"])
    x = problems.line_points(c["n"], pb.L)
    keys, step, period = problems.line_keys(x, pb.L)
    line = oracle.Line(keys, period, step, pb.gamma, periodic=pb.periodic, walls=pb.walls, limiter=LIMITER[c["limiter"]])
    V = np.array([line.volume(s) for s in range(c["n"])])
    return line, V


@pytest.mark.parametrize("case", GATED)
def test_line_tangent_matches_luna_jvp(case):
    c = load("luna_1d_forward")[case]
    line, V = golden_line(c)
    line.set_state(np.array(c["U0"]) * V[:, None])
    dq = np.array(c["dU"]) * V[:, None]
    for _ in range(c["steps"]):
        dq = line.step_tangent(c["dt"], dq)
    JdU = dq / V[:, None]
    ref = np.array(c["JdU"])
    assert np.max(np.abs(JdU - ref)) / np.max(np.abs(ref)) < 1e-8
    # the tangent's own forward is the step: U after 5 steps
    assert np.max(np.abs(line.state / V[:, None] - np.array(c["U"]))) / np.max(np.abs(c["U"])) < 1e-12


@pytest.mark.parametrize("case", GATED)
def test_line_adjoint_matches_luna_vjp(case):
    c = load("luna_1d_forward")[case]
    line, V = golden_line(c)
    q = np.array(c["U0"]) * V[:, None]
    states = []
    line.set_state(q)
    for _ in range(c["steps"]):
        states.append(line.state.copy())
        line.step(c["dt"])
    qbar = np.array(c["w"]) / V[:, None]  # <w, U> with U = q / V
    for k in reversed(range(c["steps"])):
        line.set_state(states[k])
        qbar = line.step_adjoint(c["dt"], qbar)
    JTw = qbar * V[:, None]
    ref = np.array(c["JTw"])
    assert np.max(np.abs(JTw - ref)) / np.max(np.abs(ref)) < 1e-8


def test_planted_adjoint_sign_fails_vjp():
    """Teeth: the vjp with the cotangent's sign flipped on one cell must miss LUNA's value."""
    c = load("luna_1d_forward")["luna_d_sod_64_l1"]
    line, V = golden_line(c)
    line.set_state(np.array(c["U0"]) * V[:, None])
    w = np.array(c["w"])
    w[30] = -w[30]
    qbar = line.step_adjoint(c["dt"], w / V[:, None])
    assert np.max(np.abs(qbar * V[:, None] - np.array(c["JTw"]))) > 1e-6


def rng_unit(rng, n):
    return rng.uniform(-1, 1, n)


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("floor", ["none", "density", "energy"])
def test_primitive_derivatives_fd_and_transpose(D, floor):
    rng = np.random.default_rng(7 + D)
    gamma = 5.0 / 3.0
    W = np.concatenate([[1.3], rng_unit(rng, D), [0.7]])
    U = oracle.conserved(W, gamma)
    smallr, smallc = 1e-10, 1e-10
    if floor == "density":
        smallr = 2.0  # rho floored
    if floor == "energy":
        smallc = 10.0  # e floored
    dU = rng_unit(rng, D + 2)
    h = 1e-6
    fd = (oracle.primitive(U + h * dU, gamma) - oracle.primitive(U - h * dU, gamma)) / (2 * h) if floor == "none" else None
    dW = oracle.primitive_tangent(U, dU, gamma, smallr, smallc)
    if fd is not None:
        assert np.allclose(dW, fd, rtol=1e-7, atol=1e-9)
    Wbar = rng_unit(rng, D + 2)
    Ubar = oracle.primitive_adjoint(U, Wbar, gamma, smallr, smallc)
    assert abs(Wbar @ dW - Ubar @ dU) <= 1e-14 * (1 + abs(Wbar @ dW))
    if floor == "density":
        assert dW[0] == 0 and Ubar[0] == 0
    if floor == "energy":
        assert abs(dW[D + 1] - (gamma - 1) * smallc ** 2 / gamma / (gamma - 1) * dW[0]) < 1e-12


@pytest.mark.parametrize("D", [1, 2, 3])
def test_half_step_derivatives_fd_and_transpose(D):
    rng = np.random.default_rng(11 + D)
    gamma, dt = 1.4, 0.03
    W = np.concatenate([[1.1], rng_unit(rng, D), [0.9]])
    g = rng_unit(rng, (D + 2) * D).reshape(D + 2, D)
    dW, dg = rng_unit(rng, D + 2), rng_unit(rng, (D + 2) * D).reshape(D + 2, D)
    h = 1e-6
    fd = (oracle.half_step(W + h * dW, g + h * dg, dt, gamma) - oracle.half_step(W - h * dW, g - h * dg, dt, gamma)) / (2 * h)
    t = oracle.half_step_tangent(W, dW, g, dg, dt, gamma)
    assert np.allclose(t, fd, rtol=1e-8, atol=1e-10)
    hbar = rng_unit(rng, D + 2)
    Wbar, gbar = oracle.half_step_adjoint(W, g, dt, hbar, gamma)
    assert abs(hbar @ t - (Wbar @ dW + np.sum(gbar * dg))) < 1e-13


# States chosen for each HLLC case: supersonic to the right (F_L), subsonic both ways (F*_L, F*_R), supersonic left (F_R),
# on a moving face.
HLLC_CASES = {
    "F_L": ([1.0, 3.0, 0.2, -0.1, 1.0], [0.5, 3.2, 0.1, 0.0, 0.8]),
    "star_L": ([1.0, 0.3, 0.2, -0.1, 1.0], [0.5, -0.2, 0.1, 0.0, 0.8]),
    "star_R": ([0.4, -0.3, 0.2, -0.1, 0.6], [1.5, -0.5, 0.1, 0.0, 1.8]),
    "F_R": ([1.0, -3.0, 0.2, -0.1, 1.0], [0.5, -3.2, 0.1, 0.0, 0.8]),
}


@pytest.mark.parametrize("case", sorted(HLLC_CASES))
def test_hllc_derivatives_fd_and_transpose(case):
    rng = np.random.default_rng(3)
    gamma = 5.0 / 3.0
    L, R = (np.array(x) for x in HLLC_CASES[case])
    n = np.array([0.6, 0.8, 0.0])
    w = np.array([0.05, -0.02, 0.01])
    dL, dR = rng_unit(rng, 5), rng_unit(rng, 5)
    F, dF = oracle.hllc_tangent(L, R, dL, dR, n, w, gamma)
    assert np.array_equal(F, oracle.flux(L, R, n, w, gamma))  # the same operations as the forward
    h = 1e-6
    fd = (oracle.flux(L + h * dL, R + h * dR, n, w, gamma) - oracle.flux(L - h * dL, R - h * dR, n, w, gamma)) / (2 * h)
    assert np.allclose(dF, fd, rtol=1e-7, atol=1e-8), (dF, fd)
    Fbar = rng_unit(rng, 5)
    Lbar, Rbar = oracle.hllc_adjoint(L, R, n, w, Fbar, gamma)
    assert abs(Fbar @ dF - (Lbar @ dL + Rbar @ dR)) < 1e-12 * (1 + abs(Fbar @ dF))


@pytest.mark.parametrize("walls,periodic,limiter,order", [
    (("reflective", "reflective"), False, "face", 1),
    (("outflow", "reflective"), False, "vanleer", 1),
    (("outflow", "outflow"), True, "minmod", 1),
    (("reflective", "outflow"), False, "face", 0),
    (("outflow", "outflow"), False, "none", 1),
])
def test_line_step_derivatives_dot_and_fd(walls, periodic, limiter, order):
    """A rough jittered line with each wall kind: the step's adjoint is the transpose of its tangent (dot test), and the
    tangent matches central differences of the step."""
    rng = np.random.default_rng(5)
    n = 200
    keys, step, period = problems.line_keys(problems.line_points(n, 1.0, 0.3, 2), 1.0)
    line = oracle.Line(keys, period, step, 1.4, periodic=periodic, walls=walls, limiter=limiter, order=order)
    V = np.array([line.volume(s) for s in range(n)])
    W = np.stack([1 + 0.3 * rng.uniform(size=n), 0.4 * rng.uniform(-1, 1, n), 1 + 0.3 * rng.uniform(size=n)], axis=1)
    q = np.array([oracle.conserved(w, 1.4) for w in W]) * V[:, None]
    dt, h = 2e-4, 1e-7
    dq, qbar = rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3))
    line.set_state(q)
    J = line.step_tangent(dt, dq)
    line.set_state(q)
    T = line.step_adjoint(dt, qbar)
    lhs, rhs = np.sum(qbar * J), np.sum(T * dq)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    line.set_state(q + h * dq)
    line.step(dt)
    plus = line.state.copy()
    line.set_state(q - h * dq)
    line.step(dt)
    fd = (plus - line.state) / (2 * h)
    assert np.max(np.abs(fd - J)) / np.max(np.abs(J)) < 1e-6


def mesh_state(dim, n, seed=1, amp=0.3):
    import cases

    m = cases.voronoi(n, dim, 0.25, 3)
    rng = np.random.default_rng(seed)
    N, V = m["cells"], dim + 2
    W = np.zeros((N, V))
    W[:, 0] = 1 + amp * rng.uniform(size=N)
    W[:, 1:1 + dim] = amp * rng.uniform(-1, 1, (N, dim))
    W[:, -1] = 1 + amp * rng.uniform(size=N)
    q = np.array([oracle.conserved(w, 5.0 / 3.0) for w in W]) * m["volume"][:, None]
    return m, q, rng


def forward(fv, q, dt):
    fv.set_state(q)
    fv.time_step(0.3)  # the sweep the step reuses
    fv.step(dt)
    return fv.q.copy()


@pytest.mark.parametrize("dim,n,limiter", [(3, 6, "face"), (3, 6, "none"), (2, 12, "face"), (2, 12, "none")])
def test_mesh_step_derivatives_dot_fd_conservation(dim, n, limiter):
    """The static order-1 step on a rough jittered Voronoi mesh: the tangent's own forward is the step (bits), the adjoint is
    its transpose (dot test), the tangent matches central differences to 1e-6 relative, and J^T 1 = 1 (the step conserves
    every sum of Q)."""
    m, q, rng = mesh_state(dim, n)
    N, V = q.shape
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, limiter=limiter)
    fv.set_state(q)
    dt = 0.3 * fv.time_step(0.3)
    dq, qbar = rng.uniform(-1, 1, (N, V)), rng.uniform(-1, 1, (N, V))
    fv.set_state(q)
    J = fv.step_tangent(dt, dq)
    assert np.array_equal(fv.q, forward(fv, q, dt))
    fv.set_state(q)
    T = fv.step_adjoint(dt, qbar)
    lhs, rhs = np.sum(qbar * J), np.sum(T * dq)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    h = 1e-7
    fd = (forward(fv, q + h * dq, dt) - forward(fv, q - h * dq, dt)) / (2 * h)
    assert np.max(np.abs(fd - J)) / np.max(np.abs(J)) < 1e-6
    fv.set_state(q)
    assert np.array_equal(fv.step_adjoint(dt, np.ones((N, V))), np.ones((N, V)))


def test_planted_adjoint_cell_swap_fails_dot_test():
    """Teeth: the adjoint read back with two cells' rows swapped must fail the dot test."""
    m, q, rng = mesh_state(3, 6)
    N, V = q.shape
    fv = oracle.FiniteVolume(m, 5.0 / 3.0)
    fv.set_state(q)
    dt = 0.3 * fv.time_step(0.3)
    dq, qbar = rng.uniform(-1, 1, (N, V)), rng.uniform(-1, 1, (N, V))
    fv.set_state(q)
    J = fv.step_tangent(dt, dq)
    fv.set_state(q)
    T = fv.step_adjoint(dt, qbar)
    T[[3, 17]] = T[[17, 3]]
    assert abs(np.sum(qbar * J) - np.sum(T * dq)) > 1e-6


@pytest.mark.parametrize("case", sorted(HLLC_CASES))
def test_hll_derivatives_fd_and_transpose(case):
    """HLL::flux_tangent and flux_adjoint (euler.hpp:466-540) in each case (F_L, the HLL state, F_R) on a moving face."""
    rng = np.random.default_rng(9)
    gamma = 5.0 / 3.0
    L, R = (np.array(x) for x in HLLC_CASES[case])
    n = np.array([0.6, 0.8, 0.0])
    w = np.array([0.05, -0.02, 0.01])
    dL, dR = rng_unit(rng, 5), rng_unit(rng, 5)
    F, dF = oracle.hllc_tangent(L, R, dL, dR, n, w, gamma, riemann="hll")
    assert np.array_equal(F, oracle.flux(L, R, n, w, gamma, riemann="hll"))
    h = 1e-6
    fd = (oracle.flux(L + h * dL, R + h * dR, n, w, gamma, "hll") - oracle.flux(L - h * dL, R - h * dR, n, w, gamma, "hll")) / (2 * h)
    assert np.allclose(dF, fd, rtol=1e-7, atol=1e-8), (dF, fd)
    Fbar = rng_unit(rng, 5)
    Lbar, Rbar = oracle.hllc_adjoint(L, R, n, w, Fbar, gamma, riemann="hll")
    assert abs(Fbar @ dF - (Lbar @ dL + Rbar @ dR)) < 1e-12 * (1 + abs(Fbar @ dF))


@pytest.mark.parametrize("dim,n", [(3, 6), (2, 12)])
def test_mesh_step_derivatives_hll(dim, n):
    m, q, rng = mesh_state(dim, n)
    N, V = q.shape
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, riemann="hll")
    fv.set_state(q)
    dt = 0.3 * fv.time_step(0.3)
    dq, qbar = rng.uniform(-1, 1, (N, V)), rng.uniform(-1, 1, (N, V))
    fv.set_state(q)
    J = fv.step_tangent(dt, dq)
    assert np.array_equal(fv.q, forward(fv, q, dt))
    fv.set_state(q)
    T = fv.step_adjoint(dt, qbar)
    assert abs(np.sum(qbar * J) - np.sum(T * dq)) <= 1e-13 * abs(np.sum(qbar * J))
    h = 1e-7
    fd = (forward(fv, q + h * dq, dt) - forward(fv, q - h * dq, dt)) / (2 * h)
    assert np.max(np.abs(fd - J)) / np.max(np.abs(J)) < 1e-6


def test_line_step_derivatives_hll():
    rng = np.random.default_rng(6)
    n = 120
    keys, step, period = problems.line_keys(problems.line_points(n, 1.0, 0.3, 2), 1.0)
    line = oracle.Line(keys, period, step, 1.4, walls=("reflective", "outflow"), riemann="hll")
    V = np.array([line.volume(s) for s in range(n)])
    W = np.stack([1 + 0.3 * rng.uniform(size=n), 0.4 * rng.uniform(-1, 1, n), 1 + 0.3 * rng.uniform(size=n)], axis=1)
    q = np.array([oracle.conserved(w, 1.4) for w in W]) * V[:, None]
    dt, h = 2e-4, 1e-7
    dq, qbar = rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3))
    line.set_state(q)
    J = line.step_tangent(dt, dq)
    line.set_state(q)
    T = line.step_adjoint(dt, qbar)
    assert abs(np.sum(qbar * J) - np.sum(T * dq)) <= 1e-13 * abs(np.sum(qbar * J))
    line.set_state(q + h * dq)
    line.step(dt)
    plus = line.state.copy()
    line.set_state(q - h * dq)
    line.step(dt)
    assert np.max(np.abs((plus - line.state) / (2 * h) - J)) / np.max(np.abs(J)) < 1e-6
