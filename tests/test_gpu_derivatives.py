"""Derivatives of the static step on the GPU (through the C-ABI) against the oracle's on the same inputs, and against
LUNA's jax.jvp / jax.vjp through 5 fixed steps (luna_derivative_values.hpp; the reference's gate: 1e-8, with the same
exactly-flat Sod/MC case reported and not gated, hydro_derivatives_test.cpp:1077-1081). The GPU-oracle bar is 1e-12
relative per entry (the same operation order on both sides; the test reports whether the bits agree)."""
import numpy as np
import pytest

import cases
import oracle
from golden_data import load
from luna_hydro import hydro, problems

pytestmark = pytest.mark.gpu
TOL = 1e-12
LIMITER = {0: "minmod", 1: "face", 2: "vanleer"}
UNGATED = {"luna_d_sod_64_l1"}


def say(msg):
    print(msg, flush=True)


def close(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) / scale


def lines(c):
    pb = problems.problem_1d(c["problem"])
    keys, step, period = problems.line_keys(problems.line_points(c["n"], pb.L), pb.L)
    kw = dict(periodic=pb.periodic, walls=pb.walls, limiter=LIMITER[c["limiter"]])
    gpu = hydro.Line(keys, period, step, pb.gamma, **kw)
    orc = oracle.Line(keys, period, step, pb.gamma, **kw)
    V = np.array([orc.volume(s) for s in range(c["n"])])
    return gpu, orc, V


@pytest.mark.parametrize("case", sorted(load("luna_1d_forward")))
def test_line_tangent_matches_oracle_and_luna(case):
    c = load("luna_1d_forward")[case]
    gpu, orc, V = lines(c)
    q0 = np.array(c["U0"]) * V[:, None]
    gpu.set_state(q0)
    orc.set_state(q0)
    dg = do = np.array(c["dU"]) * V[:, None]
    for k in range(c["steps"]):
        dg = gpu.step_tangent(c["dt"], dg)
        do = orc.step_tangent(c["dt"], do)
        assert close(dg, do) <= TOL, k
        assert close(gpu.state(), orc.state) <= TOL, k
    say("%s tangent: bits equal %s" % (case, np.array_equal(dg, do)))
    if case not in UNGATED:
        assert close(dg / V[:, None], np.array(c["JdU"])) < 1e-8


@pytest.mark.parametrize("case", sorted(load("luna_1d_forward")))
def test_line_adjoint_matches_oracle_and_luna(case):
    c = load("luna_1d_forward")[case]
    gpu, orc, V = lines(c)
    orc.set_state(np.array(c["U0"]) * V[:, None])
    states = []
    for _ in range(c["steps"]):
        states.append(orc.state.copy())
        orc.step(c["dt"])
    bg = bo = np.array(c["w"]) / V[:, None]
    for k in reversed(range(c["steps"])):
        gpu.set_state(states[k])
        orc.set_state(states[k])
        bg = gpu.step_adjoint(c["dt"], bg)
        bo = orc.step_adjoint(c["dt"], bo)
        assert close(bg, bo) <= TOL, k
    say("%s adjoint: bits equal %s" % (case, np.array_equal(bg, bo)))
    if case not in UNGATED:
        assert close(bg * V[:, None], np.array(c["JTw"])) < 1e-8


@pytest.mark.parametrize("walls,periodic,limiter,order,riemann", [
    (("reflective", "reflective"), False, "face", 1, "hllc"),
    (("outflow", "reflective"), False, "vanleer", 1, "hllc"),
    (("outflow", "outflow"), True, "minmod", 1, "hllc"),
    (("reflective", "outflow"), False, "face", 0, "hllc"),
    (("reflective", "outflow"), False, "face", 1, "hll"),
])
def test_line_derivatives_walls_dot_product(walls, periodic, limiter, order, riemann):
    """Rough jittered lines with each wall: GPU tangent and adjoint against the oracle's, and the dot test
    <qbar, J dq> = <J^T qbar, dq> on the GPU."""
    rng = np.random.default_rng(5)
    n = 200
    x = problems.line_points(n, 1.0, 0.3, 2)
    keys, step, period = problems.line_keys(x, 1.0)
    kw = dict(periodic=periodic, walls=walls, limiter=limiter, order=order, riemann=riemann)
    gpu = hydro.Line(keys, period, step, 1.4, **kw)
    orc = oracle.Line(keys, period, step, 1.4, **kw)
    V = np.array([orc.volume(s) for s in range(n)])
    W = np.stack([1 + 0.3 * rng.uniform(size=n), 0.4 * rng.uniform(-1, 1, n), 1 + 0.3 * rng.uniform(size=n)], axis=1)
    q = np.array([oracle.conserved(w, 1.4) for w in W]) * V[:, None]
    dt = 2e-4
    dq, qbar = rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3))
    gpu.set_state(q)
    orc.set_state(q)
    Jg = gpu.step_tangent(dt, dq)
    Jo = orc.step_tangent(dt, dq)
    assert close(Jg, Jo) <= TOL
    gpu.set_state(q)
    orc.set_state(q)
    Tg = gpu.step_adjoint(dt, qbar)
    To = orc.step_adjoint(dt, qbar)
    assert close(Tg, To) <= TOL
    lhs, rhs = np.sum(qbar * Jg), np.sum(Tg * dq)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    say("walls %s periodic %s %s order %d: tangent bits %s, adjoint bits %s, dot %.3e" % (walls, periodic, limiter, order,
        np.array_equal(Jg, Jo), np.array_equal(Tg, To), abs(lhs - rhs)))


def mesh_case(dim, n, seed=1, amp=0.3):
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


@pytest.mark.parametrize("dim,n,limiter,riemann", [(3, 10, "face", "hllc"), (3, 8, "none", "hllc"), (3, 8, "face", "hll"),
                                                   (2, 24, "face", "hllc"), (2, 24, "none", "hll")])
def test_mesh_derivatives_match_oracle(dim, n, limiter, riemann):
    """The static order-1 step's tangent and adjoint on a rough jittered Voronoi mesh: the GPU's against the oracle's
    (1e-12 relative), the GPU adjoint as the transpose of the GPU tangent (dot test), and J^T 1 = 1."""
    m, q, rng = mesh_case(dim, n)
    N, V = q.shape
    gpu = hydro.FiniteVolume(m, 5.0 / 3.0, limiter=limiter, riemann=riemann)
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, limiter=limiter, riemann=riemann)
    orc.set_state(q)
    dt = 0.3 * orc.time_step(0.3)
    dq, qbar = rng.uniform(-1, 1, (N, V)), rng.uniform(-1, 1, (N, V))
    gpu.set_state(q)
    orc.set_state(q)
    Jg, Jo = gpu.step_tangent(dt, dq), orc.step_tangent(dt, dq)
    assert close(Jg, Jo) <= TOL
    assert close(gpu.state(), orc.q) <= TOL
    gpu.set_state(q)
    orc.set_state(q)
    Tg, To = gpu.step_adjoint(dt, qbar), orc.step_adjoint(dt, qbar)
    assert close(Tg, To) <= TOL
    assert np.array_equal(gpu.state(), q)  # the adjoint leaves Q
    lhs, rhs = np.sum(qbar * Jg), np.sum(Tg * dq)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert np.array_equal(gpu.step_adjoint(dt, np.ones((N, V))), np.ones((N, V)))
    say("%dD n=%d %s %s: tangent bits %s, adjoint bits %s, dot %.2e" % (dim, n, limiter, riemann, np.array_equal(Jg, Jo), np.array_equal(Tg, To),
                                                                   abs(lhs - rhs) / abs(lhs)))


def test_mesh_adjoint_device_rows():
    """The adjoint on device rows (a torch tensor on the GPU, in place) equals the host-row call."""
    torch = pytest.importorskip("torch")
    m, q, rng = mesh_case(3, 8)
    N, V = q.shape
    gpu = hydro.FiniteVolume(m, 5.0 / 3.0)
    gpu.set_state(q)
    dt = 0.3 * gpu.time_step(0.3)
    qbar = rng.uniform(-1, 1, (N, V))
    host = gpu.step_adjoint(dt, qbar)
    dev = torch.tensor(qbar, device="cuda")
    torch.cuda.synchronize()
    gpu.step_adjoint(dt, dev)
    assert np.array_equal(dev.cpu().numpy(), host)


def test_autograd_rollout_matches_fd_and_oracle():
    """A 12-step static rollout through torch.autograd (checkpoints every 5 steps) over the adjoint kernels: the gradient of
    L = 1/2 sum_i V_i (rho_i - rho*_i)^2 in Q0 against central differences along random directions (1e-6 relative) and
    against the oracle's adjoint chain."""
    torch = pytest.importorskip("torch")
    from luna_hydro import autograd

    m, q, rng = mesh_case(3, 8, amp=0.05)
    N, V = q.shape
    fv = hydro.FiniteVolume(m, 5.0 / 3.0)
    fv.set_state(q)
    dt = 0.3 * fv.time_step(0.3)
    dts = [dt] * 12
    vol = torch.tensor(m["volume"], device="cuda")
    target = torch.tensor(q[:, 0] / m["volume"] * (1 + 0.01 * rng.uniform(-1, 1, N)), device="cuda")
    solver = autograd.Solver(fv)

    def loss(qT):
        return 0.5 * torch.sum(vol * (qT[:, 0] / vol - target) ** 2)

    q0 = torch.tensor(q, device="cuda", requires_grad=True)
    L = loss(autograd.rollout(q0, solver, dts, every=5))
    L.backward()
    grad = q0.grad.detach().cpu().numpy()
    # the oracle's chain: the same steps forward, then its adjoint backward from dL/dqT
    orc = oracle.FiniteVolume(m, 5.0 / 3.0)
    states = []
    orc.set_state(q)
    for d in dts:
        states.append(orc.q.copy())
        orc.time_step(0.3)
        orc.step(d)
    qT = orc.q.copy()
    bar = np.zeros_like(qT)
    V_ = m["volume"]
    bar[:, 0] = (qT[:, 0] / V_ - target.cpu().numpy())  # dL/dq_rho = V (rho - rho*) / V
    for k in reversed(range(len(dts))):
        orc.set_state(states[k])
        bar = orc.step_adjoint(dts[k], bar)
    assert close(grad, bar) < 1e-10
    # central differences of the GPU rollout along random directions
    worst = 0.0
    for s in range(4):
        d = torch.tensor(rng.uniform(-1, 1, (N, V)) * np.abs(q), device="cuda")
        h = 1e-6
        with torch.no_grad():
            Lp = loss(autograd.rollout(torch.tensor(q, device="cuda") + h * d, solver, dts, every=5)).item()
            Lm = loss(autograd.rollout(torch.tensor(q, device="cuda") - h * d, solver, dts, every=5)).item()
        fd = (Lp - Lm) / (2 * h)
        ad = float(np.sum(grad * d.cpu().numpy()))
        worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-30))
    say("autograd rollout: oracle %.2e, worst FD rel %.2e" % (close(grad, bar), worst))
    assert worst < 1e-6


def test_time_step_tangent_is_time_step_for_the_tangent():
    """time_step_tangent(courant, dq) (finite_volume.hpp:1941-1964) gives time_step()'s dt, and the step_tangent after it
    the same J dq; a moving mesh refuses it as the other derivatives."""
    m = cases.voronoi(8, 3)
    q = cases.rough_state(m)
    dq = np.random.default_rng(3).standard_normal(q.shape) * 1e-3
    a = hydro.FiniteVolume(m, 5.0 / 3.0)
    b = hydro.FiniteVolume(m, 5.0 / 3.0)
    c = hydro.FiniteVolume(m, 5.0 / 3.0, motion="moving")
    try:
        a.set_state(q)
        b.set_state(q)
        c.set_state(q)
        dt = a.time_step(0.3)
        assert b.time_step_tangent(0.3, dq) == dt
        assert np.array_equal(b.gradients(), a.gradients())  # the fit's tangent pass wrote the step's gradients
        assert np.array_equal(a.step_tangent(dt, dq.copy()), b.step_tangent(dt, dq.copy()))
        assert np.array_equal(a.state(), b.state())
        # a plain step after time_step_tangent steps as after time_step
        a.set_state(q)
        b.set_state(q)
        a.step(a.time_step(0.3))
        b.time_step_tangent(0.3, dq)
        b.step(dt)
        assert np.array_equal(a.state(), b.state())
        # a state set in between drops the kept rows: step_tangent takes the fit's tangent itself
        a.set_state(q)
        b.time_step_tangent(0.3, 2 * dq)
        b.set_state(q)
        assert a.time_step(0.3) == b.time_step(0.3)
        assert np.array_equal(a.step_tangent(dt, dq.copy()), b.step_tangent(dt, dq.copy()))
        with pytest.raises(RuntimeError):
            c.time_step_tangent(0.3, dq)
    finally:
        a.close()
        b.close()
        c.close()
