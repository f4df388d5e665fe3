"""The GPU's reverse of the moving step (luna_fv_moving_adjoint, order 1, fixed topology) against the oracle's (which central
differences pin, tests/test_oracle_moving_adjoint.py), on the same meshes with the mesh library's geometry derivative: 2D and
3D, HLLC and HLL, the grid steps held or free; and a uniform Q cotangent passes through unchanged (conservation)."""
import numpy as np
import pytest

import oracle
from luna_hydro import hydro, mesh, problems

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0


def _case(D, n, seed, riemann):
    rng = np.random.default_rng(seed)
    m0 = mesh.build(problems.jittered_lattice(n, D, 0.25, seed), 1.0)
    N, V = m0["cells"], D + 2
    rho = 1 + 0.2 * rng.uniform(-1, 1, N)
    P = 1 + 0.2 * rng.uniform(-1, 1, N)
    vel = 0.3 + 0.2 * rng.uniform(-1, 1, (N, D))
    vol = m0["volume"]
    q = np.zeros((N, V))
    q[:, 0] = rho * vol
    q[:, 1:1 + D] = (rho[:, None] * vel) * vol[:, None]
    q[:, V - 1] = (P / (GAMMA - 1) + 0.5 * rho * np.sum(vel ** 2, axis=1)) * vol
    fv = hydro.FiniteVolume(m0, GAMMA, riemann=riemann, motion="moving")
    fv.set_state(q)
    dt = fv.time_step(0.3)
    m1 = mesh.build(fv.moving_begin(dt), 1.0)
    fv.moving_finish(m1, dt)
    return rng, fv, m0, m1, q, dt


def _oracle_adjoint(m0, m1, q, dt, qbar, xbar, fixed, riemann):
    fvo = oracle.FiniteVolume(m0, GAMMA, moving=True, riemann=riemann)
    fvo.set_state(q)
    g0, g1 = mesh.GeometryDerivative(m0), mesh.GeometryDerivative(m1)
    return fvo.moving_adjoint(dt, qbar, xbar, m1, lambda *a: g0.adjoint(*a), lambda *a: g1.adjoint(*a), steps_fixed=fixed)


@pytest.mark.parametrize("D,n,riemann", [(2, 12, "hllc"), (2, 12, "hll"), (3, 6, "hllc"), (3, 8, "hll")])
@pytest.mark.parametrize("fixed", [True, False])
def test_moving_adjoint_matches_oracle(D, n, riemann, fixed):
    print("moving adjoint D=%d n=%d %s steps_fixed=%s" % (D, n, riemann, fixed), flush=True)
    rng, fv, m0, m1, q, dt = _case(D, n, 11, riemann)
    N, V = m0["cells"], D + 2
    qbar, xbar = rng.normal(size=(N, V)), rng.normal(size=(N, D))
    fv.set_state(q)
    qg, xg = fv.moving_adjoint(dt, m0, m1, qbar, xbar, steps_fixed=fixed)
    qo, xo = _oracle_adjoint(m0, m1, q, dt, qbar, xbar, fixed, riemann)
    for got, want, name in ((qg, qo, "qbar"), (xg, xo, "xbar")):
        scale = np.max(np.abs(want))
        err = np.max(np.abs(got - want)) / scale
        print("  %s: max |GPU - oracle| / max |oracle| = %.2e" % (name, err), flush=True)
        assert err < 1e-10, (name, err)
    # the handle is back on t^n with Q as set
    assert np.array_equal(fv.state(), q)
    fv.close()


def test_moving_adjoint_uniform_cotangent():
    rng, fv, m0, m1, q, dt = _case(3, 6, 5, "hllc")
    N, V = m0["cells"], 5
    for v in range(V):
        qbar = np.zeros((N, V))
        qbar[:, v] = 1.0
        fv.set_state(q)
        qb, xb = fv.moving_adjoint(dt, m0, m1, qbar, np.zeros((N, 3)))
        assert np.array_equal(qb, qbar) and not np.any(xb)
    fv.close()


def _oracle_run(q0, x0, dts, held=None, riemann="hllc"):
    """The oracle's moving steps from (Q, x): per step its input Q and mesh, the steps taken; the final (Q, x, mesh)."""
    m = mesh.build(x0, 1.0)
    fvo = oracle.FiniteVolume(m, GAMMA, moving=True, riemann=riemann)
    fvo.set_state(q0)
    qs, meshes, steps = [], [m], []
    for k, dt in enumerate(dts):
        qs.append(fvo.q.copy())
        fvo.time_step(0.3)
        if held is not None:
            fvo.c.hold_steps = 1
            fvo.steps[...] = held[k]
        x1 = fvo.moving_begin(dt)
        steps.append(fvo.steps.copy())
        m1 = mesh.build(x1, 1.0)
        fvo.moving_finish(m1, dt)
        meshes.append(m1)
    return fvo.q.copy(), meshes, qs, steps


def _misfit(q, m, target):
    rho = q[:, 0] / m["volume"]
    return 0.5 * np.sum(m["volume"] * (rho - target) ** 2)


def _rollout_gradient(fv, q0, x0, dts, target, fixed, every):
    torch = pytest.importorskip("torch")
    from luna_hydro import autograd

    solver = autograd.MovingSolver(fv, steps_fixed=fixed)
    qt = torch.tensor(q0, requires_grad=True)
    xt = torch.tensor(x0, requires_grad=True)
    qTt, xTt = autograd.moving_rollout(qt, xt, solver, dts, every)
    autograd.density_misfit(qTt, xTt, torch.tensor(target)).backward()
    return qTt.detach().numpy(), qt.grad.numpy(), xt.grad.numpy()


def _fd(q0, x0, dts, steps, target, riemann, gq, gx, rng, step, trials=3):
    worst = 0.0
    for s in range(trials):
        dq = rng.normal(size=q0.shape) * np.abs(q0) * 1e-7
        dx = rng.integers(-100, 101, size=x0.shape).astype(float) * step * (s > 0)
        L = []
        for sign in (1, -1):
            q, ms, _, _ = _oracle_run(q0 + sign * dq, x0 + sign * dx, dts, held=steps, riemann=riemann)
            L.append(_misfit(q, ms[-1], target))
        fd = (L[0] - L[1]) / 2
        worst = max(worst, abs(fd - (np.sum(gq * dq) + np.sum(gx * dx))) / abs(fd))
    return worst


def test_moving_rollout_acoustic_wave():
    """The moving step's rollout through torch.autograd on a small case of the smooth acoustic wave (rho = 1 + 1e-3
    sin(2 pi (x + y)), p = 3/5, gamma 5/3, HLLC; jittered 6^3 moving mesh, 9 steps, checkpoints every 4), loss the density
    misfit against a seeded perturbed target: the gradient in (Q0, x0) equals the oracle's chain of step adjoints (steps
    free; the rollout's own steps equal the oracle's bit for bit), and checkpointing changes nothing. Central differences
    are reported: on this smooth state a perturbation crosses the limiter's ties (p uniform, v = 0: every clip at its kink)
    and, once the faces move with the flow, HLLC's contact switch at S* = w, as the reference's own multi-step HLLC check
    reports (hydro_derivatives_test.cpp:1641-1643); the gated check is the next test's."""
    rng = np.random.default_rng(9)
    m0 = mesh.build(problems.jittered_lattice(6, 3, 0.25, 9), 1.0)
    q0, x0 = problems.acoustic_state(m0), m0["positions"].copy()
    fv = hydro.FiniteVolume(m0, GAMMA, motion="moving")
    fv.set_state(q0)
    dts = [0.5 * fv.time_step(0.3)] * 9
    qT, meshes, qs, steps = _oracle_run(q0, x0, dts)
    target = qT[:, 0] / meshes[-1]["volume"] * (1 + 1e-3 * rng.uniform(-1, 1, len(qT)))
    qT4, gq, gx = _rollout_gradient(fv, q0, x0, dts, target, False, 4)
    qT1, gq1, gx1 = _rollout_gradient(fv, q0, x0, dts, target, False, 1)
    assert np.array_equal(qT4, qT) and np.array_equal(qT1, qT)
    assert np.array_equal(gq, gq1) and np.array_equal(gx, gx1)
    mT = meshes[-1]
    rho = qT[:, 0] / mT["volume"]
    qbar = np.zeros_like(qT)
    qbar[:, 0] = rho - target
    gT = mesh.GeometryDerivative(mT)
    xbar = gT.adjoint(0.5 * (target ** 2 - rho ** 2), np.zeros((len(qT), 3)), np.zeros((mT["faces"], 3)), np.zeros((mT["faces"], 3)))
    for k in reversed(range(len(dts))):
        qbar, xbar = _oracle_adjoint(meshes[k], meshes[k + 1], qs[k], dts[k], qbar, xbar, False, "hllc")
    err_q = np.max(np.abs(gq - qbar)) / np.max(np.abs(qbar))
    err_x = np.max(np.abs(gx - xbar)) / np.max(np.abs(xbar))
    print("moving rollout (acoustic wave, 9 steps): GPU vs the oracle's chain: Q %.2e, x %.2e" % (err_q, err_x), flush=True)
    assert err_q < 1e-10 and err_x < 1e-10
    _, hq, hx = _rollout_gradient(fv, q0, x0, dts, target, True, 4)
    worst = _fd(q0, x0, dts, steps, target, "hllc", hq, hx, rng, m0["step"], trials=2)
    print("moving rollout (acoustic wave), steps held: |FD - adjoint| / |FD| %.2e (reported)" % worst, flush=True)
    assert worst < 0.1
    fv.close()


@pytest.mark.parametrize("steps_n", [1, 2])
def test_moving_rollout_central_differences(steps_n):
    """The rollout's gradient (HLL, the grid steps held, every step a checkpoint) against central differences of the
    oracle's forward with the same steps held, on a rough random state (no tie within reach of the perturbation), along Q
    directions and whole-grid-step position directions: 1e-6."""
    rng = np.random.default_rng(21)
    m0 = mesh.build(problems.jittered_lattice(6, 3, 0.25, 21), 1.0)
    N = m0["cells"]
    W = np.zeros((N, 5))
    W[:, 0] = 1 + 0.2 * rng.uniform(-1, 1, N)
    W[:, 1:4] = 0.2 * rng.uniform(-1, 1, (N, 3))
    W[:, 4] = 0.6 * (1 + 0.2 * rng.uniform(-1, 1, N))
    q0, x0 = problems.conserved_rows(W, GAMMA) * m0["volume"][:, None], m0["positions"].copy()
    fv = hydro.FiniteVolume(m0, GAMMA, riemann="hll", motion="moving")
    fv.set_state(q0)
    dts = [0.5 * fv.time_step(0.3)] * steps_n
    qT, meshes, qs, steps = _oracle_run(q0, x0, dts, riemann="hll")
    target = qT[:, 0] / meshes[-1]["volume"] * (1 + 1e-3 * rng.uniform(-1, 1, N))
    _, gq, gx = _rollout_gradient(fv, q0, x0, dts, target, True, 1)
    worst = _fd(q0, x0, dts, steps, target, "hll", gq, gx, rng, m0["step"])
    print("moving rollout (HLL, %d steps, steps held): |FD - adjoint| / |FD| %.2e" % (steps_n, worst), flush=True)
    assert worst < 1e-6
    fv.close()


@pytest.mark.parametrize("D,n,riemann", [(2, 12, "hllc"), (3, 6, "hllc"), (3, 6, "hll")])
def test_order0_moving_adjoint_matches_oracle(D, n, riemann):
    """The order-0 moving step's reverse (finite_volume.hpp:2490-2598) against the oracle's, which central differences pin."""
    print("order-0 moving adjoint D=%d n=%d %s" % (D, n, riemann), flush=True)
    rng = np.random.default_rng(5)
    m0 = mesh.build(problems.jittered_lattice(n, D, 0.25, 5), 1.0)
    N, V = m0["cells"], D + 2
    q = np.abs(rng.normal(size=(N, V))) * 0.1
    q[:, 0] = (1 + 0.2 * rng.uniform(-1, 1, N)) * m0["volume"]
    q[:, V - 1] = (2 + 0.2 * rng.uniform(-1, 1, N)) * m0["volume"]
    fv = hydro.FiniteVolume(m0, GAMMA, order=0, riemann=riemann, motion="moving")
    fv.set_state(q)
    dt = fv.time_step(0.3)
    qbar, xbar = rng.normal(size=(N, V)), rng.normal(size=(N, D))
    qg, xg = fv.moving_adjoint(dt, m0, m0, qbar, xbar)
    fvo = oracle.FiniteVolume(m0, GAMMA, order=0, moving=True, riemann=riemann)
    fvo.set_state(q)
    g0 = mesh.GeometryDerivative(m0)
    qo, xo = fvo.moving_adjoint(dt, qbar, xbar, m0, lambda *a: g0.adjoint(*a), lambda *a: None)
    for got, want, name in ((qg, qo, "qbar"), (xg, xo, "xbar")):
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print("  %s: max |GPU - oracle| / max |oracle| = %.2e" % (name, err), flush=True)
        assert err < 1e-10, (name, err)
    fv.close()


def test_acoustic_driver_runs_and_repeats(tmp_path):
    """The acoustic-wave adjoint driver (python -m luna_hydro.acoustic) on a small case: its JSON line, and the same
    gradient bits from a second run (the checkpointed rollout and its reverse are deterministic)."""
    import json

    from luna_hydro import acoustic

    outs = []
    for k in range(2):
        out, npz = tmp_path / ("a%d.json" % k), tmp_path / ("a%d.npz" % k)
        acoustic.main(["--n", "6", "--steps", "6", "--every", "4", "--directions", "0:1", "--out", str(out), "--save", str(npz)])
        outs.append((json.loads(out.read_text()), np.load(npz)))
    (j0, g0), (j1, g1) = outs
    assert j0["cells"] == 216 and len(j0["directions"]) == 1 and j0["loss"] == j1["loss"]
    assert np.array_equal(g0["g_rho"], g1["g_rho"]) and np.array_equal(g0["g_vel"], g1["g_vel"])
    assert np.isfinite(j0["directions"][0]["fd"]) and np.isfinite(j0["directions"][0]["adjoint"])


def test_handles_release_their_device_memory():
    """A moving handle that ran the step, the moving reverse and the moving tangent, then closed, leaves the device's free
    memory where it found it (hipMemGetInfo through torch, after one warm-up round for the runtime's own first-use
    allocations); likewise a static handle after the static derivatives."""
    torch = pytest.importorskip("torch")

    def round_trip():
        m0 = mesh.build(problems.jittered_lattice(10, 3, 0.25, 4), 1.0)
        q = problems.acoustic_state(m0)
        fv = hydro.FiniteVolume(m0, GAMMA, motion="moving")
        fv.set_state(q)
        dt = fv.time_step(0.3)
        m1 = mesh.build(fv.moving_begin(dt), 1.0)
        fv.moving_finish(m1, dt)
        fv.set_geometry(m0)
        fv.set_state(q)
        N = m0["cells"]
        fv.moving_adjoint(dt, m0, m1, np.ones((N, 5)), np.zeros((N, 3)))
        fv.moving_tangent(dt, m0, m1, np.ones((N, 5)), np.zeros((N, 3)))
        fv.close()
        st = hydro.FiniteVolume(m0, GAMMA)
        st.set_state(q)
        dt = st.time_step(0.3)
        st.step_tangent(dt, np.ones((N, 5)))
        st.step_adjoint(dt, np.ones((N, 5)))
        st.close()
        hydro.synchronize()

    round_trip()
    before = torch.cuda.mem_get_info()[0]
    round_trip()
    after = torch.cuda.mem_get_info()[0]
    print("device free memory before %d, after %d" % (before, after), flush=True)
    assert after == before
