"""Pins the oracle's 1D step (finite_volume_1d.hpp) on LUNA's golden vectors: final states at N = 16 with the MC (face)
limiter, exact step counts (hydro1d_test.cpp:266-295, luna_1d_values.hpp), and the forward states after 5 fixed steps
with minmod, MC and van Leer (hydro_derivatives_test.cpp:991-1059, luna_derivative_values.hpp); plus Sod against the
exact Riemann solution within LUNA's L1 bounds (hydro1d_test.cpp:232-264)."""
import numpy as np
import pytest

import exact_riemann
import oracle
from golden_data import load
from luna_hydro import problems

LIMITER = {0: "minmod", 1: "face", 2: "vanleer"}


def make_line(name, n, limiter="face", jitter=0.0, seed=1):
    pb = problems.problem_1d(name)
    x = problems.line_points(n, pb.L, jitter, seed)
    keys, step, period = problems.line_keys(x, pb.L)
    line = oracle.Line(keys, period, step, pb.gamma, periodic=pb.periodic, walls=pb.walls, limiter=limiter)
    line.set_state(problems.line_initial_state(pb, keys, step, period))
    return pb, line, keys, step


def run(line, t_end, courant, rule="luna"):
    t, steps = 0.0, 0
    while t < t_end * (1 - 1e-12):
        dt = min(line.time_step(courant, rule), t_end - t)
        line.step(dt)
        t += dt
        steps += 1
    return steps


def cell_primitives(line, gamma):
    q = line.state
    out = []
    for s in range(line.n):
        V = line.volume(s)
        out.append(oracle.primitive(q[s] / V, gamma))
    return np.array(out)


@pytest.mark.parametrize("name", ["sod", "shu-osher", "woodward-colella"])
def test_luna_final_states(name):
    gold = load("luna_1d_final")[name]
    pb, line, _, _ = make_line(name, 16)
    steps = run(line, pb.t_end, 0.8)
    W = cell_primitives(line, pb.gamma)
    ref = np.array(gold["W"])
    worst = np.max(np.abs(W - ref) / (1e-3 + np.abs(ref)))
    assert steps == gold["steps"] and worst < 1e-12, (steps, worst)


@pytest.mark.parametrize("case", sorted(load("luna_1d_forward")))
def test_luna_forward_five_steps(case):
    c = load("luna_1d_forward")[case]
    pb = problems.problem_1d(c["problem"])
    x = problems.line_points(c["n"], pb.L)
    keys, step, period = problems.line_keys(x, pb.L)
    line = oracle.Line(keys, period, step, pb.gamma, periodic=pb.periodic, walls=pb.walls, limiter=LIMITER[c["limiter"]])
    V = np.array([line.volume(s) for s in range(c["n"])])
    line.set_state(np.array(c["U0"]) * V[:, None])
    for _ in range(c["steps"]):
        line.step(c["dt"])
    U = line.state / V[:, None]
    ref = np.array(c["U"])
    assert np.max(np.abs(U - ref)) / np.max(np.abs(ref)) < 1e-12


def test_planted_limiter_swap_fails_golden():
    """Teeth: running the minmod case with van Leer must miss LUNA's minmod forward state."""
    c = load("luna_1d_forward")["luna_d_sod_64_l0"]
    pb = problems.problem_1d("sod")
    keys, step, period = problems.line_keys(problems.line_points(64, 1.0), 1.0)
    line = oracle.Line(keys, period, step, pb.gamma, walls=pb.walls, limiter="vanleer")
    V = np.array([line.volume(s) for s in range(64)])
    line.set_state(np.array(c["U0"]) * V[:, None])
    for _ in range(5):
        line.step(c["dt"])
    assert np.max(np.abs(line.state / V[:, None] - np.array(c["U"]))) / np.max(np.abs(c["U"])) > 1e-6


@pytest.mark.parametrize("n", [128, 256])
def test_sod_exact_l1_within_luna_bounds(n):
    bounds = load("sod_l1_bounds")[str(n)]
    pb, line, keys, step = make_line("sod", n)
    run(line, 0.25, 0.8)
    W = cell_primitives(line, pb.gamma)
    solver = exact_riemann.Solver((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    l1 = np.zeros(3)
    for s in range(n):
        V = line.volume(s)
        xc = keys[s] * step + line.centroid(s)
        e = solver.average(xc - 0.5 * V, xc + 0.5 * V, 0.5, 0.25)
        l1 += V * np.abs(W[s] - np.array(e))
    assert np.all(l1 <= np.array(bounds)), l1


def test_fused_equals_unfused_static():
    pb, a, _, _ = make_line("sod", 256, limiter="minmod")
    _, b, _, _ = make_line("sod", 256, limiter="minmod")
    dt_a = a.time_step(0.4, "luna")
    dt_b = dt_a
    for _ in range(20):
        a.step(dt_a)
        dt_a = a.time_step(0.4, "luna")
        dt_b = b.step_and_time_step(dt_b, 0.4)
        assert dt_a == dt_b
    assert np.array_equal(a.state, b.state)


def test_conservation_and_uniform_periodic():
    pb = problems.problem_1d("sound-right")
    x = problems.line_points(200, 1.0, 0.8, 1)
    keys, step, period = problems.line_keys(x, 1.0)
    line = oracle.Line(keys, period, step, 1.4, periodic=True)
    W = np.array([1.3, 0.7, 0.4])
    U = oracle.conserved(W, 1.4)
    V = np.array([line.volume(s) for s in range(200)])
    line.set_state(U[None, :] * V[:, None])
    for _ in range(20):
        line.step(line.time_step(0.3, "arepo"))
    Wn = cell_primitives(line, 1.4)
    assert np.max(np.abs(Wn - W) / np.abs(W)) < 1e-13


def test_exact_riemann_star_state_toro_table():
    """Toro (2009) Table 4.3, test 1 (Sod): p* = 0.30313, u* = 0.92745."""
    s = exact_riemann.Solver((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    assert abs(s.p - 0.30313) < 5e-6 and abs(s.u - 0.92745) < 5e-6
