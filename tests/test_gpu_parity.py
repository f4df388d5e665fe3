"""Parity of the HIP path (through the C-ABI of libluna_hydro.so) with the CPU oracle on the same seeded inputs.

Bar: every conserved field after every step within 1e-12 relative of the oracle (the same operation order and no
contraction on both sides, so most cases agree to the bit; the test reports the bits), dt and fallback counts equal.
"""
import math

import numpy as np
import pytest

import cases
import oracle
from golden_data import load
from luna_hydro import hydro, mesh, problems

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(msg):
    print(msg, flush=True)


def static_pair(m, q, gamma=5.0 / 3.0, **kw):
    gpu = hydro.FiniteVolume(m, gamma, **kw)
    orc = oracle.FiniteVolume(m, gamma, **{k: v for k, v in kw.items() if k != "motion"})
    gpu.set_state(q)
    orc.set_state(q)
    return gpu, orc


@pytest.mark.parametrize("dim,n,kw", [
    (3, 12, {}),
    (3, 10, {"riemann": "hll"}),
    (3, 10, {"limiter": "none"}),
    (3, 10, {"order": 0}),
    (2, 24, {}),
    (2, 24, {"step_rule": "luna"}),
])
def test_static_steps_match_oracle(dim, n, kw):
    say("static parity dim=%d n=%d %s" % (dim, n, kw))
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    gpu, orc = static_pair(m, q, **kw)
    for s in range(4):
        dt_g, dt_o = gpu.time_step(0.3), orc.time_step(0.3)
        assert dt_g == dt_o, (s, dt_g, dt_o)
        if kw.get("order", 1) == 1:
            assert cases.max_relative(gpu.gradients(), orc.grad) <= TOL or np.array_equal(gpu.gradients(), orc.grad)
        gpu.step(dt_g)
        orc.step(dt_o)
        qg, qo = gpu.state(), orc.q
        assert np.max(np.abs(qg - qo) / np.maximum(np.abs(qo), 1e-300)) <= TOL, s
        assert gpu.fallbacks() == orc.fallbacks()
    say("  bits equal: %s" % np.array_equal(gpu.state(), orc.q))


def test_fallbacks_counted_like_oracle():
    m = cases.voronoi(10, 3)
    q = cases.rough_state(m, noise=0.95, blast=1e4)
    gpu, orc = static_pair(m, q)
    total = 0
    for _ in range(3):
        dt = orc.time_step(0.3)
        assert gpu.time_step(0.3) == dt
        gpu.step(dt)
        orc.step(dt)
        assert gpu.fallbacks() == orc.fallbacks()
        total += orc.fallbacks()
    say("fallbacks over 3 steps: %d" % total)


def moving_pair(m, q, gamma=5.0 / 3.0, **kw):
    gpu = hydro.FiniteVolume(m, gamma, motion="moving", **kw)
    orc = oracle.FiniteVolume(m, gamma, moving=True, **kw)
    gpu.set_state(q)
    orc.set_state(q)
    return gpu, orc


@pytest.mark.parametrize("dim,n", [(3, 10), (2, 24)])
def test_moving_steps_match_oracle(dim, n):
    say("moving parity dim=%d n=%d" % (dim, n))
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    gpu, orc = moving_pair(m, q)
    for s in range(3):
        dt = orc.time_step(0.3)
        assert gpu.time_step(0.3) == dt
        assert np.array_equal(gpu.point_velocities(), orc.w)
        xg = gpu.moving_begin(dt)
        xo = orc.moving_begin(dt)
        assert np.array_equal(xg, xo), s
        m = mesh.build(xg, 1.0)
        gpu.moving_finish(m, dt)
        orc.moving_finish(m, dt)
        qg, qo = gpu.state(), orc.q
        assert np.max(np.abs(qg - qo) / np.maximum(np.abs(qo), 1e-300)) <= TOL, s
        assert gpu.fallbacks() == orc.fallbacks()
    say("  bits equal: %s" % np.array_equal(gpu.state(), orc.q))


@pytest.mark.parametrize("dim,n,kw", [(3, 10, {}), (2, 24, {}), (3, 10, {"riemann": "hll"})])
def test_moving_order0_steps_match_oracle(dim, n, kw):
    """Order 0 on a moving mesh (advance<0, true>() and move() off the grid, finite_volume.hpp:216-225)."""
    say("moving order 0 parity dim=%d n=%d %s" % (dim, n, kw))
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    gpu, orc = moving_pair(m, q, order=0, **kw)
    try:
        for s in range(3):
            dt = orc.time_step(0.3)
            assert gpu.time_step(0.3) == dt
            assert np.array_equal(gpu.point_velocities(), orc.w)
            xg = gpu.moving_begin(dt)
            assert np.array_equal(xg, orc.moving_begin(dt)), s
            m = mesh.build(xg, 1.0)
            gpu.moving_finish(m, dt)
            orc.moving_finish(m, dt)
            assert np.array_equal(gpu.state(), orc.q), s
            assert gpu.fallbacks() == orc.fallbacks()
    finally:
        gpu.close()


def line_pair(name, n, limiter):
    pb = problems.problem_1d(name)
    keys, step, period = problems.line_keys(problems.line_points(n, pb.L), pb.L)
    kw = dict(periodic=pb.periodic, walls=pb.walls, limiter=limiter)
    gpu = hydro.Line(keys, period, step, pb.gamma, **kw)
    orc = oracle.Line(keys, period, step, pb.gamma, **kw)
    q = problems.line_initial_state(pb, keys, step, period)
    gpu.set_state(q)
    orc.set_state(q)
    return pb, gpu, orc


@pytest.mark.parametrize("name", ["sod", "shu-osher", "woodward-colella"])
def test_line_luna_final_states(name):
    gold = load("luna_1d_final")[name]
    pb, gpu, _ = line_pair(name, 16, "face")
    t, steps = 0.0, 0
    while t < pb.t_end * (1 - 1e-12):
        dt = min(gpu.time_step(0.8, "luna"), pb.t_end - t)
        gpu.step(dt)
        t += dt
        steps += 1
    q = gpu.state()
    V = np.array([orc_volume for orc_volume in line_volumes(pb, 16)])
    W = np.array([oracle.primitive(q[s] / V[s], pb.gamma) for s in range(16)])
    ref = np.array(gold["W"])
    assert steps == gold["steps"]
    assert np.max(np.abs(W - ref) / (1e-3 + np.abs(ref))) < 1e-12


def line_volumes(pb, n):
    keys, step, period = problems.line_keys(problems.line_points(n, pb.L), pb.L)
    line = oracle.Line(keys, period, step, pb.gamma, periodic=pb.periodic, walls=pb.walls)
    return [line.volume(s) for s in range(n)]


@pytest.mark.parametrize("limiter", ["minmod", "face", "vanleer", "none"])
def test_line_sod_1024_fused_matches_oracle(limiter):
    """The 1D Sod configuration: N = 1024, CFL 0.4, LUNA's rule fused, to t = 0.2; every step compared."""
    pb, gpu, orc = line_pair("sod", 1024, limiter)
    dt_g, dt_o = gpu.time_step(0.4, "luna"), orc.time_step(0.4, "luna")
    assert dt_g == dt_o
    t, steps = 0.0, 0
    while t < 0.2 * (1 - 1e-12):
        dt = min(dt_o, 0.2 - t)
        dt_g = gpu.step_and_time_step(dt, 0.4)
        dt_o = orc.step_and_time_step(dt, 0.4)
        assert dt_g == dt_o, steps
        t += dt
        steps += 1
    qg, qo = gpu.state(), orc.state
    assert np.max(np.abs(qg - qo) / np.maximum(np.abs(qo), 1e-300)) <= TOL
    say("sod 1024 %s: %d steps, bits equal %s" % (limiter, steps, np.array_equal(qg, qo)))


@pytest.mark.parametrize("case", sorted(load("luna_1d_forward")))
def test_line_luna_forward(case):
    c = load("luna_1d_forward")[case]
    pb = problems.problem_1d(c["problem"])
    keys, step, period = problems.line_keys(problems.line_points(c["n"], pb.L), pb.L)
    lim = {0: "minmod", 1: "face", 2: "vanleer"}[c["limiter"]]
    gpu = hydro.Line(keys, period, step, pb.gamma, periodic=pb.periodic, walls=pb.walls, limiter=lim)
    V = np.array(line_volumes(pb, c["n"]))
    gpu.set_state(np.array(c["U0"]) * V[:, None])
    for _ in range(c["steps"]):
        gpu.step(c["dt"])
    U = gpu.state() / V[:, None]
    assert np.max(np.abs(U - np.array(c["U"]))) / np.max(np.abs(c["U"])) < 1e-12


def test_sedov_128_first_step_matches_oracle():
    """The 3D Sedov configuration at its full size (128^3, jitter 0.25): the first steps against the oracle, and
    conservation."""
    say("building 128^3 mesh")
    m = cases.voronoi(128, 3, 0.25, 1)
    q, deposited = problems.sedov_state(m)
    say("sedov: %d cells deposited, %d faces" % (deposited, m["faces"]))
    gpu, orc = static_pair(m, q)
    before = [math.fsum(q[:, v]) for v in range(5)]
    for s in range(2):
        dt = orc.time_step(0.3)
        assert gpu.time_step(0.3) == dt
        gpu.step(dt)
        orc.step(dt)
        qg = gpu.state()
        assert np.max(np.abs(qg - orc.q) / np.maximum(np.abs(orc.q), 1e-300)) <= TOL
        say("  step %d dt=%.6g bits equal %s" % (s, dt, np.array_equal(qg, orc.q)))
    after = [math.fsum(qg[:, v]) for v in range(5)]
    assert abs(after[0] - before[0]) / before[0] < 1e-14
    assert abs(after[4] - before[4]) / before[4] < 1e-14


def test_set_step_rule_release_mesh_and_mesh_stats():
    """set_step_rule (finite_volume.hpp:181) gives the dt of a handle made with that rule; release_mesh refuses a moving mesh
    (:134-139); mesh_stats counts the meshes a moving run takes after its first."""
    m = cases.voronoi(24, 2)
    q = cases.rough_state(m)
    a = hydro.FiniteVolume(m, 5.0 / 3.0)
    b = hydro.FiniteVolume(m, 5.0 / 3.0, step_rule="luna")
    c = hydro.FiniteVolume(m, 5.0 / 3.0, motion="moving")
    try:
        for fv in (a, b, c):
            fv.set_state(q)
        a.set_step_rule("luna")
        assert a.time_step(0.8) == b.time_step(0.8)
        with pytest.raises(ValueError):
            hydro.FiniteVolume(cases.voronoi(6, 3), 5.0 / 3.0).set_step_rule("luna")
        a.release_mesh()
        with pytest.raises(RuntimeError):
            c.release_mesh()
        assert c.mesh_updates() == 0
        mm = m
        for _ in range(2):
            dt = c.time_step(0.3)
            mm = mesh.build(c.moving_begin(dt), 1.0)
            c.moving_finish(mm, dt)
        assert c.mesh_updates() == 2
    finally:
        a.close()
        b.close()
        c.close()


@pytest.mark.parametrize("moving", [False, True])
def test_sedov_256_first_steps_match_oracle(moving):
    """The 256^3 Sedov configuration (static and moving, the HLL fallback on for the moving one as that configuration
    asks): the first step against the oracle, to the bit, and conservation."""
    say("building 256^3 mesh")
    m = cases.voronoi(256, 3, 0.25, 1)
    q, deposited = problems.sedov_state(m)
    say("sedov 256: %d cells, %d faces, blast in %d cells" % (m["cells"], m["faces"], deposited))
    kappa = 1.0 if moving else 0.0
    gpu = hydro.FiniteVolume(m, 5.0 / 3.0, motion="moving" if moving else "static", hll_compression=kappa)
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, moving=moving, hll_compression=kappa)
    gpu.set_state(q)
    orc.set_state(q)
    try:
        dt = orc.time_step(0.3)
        assert gpu.time_step(0.3) == dt
        if moving:
            xg = gpu.moving_begin(dt)
            assert np.array_equal(xg, orc.moving_begin(dt))
            m1 = mesh.build(xg, 1.0)
            gpu.moving_finish(m1, dt)
            orc.moving_finish(m1, dt)
        else:
            gpu.step(dt)
            orc.step(dt)
        qg = gpu.state()
        assert np.array_equal(qg, orc.q)
        for v in (0, 4):
            assert abs(math.fsum(qg[:, v]) - math.fsum(q[:, v])) <= 1e-14 * math.fsum(q[:, v])
        say("  sedov 256 %s: dt=%.6g bits equal" % ("moving" if moving else "static", dt))
    finally:
        gpu.close()
