"""The GPU's moving line (hydro.MovingLine through the C-ABI) against the oracle's (oracle.MovingLine), bit for bit: each
step's dt, Q, positions, the line's keys and cells, the cells the walls supply and the sorts, on problems that exercise
every path: inflow walls that supply cells (Shu-Osher, Toro 5, LA streams), reflective walls (Noh), outflow (Sod), a
periodic line (the advected sound wave), HLL with minmod, order 0, LUNA's fused rule, and a uniform flow whose points
cross the periodic seam (each crossing re-sorts the line on the device)."""
import numpy as np
import pytest

import oracle
from luna_hydro import hydro, problems

pytestmark = pytest.mark.gpu


def _pair(name, n, **kw):
    pb = problems.problem_1d(name)
    step, period = problems.line_grid(pb.L)
    walls, inflow = problems.moving_walls(pb)
    x = pb.L * (np.arange(n) + 0.5) / n
    args = dict(periodic=pb.periodic, walls=walls, inflow=inflow, spacing=pb.L / n)
    args.update(kw)
    ref = oracle.MovingLine(x, pb.L, period, pb.gamma, **args)
    gpu = hydro.MovingLine(x, pb.L, period, pb.gamma, **args)
    keys, ids = ref.line()
    gk, gi = gpu.line()
    assert np.array_equal(keys, gk) and np.array_equal(ids, gi)
    a, b, vol = problems.line_cells(keys, step, period, pb.periodic)
    q = np.empty((n, 3))
    for s in range(n):
        u = pb.average(a[s] + pb.origin, b[s] + pb.origin)
        q[ids[s]] = [u[k] * vol[s] for k in range(3)]
    ref.set_state(q)
    gpu.set_state(q)
    return pb, ref, gpu


def _same(ref, gpu, what):
    assert ref.n == gpu.n, (what, ref.n, gpu.n)
    assert np.array_equal(ref.state(), gpu.state()), (what, np.max(np.abs(ref.state() - gpu.state())))
    assert np.array_equal(ref.positions(), gpu.positions()), what
    rk, ri = ref.line()
    gk, gi = gpu.line()
    assert np.array_equal(rk, gk) and np.array_equal(ri, gi), what


CASES = [("shu-osher", 256, 400, {}), ("toro-5", 128, 300, {}), ("la-streams", 128, 400, {}), ("noh", 128, 300, {}),
         ("sod", 128, 200, {}), ("sound-advected", 128, 200, {}), ("toro-1", 128, 200, {"riemann": "hll", "limiter": "minmod"}),
         ("shu-osher", 128, 200, {"order": 0}), ("toro-3", 128, 200, {"regularise": False})]


@pytest.mark.parametrize("name,n,steps,kw", CASES, ids=["%s-%d-%s" % (c[0], c[1], "-".join("%s=%s" % i for i in c[3].items()) or "default")
                                                          for c in CASES])
def test_moving_steps_match_oracle(name, n, steps, kw):
    print("moving line %s N=%d %s" % (name, n, kw), flush=True)
    pb, ref, gpu = _pair(name, n, **kw)
    t = 0.0
    for k in range(steps):
        dt = ref.time_step(0.4, "arepo")
        assert dt == gpu.time_step(0.4, "arepo"), (k, dt)
        dt = min(dt, pb.t_end - t)
        if not dt > 0:
            break
        ref.step(dt)
        gpu.step(dt)
        t += dt
        if k % 50 == 0 or k == steps - 1:
            _same(ref, gpu, (name, k))
    _same(ref, gpu, (name, "end"))
    assert ref.cells_added() == gpu.cells_added() and ref.sorts() == gpu.sorts()
    if name in ("shu-osher", "toro-5", "la-streams"):
        assert gpu.cells_added() > 0


@pytest.mark.parametrize("name,n", [("shu-osher", 256), ("sound-advected", 128)])
def test_moving_fused_luna_rule_matches_oracle(name, n):
    print("moving line fused %s N=%d" % (name, n), flush=True)
    pb, ref, gpu = _pair(name, n)
    dt_r = ref.time_step(0.8, "luna")
    dt_g = gpu.time_step(0.8, "luna")
    t = 0.0
    for k in range(300):
        assert dt_r == dt_g, (k, dt_r, dt_g)
        dt = min(dt_r, pb.t_end - t)
        if not dt > 0:
            break
        dt_r = ref.step_and_time_step(dt, 0.8)
        dt_g = gpu.step_and_time_step(dt, 0.8)
        t += dt
    _same(ref, gpu, (name, "fused"))


def test_moving_seam_crossings_sort_on_device():
    """A uniform flow on a periodic line of 256 equal cells (every generator on a grid point): points cross the seam,
    each crossing re-sorts (radix sort on the device); the state stays exactly uniform and equals the oracle's."""
    n, L = 256, 1.0
    step, period = problems.line_grid(L)
    x = L * (np.arange(n) + 0.5) / n
    W = (1.3, 0.37, 0.9)
    q = np.tile([W[0] * L / n, W[0] * W[1] * L / n, (W[2] / 0.4 + 0.5 * W[0] * W[1] ** 2) * L / n], (n, 1))
    ref = oracle.MovingLine(x, L, period, 1.4, periodic=True)
    gpu = hydro.MovingLine(x, L, period, 1.4, periodic=True)
    ref.set_state(q)
    gpu.set_state(q)
    for _ in range(60):
        dt = ref.time_step(0.4, "arepo")
        assert dt == gpu.time_step(0.4, "arepo")
        ref.step(dt)
        gpu.step(dt)
    _same(ref, gpu, "seam")
    assert gpu.sorts() > 0 and gpu.sorts() == ref.sorts()
    assert np.array_equal(gpu.state(), q)


def test_moving_rejects_bad_input_and_static_only_calls():
    with pytest.raises(ValueError):
        hydro.MovingLine([0.1, 0.5, 1.5], 1.0, 1 << 30, 1.4)  # a position outside [0, side]
    with pytest.raises(ValueError):
        hydro.MovingLine([0.1, 0.5, 0.9], 1.0, 1 << 30, 1.4, order=2)
    line = hydro.MovingLine([0.1, 0.5, 0.9], 1.0, 1 << 30, 1.4)
    line.set_state(np.tile([1.0 / 3, 0.0, 2.5 / 3], (3, 1)))
    with pytest.raises(RuntimeError):
        line.step_tangent(1e-3, np.zeros((3, 3)))
    line.close()
