"""The 1D Sod tube as its configuration states it (1024 equal cells on [0, 1], gamma 1.4, (rho, v, p) = (1, 0, 1) |
(0.125, 0, 0.1) split at 0.5, outflow walls, HLLC, minmod, CFL 0.4 with LUNA's rule and the fused step, t = 0.2) on the
GPU: the final state equals the oracle's to the bit, and the L1 errors of rho, v and p against the exact Riemann
solution's cell averages (Toro ch. 4, tests' own solver) are printed, with the microseconds per step, and stay under
the L1 that LUNA reaches on the same tube at N = 256 with its own limiter at t = 0.25 (the gates' toro-1 row)."""
import time

import numpy as np
import pytest

import exact_riemann
import gates_1d
import oracle
from luna_hydro import hydro, hydro1d, problems

pytestmark = pytest.mark.gpu
N, COURANT, T_END = 1024, 0.4, 0.2


def cells_of(pb, q):
    keys, step, period = problems.line_keys(problems.line_points(N, pb.L), pb.L)
    a, b, vol = problems.line_cells(keys, step, period, pb.periodic)
    return sorted([(a[s] + pb.origin, b[s] + pb.origin, oracle.primitive(q[s] / vol[s], pb.gamma)) for s in range(N)],
                  key=lambda c: c[0])


def l1(pb, cells):
    solver = exact_riemann.Solver(pb.left, pb.right, pb.gamma)
    total = [0.0, 0.0, 0.0]
    for a, b, W in cells:
        e = solver.average(a, b, pb.x0, T_END)
        for v in range(3):
            total[v] += (b - a) * abs(W[v] - e[v])
    return total


def test_sod_config_matches_oracle_and_exact_solution():
    print("1D Sod, N = 1024, minmod, CFL 0.4, LUNA rule fused, t = 0.2", flush=True)
    pb, line = hydro1d.make_line("sod", N, limiter="minmod")
    t, steps = 0.0, 0
    a = time.perf_counter()
    nxt = line.time_step(COURANT, "luna")
    while t < T_END * (1 - 1e-12):
        dt = min(nxt, T_END - t)
        nxt = line.step_and_time_step(dt, COURANT)
        t += dt
        steps += 1
    hydro.synchronize()
    wall = time.perf_counter() - a
    q = np.array(line.state())
    line.close()
    _, ref = hydro1d.make_line("sod", N, limiter="minmod", engine=oracle.Line)
    t2, steps2 = 0.0, 0
    while t2 < T_END * (1 - 1e-12):
        dt = min(ref.time_step(COURANT, "luna"), T_END - t2)
        ref.step(dt)
        t2 += dt
        steps2 += 1
    qo = np.asarray(ref.state() if callable(getattr(ref, "state")) else ref.state)
    assert steps == steps2 and np.array_equal(q, qo)
    err = l1(pb, cells_of(pb, q))
    print("  %d steps, %.1f us/step (wall, launches included); L1 rho %.4e, v %.4e, p %.4e" % (steps, 1e6 * wall / steps, *err),
          flush=True)
    luna_256 = next(r for r in gates_1d.GATES["rows"] if r["problem"] == "toro-1" and r["n"] == 256)["luna"]
    assert all(err[v] < luna_256[v] for v in range(3)), (err, luna_256)
