"""The 1D ladder gates of luna-cpp (src/test/luna_1d_gates.hpp, kept as tests/golden/luna_1d_gates.json) as a check of
the static line of libluna_hydro: each problem on N uniform cells to its t_end under LUNA's rule at CFL 0.8 (the reference's
hydro1d_cases_test.cpp:46-110 simulate()), then the L1 errors of rho, v, P with the rules of LUNA's tests
(:164-221: exact cell averages of the Riemann solution, Noh's exact profile away from the wall, the linear waves from
their +eps and -eps runs, the simple wave from its characteristics, three problems against stored fine-run densities).
A row passes when each L1 equals LUNA's to the row's tolerance (or LUNA's is below the row's floor). The moving gates
(:314-327) compare with the reference's own moving runs (AREPO's rule at 0.4, cells supplied at open walls): the run
reaches t_end and each L1 is at most 1.3 times theirs.

The runs go through the oracle (CPU) or the GPU: `engine` is a constructor with the oracle.Line signature (static) or
the oracle.MovingLine signature (moving).
"""
import math

import numpy as np

import exact_riemann
import oracle
from golden_data import load
from luna_hydro import problems

GATES = load("luna_1d_gates")
REFERENCES = {r["problem"]: r for r in GATES["references"]}
LIMITERS = {"luna": "face", "minmod": "minmod", "vanleer": "vanleer"}
STEPS_PER_CELL = 200  # a run this long has stalled (the reference's rule)


def simulate(engine, name, n, limiter="face", amplitude=0.0):
    """cells [(a, b, W)] sorted by a, t reached, steps."""
    pb = problems.problem_1d(name, amplitude)
    keys, step, period = problems.line_keys(problems.line_points(n, pb.L), pb.L)
    line = engine(keys, period, step, pb.gamma, periodic=pb.periodic, walls=pb.walls, order=1, limiter=limiter)
    line.set_state(problems.line_initial_state(pb, keys, step, period))
    t, steps = 0.0, 0
    while t < pb.t_end * (1 - 1e-12) and steps < STEPS_PER_CELL * n:
        dt = min(line.time_step(0.8, "luna"), pb.t_end - t)
        if not (dt > 1e-15 * pb.t_end):
            break
        line.step(dt)
        t += dt
        steps += 1
    q = np.asarray(line.state() if callable(getattr(line, "state")) else line.state)
    a, b, vol = problems.line_cells(keys, step, period, pb.periodic)
    cells = []
    for s in range(n):
        W = oracle.primitive(q[s] / vol[s], pb.gamma)
        cells.append((a[s] + pb.origin, b[s] + pb.origin, W))
    cells.sort(key=lambda c: c[0])
    if hasattr(line, "close"):
        line.close()
    return cells, t, steps


def simulate_moving(engine, name, n, limiter="face", amplitude=0.0):
    """A moving run as the reference's simulate() with Motion::moving (hydro1d_cases_test.cpp:48-110): generators at
    L (i + 1/2) / n, the problem's moving walls (an inflow wall where gas enters supersonically), cells supplied every
    L / n, AREPO's rule at 0.4. `engine` has the oracle.MovingLine signature. cells [(a, b, W)] sorted by a, t, steps."""
    pb = problems.problem_1d(name, amplitude)
    step, period = problems.line_grid(pb.L)
    x = pb.L * (np.arange(n) + 0.5) / n
    walls, inflow = problems.moving_walls(pb)
    line = engine(x, pb.L, period, pb.gamma, periodic=pb.periodic, walls=walls, inflow=inflow, order=1, limiter=limiter,
                  spacing=pb.L / n)
    keys, ids = line.line()
    a, b, vol = problems.line_cells(keys, step, period, pb.periodic)
    q = np.empty((n, 3))
    for s in range(n):
        u = pb.average(a[s] + pb.origin, b[s] + pb.origin)
        q[ids[s]] = [u[k] * vol[s] for k in range(3)]
    line.set_state(q)
    t, steps = 0.0, 0
    while t < pb.t_end * (1 - 1e-12) and steps < STEPS_PER_CELL * n:
        dt = min(line.time_step(0.4, "arepo"), pb.t_end - t)
        if not (dt > 1e-15 * pb.t_end):
            break
        line.step(dt)
        t += dt
        steps += 1
    q = line.state()
    keys, ids = line.line()
    a, b, vol = problems.line_cells(keys, step, period, pb.periodic)
    cells = [(a[s] + pb.origin, b[s] + pb.origin, oracle.primitive(q[ids[s]] / vol[s], pb.gamma)) for s in range(len(keys))]
    cells.sort(key=lambda c: c[0])
    if hasattr(line, "close"):
        line.close()
    return cells, t, steps


def primitive_of(U, gamma):
    u = U[1] / U[0] if U[0] > 0 else 0.0
    return (U[0], u, (gamma - 1) * (U[2] - 0.5 * U[0] * u * u))


def simple_at(x, t, amp):
    """The isentropic simple wave at (x, t), along its characteristics (hydro1d_cases_test.cpp:120-131)."""
    two_pi = 2 * math.pi
    x0 = x - t - math.floor(x - t)
    for _ in range(60):
        c = 1 + amp * math.sin(two_pi * x0)
        f = x0 + (6 * c - 5) * t - x
        f -= round(f)
        x0 -= f / (1 + 6 * t * amp * two_pi * math.cos(two_pi * x0))
    c = 1 + amp * math.sin(two_pi * x0)
    rho = c ** 5
    return (rho, 5 * (c - 1), rho ** 1.4 / 1.4)


class Exact:
    """The exact primitives of a cell [a, b] at t_end (a conservative average)."""

    def __init__(self, name, eps):
        self.name, self.pb = name, problems.problem_1d(name, eps)
        self.solver = None
        if self.pb.left is not None and self.pb.right is not None:
            self.solver = exact_riemann.Solver(self.pb.left, self.pb.right, self.pb.gamma)

    def __call__(self, a, b):
        name, pb = self.name, self.pb
        if name == "noh":
            g = 5.0 / 3.0
            D, rho_s = (g - 1) / 2, (g + 1) / (g - 1)
            p_s = (g - 1) / 2 * rho_s
            f = min(max((D * pb.t_end - a) / (b - a), 0.0), 1.0)
            return (f * rho_s + (1 - f), -(1 - f), f * p_s + (1 - f) * 1e-6)
        if name in ("sound-right", "sound-left", "entropy"):
            return primitive_of(pb.average(a, b), pb.gamma)
        if name == "sound-advected":
            x, t, k, c = 0.5 * (a + b), pb.t_end, 2 * math.pi, math.sqrt(pb.gamma)
            amp = pb.average(0.25, 0.25)[0] - 1
            s0, sp, sm = amp * math.sin(k * (x - t)), amp * math.sin(k * (x - (1 + c) * t)), amp * math.sin(k * (x - (1 - c) * t))
            ap, am = 0.5 * (1 + c) / (c * c), 0.5 * (1 - c) / (c * c)
            return (1 + (1 - 1 / (c * c)) * s0 + ap * sp + am * sm, 1 + c * (ap * sp - am * sm), 1 + c * c * (ap * sp + am * sm))
        if name == "simple-wave":
            U = problems.gauss_average(a, b, lambda x: problems._conserved1(simple_at(x, pb.t_end, 0.1), 1.4))
            return primitive_of(U, 1.4)
        return self.solver.average(a, b, pb.x0, pb.t_end)


def is_linear(name):
    return name in ("sound-right", "sound-left", "entropy", "sound-advected")


def l1_of(name, cells, minus=None, eps=problems.LINEAR_AMPLITUDE):
    """Sum over cells of V |W - exact| per variable (hydro1d_cases_test.cpp:164-188)."""
    plus = Exact(name, eps if is_linear(name) else 0.0)
    negative = Exact(name, -eps) if minus is not None else None
    total = [0.0, 0.0, 0.0]
    for i, (a, b, W) in enumerate(cells):
        V = b - a
        if name == "noh" and (i < 4 or 0.5 * (a + b) >= 0.3):
            continue
        e = plus(a, b)
        err = [W[v] - e[v] for v in range(3)]
        if minus is not None:
            ma, mb, mW = minus[i]
            em = negative(ma, mb)
            err = [(err[v] - (mW[v] - em[v])) / (2 * eps) for v in range(3)]
        for v in range(3):
            if v == 1 and name == "la-rcvcr" and not (e[0] > 0):
                continue
            if v == 1 and name in ("la-near-vacuum", "la-vacuum") and not (e[0] > 1e-3):
                continue
            total[v] += V * abs(err[v])
    return total


def reference_l1(cells, ref):
    """rho binned on the stored fine run's bins, L1 against it (hydro1d_cases_test.cpp:190-201)."""
    bins, rho = ref["bins"], ref["rho"]
    avg = [0.0] * bins
    lo, hi = cells[0][0], cells[-1][1]
    h = (hi - lo) / bins
    for a, b, W in cells:
        k = min(bins - 1, int((0.5 * (a + b) - lo) / h))
        avg[k] += (b - a) * W[0] / h
    return [sum(h * abs(avg[k] - rho[k]) for k in range(bins)), float("nan"), float("nan")]


def measure(engine, name, n, limiter="face", moving=False):
    """(L1 per variable, t reached) of a problem at N (hydro1d_cases_test.cpp:210-228), static or moving."""
    run = simulate_moving if moving else simulate
    if is_linear(name):
        plus, t, _ = run(engine, name, n, limiter, problems.LINEAR_AMPLITUDE)
        minus, _, _ = run(engine, name, n, limiter, -problems.LINEAR_AMPLITUDE)
        return l1_of(name, plus, minus), t
    cells, t, _ = run(engine, name, n, limiter)
    if name in REFERENCES:
        return reference_l1(cells, REFERENCES[name]), t
    return l1_of(name, cells), t


def check_row(engine, row, series):
    """[(variable, L1, LUNA's, |ratio - 1|, ok)] of one gate row and series ("luna", "minmod", "vanleer")."""
    got, _ = measure(engine, row["problem"], row["n"], LIMITERS[series])
    out = []
    for v in range(3):
        luna = row[series][v]
        if luna is None:
            continue
        ratio = abs(got[v] / luna - 1)
        out.append((v, got[v], luna, ratio, ratio < row["tolerance"] or luna < row["floor"]))
    return out


def static_cases(max_n=None):
    """(row, series) of every static gate (the problems LUNA ran), optionally up to N = max_n."""
    cases = []
    for row in GATES["rows"]:
        if not row["run_luna"] or (max_n and row["n"] > max_n):
            continue
        for series in ("luna", "minmod", "vanleer"):
            if row[series][0] is not None:
                cases.append((row, series))
    return cases


def check_moving_row(engine, row):
    """[(variable, L1, the reference's moving L1, ok)] of one moving gate (hydro1d_cases_test.cpp:314-327): the run reaches
    t_end and each L1 is at most 1.3 times the reference's own moving run's; plus whether t_end was reached."""
    pb = problems.problem_1d(row["problem"])
    got, t = measure(engine, row["problem"], row["n"], "face", moving=True)
    reached = t >= pb.t_end * (1 - 1e-12)
    out = []
    for v in range(3):
        ref = row["moving"][v]
        if ref is None:
            continue
        out.append((v, got[v], ref, got[v] <= 1.3 * ref))
    return out, reached


def moving_cases(max_n=None):
    """The moving gates (rows with a measured moving run and no open issue), optionally up to N = max_n."""
    return [row for row in GATES["rows"] if row["issue"] == 0 and row["moving"][0] is not None and not (max_n and row["n"] > max_n)]
