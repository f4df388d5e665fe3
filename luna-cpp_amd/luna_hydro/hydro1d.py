"""The 1D driver, as luna-cpp's hydro_1d runs a line (src/bench/hydro_1d.cpp:126-236, 238-330): n generators on the
problem's interval, the exact cell averages of its initial state, time_step + step (or the fused step_and_time_step
under LUNA's rule) to t_end, repeated, with its report lines (`run: r wall_s=... hash=...` per repeat, FNV-1a over the
bits of Q, and on a moving line then of the positions, from the standard offset basis; then `ours:`), so runs compare
line by line. --moving: the points move (AREPO's rule at 0.4 by default), open walls supply cells every
inflow-spacing x L / N, an inflow wall where the problem's gas enters supersonically.

    python -m luna_hydro.hydro1d sod 1024 [--moving] [--riemann hllc|hll] [--order 0|1] [--dt luna|arepo] [--courant C]
                                 [--limiter face|none|minmod|vanleer] [--jitter J] [--t-end T] [--steps K] [--unfused]
                                 [--repeats R] [--inflow-spacing S] [--no-regularise]

The jitter draws are numpy's MT19937, not the reference's std::mt19937_64 bits.
"""
import argparse
import sys
import time

import numpy as np

from . import hydro, problems
from .sedov import bits_hash

FNV_BASIS = 14695981039346656037  # hydro_1d.cpp: r.hash


def make_line(name, n, riemann="hllc", order=1, limiter="face", jitter=0.0, moving=False, inflow_spacing=1.0, regularise=True,
              engine=None):
    """The line of a run and its initial Q (by cell): hydro.Line / hydro.MovingLine, or `engine` (a constructor with the
    same signature, e.g. the oracle's)."""
    pb = problems.problem_1d(name)
    x = problems.line_points(n, pb.L, jitter)
    if moving:
        step, period = problems.line_grid(pb.L)
        walls, inflow = problems.moving_walls(pb)
        line = (engine or hydro.MovingLine)(x, pb.L, period, pb.gamma, periodic=pb.periodic, walls=walls, inflow=inflow, order=order,
                                            limiter=limiter, riemann=riemann, spacing=inflow_spacing * pb.L / n, regularise=regularise)
        keys, ids = line.line()
        a, b, vol = problems.line_cells(keys, step, period, pb.periodic)
        q0 = np.empty((n, 3))
        for s in range(n):
            u = pb.average(a[s] + pb.origin, b[s] + pb.origin)
            q0[ids[s]] = [u[k] * vol[s] for k in range(3)]
    else:
        keys, step, period = problems.line_keys(x, pb.L)
        q0 = problems.line_initial_state(pb, keys, step, period)
        line = (engine or hydro.Line)(keys, period, step, pb.gamma, periodic=pb.periodic, walls=pb.walls, order=order, limiter=limiter,
                                      riemann=riemann)
    line.set_state(q0)
    return pb, line


def simulate(name, n, riemann="hllc", order=1, rule="luna", courant=0.8, limiter="face", jitter=0.0, t_end=None,
             max_steps=0, fused=True, moving=False, inflow_spacing=1.0, regularise=True):
    """One run; returns (steps, t, wall_s, hash, final Q, cell updates, final cells, sorts)."""
    pb, line = make_line(name, n, riemann, order, limiter, jitter, moving, inflow_spacing, regularise)
    t_end = pb.t_end if t_end is None else t_end
    try:
        fused = fused and rule == "luna"
        t, steps, updates = 0.0, 0, 0
        a = time.perf_counter()
        nxt = line.time_step(courant, rule) if fused else 0.0
        while t < t_end * (1 - 1e-12) and (max_steps == 0 or steps < max_steps):
            dt = nxt if fused else line.time_step(courant, rule)
            if not (dt > 1e-15 * t_end):
                break
            dt = min(dt, t_end - t)
            updates += line.n if moving else n
            if fused:
                nxt = line.step_and_time_step(dt, courant)
            else:
                line.step(dt)
            t += dt
            steps += 1
        hydro.synchronize()
        wall = time.perf_counter() - a
        q = line.state()
        h = bits_hash(q, FNV_BASIS)
        cells, sorts = n, 0
        if moving:
            h = bits_hash(line.positions(), h)  # hydro_1d.cpp:215-218: then the positions
            cells, sorts = line.n, line.sorts()
    finally:
        line.close()
    return steps, t, wall, h, q, updates, cells, sorts


def run(name, n, riemann="hllc", order=1, rule=None, courant=None, limiter="face", jitter=0.0, t_end=None, max_steps=0,
        fused=True, repeats=2, moving=False, inflow_spacing=1.0, regularise=True, log=print):
    rule = rule or ("arepo" if moving else "luna")  # hydro_1d.cpp:253
    courant = courant if courant is not None else (0.4 if rule == "arepo" else 0.8)  # finite_volume_1d.hpp:20
    best, same, first = None, True, None
    for r in range(repeats):
        res = simulate(name, n, riemann, order, rule, courant, limiter, jitter, t_end, max_steps, fused, moving, inflow_spacing,
                       regularise)
        steps, t, wall, h, q, updates = res[:6]
        log("run: %d wall_s=%g M_cell_updates_per_s=%g hash=%d" % (r, wall, updates / wall / 1e6, h))
        same = same and (first is None or h == first)
        first = h if first is None else first
        if best is None or wall < best[2]:
            best = res
    steps, t, wall, h, q, updates, cells, sorts = best
    pb = problems.problem_1d(name)
    reached = int(t >= (pb.t_end if t_end is None else t_end) * (1 - 1e-12))
    log("ours: problem=%s N=%d mesh=%s riemann=%s order=%d dt=%s courant=%g jitter=%g steps=%d t=%g reached=%d sorts=%d "
        "wall_s=%g final_cells=%d cell_updates=%g M_cell_updates_per_s=%g hash=%d same_bits=%d backend=HIP"
        % (name, n, "moving" if moving else "static", riemann, order, rule, courant, jitter, steps, t, reached, sorts, wall, cells,
           float(updates), updates / wall / 1e6, h, int(same)))
    return {"steps": steps, "t": t, "wall_s": wall, "hash": h, "same_bits": same, "q": q, "cells": cells, "sorts": sorts}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("problem")
    ap.add_argument("n", type=int)
    ap.add_argument("--riemann", choices=["hllc", "hll"], default="hllc")
    ap.add_argument("--order", type=int, default=1)
    ap.add_argument("--moving", action="store_true")
    ap.add_argument("--dt", choices=["luna", "arepo"], default=None)
    ap.add_argument("--courant", type=float, default=None)
    ap.add_argument("--limiter", choices=["face", "none", "minmod", "vanleer"], default="face")
    ap.add_argument("--jitter", type=float, default=0.0)
    ap.add_argument("--t-end", type=float, default=None)
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--unfused", action="store_true")
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--inflow-spacing", type=float, default=1.0)
    ap.add_argument("--no-regularise", action="store_true")
    a = ap.parse_args(argv)
    run(a.problem, a.n, a.riemann, a.order, a.dt, a.courant, a.limiter, a.jitter, a.t_end, a.steps, not a.unfused, a.repeats, a.moving,
        a.inflow_spacing, not a.no_regularise)
    return 0


if __name__ == "__main__":
    sys.exit(main())
