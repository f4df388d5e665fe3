"""The 3D acoustic-wave adjoint run: the gradient of a density misfit after many moving steps with respect to the initial
density and velocity, through the checkpointed rollout of the moving step's reverse (luna_fv_moving_adjoint per step, the
mesh library's geometry derivative per mesh), checked against central differences of the forward along seeded random
directions. The reference's counterpart is its multi-step moving derivative check (src/test/hydro_derivatives_test.cpp,
`moving2d`/`moving3d` cases: forward with checkpoints, the reverse backwards along it, central differences reported).

    python -m luna_hydro.acoustic --n 32 [--steps 200] [--every 20] [--directions 0:32] [--h 1e-6]
                                  [--riemann hllc|hll] [--steps-fixed] [--out FILE.json] [--save FILE.npz]

Setup: a jittered periodic lattice of n^3 generators (jitter 0.25, seed 1); rho = 1 + 1e-3 sin(2 pi k.x), k = (1, 1, 0),
v = 0, p = 3/5, gamma 5/3 (sound speed 1); every step takes the dt of the AREPO rule at CFL 0.3 on the initial state
(fixed); the target density is the forward result of the same run from a seeded perturbed density (1e-4 seeded normal
noise); L = 1/2 sum_i V_i (rho_i - rho*_i)^2 on the final mesh. The gradient is taken in (rho_0, v_0) at fixed p_0 and
fixed initial positions. Direction j (seeded by j) perturbs rho_0 and v_0 by h times standard normal values; its central
difference (L(+) - L(-)) / 2h is compared with the gradient's dot product. Prints one JSON line.

What bounds the size: each step's mesh comes from the host Voronoi producer, and each adjoint step needs the geometry
derivative (host) of two meshes; the report states both times per step.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import hydro, mesh, problems

GAMMA = 5.0 / 3.0
PRESSURE = 3.0 / 5.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def conserved(rho, vel, volume):
    """Q rows of (rho, v, p = 3/5) on cells of the given volumes (numpy or torch, whichever rho is)."""
    ke = 0.5 * rho * (vel ** 2).sum(axis=1) if isinstance(rho, np.ndarray) else 0.5 * rho * (vel ** 2).sum(dim=1)
    energy = PRESSURE / (GAMMA - 1) + ke
    cols = [rho] + [rho * vel[:, d] for d in range(vel.shape[1])] + [energy]
    if isinstance(rho, np.ndarray):
        return np.stack(cols, axis=1) * volume[:, None]
    import torch

    return torch.stack(cols, dim=1) * volume[:, None]


class Run:
    def __init__(self, n, riemann, seed=1):
        self.m0 = mesh.build(problems.jittered_lattice(n, 3, 0.25, seed), 1.0)
        self.N = self.m0["cells"]
        q0 = problems.acoustic_state(self.m0, GAMMA)
        self.rho0 = q0[:, 0] / self.m0["volume"]
        self.vel0 = np.zeros((self.N, 3))
        self.riemann = riemann
        self.fv = hydro.FiniteVolume(self.m0, GAMMA, motion="moving", riemann=riemann)
        self.fv.set_state(q0)
        self.dt = self.fv.time_step(0.3)
        self.producer_s = 0.0
        self.hydro_s = 0.0

    def forward(self, rho, vel, steps):
        """Q and the mesh after `steps` moving steps of dt from (rho, v) on the initial mesh."""
        fv = self.fv
        fv.set_geometry(self.m0)
        fv.set_state(conserved(rho, vel, self.m0["volume"]))
        m = self.m0
        for _ in range(steps):
            a = time.perf_counter()
            x1 = fv.moving_begin(self.dt)
            b = time.perf_counter()
            m = mesh.build(x1, 1.0)
            c = time.perf_counter()
            fv.moving_finish(m, self.dt)
            fv.fallbacks()
            d = time.perf_counter()
            self.hydro_s += (b - a) + (d - c)
            self.producer_s += c - b
        return fv.state(), m

    @staticmethod
    def misfit(q, m, target):
        rho = q[:, 0] / m["volume"]
        return 0.5 * math.fsum(m["volume"] * (rho - target) ** 2)

    def close(self):
        self.fv.close()


def direction(j, N):
    rng = np.random.Generator(np.random.MT19937(1000 + j))
    return rng.normal(size=N), rng.normal(size=(N, 3))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--directions", default="0:32", help="first:last (seeded directions first .. last-1)")
    ap.add_argument("--h", type=float, default=1e-6)
    ap.add_argument("--riemann", choices=["hllc", "hll"], default="hllc")
    ap.add_argument("--steps-fixed", action="store_true", help="hold the whole grid steps in the reverse")
    ap.add_argument("--out", default="")
    ap.add_argument("--save", default="", help="write the gradient to this .npz")
    args = ap.parse_args(argv)

    import torch

    from . import autograd

    first, last = (int(s) for s in args.directions.split(":"))
    t_start = time.perf_counter()
    run = Run(args.n, args.riemann)
    N = run.N
    log("acoustic wave %d^3 (%d cells, %d faces), %s, dt %.6g, %d steps, checkpoints every %d"
        % (args.n, N, run.m0["faces"], args.riemann, run.dt, args.steps, args.every))

    noise = np.random.Generator(np.random.MT19937(7)).normal(size=N)
    qT, mT = run.forward(run.rho0 * (1 + 1e-4 * noise), run.vel0, args.steps)
    target = qT[:, 0] / mT["volume"]
    t_target = time.perf_counter()
    log("target: %.1f s" % (t_target - t_start))

    # The gradient: Q0(rho0, v0) -> checkpointed moving rollout -> misfit, reverse through torch.autograd.
    run.fv.set_geometry(run.m0)
    solver = autograd.MovingSolver(run.fv, steps_fixed=args.steps_fixed)
    rho_t = torch.tensor(run.rho0, requires_grad=True)
    vel_t = torch.tensor(run.vel0, requires_grad=True)
    vol = torch.tensor(run.m0["volume"])
    q0 = conserved(rho_t, vel_t, vol)
    x0 = torch.tensor(run.m0["positions"])
    f0 = time.perf_counter()
    qT_t, xT_t = autograd.moving_rollout(q0, x0, solver, [run.dt] * args.steps, args.every)
    f1 = time.perf_counter()
    loss = autograd.density_misfit(qT_t, xT_t, torch.tensor(target))
    loss.backward()
    f2 = time.perf_counter()
    g_rho, g_vel = rho_t.grad.numpy().copy(), vel_t.grad.numpy().copy()
    L0 = float(loss.detach())
    log("gradient: forward %.1f s, reverse %.1f s; L = %.6e, |dL/drho| %.3e, |dL/dv| %.3e"
        % (f1 - f0, f2 - f1, L0, np.linalg.norm(g_rho), np.linalg.norm(g_vel)))
    if args.save:
        np.savez(args.save, g_rho=g_rho, g_vel=g_vel, L=L0)

    # Central differences along seeded directions.
    rows = []
    run.producer_s = run.hydro_s = 0.0
    d0 = time.perf_counter()
    for j in range(first, last):
        u_rho, u_vel = direction(j, N)
        L = []
        for sign in (1, -1):
            q, m = run.forward(run.rho0 + sign * args.h * u_rho, run.vel0 + sign * args.h * u_vel, args.steps)
            L.append(run.misfit(q, m, target))
        fd = (L[0] - L[1]) / (2 * args.h)
        ad = float(np.dot(g_rho, u_rho) + np.sum(g_vel * u_vel))
        rel = abs(fd - ad) / max(abs(fd), 1e-300)
        rows.append({"direction": j, "fd": fd, "adjoint": ad, "rel": rel})
        log("  direction %d: central %.9e, adjoint %.9e, |diff| / |central| %.2e" % (j, fd, ad, rel))
    d1 = time.perf_counter()
    runs = 2 * (last - first) * args.steps
    rels = [r["rel"] for r in rows]
    out = {
        "workload": "3D acoustic wave, moving mesh, %d^3 jittered (0.25, seed 1), %d steps, %s, checkpoints every %d"
                    % (args.n, args.steps, args.riemann, args.every),
        "cells": N, "faces": run.m0["faces"], "dt": run.dt, "steps_fixed": args.steps_fixed, "h": args.h, "loss": L0,
        "gradient_norm": {"rho": float(np.linalg.norm(g_rho)), "v": float(np.linalg.norm(g_vel))},
        "directions": rows,
        "rel_max": max(rels) if rels else None, "rel_median": float(np.median(rels)) if rels else None,
        "seconds": {"target": t_target - t_start, "rollout_forward": f1 - f0, "reverse": f2 - f1,
                    "central_differences": d1 - d0,
                    "per_forward_step_producer": run.producer_s / max(runs, 1),
                    "per_forward_step_hydro": run.hydro_s / max(runs, 1)},
    }
    run.close()
    line = json.dumps(out)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
