"""The 2D Kelvin-Helmholtz driver, as luna-cpp's hydro_kh2d runs it (src/bench/hydro_kh2d.cpp:186-330): a periodic
Voronoi mesh of n^2 generators, the shear layer, time_step + step to t_end with an output every 0.02 (the mode
amplitude M of v_y's sin/cos(4 pi x) mode and the peak kinetic energy density K, :86-102), then its report lines
(`totals:`, `fallbacks:`, `run:` with FNV-1a over the bits of Q from the standard offset basis, :274-291, and `ours:`),
so runs compare line by line.

    python -m luna_hydro.kh2d 128 [--moving] [--ic smooth|sharp] [--dt luna|arepo] [--courant C] [--t-end T]
                               [--limiter face|none|learned] [--jitter J] [--seed S] [--steps K] [--out PREFIX]
                               [--precision default|double|grad32|grad16] [--order 0|1] [--riemann hllc|hll]

Defaults: luna-cpp's smooth McNally layer and LUNA's rule at 0.8 (AREPO's at 0.3), t_end 1.5, --precision default
(DefaultPrecision<2>, grad32, as hydro_kh2d.cpp:330-337; double with the learned limiter). Mesh: this project's
jittered lattice (jitter J spacings per axis, uniform in [-J, J], numpy MT19937 seed S, Morton order); the reference
draws its own jitter with std::mt19937_64, so the meshes, and with them the hashes, are this project's own.
"""
import argparse
import math
import sys
import time

import numpy as np

from . import hydro, mesh, problems
from .sedov import bits_hash

GAMMA = 5.0 / 3.0
FNV_BASIS = 14695981039346656037  # hydro_kh2d.cpp:276
OUTPUT_EVERY = 0.02


def mode_amplitude(q, positions, centroid, volume):
    """(M, K) of a state (hydro_kh2d.cpp:86-102): the amplitude of v_y's sin/cos(4 pi x) mode, weighted by
    exp(-4 pi |h - 1/4|) about the layers (h the distance to y = 0 folded about 1/2, at the cell centroids), and the
    peak kinetic energy density of v_y."""
    x = np.asarray(positions).reshape(-1, 2) + np.asarray(centroid).reshape(-1, 2)
    px = x[:, 0]
    y = x[:, 1] - np.floor(x[:, 1])
    vy = q[:, 2] / q[:, 0]
    h = np.where(y < 0.5, y, 1.0 - y)
    w = np.exp(-4 * np.pi * np.abs(h - 0.25)) * volume
    s = math.fsum(vy * np.sin(4 * np.pi * px) * w)
    c = math.fsum(vy * np.cos(4 * np.pi * px) * w)
    d = math.fsum(w)
    K = float(np.max(0.5 * q[:, 2] * q[:, 2] / q[:, 0] / volume))
    return 2 * math.hypot(s / d, c / d), K


def run(n=128, moving=False, ic="smooth", t_end=1.5, rule="luna", courant=None, limiter="face", params=None, jitter=0.25,
        seed=1, steps=0, count_fallbacks=False, out="", log=print, precision="double", order=1, riemann="hllc"):
    """One run; returns its report (series of (t, M, K), hash, totals, steps, wall) after printing the lines."""
    courant = courant if courant is not None else (0.3 if rule == "arepo" else 0.8)
    t0 = time.perf_counter()
    m = mesh.build(problems.jittered_lattice(n, 2, jitter, seed), 1.0)
    build_ms = 1e3 * (time.perf_counter() - t0)
    state = problems.kelvin_helmholtz_smooth_state if ic == "smooth" else problems.kelvin_helmholtz_state
    q = state(m, GAMMA)
    N = m["cells"]
    fv = hydro.FiniteVolume(m, GAMMA, order=order, limiter=limiter, riemann=riemann, motion="moving" if moving else "static",
                            step_rule=rule, precision=hydro.policy(precision, 2) if order == 1 else "double")
    try:
        if limiter == "learned":
            from . import learned

            fv.set_limiter_network(learned.default_params() if params is None else params)
        fv.set_state(q)
        series = [(0.0,) + mode_amplitude(q, m["positions"], m["centroid"], m["volume"])]
        t, count, wall, fallbacks = 0.0, 0, 0.0, 0
        outputs = 1 if steps > 0 else int(round(t_end / OUTPUT_EVERY))
        for k in range(1, outputs + 1):
            target = 1e300 if steps > 0 else k * OUTPUT_EVERY
            a = time.perf_counter()
            while t < target * (1 - 1e-13) and (steps == 0 or count < steps):
                dt = min(fv.time_step(courant), target - t)
                if not (dt > 0):
                    raise RuntimeError("kh: dt %g at t = %g" % (dt, t))
                if moving:
                    x = fv.moving_begin(dt)
                    wall += time.perf_counter() - a
                    m = mesh.build(x, 1.0)  # the mesh producer: outside the timed hydro part
                    a = time.perf_counter()
                    fv.moving_finish(m, dt)
                else:
                    fv.step(dt)
                if count_fallbacks:
                    fallbacks += fv.fallbacks()
                t = target if dt == target - t else t + dt
                count += 1
            hydro.synchronize()
            wall += time.perf_counter() - a
            if steps == 0:
                series.append((t,) + mode_amplitude(fv.state(), m["positions"], m["centroid"], m["volume"]))
        q = fv.state()
    finally:
        fv.close()
    h = bits_hash(q, FNV_BASIS)
    total = [math.fsum(q[:, v]) for v in range(4)]
    rate = N * count / wall * 1e-6 if wall > 0 else 0.0
    log("totals: t=%.6g mass=%.15e x_momentum=%.15e y_momentum=%.15e energy=%.15e" % (t, *total))
    if count_fallbacks:
        log("fallbacks: total=%d steps=%d" % (fallbacks, count))
    log("run: wall_s=%.4f steps=%d M_cell_updates_per_s=%.1f hash=%016x" % (wall, count, rate, h))
    log("ours: N=%d mesh=%s order=%d solver=%s limiter=%s dt=%s courant=%g jitter=%g ic=%s precision=%s steps=%d wall_s=%.4f "
        "build_ms=%.2f M_cell_updates_per_s=%.1f M_final=%.8e" % (n, "moving" if moving else "static", order, riemann, limiter, rule,
                                                                   courant, jitter, ic, precision, count, wall, build_ms, rate,
                                                                   series[-1][1]))
    if out:
        with open(out + ".txt", "w") as f:
            for row in series:
                f.write("%.17g %.17g %.17g\n" % row)
    return {"series": series, "hash": h, "totals": total, "steps": count, "wall_s": wall, "t": t, "fallbacks": fallbacks}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("n", type=int)
    ap.add_argument("--moving", action="store_true")
    ap.add_argument("--ic", choices=["smooth", "sharp"], default="smooth")
    ap.add_argument("--dt", choices=["luna", "arepo"], default="luna")
    ap.add_argument("--courant", type=float, default=None)
    ap.add_argument("--t-end", type=float, default=1.5)
    ap.add_argument("--limiter", choices=["face", "none", "learned"], default="face")
    ap.add_argument("--jitter", type=float, default=0.25)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--count-fallbacks", action="store_true")
    ap.add_argument("--out", default="")
    ap.add_argument("--precision", choices=["default", "double", "grad32", "grad16"], default="default",
                    help="default: the reference's DefaultPrecision<2> (grad32); the learned limiter needs double")
    ap.add_argument("--order", type=int, choices=[0, 1], default=1)
    ap.add_argument("--riemann", choices=["hllc", "hll"], default="hllc")
    a = ap.parse_args(argv)
    if a.limiter == "learned" and a.precision == "default":
        a.precision = "double"
    run(a.n, a.moving, a.ic, a.t_end, a.dt, a.courant, a.limiter, None, a.jitter, a.seed, a.steps, a.count_fallbacks, a.out,
        precision=a.precision, order=a.order, riemann=a.riemann)
    return 0


if __name__ == "__main__":
    sys.exit(main())
