"""The 3D Sedov blast driver, as luna-cpp's hydro_sedov runs it (src/bench/hydro_sedov.cpp:300-406): a jittered periodic
Voronoi mesh, the volume-weighted blast, time_step + step to t_end with snapshot times hit exactly, then its report lines
(`run:`, `fallbacks:`, `conservation:`, `phases_ms_median:`, `checksum:` with FNV-1a over the bits of Q, positions and
volume) so runs compare line by line.

    python -m luna_hydro.sedov --n 64 [--moving] [--t-end 0.05] [--courant 0.3] [--snapshots 10] [--out DIR] [--max-steps K]
                               [--precision default|double|grad32|grad16] [--order 0|1] [--limiter face|none]
                               [--riemann hllc|hll] [--hll-compression K]

--precision as the reference's (hydro_sedov.cpp:409-428): default is its DefaultPrecision<3>, grad16.
"""
import argparse
import ctypes
import math
import os
import subprocess
import sys
import time

import numpy as np

from . import hydro, mesh, problems, snapshot

GAMMA = 5.0 / 3.0
FNV_START = 1469598103934665603  # the start value of the reference's checksum (hydro_sedov.cpp:162)


def bits_hash(a, h=FNV_START):
    """FNV-1a over the bits of every value of a (row-major), chained from h (hydro_sedov.cpp:159-174)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    snapshot.lib().luna_fnv1a.restype = ctypes.c_ulonglong
    return snapshot.lib().luna_fnv1a(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), ctypes.c_ulonglong(h))


def totals(q):
    return {"mass": math.fsum(q[:, 0]), "px": math.fsum(q[:, 1]), "py": math.fsum(q[:, 2]), "pz": math.fsum(q[:, 3]),
            "energy": math.fsum(q[:, 4])}


# The Sedov-Taylor radius R = xi0 (E t^2 / rho)^(1/5) of a point blast, xi0 = 1.15167 for gamma = 5/3 (Sedov 1959;
# the constant of Kamm & Timmes's solution for a spherical blast in a uniform medium).
XI0 = 1.15167


def shock_radius(q, x, volume, h, t, energy=1.0, rho=1.0):
    """(the radius of the peak of the shell-averaged density, the radius beyond it where the average falls to half way
    between the peak and the ambient density, the Sedov-Taylor radius) of the blast about (1/2, 1/2, 1/2): shells of
    width h, volume-weighted averages of rho = Q_0 / V at the generators' periodic distances."""
    d = x - 0.5
    d -= np.round(d)
    r = np.sqrt(np.sum(d * d, axis=1))
    shell = np.minimum((r / h).astype(int), int(0.5 / h))
    mass = np.bincount(shell, weights=q[:, 0])
    vol = np.bincount(shell, weights=volume)
    avg = np.where(vol > 0, mass / np.maximum(vol, 1e-300), rho)
    k = int(np.argmax(avg))
    half = rho + 0.5 * (avg[k] - rho)
    j = k
    while j + 1 < len(avg) and avg[j + 1] > half:
        j += 1
    # the half-way crossing between shells j and j + 1, linearly
    front = (j + 0.5) * h
    if j + 1 < len(avg) and avg[j] != avg[j + 1]:
        front = (j + 0.5 + (avg[j] - half) / (avg[j] - avg[j + 1])) * h
    return (k + 0.5) * h, front, XI0 * (energy * t * t / rho) ** 0.2


def commit():
    try:
        here = os.path.dirname(os.path.abspath(__file__))
        rev = subprocess.run(["git", "-C", here, "rev-parse", "--short", "HEAD"], capture_output=True, text=True, timeout=10)
        dirty = subprocess.run(["git", "-C", here, "status", "--porcelain"], capture_output=True, text=True, timeout=10)
        return rev.stdout.strip(), bool(dirty.stdout.strip())
    except (OSError, subprocess.SubprocessError):
        return "", False


class _Engine:
    """hydro.FiniteVolume (the GPU) or an object with its methods; the moving step needs the mesh producer between
    moving_begin and moving_finish."""

    def __init__(self, fv, moving):
        self.fv, self.moving = fv, moving

    def step(self, dt):
        if not self.moving:
            self.fv.step(dt)
            return None
        x = self.fv.moving_begin(dt)
        m = mesh.build(x, 1.0)
        self.fv.moving_finish(m, dt)
        return m


def run(n=64, moving=False, t_end=0.05, courant=0.3, snapshots=10, out="", max_steps=0, jitter=0.25, seed=1, log=print,
        precision="double", order=1, limiter="face", riemann="hllc", hll_compression=0.0):
    m = mesh.build(problems.jittered_lattice(n, 3, jitter, seed), 1.0)
    q, deposited = problems.sedov_state(m)
    N = m["cells"]
    fv = hydro.FiniteVolume(m, GAMMA, order=order, limiter=limiter, riemann=riemann, motion="moving" if moving else "static",
                            precision=hydro.policy(precision, 3) if order == 1 else "double", hll_compression=hll_compression)
    fv.set_state(q)
    engine = _Engine(fv, moving)
    before = totals(q)
    rev, dirty = commit()
    current = m

    def write(index, t):
        if not out:
            return
        x = fv.positions() if moving else np.asarray(current["positions"]).reshape(-1, 3)
        snapshot.write(out, index, t, x, current["volume"], fv.state(), GAMMA, commit=rev, dirty=dirty)

    write(0, 0.0)
    walls, steps, fallbacks, most, with_fallbacks, t, nxt = [], 0, 0, 0, 0, 0.0, 1
    fv.profile(True)
    start = time.perf_counter()
    said = start
    while t < t_end and (max_steps == 0 or steps < max_steps):
        if time.perf_counter() - said > 20.0:  # a progress line every 20 s of wall (moving runs wait on the host producer)
            said = time.perf_counter()
            print("progress: steps=%d t=%.6g" % (steps, t), file=sys.stderr, flush=True)
        a = time.perf_counter()
        dt = fv.time_step(courant)
        t_next = t_end * nxt / snapshots
        snap = t + dt >= t_next
        if snap:
            dt = t_next - t
        new = engine.step(dt)
        if new is not None:
            current = new
        step_fallbacks = fv.fallbacks()  # fences: the step is done
        walls.append((time.perf_counter() - a) * 1e3)
        fallbacks += step_fallbacks
        most = max(most, step_fallbacks)
        with_fallbacks += step_fallbacks > 0
        t = t_next if snap else t + dt
        steps += 1
        if snap:
            if out or nxt == snapshots:
                write(nxt, t)
            nxt += 1
    wall_s = time.perf_counter() - start
    qf = fv.state()
    after = totals(qf)
    median = float(np.median(walls)) if walls else float("nan")
    kernels = fv.kernel_times()
    log("run: precision=%s order=%d update=face steps=%d limiter=%s riemann=%s mesh=%s blast=%d t=%.17g wall_s=%.6g "
        "s_per_step=%.6g step_wall_ms_median=%.6g G_cell_updates_per_s=%.6g" % (precision, order, steps, limiter, riemann,
                                                                                "moving" if moving else "static", deposited, t,
                                                                                wall_s, wall_s / max(steps, 1), median,
                                                                                N / (median * 1e-3) / 1e9))
    log("fallbacks: total=%d steps_with=%d most_in_a_step=%d face_sides_per_step=%d" % (fallbacks, with_fallbacks, most, 2 * fv.faces()))
    log("phases_ms_median: " + " ".join("%s=%.4g" % (k, v[1] / v[0]) for k, v in sorted(kernels.items())))
    scale = before["mass"] * math.sqrt(2 * before["energy"] / before["mass"])
    log("conservation: mass=%.3e energy=%.3e momentum=(%.3e,%.3e,%.3e) momentum_scale=%.6g" % (
        (after["mass"] - before["mass"]) / before["mass"], (after["energy"] - before["energy"]) / before["energy"],
        after["px"] - before["px"], after["py"] - before["py"], after["pz"] - before["pz"], scale))
    x = fv.positions() if moving else np.asarray(current["positions"]).reshape(-1, 3)
    log("checksum: steps=%d Q=%x x=%x volume=%x" % (steps, bits_hash(qf), bits_hash(x), bits_hash(current["volume"])))
    peak, front, analytic = shock_radius(qf, x, current["volume"], 1.0 / n, t)
    log("shock: t=%.6g peak_radius=%.5g half_max_radius=%.5g sedov_taylor_radius=%.5g" % (t, peak, front, analytic))
    fv.close()
    return {"steps": steps, "t": t, "q": qf, "x": x, "volume": current["volume"]}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--moving", action="store_true")
    ap.add_argument("--t-end", type=float, default=0.05)
    ap.add_argument("--courant", type=float, default=0.3)
    ap.add_argument("--snapshots", type=int, default=10)
    ap.add_argument("--out", default="")
    ap.add_argument("--max-steps", type=int, default=0)
    ap.add_argument("--precision", choices=["default", "double", "grad32", "grad16"], default="default",
                    help="default: the reference's DefaultPrecision<3> (grad16)")
    ap.add_argument("--order", type=int, choices=[0, 1], default=1)
    ap.add_argument("--limiter", choices=["face", "none"], default="face")
    ap.add_argument("--riemann", choices=["hllc", "hll"], default="hllc")
    ap.add_argument("--hll-compression", type=float, default=0.0, help="the HLL fallback at extreme compression (0: off)")
    a = ap.parse_args(argv)
    run(a.n, a.moving, a.t_end, a.courant, a.snapshots, a.out, a.max_steps, precision=a.precision, order=a.order, limiter=a.limiter,
        riemann=a.riemann, hll_compression=a.hll_compression)


if __name__ == "__main__":
    sys.exit(main())
