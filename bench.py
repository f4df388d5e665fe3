"""Benchmark of the hydro step on one MI355X: cell updates/s of the 3D Sedov blast on a 128^3 jittered Voronoi mesh,
moving (the headline: the hydro part of the two-pass moving step, the mesh producer excluded and reported apart) and
static, FP64, with the dominant kernel's HBM roofline fraction from HIP events and the CPU restatement timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--static-steps S] [--no-cpu] [--no-learned]

One JSON line on stdout (the last line); progress on stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "luna-cpp_amd"))

import numpy as np  # noqa: E402

from luna_hydro import hydro, mesh, problems  # noqa: E402

GAMMA = 5.0 / 3.0
COURANT = 0.3
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell of each kernel: every array a phase needs read once and each output written once (unique
# bytes; FP64 values, int32 indices), as functions of the mesh's faces and list entries per cell (f = F/N, e = E/N = 2f;
# luna-cpp quotes its figures at f = 7.95, the jittered 128^3 mesh here has 7.775).
#   sweep+dt: Q-derived rows W (40), cell geometry (V 8, c 24, x 24, CSR 4), list entries 4e, face records 56f, writes
#             g 120 and W 40 -> 260 + 4e + 56f static (768.8 at f = 7.95); moving also writes w (24) and keeps c^n (24)
#   keep_centroids (moving): x read and snapped (48), w (24), grid steps written (12) -> 84
#   flux:     faces 56f, W + g + c + x of the cells 208, writes the flux rows 40f -> 208 + 96f; moving pass 1 also reads
#             the grid steps (int32, 12); pass 2 the steps and the old centroids (24)
#   update:   CSR 4, list 4e, flux rows 40f, Q read and written 80 -> 84 + 4e + 40f; moving pass 1 with the drift (x r/w
#             48, grid steps 12)
# and of the geometry pass of a moving step (the new mesh's derived layouts from the producer's arrays):
#   cells:  the producer's V, c, x and list start in (60), the 64 B geometry half of the cell row, the count and a dense
#           copy of V out -> 136
#   pairs:  owner and neighbour checked (8 a face) -> 8f (the flux kernels read the face arrays in place)
#   slots:  the producer's list (4e), each face's owner, neighbour, A and centroid once (56f), the cells' x and c once
#           (48), 40 B per list entry out (the face id with its owner bit, the other cell, |A|, the face offset)
#           -> 48 + 44e + 56f
def algo_bytes(f, e):
    return {
        "static": {"sweep": 260 + 4 * e + 56 * f, "flux": 208 + 96 * f, "update": 84 + 4 * e + 40 * f},
        "moving": {"sweep": 308 + 4 * e + 56 * f, "keep_centroids": 84.0, "flux_pass1": 220 + 96 * f,
                   "update_pass1": 144 + 4 * e + 40 * f, "flux_pass2": 244 + 96 * f, "update_pass2": 84 + 4 * e + 40 * f},
        "geometry": {"geometry_cells": 136.0, "geometry_pairs": 8 * f, "geometry_slots": 48 + 44 * e + 56 * f},
    }

TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_sedov128.md")


def measured_traffic(phase, n):
    """HBM bytes per launch of a phase ("moving sweep", "static flux", ...) at n = 128 from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (its table rows: | phase | kernel | ... | bytes per launch |, the counters already
    scaled by the plain-copy calibration); None when the summary is absent or the workload is another size."""
    if n != 128 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) >= 2 and cols[0] == phase:
                try:
                    return float(cols[-1])
                except ValueError:
                    return None
    return None


VALU_FILE = os.path.join(ROOT, "profiles", "valu_counters_sedov128.md")
SIMDS, CLOCK_HZ = 1024, 2.4e9  # 256 CUs x 4 SIMDs; a wave64 VALU instruction issues in 4 cycles of one SIMD


def valu_floor_ms(kernel_prefix, n, cells):
    """The time a kernel would take issuing only its VALU instructions, every SIMD busy (instructions per wave from the
    committed rocprofv3 SQ_INSTS_VALU summary's VALU-busy table, at n = 128): the bound of the FP64 arithmetic the
    kernel must do. None when the summary is absent or the workload is another size."""
    if n != 128 or not os.path.exists(VALU_FILE):
        return None
    with open(VALU_FILE) as f:
        for line in f:
            cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
            if len(cols) == 5 and cols[0].startswith(kernel_prefix):
                try:
                    per_wave = float(cols[2])
                except ValueError:
                    return None
                return per_wave * (cells / 64.0) * 4 / SIMDS / CLOCK_HZ * 1e3
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def dominant(times, table):
    best = max((k for k in times if k in table), key=lambda k: times[k][1])
    launches, total = times[best]
    return best, launches, total / launches


def run_static(m, q, steps, warmup, precision="double", hll=0.0, update="per_face", lanes=1):
    fv = hydro.FiniteVolume(m, GAMMA, precision=precision, hll_compression=hll, update=update, sweep_lanes=lanes)
    fv.set_state(q)
    for _ in range(warmup):
        fv.step(fv.time_step(COURANT))
    fv.fallbacks()
    hydro.synchronize()
    fv.profile(True)
    walls = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        dt = fv.time_step(COURANT)
        fv.step(dt)
        fv.fallbacks()  # fences: the step is done (as hydro_sedov's step wall)
        walls.append(time.perf_counter() - a)
    hydro.synchronize()
    total = time.perf_counter() - t0
    times = fv.kernel_times()
    fv.close()
    return total, walls, times


def run_moving(m, q, steps, warmup, precision="double", hll=0.0, update="per_face", lanes=1):
    """Moving steps; the hydro part is timed as the reference's: time_step, the first pass and drift, the new mesh's
    geometry pass (luna_fv_set_geometry_device on the producer's arrays in device memory, as the reference reads
    voronoi-cpp's device views in its sweep and primitives), the second pass. The mesh producer (host) and the copies
    that stand in for a GPU producer's output (positions to the host, the new mesh's arrays to the device) are excluded
    and reported apart."""
    fv = hydro.FiniteVolume(hydro.DeviceMesh(m), GAMMA, motion="moving", precision=precision, hll_compression=hll,
                            update=update, sweep_lanes=lanes)
    fv.set_state(q)
    mesh_s, hydro_s, geo_s = [], [], []

    def one():
        a = time.perf_counter()
        dt = fv.time_step(COURANT)
        fv.moving_begin(dt, positions=False)
        b = time.perf_counter()
        nm = mesh.build(fv.positions(), 1.0)
        dm = hydro.DeviceMesh(nm)
        c = time.perf_counter()
        fv.set_geometry(dm)
        e = time.perf_counter()
        fv.moving_finish(None, dt)
        fv.fallbacks()  # fences
        d = time.perf_counter()
        return (b - a) + (d - c), c - b, e - c

    for _ in range(warmup):
        one()
    hydro.synchronize()
    fv.profile(True)
    t0 = time.perf_counter()
    for k in range(steps):
        h, mm, g = one()
        hydro_s.append(h)
        mesh_s.append(mm)
        geo_s.append(g)
        log("  moving step %d: hydro %.3f ms (geometry pass %.3f ms), mesh producer + copies %.2f s" % (k, 1e3 * h, 1e3 * g, mm))
    hydro.synchronize()
    total = time.perf_counter() - t0
    times = fv.kernel_times()
    fv.close()
    return total, hydro_s, mesh_s, times, geo_s


def cpu_baseline(m, q, static_steps=40, moving_steps=10, warmup=2):
    """The oracle (CPU restatement of luna-cpp's FP64 algorithm, C + OpenMP) on the same mesh and state, on this host's
    cores (OpenMP's own placement: on the MI355X's host, whose affinity mask is wider than the process's CPU share,
    OMP_PROC_BIND=close put all threads on one core, so it is not set here): the static step's median wall over `static_steps` after `warmup`, as
    hydro_sedov's run line; the moving step's hydro part (mesh producer excluded) over `moving_steps`; and one round of
    each on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = oracle.max_threads()

    def static_walls(count, warm, budget):
        fv = oracle.FiniteVolume(m, GAMMA)
        fv.set_state(q)
        walls = []
        t0 = time.perf_counter()
        for k in range(warm + count):
            a = time.perf_counter()
            fv.step(fv.time_step(COURANT))
            if k >= warm:
                walls.append(time.perf_counter() - a)
            if (k + 1) % 10 == 0 or k + 1 == warm + count:
                log("  cpu static step %d: %.1f ms" % (k, 1e3 * (time.perf_counter() - a)))
            if time.perf_counter() - t0 > budget and walls:
                break
        return walls

    def moving_walls(count, budget):
        fv = oracle.FiniteVolume(m, GAMMA, moving=True)
        fv.set_state(q)
        walls = []
        t0 = time.perf_counter()
        for k in range(count):
            a = time.perf_counter()
            dt = fv.time_step(COURANT)
            x = fv.moving_begin(dt)
            b = time.perf_counter()
            nm = mesh.build(x, 1.0)
            c = time.perf_counter()
            fv.moving_finish(nm, dt)
            walls.append((b - a) + (time.perf_counter() - c))
            log("  cpu moving step %d: hydro %.1f ms" % (k, 1e3 * walls[-1]))
            if time.perf_counter() - t0 > budget:
                break
        return walls

    log("  cpu: %d threads, OMP_PROC_BIND=%s" % (threads, os.environ.get("OMP_PROC_BIND")))
    s_walls = static_walls(static_steps, warmup, 30.0)
    m_walls = moving_walls(moving_steps, 40.0)
    oracle.set_threads(1)
    s1 = static_walls(1, 0, 0.0)
    m1 = moving_walls(1, 0.0)
    oracle.set_threads(threads)
    N = m["cells"]
    return {"value": N / float(np.median(m_walls)), "unit": "cell updates/s", "cores": threads, "kind": "port",
            "static": N / float(np.median(s_walls)), "static_1_thread": N / s1[0], "moving_1_thread": N / m1[0],
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "sample": "oracle/ (C + OpenMP restatement of luna-cpp's FP64 algorithm) on the same %d^3 Sedov mesh and state: "
                      "value = moving hydro part (mesh producer excluded), median of %d steps; static = median of %d steps "
                      "after %d warm-ups (each leg capped at about 30-40 s of wall); one step each on 1 thread"
                      % (round(N ** (1 / 3)), len(m_walls), len(s_walls), warmup)}


def derivative_costs(m, q, repeats=5):
    """Wall per call of the static step's derivatives against the forward step on the same 128^3 state (device rows):
    forward = time_step + step; tangent = time_step_tangent + step_tangent; adjoint = step_adjoint (its own sweep of Q)."""
    import torch

    fv = hydro.FiniteVolume(m, GAMMA)
    rng = np.random.Generator(np.random.MT19937(3))
    dq = torch.as_tensor(rng.standard_normal(q.shape) * 1e-3 * np.abs(q)).cuda()
    out = {}
    for name in ("forward", "tangent", "adjoint"):
        walls = []
        for k in range(repeats + 1):
            fv.set_state(q)
            hydro.synchronize()
            a = time.perf_counter()
            if name == "forward":
                fv.step(fv.time_step(COURANT))
                fv.fallbacks()
            elif name == "tangent":
                fv.step_tangent(fv.time_step_tangent(COURANT, dq), dq)
            else:
                fv.step_adjoint(2e-6, dq)
            hydro.synchronize()
            if k:
                walls.append(time.perf_counter() - a)
        out[name + "_ms"] = 1e3 * float(np.median(walls))
    fv.close()
    out["tangent_over_forward"] = out["tangent_ms"] / out["forward_ms"]
    out["adjoint_over_forward"] = out["adjoint_ms"] / out["forward_ms"]
    out["reference_a100"] = {"forward_ms": 13.0, "tangent_over_forward": 1.86, "adjoint_over_forward": 2.46,
                             "source": "research_notes/autodiff_l2_2026-11-20.md:36-45"}
    return out


def sod_1d():
    """The 1D Sod tube of luna-cpp's hydro_1d (N = 1024, minmod, HLLC, CFL 0.4, LUNA's rule, fused step, t = 0.2):
    steps and microseconds per step on the GPU (launch-latency bound at 72 KB of state)."""
    from luna_hydro import hydro1d

    best = None
    for _ in range(3):
        steps, t, wall, h, *_ = hydro1d.simulate("sod", 1024, limiter="minmod", courant=0.4, t_end=0.2)
        best = wall if best is None else min(best, wall)
    return {"workload": "1D Sod, 1024 cells, minmod, HLLC, CFL 0.4, LUNA rule, fused step, t = 0.2", "steps": steps,
            "us_per_step": 1e6 * best / steps, "steps_per_s": steps / best, "hash": "%016x" % h}


def learned_vs_plain(log):
    """The learned limiter (the shipped network) against the plain face-centroid limiter: density L1 against a plain
    run at 4x the resolution, and GPU-seconds, on the smooth KH layer at 64^2 to t = 1 (static mesh, LUNA rule CFL 0.8)."""
    from luna_hydro import learned

    rows, ref = learned.curve(learned.KH_SMOOTH, [64], 99, 1.0, learned.default_params(), 256, log=log)
    out = {"workload": "smooth Kelvin-Helmholtz layer, 64^2 jittered static mesh, t = 1, reference: plain at 256^2",
           "reference_gpu_s": ref["seconds"]}
    for r in rows:
        out[r["limiter"]] = {"l1": r["l1"], "gpu_s": r["seconds"], "steps": r["steps"],
                             "l1_per_gpu_s": r["l1"] / r["seconds"]}
    out["l1_ratio"] = out["learned"]["l1"] / out["plain"]["l1"]
    out["gpu_s_ratio"] = out["learned"]["gpu_s"] / out["plain"]["gpu_s"]
    return out


def calibrate_copy(mib):
    """A plain device-to-device copy of `mib` MiB (the runtime's blit kernel), three times: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch against the known bytes give the counters' scale."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration_copy_bytes": n, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--static-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-limiter comparison")
    ap.add_argument("--no-precision", action="store_true", help="skip the grad32/grad16 policy runs")
    ap.add_argument("--no-derivatives", action="store_true", help="skip the derivative cost runs")
    ap.add_argument("--no-sod", action="store_true", help="skip the 1D Sod run")
    ap.add_argument("--no-update-modes", action="store_true", help="skip the per-cell update (Update::per_cell) runs")
    ap.add_argument("--no-lanes", action="store_true", help="skip the LaneSweep (2, 4, 8 lanes per cell) runs")
    ap.add_argument("--hll-compression", type=float, default=0.0,
                    help="the HLL fallback at extreme compression in the FP64 runs (threshold; 0: off, the reference's scheme)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (scale of the rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: one GPU only; the step has no multi-GPU path (--gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    n = args.n
    log("building the %d^3 jittered Voronoi mesh (jitter 0.25, seed 1)" % n)
    t = time.perf_counter()
    m = mesh.build(problems.jittered_lattice(n, 3, 0.25, 1), 1.0)
    build_s = time.perf_counter() - t
    N, F = m["cells"], m["faces"]
    q, deposited = problems.sedov_state(m)
    log("mesh: %d cells, %d faces (F/N %.3f), built in %.1f s; blast in %d cells" % (N, F, F / N, build_s, deposited))

    log("static: %d warm-up + %d timed steps" % (args.warmup, args.static_steps))
    s_total, s_walls, s_times = run_static(m, q, args.static_steps, args.warmup, hll=args.hll_compression)
    static_rate = N * args.static_steps / s_total
    log("static: %.3f ms/step, %.3g cell updates/s; kernels %s" % (1e3 * s_total / args.static_steps, static_rate,
                                                                    {k: round(v[1] / v[0], 4) for k, v in s_times.items()}))

    log("moving: %d warm-up + %d timed steps" % (args.warmup, args.steps))
    m_total, hydro_s, mesh_s, m_times, geo_s = run_moving(m, q, args.steps, args.warmup, hll=args.hll_compression)
    hydro_total = float(np.sum(hydro_s))
    value = N * args.steps / hydro_total
    log("moving: hydro %.3f ms/step, %.3g cell updates/s; kernels %s" % (1e3 * hydro_total / args.steps, value,
                                                                          {k: round(v[1] / v[0], 4) for k, v in m_times.items()}))

    ALGO = algo_bytes(F / N, 2.0 * F / N)
    name, launches, avg_ms = dominant(m_times, ALGO["moving"])
    per_cell = ALGO["moving"][name]
    achieved = per_cell * N / (avg_ms * 1e-3) / 1e9
    sname, _, savg = dominant(s_times, ALGO["static"])
    s_ach = ALGO["static"][sname] * N / (savg * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, %d threads)" % (os.cpu_count() or 1))
        cpu = cpu_baseline(m, q)
        log("cpu: %.3g cell updates/s" % cpu["value"])

    policies = None
    if not args.no_precision:
        # The reference's reduced-precision policies (its published A100 numbers are grad16, its 3D default): the same
        # mesh and state, static and moving, reported beside the FP64 headline.
        policies = {}
        for p in ("grad32", "grad16"):
            t_s, w_s, k_s = run_static(m, q, args.static_steps, args.warmup, p)
            t_m, h_m, _, k_m, _ = run_moving(m, q, max(2, args.steps // 2), args.warmup, p)
            h_total = float(np.sum(h_m))
            policies[p] = {"static": N * args.static_steps / t_s, "static_ms_per_step": 1e3 * t_s / args.static_steps,
                           "moving": N * len(h_m) / h_total, "moving_hydro_ms_per_step": 1e3 * h_total / len(h_m),
                           "static_kernels_ms_avg": {k: v[1] / v[0] for k, v in k_s.items()},
                           "moving_kernels_ms_avg": {k: v[1] / v[0] for k, v in k_m.items()}}
            log("%s: static %.3g, moving %.3g cell updates/s" % (p, policies[p]["static"], policies[p]["moving"]))

    modes = None
    if not args.no_update_modes:
        # The reference's other update schedule (Update::per_cell, hydro::fluxes_and_update): each cell computes the
        # fluxes of all its faces and updates only itself; the same bits as per_face.
        t_s, _, k_s = run_static(m, q, args.static_steps, args.warmup, update="per_cell")
        t_m, h_m, _, k_m, _ = run_moving(m, q, max(2, args.steps // 2), args.warmup, update="per_cell")
        h_total = float(np.sum(h_m))
        modes = {"per_cell": {"static": N * args.static_steps / t_s, "static_ms_per_step": 1e3 * t_s / args.static_steps,
                              "moving": N * len(h_m) / h_total, "moving_hydro_ms_per_step": 1e3 * h_total / len(h_m),
                              "static_kernels_ms_avg": {k: v[1] / v[0] for k, v in k_s.items()},
                              "moving_kernels_ms_avg": {k: v[1] / v[0] for k, v in k_m.items()}}}
        log("per_cell update: static %.3g, moving %.3g cell updates/s" % (modes["per_cell"]["static"], modes["per_cell"]["moving"]))

    lanes = None
    if not args.no_lanes:
        # The reference's LaneSweep (LUNA_SWEEP_LANES): L lanes per cell, other bits than the one-lane sweep's.
        lanes = {}
        for L in (2, 4, 8):
            t_s, _, k_s = run_static(m, q, args.static_steps, args.warmup, lanes=L)
            t_m, h_m, _, k_m, _ = run_moving(m, q, max(2, args.steps // 2), args.warmup, lanes=L)
            h_total = float(np.sum(h_m))
            lanes[L] = {"static": N * args.static_steps / t_s, "static_ms_per_step": 1e3 * t_s / args.static_steps,
                        "moving": N * len(h_m) / h_total, "moving_hydro_ms_per_step": 1e3 * h_total / len(h_m),
                        "static_sweep_ms": k_s["sweep"][1] / k_s["sweep"][0], "moving_sweep_ms": k_m["sweep"][1] / k_m["sweep"][0]}
            log("LaneSweep %d lanes: static sweep %.3f ms, moving sweep %.3f ms" % (L, lanes[L]["static_sweep_ms"], lanes[L]["moving_sweep_ms"]))

    derivs = None
    if not args.no_derivatives:
        log("derivative costs of the static step (forward, tangent, adjoint)")
        derivs = derivative_costs(m, q)
        log("derivatives: %s" % {k: v for k, v in derivs.items() if k != "reference_a100"})

    sod = None
    if not args.no_sod:
        sod = sod_1d()
        log("1D Sod: %s" % sod)

    learned_cmp = None
    if not args.no_learned:
        log("learned limiter vs plain (smooth KH 64^2, reference 256^2)")
        learned_cmp = learned_vs_plain(log)

    step_bytes = sum(ALGO["moving"].values())
    out = {
        "metric": "cell updates/s at 128^3 moving mesh (hydro part with the per-step geometry pass, mesh producer excluded)",
        "value": value,
        "unit": "cell updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * hydro_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded 3D Sedov blast on a jittered lattice (jitter 0.25, seed 1), E = 1 in r < 2h",
        "config": {"workload": "3D Sedov blast, moving Voronoi mesh, %d^3 = %d cells, %d faces, HLLC, MUSCL-Hancock "
                               "two-pass moving step, CFL 0.3, FP64" % (n, N, F), "cells": N, "faces": F,
                   "parallelism": "single GPU", "hll_compression": args.hll_compression},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("moving " + name, n),
                     "algorithmic_bytes_per_cell": per_cell, "avg_ms": avg_ms, "launches": launches,
                     "valu_floor_ms": valu_floor_ms("k_sweep<3, true, true, 1, 0, 0>", n, N) if name == "sweep" else None,
                     "note": "the FP64 sweep's arithmetic (7430-8396 VALU instructions a cell, 46 IEEE divisions in its solve) "
                             "bounds it at its VALU floor before HBM: at the floor it would move its algorithmic bytes at "
                             "about half of 8 TB/s"},
        "cpu_baseline": cpu,
        "learned_vs_plain": learned_cmp,
        "precision_policies": policies,
        "update_modes": modes,
        "sweep_lanes": lanes,
        "step_roofline": {"algorithmic_bytes_per_cell": step_bytes,
                          "achieved_GBs": step_bytes * N * args.steps / hydro_total / 1e9,
                          "frac": step_bytes * N * args.steps / hydro_total / 1e9 / HBM_PEAK_GBS},
        "mesh": {"build_s_first": build_s, "producer_and_copies_s_per_step": float(np.mean(mesh_s)),
                 "geometry_pass_ms_per_step": 1e3 * float(np.mean(geo_s)),
                 "geometry_roofline": {k: {"algorithmic_bytes_per_cell": b, "avg_ms": m_times[k][1] / m_times[k][0],
                                           "achieved_GBs": b * N / (m_times[k][1] / m_times[k][0] * 1e-3) / 1e9,
                                           "frac": b * N / (m_times[k][1] / m_times[k][0] * 1e-3) / 1e9 / HBM_PEAK_GBS}
                                       for k, b in ALGO["geometry"].items() if k in m_times},
                 "note": "value includes the geometry pass on the device (luna_fv_set_geometry_device: cell rows, "
                         "slot lists with |A| and face offsets, radii; the face arrays read in place); excluded: the Voronoi producer (C++/OpenMP, "
                         "host) and the copies standing in for a GPU producer (positions to the host, its arrays to the device)"},
        "derivatives": derivs,
        "sod_1d": sod,
        "wall_s_moving_loop": m_total,
        "kernels_ms_avg": {k: v[1] / v[0] for k, v in m_times.items()},
        "static": {"value": static_rate, "unit": "cell updates/s", "ms_per_step": 1e3 * s_total / args.static_steps,
                   "step_median_ms": 1e3 * float(np.median(s_walls)), "kernel": sname, "achieved_GBs": s_ach,
                   "frac": s_ach / HBM_PEAK_GBS, "step_frac": sum(ALGO["static"].values()) * static_rate / 1e9 / HBM_PEAK_GBS,
                   "kernels_ms_avg": {k: v[1] / v[0] for k, v in s_times.items()}},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
