"""The learned slope limiter (no counterpart in luna-cpp): its trainer on the GPU, the coarse/fine training pairs that the
plain solver generates, and its evaluation against the plain method.

    from luna_hydro import learned, hydro
    x, t, w = learned.pairs(learned.KH, n=64, seeds=[1, 2], marks=[0.5, 1.0])   # features, target fractions, weights
    params, history = learned.train(x, t, w, steps=3000)
    fv = hydro.FiniteVolume(mesh, 5 / 3, limiter="learned"); fv.set_limiter_network(params)

The limiter (include/luna_hydro.h, LUNA_LIMITER_LEARNED) scales each cell's face-centroid-limited gradient of each
variable by a fraction f = sigmoid(MLP(features)) in [0, 1], except in troubled cells. A training pair is a coarse
mesh with the state restricted from a plain run at `factor` x the resolution: the features are those the coarse sweep
records on that state, and the target is the fraction that best matches the fine data inside each coarse cell, the
projection of the fine sub-cell slope g_fine on the face-centroid gradient g: t = clip(g_fine . g / |g|^2, 0, 1)
(the limiter cannot steepen g). Rows are weighted by phi max(up, down) / S (features 0, 3, 4: the size of the limited
face increments against the neighbours' range), and rows of troubled cells get weight 0.

The network's kernels (BF16 operands, FP32 accumulation on the matrix cores, csrc/hydro/limiter_kernels.hip) run the
training: forward, backward, a fixed-order reduction of the per-wavefront sums and Adam; torch only holds the device
arrays. Every routine here fails without the HIP library (there is no CPU path).
"""
import ctypes
import os
import time

import numpy as np

from . import hydro, mesh as meshlib, problems

FEATURES, HIDDEN = 16, 64
P_W1 = 0
P_B1 = P_W1 + HIDDEN * FEATURES
P_W2 = P_B1 + HIDDEN
P_B2 = P_W2 + HIDDEN * HIDDEN
P_W3 = P_B2 + HIDDEN
P_B3 = P_W3 + HIDDEN
PARAMS = P_B3 + 1
assert PARAMS == hydro.LIMITER_PARAMS

_fp = ctypes.POINTER(ctypes.c_float)


def init_params(seed=0):
    """He-normal weights, zero biases, and b3 = 3 (every fraction ~0.95 at the start: near the face-centroid limiter)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, dtype=np.float32)
    p[P_W1:P_B1] = rng.normal(0, np.sqrt(2.0 / FEATURES), HIDDEN * FEATURES)
    p[P_W2:P_B2] = rng.normal(0, np.sqrt(2.0 / HIDDEN), HIDDEN * HIDDEN)
    p[P_W3:P_B3] = rng.normal(0, np.sqrt(1.0 / HIDDEN), HIDDEN)
    p[P_B3] = 3.0
    return p


PARAMS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "limiter_params.npy")


def default_params():
    """The network this package ships: trained on the GPU by these kernels on coarse/fine pairs of the smooth KH layer
    (64^2 x 4 seeds, 128^2 x 2), the sharp one (64^2 x 2) and the Sedov blast (32^3 x 3), each against a plain run at
    4x the resolution, 8000 Adam steps (batch 65536, lr 1e-3, seed 0)."""
    p = np.load(PARAMS_FILE, allow_pickle=False)
    if p.dtype != np.float32 or p.shape != (PARAMS,):
        raise ValueError("%s: expected %d float32 parameters" % (PARAMS_FILE, PARAMS))
    return p


def bf16_value(bits):
    """The float32 values of BF16 bit patterns."""
    return (np.ascontiguousarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the learned limiter's trainer needs a GPU (there is no CPU path)")
    return torch


def device_rows(x_bits, target=None, weight=None):
    """Feature rows (uint16 BF16 bits, rows x 16) and optional targets / weights as contiguous GPU tensors."""
    torch = _torch()
    x = torch.from_numpy(np.ascontiguousarray(x_bits, dtype=np.uint16).view(np.int16)).cuda()
    out = [x]
    for a in (target, weight):
        if a is not None:
            out.append(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda())
    torch.cuda.synchronize()
    return out if len(out) > 1 else out[0]


class Trainer:
    """luna_mlp: the network's FP32 parameters with Adam, and its BF16 copy, on the GPU."""

    def __init__(self, params=None, seed=0):
        self._torch = _torch()
        h = ctypes.c_void_p()
        hydro._check(hydro.lib().luna_mlp_create(ctypes.byref(h)), "luna_mlp_create")
        self._h = h
        self.set_params(init_params(seed) if params is None else params)

    def close(self):
        if getattr(self, "_h", None):
            hydro.lib().luna_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != PARAMS:
            raise ValueError("expected %d parameters" % PARAMS)
        hydro._check(hydro.lib().luna_mlp_set_params(self._h, p.ctypes.data_as(_fp), PARAMS), "luna_mlp_set_params")

    def params(self):
        p = np.empty(PARAMS, dtype=np.float32)
        hydro._check(hydro.lib().luna_mlp_get_params(self._h, p.ctypes.data_as(_fp), PARAMS), "luna_mlp_get_params")
        return p

    def _ready(self, *tensors):
        for t in tensors:
            if not (t.is_cuda and t.is_contiguous()):
                raise ValueError("contiguous GPU tensors expected")
        self._torch.cuda.synchronize()  # the kernels run on the trainer's own stream

    def forward(self, x):
        """sigmoid(z) of each feature row (x: GPU int16 tensor, rows x 16 BF16 bits) as a GPU float32 tensor."""
        self._ready(x)
        rows = x.shape[0]
        out = self._torch.empty(rows, dtype=self._torch.float32, device=x.device)
        hydro._check(hydro.lib().luna_mlp_forward(self._h, ctypes.c_void_p(x.data_ptr()), rows, ctypes.c_void_p(out.data_ptr())),
                     "luna_mlp_forward")
        return out

    def gradient(self, x, target, weight, weight_sum=None):
        """(loss, gradient in the parameter order) of a batch."""
        self._ready(x, target, weight)
        ws = float(weight.sum().item()) if weight_sum is None else float(weight_sum)
        g = np.empty(PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        hydro._check(hydro.lib().luna_mlp_gradient(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(target.data_ptr()),
                                                   ctypes.c_void_p(weight.data_ptr()), x.shape[0], ctypes.c_double(ws),
                                                   g.ctypes.data_as(_fp), ctypes.byref(loss)), "luna_mlp_gradient")
        return float(loss.value), g

    def step(self, x, target, weight, weight_sum, lr, want_loss=False):
        """One Adam step on a batch (inputs ready on the GPU); returns the loss before it if want_loss (blocks)."""
        loss = ctypes.c_float()
        hydro._check(hydro.lib().luna_mlp_train_step(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(target.data_ptr()),
                                                     ctypes.c_void_p(weight.data_ptr()), x.shape[0], ctypes.c_double(weight_sum),
                                                     ctypes.c_float(lr), ctypes.byref(loss) if want_loss else None),
                     "luna_mlp_train_step")
        return float(loss.value) if want_loss else None

    def synchronize(self):
        hydro._check(hydro.lib().luna_mlp_synchronize(self._h), "luna_mlp_synchronize")


def train(x_bits, target, weight, steps=3000, batch=65536, lr=1e-3, seed=0, params=None, log=None, every=500):
    """Adam on the weighted squared error of sigmoid(MLP(x)) against the targets, mini-batches drawn without
    replacement (seeded); returns (parameters, [(step, loss of the whole set)])."""
    torch = _torch()
    x, t, w = device_rows(x_bits, target, weight)
    rows = x.shape[0]
    batch = min(batch, rows)
    trainer = Trainer(params=params, seed=seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    total = float(w.sum().item())
    history = []

    def whole():
        return trainer.gradient(x, t, w, total)[0]

    history.append((0, whole()))
    perm, at = torch.randperm(rows, device="cuda", generator=gen), 0
    for k in range(1, steps + 1):
        if at + batch > rows:
            perm, at = torch.randperm(rows, device="cuda", generator=gen), 0
        idx = perm[at:at + batch]
        at += batch
        xb, tb, wb = x[idx].contiguous(), t[idx].contiguous(), w[idx].contiguous()
        wsum = float(wb.sum().item())  # also waits for torch's gathers
        if wsum > 0:
            trainer.step(xb, tb, wb, wsum, lr)
            trainer.synchronize()  # before torch may reuse the batch's memory
        if k % every == 0 or k == steps:
            trainer.synchronize()
            history.append((k, whole()))
            if log:
                log("  train step %d: loss %.6g" % history[-1])
    p = trainer.params()
    trainer.close()
    return p, history


# ------------------------------------------------------------------------------------------------ training pairs


def primitives(q, volume, gamma, D):
    """W (N x V) of the extensive Q (host, numpy)."""
    U = np.asarray(q) / np.asarray(volume)[:, None]
    rho = U[:, 0]
    v = U[:, 1:D + 1] / rho[:, None]
    P = (gamma - 1.0) * (U[:, D + 1] - 0.5 * rho * (v * v).sum(axis=1))
    return np.column_stack([rho, v, P])


def run_to(fv, marks, courant, log=None):
    """Steps a static solver to each time in marks (the last step of each stretch shortened to land on it); returns the
    states at the marks and the number of steps."""
    t, out, steps = 0.0, [], 0
    for mark in marks:
        while t < mark * (1 - 1e-12):
            dt = fv.time_step(courant)
            if not (dt > 0):
                raise RuntimeError("dt <= 0")
            dt = min(dt, mark - t)
            fv.step(dt)
            t += dt
            steps += 1
        out.append(fv.state())
        if log:
            log("  t = %.3f after %d steps" % (t, steps))
    return out, steps


def owners(coarse, fine):
    """The coarse cell holding each fine generator: the nearest coarse generator (periodic box)."""
    from scipy.spatial import cKDTree

    D = coarse["dim"]
    side = np.asarray(coarse["side"][:D], dtype=np.float64)
    xc = np.asarray(coarse["positions"]).reshape(-1, D) % side
    xf = np.asarray(fine["positions"]).reshape(-1, D) % side
    return cKDTree(xc, boxsize=side).query(xf)[1]


def restrict(q_fine, fine, coarse, owner):
    """The fine state on the coarse cells: conserved densities averaged over the fine cells each coarse cell holds,
    times the coarse volume (so the coarse densities are the fine averages). A coarse cell that holds no fine generator
    (a small cell of a distorted moving mesh, against a reference only twice as fine) takes the densities of the fine
    cell nearest to its generator."""
    Nc, V = coarse["cells"], q_fine.shape[1]
    vol_f = np.asarray(fine["volume"])
    mass = np.zeros((Nc, V))
    for v in range(V):
        mass[:, v] = np.bincount(owner, weights=q_fine[:, v], minlength=Nc)
    held = np.bincount(owner, weights=vol_f, minlength=Nc)
    empty = np.flatnonzero(held <= 0)
    if len(empty):
        from scipy.spatial import cKDTree

        D = coarse["dim"]
        side = np.asarray(coarse["side"][:D], dtype=np.float64)
        xf = np.asarray(fine["positions"]).reshape(-1, D) % side
        xc = np.asarray(coarse["positions"]).reshape(-1, D)[empty] % side
        near = cKDTree(xf, boxsize=side).query(xc)[1]
        mass[empty] = q_fine[near]
        held[empty] = vol_f[near]
    return mass / held[:, None] * np.asarray(coarse["volume"])[:, None]


def sub_cell_slopes(W_fine, fine, coarse, owner):
    """Per coarse cell and variable, the volume-weighted least-squares slope (with intercept) of the fine primitives
    against the fine generators' positions (periodic separations from the coarse generator): N x V x D."""
    D, Nc = coarse["dim"], coarse["cells"]
    side = np.asarray(coarse["side"][:D], dtype=np.float64)
    xc = np.asarray(coarse["positions"]).reshape(-1, D)
    xf = np.asarray(fine["positions"]).reshape(-1, D)
    d = xf - xc[owner]
    d -= side * np.round(d / side)
    w = np.asarray(fine["volume"])
    V = W_fine.shape[1]

    def s(vals):
        return np.bincount(owner, weights=vals, minlength=Nc)

    S0 = s(w)
    S1 = np.stack([s(w * d[:, a]) for a in range(D)], axis=1)
    S2 = np.stack([np.stack([s(w * d[:, a] * d[:, b]) for b in range(D)], axis=1) for a in range(D)], axis=1)
    T0 = np.stack([s(w * W_fine[:, v]) for v in range(V)], axis=1)
    T1 = np.stack([np.stack([s(w * d[:, a] * W_fine[:, v]) for v in range(V)], axis=1) for a in range(D)], axis=1)  # Nc x D x V
    C = S2 - S1[:, :, None] * S1[:, None, :] / S0[:, None, None]
    B = T1 - S1[:, :, None] * T0[:, None, :] / S0[:, None, None]
    C += np.eye(D)[None] * 1e-30
    g = np.linalg.solve(C, B)  # Nc x D x V
    return np.transpose(g, (0, 2, 1))


def pair_rows(coarse, q_coarse, g_fine, gamma, motion="static"):
    """Feature rows, targets and weights of one coarse state whose ideal slopes are g_fine (N x V x D)."""
    D, N = coarse["dim"], coarse["cells"]
    V = D + 2
    fv = hydro.FiniteVolume(coarse, gamma, limiter="learned", motion=motion, step_rule="luna" if D == 2 else "arepo")
    try:
        fv.set_state(q_coarse)
        fv.time_step(1.0)
        x, troubled = fv.limiter_features()
        g = np.empty((N, V * D))
        hydro._check(hydro.lib().luna_fv_gradients(fv._h, g.ctypes.data_as(hydro._dp)), "gradients")
    finally:
        fv.close()
    g = g.reshape(N * V, D)
    gf = g_fine.reshape(N * V, D)
    gg = (g * g).sum(axis=1)
    t = np.where(gg > 0, (gf * g).sum(axis=1) / np.where(gg > 0, gg, 1.0), 1.0)
    t = np.clip(t, 0.0, 1.0)
    f = bf16_value(x)
    w = np.clip(f[:, 0] * np.maximum(f[:, 3], f[:, 4]), 0.0, 1.0).astype(np.float64)
    w[np.repeat(troubled.astype(bool), V)] = 0.0
    w[gg == 0] = 0.0
    return x, t.astype(np.float32), w.astype(np.float32)


class Problem:
    """A benchmark problem as the training and the evaluation use it: its dimension, CFL number (LUNA rule in 2D,
    AREPO in 3D), the jittered mesh of n generators per axis and seed, and the initial state on a mesh of resolution n."""

    def __init__(self, name, dim, courant, state):
        self.name, self.dim, self.courant, self._state = name, dim, courant, state

    def mesh(self, n, seed, jitter=0.25):
        return meshlib.build(problems.jittered_lattice(n, self.dim, jitter, seed), 1.0)

    def state(self, m, n, gamma=5.0 / 3.0):
        return self._state(m, n, gamma)


# The Kelvin-Helmholtz shear layer (2D; sharp, and luna-cpp's smooth one) and the Sedov blast (3D) with its energy in a
# fixed physical radius of 1/16 (two cells at 32^3), so that every resolution starts from the same flow.
KH = Problem("kelvin-helmholtz", 2, 0.8, lambda m, n, gamma: problems.kelvin_helmholtz_state(m, gamma))
KH_SMOOTH = Problem("kelvin-helmholtz-smooth", 2, 0.8, lambda m, n, gamma: problems.kelvin_helmholtz_smooth_state(m, gamma))
SEDOV = Problem("sedov", 3, 0.3, lambda m, n, gamma: problems.sedov_state(m, gamma, radius_cells=n / 16.0, n=n)[0])


def pairs(problem, n=64, seeds=(1,), marks=(0.5, 1.0), factor=4, gamma=5.0 / 3.0, log=None):
    """Training rows of a problem: for each seed, a plain static run at factor x n per axis to each mark (and its
    initial state), restricted to the n-mesh of the same seed."""
    xs, ts, ws = [], [], []
    for seed in seeds:
        coarse = problem.mesh(n, seed)
        fine = problem.mesh(factor * n, seed + 1000)
        owner = owners(coarse, fine)
        q0 = problem.state(fine, factor * n, gamma)
        fv = hydro.FiniteVolume(fine, gamma, step_rule="luna" if problem.dim == 2 else "arepo")
        try:
            fv.set_state(q0)
            states, _ = run_to(fv, list(marks), problem.courant, log)
        finally:
            fv.close()
        for q in [q0] + states:
            W_f = primitives(q, fine["volume"], gamma, problem.dim)
            qc = restrict(q, fine, coarse, owner)
            x, t, w = pair_rows(coarse, qc, sub_cell_slopes(W_f, fine, coarse, owner), gamma)
            xs.append(x)
            ts.append(t)
            ws.append(w)
        if log:
            log("  %s seed %d: %d rows so far" % (problem.name, seed, sum(len(a) for a in ts)))
    return np.concatenate(xs), np.concatenate(ts), np.concatenate(ws)


# ------------------------------------------------------------------------------------------------ evaluation


def run_timed(m, q0, t_end, gamma, courant, limiter="face", params=None, motion="static", log=None):
    """A run to t_end; returns (final state, final mesh, GPU-side seconds of the hydro calls (the mesh producer
    excluded), steps). The timed calls are fenced, so the seconds are the GPU work and its launches."""
    D = m["dim"]
    fv = hydro.FiniteVolume(m, gamma, limiter=limiter, motion=motion, step_rule="luna" if D == 2 else "arepo")
    try:
        if params is not None:
            fv.set_limiter_network(params)
        fv.set_state(q0)
        t, steps, busy, current = 0.0, 0, 0.0, m
        said = time.perf_counter()
        while t < t_end * (1 - 1e-12):
            if log and time.perf_counter() - said > 20.0:  # progress every 20 s of wall (moving runs wait on the host)
                said = time.perf_counter()
                log("    progress: steps=%d t=%.6g" % (steps, t))
            a = time.perf_counter()
            dt = min(fv.time_step(courant), t_end - t)
            if motion == "static":
                fv.step(dt)
                fv.fallbacks()
                busy += time.perf_counter() - a
            else:
                x = fv.moving_begin(dt)
                busy += time.perf_counter() - a
                current = meshlib.build(x, np.asarray(m["side"][:D]))
                a = time.perf_counter()
                fv.moving_finish(current, dt)
                fv.fallbacks()
                busy += time.perf_counter() - a
            t += dt
            steps += 1
        return fv.state(), current, busy, steps
    finally:
        fv.close()


def density_l1(q, m, q_ref):
    """sum_i V_i |rho_i - rho_ref_i| / sum_i V_i on the mesh m (q, q_ref extensive on m)."""
    vol = np.asarray(m["volume"])
    return float(np.sum(np.abs(q[:, 0] - q_ref[:, 0])) / np.sum(vol))


def curve(problem, ns, seed, t_end, params, ref_n, gamma=5.0 / 3.0, motion="static", log=None, limiters=("plain", "learned")):
    """Error against cost: plain and learned runs at each n in ns (meshes of `seed`) against one plain static reference
    run at ref_n (another seed's mesh), restricted to each run's final mesh. Returns rows {"n", "limiter", "l1",
    "seconds", "steps"} and the reference's cost."""
    fine = problem.mesh(ref_n, seed + 1000)
    fq, fm, fine_s, fine_steps = run_timed(fine, problem.state(fine, ref_n, gamma), t_end, gamma, problem.courant, log=log)
    if log:
        log("  %s reference n=%d t=%g: %.2f GPU-s, %d steps" % (problem.name, ref_n, t_end, fine_s, fine_steps))
    rows = []
    for n in ns:
        coarse = problem.mesh(n, seed)
        q0 = problem.state(coarse, n, gamma)
        for name, limiter, p in (("plain", "face", None), ("learned", "learned", params)):
            if name not in limiters:
                continue
            q, cm, secs, steps = run_timed(coarse, q0, t_end, gamma, problem.courant, limiter, p, motion, log)
            l1 = density_l1(q, cm, restrict(fq, fm, cm, owners(cm, fm)))
            rows.append({"n": n, "limiter": name, "l1": l1, "seconds": secs, "steps": steps})
            if log:
                log("  %s %s n=%d: L1 %.5g in %.3f GPU-s (%d steps)" % (problem.name, name, n, l1, secs, steps))
    return rows, {"n": ref_n, "seconds": fine_s, "steps": fine_steps}


def speedup_at_equal_error(rows):
    """For each learned run, the plain method's GPU-seconds to reach the same L1 (log-log interpolation along the plain
    runs, which must bracket it) divided by the learned run's: > 1 means the learned limiter gets there sooner. None
    where the plain runs do not bracket the learned error."""
    plain = sorted((r["seconds"], r["l1"]) for r in rows if r["limiter"] == "plain")
    out = {}
    for r in rows:
        if r["limiter"] != "learned":
            continue
        val = None
        for (s0, e0), (s1, e1) in zip(plain, plain[1:]):
            if min(e0, e1) <= r["l1"] <= max(e0, e1) and e0 != e1:
                a = (np.log(r["l1"]) - np.log(e0)) / (np.log(e1) - np.log(e0))
                val = float(np.exp(np.log(s0) + a * (np.log(s1) - np.log(s0))) / r["seconds"])
                break
        out[r["n"]] = val
    return out


# ------------------------------------------------------------------------------------------------ command line

PROBLEMS = {"kh": KH, "kh-smooth": KH_SMOOTH, "sedov": SEDOV}


def main(argv=None):
    """python -m luna_hydro.learned train [--out FILE.npy] [--steps K] [--quick]
       python -m luna_hydro.learned curve --problem kh|kh-smooth|sedov --ns 32,64 --t-end T --ref-n R [--moving] [--params FILE.npy]

    train: a network on the shipped one's problem mix (coarse/fine pairs of the smooth KH layer at 64^2 x 4 seeds and
    128^2 x 2, the sharp one at 64^2 x 2, Sedov at 32^3 x 3, each against a plain run at 4x the resolution, states at
    t = 0, 1, 2 (KH) and 0, 0.02, 0.05 (Sedov); 8000 Adam steps, batch 65536, lr 1e-3, seed 0); --quick: one 64^2
    smooth KH seed and 500 steps. curve: plain and learned runs at each n
    against a plain static run at ref_n, with the GPU-seconds and the speed-up at equal error. Prints one JSON line."""
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description=main.__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("what", choices=["train", "curve"])
    ap.add_argument("--out", default="")
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--problem", choices=sorted(PROBLEMS), default="kh-smooth")
    ap.add_argument("--ns", default="32,64")
    ap.add_argument("--t-end", type=float, default=1.0)
    ap.add_argument("--ref-n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=99)
    ap.add_argument("--moving", action="store_true")
    ap.add_argument("--params", default="")
    ap.add_argument("--limiters", default="plain,learned")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if a.what == "train":
        t0 = time.perf_counter()
        if a.quick:
            sets = [(KH_SMOOTH, 64, (1,), (0.5, 1.0))]
            steps = min(a.steps, 500)
        else:
            sets = [(KH_SMOOTH, 64, (1, 2, 3, 4), (1.0, 2.0)), (KH_SMOOTH, 128, (5, 6), (1.0, 2.0)),
                    (KH, 64, (7, 8), (1.0, 2.0)), (SEDOV, 32, (9, 10, 11), (0.02, 0.05))]
            steps = a.steps
        xs, ts, ws = [], [], []
        for problem, n, seeds, marks in sets:
            x, t, w = pairs(problem, n, seeds, marks, log=log)
            xs.append(x)
            ts.append(t)
            ws.append(w)
        x, t, w = np.concatenate(xs), np.concatenate(ts), np.concatenate(ws)
        t1 = time.perf_counter()
        log("%d rows in %.1f s; training %d steps" % (len(t), t1 - t0, steps))
        params, history = train(x, t, w, steps=steps, log=log)
        t2 = time.perf_counter()
        if a.out:
            np.save(a.out, params)
        print(json.dumps({"rows": int(len(t)), "steps": steps, "pairs_s": t1 - t0, "train_s": t2 - t1,
                          "loss": [[int(k), float(v)] for k, v in history]}), flush=True)
        return 0
    params = np.load(a.params, allow_pickle=False) if a.params else default_params()
    rows, ref = curve(PROBLEMS[a.problem], [int(s) for s in a.ns.split(",")], a.seed, a.t_end, params, a.ref_n,
                      motion="moving" if a.moving else "static", log=log, limiters=a.limiters.split(","))
    print(json.dumps({"problem": a.problem, "t_end": a.t_end, "motion": "moving" if a.moving else "static",
                      "reference": ref, "rows": rows, "speedup_at_equal_error": speedup_at_equal_error(rows)}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
