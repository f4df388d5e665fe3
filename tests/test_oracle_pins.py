"""Pins the oracle's multi-dimensional pieces by what the mathematics fixes (CPU only; the GPU is pinned to the oracle by
the parity tests):

- the MUSCL-Hancock half step (euler.hpp:186-214) in 2D and 3D: against the primitive-variable Euler Jacobians built
  here from the equations (Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 3.2), against the Euler
  equations themselves (complex-step derivatives of the exact conservative flux and of U(W)), and under rotations;
- the face velocity (finite_volume.hpp:1380-1397): the normal speed of the bisector plane of two moving generators at
  the face centroid, by central differences of that plane, on both sides of the 0.9 |d| limit;
- the regularised point velocity (:526-537, constants :74): each branch of its face-angle ramp, alone and in the sweep;
- the face-centroid limiter's maximality (:549-557): each variable's limited gradient is the unlimited one times
  min(1, min over non-tiny faces of the factor that face needs);
- the face states each flux pass hands the Riemann solver (:1306-1329, :1451-1465): static W - dt/2 A(W) g + g r, a
  moving step's first pass W + g (x_f^n - c^n), its second W - dt A(W) g + g (x_f^n+1 - c^n) in the lab frame;
- Galilean invariance of the moving step: a blast moving with the box equals the blast at rest (hydro_test.cpp:474-510).
"""
import math

import numpy as np
import pytest

import cases
import galilean
import oracle
from luna_hydro import mesh, problems

GAMMA = 5.0 / 3.0


def rng_of(seed):
    return np.random.Generator(np.random.MT19937(seed))


# ----------------------------------------------------------------------------------------------- the half step


def jacobian_rate(W, G, gamma):
    """sum_d A_d(W) G[:, d] with A_d the primitive-variable Jacobians of the Euler equations (Toro 3.2):
    rho_t + v.grad rho + rho div v = 0; v_t + (v.grad) v + grad p / rho = 0; p_t + v.grad p + gamma p div v = 0."""
    D = len(W) - 2
    P = D + 1
    rate = np.zeros(D + 2)
    for d in range(D):
        A = np.zeros((D + 2, D + 2))
        A[0, 0] = W[1 + d]
        A[0, 1 + d] = W[0]
        for k in range(D):
            A[1 + k, 1 + k] = W[1 + d]
        A[1 + d, P] = 1.0 / W[0]
        A[P, 1 + d] = gamma * W[P]
        A[P, P] = W[1 + d]
        rate += A @ G[:, d]
    return rate


def conserved_c(W, gamma):
    D = len(W) - 2
    v = W[1:D + 1]
    return np.concatenate([[W[0]], W[0] * v, [W[D + 1] / (gamma - 1) + 0.5 * W[0] * np.sum(v * v)]])


def flux_c(W, d, gamma):
    """The exact Euler flux along axis d (complex-safe)."""
    D = len(W) - 2
    v = W[1:D + 1]
    E = W[D + 1] / (gamma - 1) + 0.5 * W[0] * np.sum(v * v)
    mom = W[0] * v * v[d]
    mom = mom + np.where(np.arange(D) == d, W[D + 1], 0.0)
    return np.concatenate([[W[0] * v[d]], mom, [(E + W[D + 1]) * v[d]]])


def euler_rate(W, G, gamma):
    """-dW/dt of a linear field W + G x under the Euler equations: (dU/dW)^-1 sum_d dF_d/dW G[:, d], both derivatives by
    complex steps (exact to rounding)."""
    V = len(W)
    D = V - 2
    h = 1e-30
    dUdW = np.empty((V, V))
    for c in range(V):
        e = np.zeros(V, dtype=complex)
        e[c] = 1j * h
        dUdW[:, c] = conserved_c(W.astype(complex) + e, gamma).imag / h
    div = np.zeros(V)
    for d in range(D):
        div += flux_c(W.astype(complex) + 1j * h * G[:, d], d, gamma).imag / h
    return np.linalg.solve(dUdW, div)


def random_state(rng, D):
    W = np.concatenate([[rng.uniform(0.5, 2.0)], rng.uniform(-1, 1, D), [rng.uniform(0.5, 2.0)]])
    G = rng.uniform(-1, 1, (D + 2, D))
    return W, G


@pytest.mark.parametrize("D", [2, 3])
def test_half_step_equals_euler_jacobians_and_equations(D):
    rng = rng_of(21 + D)
    dt = 0.37
    for _ in range(200):
        W, G = random_state(rng, D)
        got = oracle.half_step(W, G, dt, GAMMA)
        scale = np.abs(W) + 0.5 * dt * np.abs(jacobian_rate(np.abs(W), np.abs(G), GAMMA))
        expect_j = W - 0.5 * dt * jacobian_rate(W, G, GAMMA)
        expect_e = W - 0.5 * dt * euler_rate(W, G, GAMMA)
        assert np.max(np.abs(got - expect_j) / scale) < 1e-14
        assert np.max(np.abs(got - expect_e) / scale) < 1e-12


@pytest.mark.parametrize("D", [2, 3])
def test_half_step_checks_have_teeth(D):
    """The two mistakes a restatement could make in the cross terms (dv_d/dx_k for dv_k/dx_d in the advection, or the
    divergence's diagonal read off the row) give a different rate on these inputs, by far more than the tolerance."""
    rng = rng_of(5)
    W, G = random_state(rng, D)
    good = jacobian_rate(W, G, GAMMA)
    P = D + 1
    swapped = np.zeros(D + 2)
    for d in range(D):
        swapped[0] += W[1 + d] * G[0, d] + W[0] * G[1 + d, d]
        for k in range(D):
            swapped[1 + k] += W[1 + d] * G[1 + d, k] + (G[P, d] / W[0] if k == d else 0.0)
        swapped[P] += GAMMA * W[P] * G[1 + d, d] + W[1 + d] * G[P, d]
    assert np.max(np.abs(swapped - good)) > 1e-3
    assert np.max(np.abs(euler_rate(W, G, GAMMA) - swapped)) > 1e-3


@pytest.mark.parametrize("D", [2, 3])
def test_half_step_rotation_covariance(D):
    rng = rng_of(31 + D)
    dt = 0.21
    for _ in range(100):
        W, G = random_state(rng, D)
        R = np.linalg.qr(rng.normal(size=(D, D)))[0]
        W2 = W.copy()
        W2[1:D + 1] = R @ W[1:D + 1]
        G2 = np.empty_like(G)
        G2[0] = R @ G[0]
        G2[D + 1] = R @ G[D + 1]
        G2[1:D + 1] = R @ G[1:D + 1] @ R.T  # J[k][d] = d v_k / d x_d
        a = oracle.half_step(W, G, dt, GAMMA)
        b = oracle.half_step(W2, G2, dt, GAMMA)
        scale = np.max(np.abs(a)) + 1
        assert abs(b[0] - a[0]) / scale < 1e-14 and abs(b[D + 1] - a[D + 1]) / scale < 1e-14
        assert np.max(np.abs(b[1:D + 1] - R @ a[1:D + 1])) / scale < 1e-14


# ----------------------------------------------------------------------------------------------- the face velocity


def plane_speed(d, x, wL, wR, eps=1e-6):
    """The speed along n = d / |d| at which the bisector plane of generators 0 and d, moving with wL and wR, passes the
    point x (which lies on it), by central differences of the plane's crossing of the line x + s n."""
    n = d / np.linalg.norm(d)

    def crossing(e):
        a, b = e * wL, d + e * wR
        dd, m = b - a, 0.5 * (a + b)
        return np.dot(m - x, dd) / np.dot(n, dd)

    return (crossing(eps) - crossing(-eps)) / (2 * eps)


def in_plane(rng, d, length):
    c = rng.normal(size=len(d))
    c -= d * (np.dot(c, d) / np.dot(d, d))
    return c * (length / np.linalg.norm(c))


@pytest.mark.parametrize("D", [2, 3])
def test_face_velocity_is_the_bisector_planes_speed(D):
    rng = rng_of(41 + D)
    significant = 0
    for trial in range(100):
        d = rng.uniform(-1, 1, D)
        dist = np.linalg.norm(d)
        wL, wR = rng.uniform(-2, 2, D), rng.uniform(-2, 2, D)
        n = d / dist
        for frac in (0.0, 0.3, 0.8, 0.9, 1.5, 4.0):
            c = in_plane(rng, d, frac * dist)
            xf = 0.5 * d + c
            wf = oracle.face_velocity(d, xf, wL, wR)
            mean = 0.5 * (wL + wR)
            # the correction to the mean velocity is along the normal
            corr = wf - mean
            assert np.linalg.norm(corr - n * np.dot(corr, n)) <= 1e-14 * (1 + np.linalg.norm(corr))
            # the normal component is the plane's speed at x_f, or at the point 0.9 |d| from the midpoint beyond the limit
            point = xf if frac <= 0.9 else 0.5 * d + c * (0.9 * dist / np.linalg.norm(c))
            expect = plane_speed(d, point, wL, wR)
            assert abs(np.dot(wf, n) - expect) < 1e-8 * (1 + abs(expect)), (trial, frac)
            significant += abs(np.dot(c, wL - wR)) / dist > 0.05
    assert significant > 300  # most cases have an off-centre term that matters (the plane turns)


# ----------------------------------------------------------------------------------------------- the point velocity


def ramp(angle):
    """The regularisation's fraction of the sound speed: 0 up to 0.75 x 2.25, linear to 0.5 at 2.25, then 0.5."""
    if angle <= 0.75 * 2.25:
        return 0.0
    if angle >= 2.25:
        return 0.5
    return 0.5 * (angle - 1.6875) / 0.5625


@pytest.mark.parametrize("D", [2, 3])
def test_regularised_point_velocity_branches(D):
    rng = rng_of(51 + D)
    known = {1.0: 0.0, 1.6875: 0.0, 1.96875: 0.25, 2.109375: 0.375, 2.25: 0.5, 3.0: 0.5, 40.0: 0.5}
    for angle, fraction in known.items():
        assert ramp(angle) == fraction
        for _ in range(20):
            v = rng.uniform(-1, 1, D)
            cs = rng.uniform(0.1, 3)
            row = np.concatenate([[1.0], v, [1.0, cs]])
            c = rng.uniform(-0.1, 0.1, D)
            w = oracle.regularised_point_velocity(row, c, angle)
            expect = v + c / np.linalg.norm(c) * (fraction * cs)
            assert np.max(np.abs(w - expect)) <= 4e-16 * (np.max(np.abs(expect)) + cs)
    # a cell whose centroid is its generator keeps the gas velocity
    assert np.array_equal(oracle.regularised_point_velocity(np.array([1, 0.5, -0.5, 0.25, 1, 2.0])[:D + 3], np.zeros(D), 3.0),
                          np.array([0.5, -0.5, 0.25])[:D])


def face_angle(D, A):
    return np.sqrt(A / np.pi) if D == 3 else 0.5 * A


def sep(m, i, j):
    d = m["positions"][j] - m["positions"][i]
    return d - np.round(d)


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_sweep_point_velocities_take_each_branch(dim, n):
    m = cases.voronoi(n, dim, jitter=0.35, seed=3)
    fv = oracle.FiniteVolume(m, GAMMA, moving=True)
    fv.set_state(cases.rough_state(m))
    fv.time_step(0.3)
    V = dim + 2
    d = m["positions"][m["neighbour"]] - m["positions"][m["owner"]]
    d -= np.round(d)
    A = np.linalg.norm(m["area"], axis=1)
    ang = face_angle(dim, A) / (0.5 * np.linalg.norm(d, axis=1))
    largest = np.zeros(m["cells"])
    np.maximum.at(largest, m["owner"], ang)
    np.maximum.at(largest, m["neighbour"], ang)
    fractions = np.array([ramp(a) for a in largest])
    c = m["centroid"]
    cs = fv.prims[:, V]
    expect = fv.prims[:, 1:dim + 1] + c / np.linalg.norm(c, axis=1)[:, None] * (fractions * cs)[:, None]
    assert np.max(np.abs(fv.w - expect)) <= 1e-14 * np.max(np.abs(expect))
    counts = [(fractions == 0).sum(), ((fractions > 0) & (fractions < 0.5)).sum(), (fractions == 0.5).sum()]
    assert min(counts) > 0, counts


# ----------------------------------------------------------------------------------------------- the limiter


def list_of(m, i):
    for k in range(m["start"][i], m["start"][i + 1]):
        e = m["list"][k]
        f = e if e >= 0 else ~e
        yield f, int(m["neighbour"][f] if e >= 0 else m["owner"][f]), e >= 0


def face_offset(m, f, i, owned):
    fc = m["face_centroid"][f] if owned else m["face_centroid"][f] + sep(m, i, m["owner"][f])
    return fc - m["centroid"][i]


def needed_factor(W, lo, hi, dp):
    """The factor one face needs for W + dp to stay in [lo, hi] (1 when it does)."""
    if dp > 0 and W + dp > hi:
        return (hi - W) / dp if hi > W else 0.0
    if dp < 0 and W + dp < lo:
        return (lo - W) / dp if lo < W else 0.0
    return 1.0


@pytest.mark.parametrize("dim,n,seed", [(3, 8, 5), (2, 16, 6)])
def test_limiter_is_the_largest_admissible_scaling(dim, n, seed):
    m = cases.voronoi(n, dim)
    V = dim + 2
    q = cases.rough_state(m, seed=seed)
    free = oracle.FiniteVolume(m, GAMMA, limiter="none")
    lim = oracle.FiniteVolume(m, GAMMA)
    for fv in (free, lim):
        fv.set_state(q)
        fv.time_step(0.3)
    W = lim.prims[:, :V]
    clipped = 0
    for i in range(m["cells"]):
        faces = list(list_of(m, i))
        areas = np.array([np.linalg.norm(m["area"][f]) for f, _, _ in faces])
        keep = areas >= 1e-6 * areas.sum()
        js = [j for (f, j, o), k in zip(faces, keep) if k]
        lo, hi = W[js].min(axis=0), W[js].max(axis=0)
        g0 = free.grad[i].reshape(V, dim)
        g = lim.grad[i].reshape(V, dim)
        offsets = [face_offset(m, f, i, o) for (f, j, o), k in zip(faces, keep) if k]
        for v in range(V):
            phi = min([1.0] + [needed_factor(W[i, v], lo[v], hi[v], float(g0[v] @ r)) for r in offsets])
            scale = np.max(np.abs(g0[v])) + 1e-300
            assert np.max(np.abs(g[v] - phi * g0[v])) <= 1e-13 * scale, (i, v, phi)
            clipped += phi < 1
    assert clipped > m["cells"] // 4  # the pin covers many clipped rows


# ----------------------------------------------------------------------------------------------- the face states


def lab_offset(m, f, i, owned, centroid):
    """x_f - c (generator i's frame) for the face centroid of mesh m and a centroid given relative to generator i."""
    fc = m["face_centroid"][f] if owned else m["face_centroid"][f] + sep(m, i, m["owner"][f])
    return fc - centroid


@pytest.mark.parametrize("dim,n", [(3, 6), (2, 12)])
def test_face_states_of_each_pass(dim, n):
    V = dim + 2
    m0 = cases.voronoi(n, dim)
    q = cases.rough_state(m0, noise=0.1, blast=2.0)
    # static: half step over dt then g . (x_f - c)
    fv = oracle.FiniteVolume(m0, GAMMA)
    fv.set_state(q)
    dt = fv.time_step(0.3)
    checked = [0, 0, 0]
    for f in range(m0["faces"]):
        L, R, nrm, wf, mag, fb = fv.face_sides(f, dt, -1)
        if fb:
            continue
        o, j = int(m0["owner"][f]), int(m0["neighbour"][f])
        assert np.allclose(nrm * mag, m0["area"][f], rtol=0, atol=1e-15 * mag) and not np.any(wf)
        for side, i, owned in ((L, o, True), (R, j, False)):
            Wi, g = fv.prims[i, :V], fv.grad[i].reshape(V, dim)
            expect = Wi - 0.5 * dt * jacobian_rate(Wi, g, GAMMA) + g @ lab_offset(m0, f, i, owned, m0["centroid"][i])
            assert np.max(np.abs(side - expect) / (np.abs(Wi) + 1e-300)) < 1e-13
        checked[0] += 1
    # moving: the first pass on the mesh of t^n, the second on the mesh of t^n+1
    fv = oracle.FiniteVolume(m0, GAMMA, moving=True)
    fv.set_state(q)
    dt = fv.time_step(0.3)
    x1 = fv.moving_begin(dt)
    unit = m0["step"] / dt
    W, G, steps = fv.prims[:, :V].copy(), fv.grad.copy(), fv.steps.copy()
    wgrid = steps * unit
    for f in range(m0["faces"]):
        L, R, nrm, wf, mag, fb = fv.face_sides(f, 0.0, 0, unit)
        if fb:
            continue
        o, j = int(m0["owner"][f]), int(m0["neighbour"][f])
        assert np.array_equal(wf, oracle.face_velocity(sep(m0, o, j), m0["face_centroid"][f], wgrid[o], wgrid[j]))
        for side, i, owned in ((L, o, True), (R, j, False)):
            g = G[i].reshape(V, dim)
            expect = W[i] + g @ lab_offset(m0, f, i, owned, m0["centroid"][i])
            assert np.max(np.abs(side - expect) / (np.abs(W[i]) + 1e-300)) < 1e-13
        checked[1] += 1
    m1 = mesh.build(x1, 1.0)
    fv.set_mesh(m1)
    moved = steps * m0["step"]  # x^n+1 - x^n, exactly
    assert np.max(np.abs(moved)) > 0
    for f in range(m1["faces"]):
        L, R, nrm, wf, mag, fb = fv.face_sides(f, dt, 1, unit)
        if fb:
            continue
        o, j = int(m1["owner"][f]), int(m1["neighbour"][f])
        assert np.array_equal(wf, oracle.face_velocity(sep(m1, o, j), m1["face_centroid"][f], wgrid[o], wgrid[j]))
        for side, i, owned in ((L, o, True), (R, j, False)):
            g = G[i].reshape(V, dim)
            # the face centroid of t^n+1 from the centroid of t^n, in the frame of generator i at t^n
            r = lab_offset(m1, f, i, owned, m0["centroid"][i]) + moved[i]
            expect = W[i] - dt * jacobian_rate(W[i], g, GAMMA) + g @ r
            assert np.max(np.abs(side - expect) / (np.abs(W[i]) + 1e-300)) < 1e-13
        checked[2] += 1
    assert min(checked) > m0["faces"] // 2, checked


# ----------------------------------------------------------------------------------------------- Galilean invariance


class _OracleRun:
    def __init__(self, m, moving):
        self.fv = oracle.FiniteVolume(m, GAMMA, moving=moving)

    def set_state(self, q):
        self.fv.set_state(q)

    def time_step(self, c):
        return self.fv.time_step(c)

    def step(self, dt):
        self.fv.step(dt)

    def moving_begin(self, dt):
        return self.fv.moving_begin(dt)

    def moving_finish(self, m, dt):
        self.fv.moving_finish(m, dt)

    def state(self):
        return self.fv.q.copy()


def test_galilean_blast_on_the_oracle():
    r = galilean.runs(_OracleRun)
    print("Galilean blast (oracle):", {k: v for k, v in r.items() if k != "states"})
    assert r["dt"] <= r["cfl"]
    # the reference's bounds for the double policy (hydro_test.cpp:683): cells 1e-6, moving profile 1e-10, and the static
    # mesh's profile error over 1e3 times the moving one's
    assert r["cells"] < 1e-6, r
    assert r["moving"] < 1e-10 and r["static"] > 1e3 * r["moving"], r


# ----------------------------------------------------------------------------------------------- the cell radius


def test_cbrt_is_correctly_rounded():
    """The radius' cube root (finite_volume.hpp:241-246), correctly rounded on both sides: equal to the cube root computed
    with 60 significant digits (Python's decimal module) and rounded to the nearest double, on volumes across the range a
    mesh gives and at exact cubes; the C library's cbrt misses it on some of them."""
    import ctypes
    from decimal import Decimal, getcontext

    getcontext().prec = 60
    fn = oracle.lib().orc_cbrt_rn
    fn.restype = ctypes.c_double
    fn.argtypes = [ctypes.c_double]
    rng = rng_of(71)
    values = np.concatenate([10.0 ** rng.uniform(-12, 2, 4000), rng.uniform(0.5, 1.0, 2000) * 2.0 ** -21,
                             np.array([1.0, 8.0, 27.0, 0.125, 2.0 ** -30, 3.0 * 2.0 ** -21 / (4 * np.pi)])])
    third = Decimal(1) / Decimal(3)
    library_misses = 0
    for v in values:
        exact = float(Decimal(float(v)) ** third)
        got = fn(float(v))
        assert got == exact, (v, got, exact)
        library_misses += np.cbrt(v) != exact
    assert fn(0.0) == 0.0 and fn(8.0) == 2.0
    print("cbrt_rn: %d values correctly rounded; numpy's cbrt misses %d of them" % (len(values), library_misses))
