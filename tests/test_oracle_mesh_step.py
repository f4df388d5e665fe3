"""Pins the oracle's 2D/3D step (finite_volume.hpp) by what the mathematics fixes: the unlimited area-weighted fit with an
intercept reproduces a linear field and equals a direct weighted least-squares solve (numpy lstsq); the face-centroid
limiter keeps every extrapolated face value within the range of the cell and its neighbours; a uniform state stays uniform; mass, momentum
and energy are conserved; the time step equals its closed form at rest; a moving mesh conserves."""
import math

import numpy as np
import pytest

import cases
import oracle
from luna_hydro import mesh, problems

GAMMA = 5.0 / 3.0


def neighbours(m, i):
    out = []
    for k in range(m["start"][i], m["start"][i + 1]):
        e = m["list"][k]
        f = e if e >= 0 else ~e
        j = m["neighbour"][f] if e >= 0 else m["owner"][f]
        out.append((f, int(j), e >= 0))
    return out


def sep(m, i, j):
    d = m["positions"][j] - m["positions"][i]
    return d - np.round(d)


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_unlimited_fit_reproduces_linear_field_and_lstsq(dim, n):
    m = cases.voronoi(n, dim)
    N = m["cells"]
    rng = np.random.Generator(np.random.MT19937(11))
    G = rng.uniform(-1, 1, (dim + 2, dim))
    fv = oracle.FiniteVolume(m, GAMMA, limiter="none")
    # a random state, then per cell compare the fit with lstsq; and a linear field on the unwrapped neighbourhood
    q = cases.rough_state(m)
    fv.set_state(q)
    fv.time_step(0.3)
    W = fv.prims[:, :dim + 2]
    for i in range(0, N, max(1, N // 40)):
        rows, rhs, wts = [], [], []
        for f, j, _ in neighbours(m, i):
            X = sep(m, i, j) + m["centroid"][j] - m["centroid"][i]
            rows.append(np.concatenate([[1.0], X]))
            rhs.append(W[j] - W[i])
            wts.append(np.linalg.norm(m["area"][f]))
        A = np.array(rows) * np.sqrt(wts)[:, None]
        B = np.array(rhs) * np.sqrt(wts)[:, None]
        sol = np.linalg.lstsq(A, B, rcond=None)[0]
        g = fv.grad[i].reshape(dim + 2, dim)
        assert np.abs(g - sol[1:].T).max() <= 1e-9 * (1 + np.abs(sol).max())
    # linear W in the centroid positions around cell i: the fit returns G (with its intercept 0)
    for i in range(0, N, max(1, N // 10)):
        grads = np.zeros((dim + 2, dim))
        rows, rhs, wts = [], [], []
        for f, j, _ in neighbours(m, i):
            X = sep(m, i, j) + m["centroid"][j] - m["centroid"][i]
            rows.append(np.concatenate([[1.0], X]))
            rhs.append(G @ X)
            wts.append(np.linalg.norm(m["area"][f]))
        sol = np.linalg.lstsq(np.array(rows) * np.sqrt(wts)[:, None], np.array(rhs) * np.sqrt(wts)[:, None], rcond=None)[0]
        assert np.abs(sol[1:].T - G).max() < 1e-10 and np.abs(sol[0]).max() < 1e-12


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_limiter_keeps_face_values_in_range(dim, n):
    m = cases.voronoi(n, dim)
    fv = oracle.FiniteVolume(m, GAMMA)
    fv.set_state(cases.rough_state(m))
    fv.time_step(0.3)
    W = fv.prims[:, :dim + 2]
    worst = 0.0
    for i in range(m["cells"]):
        nb = neighbours(m, i)
        surface = sum(np.linalg.norm(m["area"][f]) for f, _, _ in nb)
        js = [j for _, j, _ in nb]
        # the limiter scales toward W_i, so the range includes the cell's own value (a local extremum keeps its sign)
        lo, hi = np.minimum(W[js].min(axis=0), W[i]), np.maximum(W[js].max(axis=0), W[i])
        g = fv.grad[i].reshape(dim + 2, dim)
        for f, j, owned in nb:
            if np.linalg.norm(m["area"][f]) < 1e-6 * surface:
                continue
            fc = m["face_centroid"][f] if owned else m["face_centroid"][f] + sep(m, i, m["owner"][f])
            val = W[i] + g @ (fc - m["centroid"][i])
            scale = np.maximum(np.abs(W[i]), 1e-300)
            worst = max(worst, float(np.max((val - hi) / scale)), float(np.max((lo - val) / scale)))
    assert worst < 1e-13, worst


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_uniform_state_stays_uniform(dim, n):
    m = cases.voronoi(n, dim)
    W = np.array([1.3] + [0.7, -0.4, 0.2][:dim] + [0.9])
    fv = oracle.FiniteVolume(m, GAMMA)
    fv.set_state(cases.uniform_state(m, W))
    for _ in range(3):
        fv.step(fv.time_step(0.3))
    Wn = fv.prims[:, :dim + 2]
    fv.time_step(0.3)
    Wn = fv.prims[:, :dim + 2]
    assert np.max(np.abs(Wn - W) / np.abs(W)) < 1e-13


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_conservation(dim, n):
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    fv = oracle.FiniteVolume(m, GAMMA)
    fv.set_state(q)
    before = [math.fsum(q[:, v]) for v in range(dim + 2)]
    for _ in range(10):
        fv.step(fv.time_step(0.3))
    after = [math.fsum(fv.q[:, v]) for v in range(dim + 2)]
    scale = math.sqrt(2 * before[0] * before[-1])
    assert abs(after[0] - before[0]) / before[0] < 1e-14
    assert abs(after[-1] - before[-1]) / before[-1] < 1e-14
    for v in range(1, dim + 1):
        assert abs(after[v] - before[v]) / scale < 1e-14


def test_planted_owner_bit_flip_breaks_conservation():
    """Teeth: one list entry with its owner bit flipped must break conservation (the test above would fail)."""
    m = dict(cases.voronoi(8, 3))
    lst = m["list"].copy()
    lst[5] = ~lst[5]
    m["list"] = lst
    q = cases.rough_state(m)
    fv = oracle.FiniteVolume(m, GAMMA)
    fv.set_state(q)
    before = math.fsum(q[:, 0])
    fv.step(fv.time_step(0.3))
    assert abs(math.fsum(fv.q[:, 0]) - before) / before > 1e-10


def test_time_step_closed_form_at_rest():
    m = cases.voronoi(8, 3)
    rho, p, courant = 1.2, 0.8, 0.3
    fv = oracle.FiniteVolume(m, GAMMA)
    fv.set_state(cases.uniform_state(m, [rho, 0, 0, 0, p]))
    dt = fv.time_step(courant)
    c = math.sqrt(GAMMA * p / rho)
    r = np.cbrt(m["volume"] * (3.0 / (4.0 * math.pi)))
    d = np.linalg.norm(np.array([sep(m, o, j) for o, j in zip(m["owner"], m["neighbour"])]), axis=1)
    expect = min(np.min(courant * (r / c)), np.min(courant * ((d + r[m["owner"]]) / (2 * c))),
                 np.min(courant * ((d + r[m["neighbour"]]) / (2 * c))))
    assert abs(dt - expect) <= 1e-15 * expect


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_moving_mesh_conserves_and_moves_points(dim, n):
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    fv = oracle.FiniteVolume(m, GAMMA, moving=True)
    fv.set_state(q)
    before = [math.fsum(q[:, v]) for v in range(dim + 2)]
    x0 = m["positions"].copy()
    for _ in range(3):
        dt = fv.time_step(0.3)
        x = fv.moving_begin(dt)
        m = mesh.build(x, 1.0)
        assert np.array_equal(m["positions"], x)  # on the grid already
        fv.moving_finish(m, dt)
    after = [math.fsum(fv.q[:, v]) for v in range(dim + 2)]
    assert abs(after[0] - before[0]) / before[0] < 1e-14
    assert abs(after[-1] - before[-1]) / before[-1] < 1e-14
    moved = x - x0
    moved -= np.round(moved)
    assert np.abs(moved).max() > 0


def test_moving_rigid_translation_stays_uniform():
    """Bulk flow whose move over dt is a whole number of grid steps: the mesh translates, the state stays uniform
    (hydro_test.cpp:335-360's case). Jitter 0.1 keeps every cell below the regularisation's face angle."""
    m = cases.voronoi(8, 3, jitter=0.1, seed=2)
    v = np.array([1.0, 0.5, 0.25])
    W = np.array([1.1, *v, 0.7])
    fv = oracle.FiniteVolume(m, GAMMA, moving=True)
    fv.set_state(cases.uniform_state(m, W))
    dt = 2.0 ** -12
    for _ in range(3):
        fv.time_step(0.3)
        assert np.array_equal(fv.w, np.tile(v, (m["cells"], 1)))
        x = fv.moving_begin(dt)
        m = mesh.build(x, 1.0)
        fv.moving_finish(m, dt)
    Wn = fv.prims[:, :5]
    assert np.max(np.abs(Wn - W) / np.abs(W)) < 1e-13
