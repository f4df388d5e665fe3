"""The oracle's LaneSweep (finite_volume.hpp:880-1037: L lanes per cell, the list entries dealt round-robin, the lanes' sums
combined by a fixed tree, the limiter by the smallest factor over the faces). What pins it:
  - the unlimited fit equals numpy's weighted least squares with intercept (the lanes only reorder the sums), and the
    one-lane sweep's to rounding;
  - the limited gradient is the unlimited one times min(1, min over non-tiny faces of the factor that face needs, each
    from the unlimited gradient), recomputed here from the mesh arrays;
  - the combining tree is order-sensitive only in rounding: the gradients of 2, 4 and 8 lanes and the one-lane sweep's
    agree to 1e-13 of the row's scale, and the time step is the same (a minimum, exact in any order);
  - a static step keeps a uniform state uniform; static and moving steps conserve Q."""
import numpy as np
import pytest

import cases
import oracle
from test_oracle_pins import face_offset, list_of, needed_factor

GAMMA = 5.0 / 3.0


def weighted_fit(m, W, i):
    """numpy's least-squares gradient of cell i: weights |A|, rows (1, x_j + c_j - x_i - c_i) -> dW."""
    D = m["dim"]
    X, Y, w = [], [], []
    for f, j, o in list_of(m, i):
        d = np.asarray(m["positions"][j]) - np.asarray(m["positions"][i])
        d -= np.round(d)
        X.append(np.concatenate([[1.0], d + m["centroid"][j] - m["centroid"][i]]))
        Y.append(W[j] - W[i])
        w.append(np.linalg.norm(m["area"][f]))
    X, Y, s = np.array(X), np.array(Y), np.sqrt(np.array(w))[:, None]
    sol = np.linalg.lstsq(X * s, Y * s, rcond=None)[0]
    return sol[1:1 + D].T


@pytest.mark.parametrize("dim,n", [(3, 6), (2, 12)])
@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_lane_sweep_unlimited_fit_is_weighted_least_squares(dim, n, lanes):
    m = cases.voronoi(n, dim)
    V = dim + 2
    q = cases.rough_state(m, seed=3)
    one = oracle.FiniteVolume(m, GAMMA, limiter="none")
    fv = oracle.FiniteVolume(m, GAMMA, limiter="none", sweep_lanes=lanes)
    for f in (one, fv):
        f.set_state(q)
        f.time_step(0.3)
    W = fv.prims[:, :V]
    worst = 0.0
    for i in range(m["cells"]):
        g = fv.grad[i].reshape(V, dim)
        want = weighted_fit(m, W, i)
        worst = max(worst, float(np.max(np.abs(g - want)) / (np.max(np.abs(want)) + 1e-300)))
    assert worst < 1e-9, worst
    scale = np.max(np.abs(one.grad), axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(fv.grad - one.grad) / scale) < 1e-13


@pytest.mark.parametrize("dim,n,seed", [(3, 8, 5), (2, 16, 6)])
@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_lane_sweep_limiter_is_the_smallest_factor(dim, n, seed, lanes):
    m = cases.voronoi(n, dim)
    V = dim + 2
    q = cases.rough_state(m, seed=seed)
    free = oracle.FiniteVolume(m, GAMMA, limiter="none", sweep_lanes=lanes)
    lim = oracle.FiniteVolume(m, GAMMA, sweep_lanes=lanes)
    one = oracle.FiniteVolume(m, GAMMA)
    for fv in (free, lim, one):
        fv.set_state(q)
    dts = [fv.time_step(0.3) for fv in (free, lim, one)]
    assert dts[1] == dts[2]  # the same set of terms: the minimum is exact in any order
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
    assert clipped > m["cells"] // 4
    # the product of the one-lane sweep's successive clips equals the smallest factor up to rounding
    scale = np.max(np.abs(one.grad), axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(lim.grad - one.grad) / scale) < 1e-12


@pytest.mark.parametrize("moving", [False, True])
@pytest.mark.parametrize("dim,n", [(3, 6), (2, 12)])
def test_lane_sweep_step_uniform_and_conservative(dim, n, moving):
    m = cases.voronoi(n, dim)
    V = dim + 2
    W = np.array([1.3] + [0.2] * dim + [0.7])
    qu = cases.uniform_state(m, W)
    fv = oracle.FiniteVolume(m, GAMMA, sweep_lanes=4, moving=moving)
    fv.set_state(qu)
    dt = fv.time_step(0.3)
    if moving:
        from luna_hydro import mesh

        fv.moving_finish(mesh.build(fv.moving_begin(dt), 1.0), dt)
    else:
        fv.step(dt)
    if not moving:  # (moving: the regularisation moves the points off the flow, so only conservation is pinned)
        assert np.allclose(fv.q / np.asarray(m["volume"])[:, None], W_to_U(W), rtol=1e-13, atol=1e-14)
    q = cases.rough_state(m, seed=9)
    fv = oracle.FiniteVolume(m, GAMMA, sweep_lanes=4, moving=moving)
    fv.set_state(q)
    dt = fv.time_step(0.3)
    if moving:
        from luna_hydro import mesh

        fv.moving_finish(mesh.build(fv.moving_begin(dt), 1.0), dt)
    else:
        fv.step(dt)
    for v in range(V):
        assert abs(fv.q[:, v].sum() - q[:, v].sum()) <= 1e-13 * np.abs(q[:, v]).sum()


def W_to_U(W):
    D = len(W) - 2
    v = W[1:D + 1]
    return np.concatenate([[W[0]], W[0] * v, [W[D + 1] / (GAMMA - 1) + 0.5 * W[0] * np.sum(v * v)]])
