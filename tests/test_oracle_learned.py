"""The learned limiter's oracle (CPU): the features the sweep records, the admissibility of any fraction in [0, 1], the
BF16 rounding, and the network's forward and backward pass (oracle/mlp.py). The limiter has no counterpart in luna-cpp,
so these are pinned by what the mathematics fixes: finite differences of the unrounded network, the identity
phi x unlimited = limited, the face-centroid bound, IEEE rounding's known answers, and a planted mistake that the
checks must catch."""
import ctypes

import numpy as np
import pytest

import cases
import mlp
import oracle
from luna_hydro import problems

GAMMA = 5.0 / 3.0


def kh_state(m):
    return problems.kelvin_helmholtz_state(m, GAMMA)


def swept(m, q, limiter, **kw):
    fv = oracle.FiniteVolume(m, GAMMA, limiter=limiter, **kw)
    fv.set_state(q)
    fv.time_step(0.3)
    return fv


def test_bf16_known_answers():
    vals = np.array([1.0, -2.0, 0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, 65504.0, np.inf, np.nan])
    bits = mlp.bf16_bits(vals)
    # 1 = 0x3F80 and the ulp at 1 is 2^-7; a tie goes to the even mantissa (down for 1 + 2^-8, up for 1 + 2^-7 + 2^-8);
    # 1 + 0.75 ulp rounds up
    assert list(bits[:8]) == [0x3F80, 0xC000, 0x0000, 0x3F80, 0x3F81, 0x3F82, 0x4780, 0x7F80]
    assert bits[8] & 0x7FC0 == 0x7FC0  # quiet NaN
    assert np.array_equal(mlp.bf16_value(bits[:3]), [1.0, -2.0, 0.0])


def test_bf16_c_and_numpy_agree():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 1, 2000) * 10.0 ** rng.integers(-6, 6, 2000),
                        (1 + (np.arange(256) + 0.5) * 2.0 ** -7)])  # exact ties
    f = oracle.lib().orc_bf16_bits
    f.restype = ctypes.c_ushort
    c = np.array([f(ctypes.c_double(v)) for v in x], dtype=np.uint16)
    assert np.array_equal(c, mlp.bf16_bits(x))


@pytest.mark.parametrize("dim,n", [(2, 12), (3, 5)])
def test_learned_sweep_keeps_face_gradients(dim, n):
    """Recording the features leaves the face-centroid limiter's gradients bit for bit."""
    m = cases.voronoi(n, dim)
    q = kh_state(m) if dim == 2 else cases.rough_state(m)
    a, b = swept(m, q, "face"), swept(m, q, "learned")
    assert np.array_equal(a.grad, b.grad)


@pytest.mark.parametrize("dim,n", [(2, 12), (3, 5)])
def test_features_phi_and_flags(dim, n):
    """Feature 0 is the factor that maps the unlimited gradient onto the limited one; flags 13/14 mark density and
    pressure; clip fractions lie in [0, 1]; every feature is within the clamp."""
    m = cases.voronoi(n, dim)
    q = kh_state(m) if dim == 2 else cases.rough_state(m)
    lim, free = swept(m, q, "learned"), swept(m, q, "none")
    N, V = m["cells"], dim + 2
    f = mlp.bf16_value(lim.features).reshape(N, V, 16)
    phi = f[:, :, 0]
    assert np.all((phi >= 0) & (phi <= 1))
    g, g0 = lim.grad.reshape(N, V, dim), free.grad.reshape(N, V, dim)
    # phi is rounded to BF16 (8 bits): limited = phi x unlimited to that precision
    scale = np.abs(g0).max(axis=2) + 1e-300
    err = np.abs(g - phi[:, :, None] * g0).max(axis=2) / scale
    assert err.max() < 2 ** -8
    assert np.any(phi < 1) and np.any(phi == 1)
    assert np.all(f[:, 0, 13] == 1) and np.all(f[:, 1:, 13] == 0)
    assert np.all(f[:, V - 1, 14] == 1) and np.all(f[:, :V - 1, 14] == 0)
    assert np.all((f[:, :, 6] >= 0) & (f[:, :, 6] <= 1))
    assert np.all(np.abs(f) <= 8)
    # features 7-12 are per cell: the same in every variable's row; 7 is the smallest phi (rounding is monotonic)
    for k in range(7, 13):
        assert np.all(f[:, :, k] == f[:, :1, k])
    assert np.array_equal(f[:, 0, 7], phi.min(axis=1))


def test_troubled_cells_at_the_blast():
    """The shock indicator fires where a strong compression meets a pressure jump: the Sedov blast after a few steps
    (its edge), and never in a uniform state."""
    m = cases.voronoi(10, 3)
    q, _ = problems.sedov_state(m, n=10)
    fv = oracle.FiniteVolume(m, GAMMA, limiter="learned")
    fv.set_state(q)
    for _ in range(6):
        fv.step(fv.time_step(0.3))
    fv.time_step(0.3)
    assert 0 < fv.troubled.sum() < m["cells"]
    calm = swept(m, cases.uniform_state(m, [1.0, 0.1, 0.0, 0.0, 1.0]), "learned")
    assert calm.troubled.sum() == 0


@pytest.mark.parametrize("dim,n", [(2, 12), (3, 5)])
def test_any_fraction_stays_within_the_face_bound(dim, n):
    """With fractions in [0, 1] every face-centroid value W_i + f g.(x_f - c_i) stays within the range of the cell and
    its neighbours (the face-centroid limiter's bound), checked face by face on the oracle's geometry (faces under 1e-6
    of the cell's surface, which the limiter skips, excepted)."""
    m = cases.voronoi(n, dim)
    q = kh_state(m) if dim == 2 else cases.rough_state(m)
    fv = swept(m, q, "learned")
    N, V = m["cells"], dim + 2
    rng = np.random.default_rng(2)
    fv.apply_fractions(rng.uniform(0, 1, N * V).astype(np.float32))
    g = fv.grad.reshape(N, V, dim)
    W = fv.prims[:, :V]
    x = np.asarray(m["positions"]).reshape(N, dim)
    fc = np.asarray(m["face_centroid"]).reshape(-1, dim)
    c = np.asarray(m["centroid"]).reshape(N, dim)
    owner, nb = np.asarray(m["owner"]), np.asarray(m["neighbour"])
    lo, hi = np.full((N, V), np.inf), np.full((N, V), -np.inf)
    np.minimum.at(lo, owner, W[nb])
    np.minimum.at(lo, nb, W[owner])
    np.maximum.at(hi, owner, W[nb])
    np.maximum.at(hi, nb, W[owner])
    lo, hi = np.minimum(lo, W), np.maximum(hi, W)
    mag = np.linalg.norm(np.asarray(m["area"]).reshape(-1, dim), axis=1)
    surface = np.bincount(owner, mag, N) + np.bincount(nb, mag, N)
    worst = 0.0
    for side, cell, other in ((0, owner, nb), (1, nb, owner)):
        keep = mag >= 1e-6 * surface[cell]
        rel = fc.copy()
        if side == 1:  # the face centroid is stored relative to the owner: move it to this cell's generator
            d = x[other] - x[cell]
            d -= np.round(d)
            rel = rel + d
        off = rel - c[cell]
        val = W[cell] + np.einsum("fvd,fd->fv", g[cell], off)
        span = hi[cell] - lo[cell] + 1e-300
        over = np.maximum((val - hi[cell]) / span, (lo[cell] - val) / span)[keep]
        worst = max(worst, float(over.max()))
    assert worst < 1e-12


def test_mlp_gradient_matches_finite_differences():
    """The backward pass of the unrounded network against central differences of its loss; the rounded (device) one
    against the unrounded within BF16's precision."""
    rng = np.random.default_rng(0)
    p = mlp.init_params(1).astype(np.float64)
    x = mlp.bf16_bits(rng.normal(0, 1, (64, 16)))
    t, w = rng.uniform(0, 1, 64), rng.uniform(0.5, 1, 64)
    loss, g = mlp.gradient(p, x, t, w, w.sum(), rounding=False)
    h = 1e-6
    for k in rng.choice(mlp.PARAMS, 60, replace=False):
        e = np.zeros(mlp.PARAMS)
        e[k] = h
        fd = (mlp.gradient(p + e, x, t, w, w.sum(), rounding=False)[0] - mlp.gradient(p - e, x, t, w, w.sum(), rounding=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * (abs(g[k]) + 1e-6), k
    loss_r, g_r = mlp.gradient(p, x, t, w, w.sum())
    assert abs(loss_r - loss) < 1e-2 * loss
    assert np.linalg.norm(g_r - g) < 2e-2 * np.linalg.norm(g)


def test_planted_transpose_is_caught():
    """A backward pass with W2 transposed (a plausible slip) fails the finite-difference check."""
    rng = np.random.default_rng(1)
    p = mlp.init_params(2).astype(np.float64)
    x = mlp.bf16_bits(rng.normal(0, 1, (32, 16)))
    t, w = rng.uniform(0, 1, 32), np.ones(32)
    W1, b1, W2, b2, w3, b3 = mlp.unpack(p, rounding=False)
    f = mlp.forward(p, x, rounding=False)
    sig = f["fraction"]
    dz = 2.0 * w * (sig - t) * sig * (1 - sig) / w.sum()
    dh2 = np.where(f["pre2"] > 0, dz[:, None] * w3[None, :], 0.0)
    dh1_wrong = np.where(f["pre1"] > 0, dh2 @ W2.T, 0.0)
    wrong = (dh1_wrong.T @ f["x"]).ravel()
    _, right = mlp.gradient(p, x, t, w, w.sum(), rounding=False)
    assert np.linalg.norm(wrong - right[mlp.P_W1:mlp.P_B1]) > 0.1 * np.linalg.norm(right[mlp.P_W1:mlp.P_B1])


def test_forward_fraction_in_unit_interval():
    rng = np.random.default_rng(4)
    p = mlp.init_params(5)
    x = mlp.bf16_bits(rng.normal(0, 3, (500, 16)))
    f = mlp.forward(p, x)["fraction"]
    assert np.all((f > 0) & (f < 1))


def test_shipped_network_file():
    """The shipped parameters load without executing anything from the file, in the documented layout."""
    from luna_hydro import learned

    p = learned.default_params()
    assert p.dtype == np.float32 and p.shape == (mlp.PARAMS,) and np.all(np.isfinite(p))
    f = mlp.forward(p, mlp.bf16_bits(np.zeros((4, 16))))["fraction"]
    assert np.all((f > 0) & (f < 1))


def test_restrict_fills_a_coarse_cell_without_fine_generators():
    """learned.restrict: volume-weighted averages of the fine cells each coarse cell holds; a coarse cell that holds none
    takes the densities of the fine cell nearest to its generator (periodic box)."""
    from luna_hydro import learned

    coarse = {"dim": 2, "cells": 3, "side": [1.0, 1.0, 1.0], "positions": np.array([[0.1, 0.1], [0.6, 0.6], [0.95, 0.1]]),
              "volume": np.array([0.5, 0.4, 0.1])}
    fine = {"dim": 2, "cells": 4, "side": [1.0, 1.0, 1.0],
            "positions": np.array([[0.1, 0.2], [0.2, 0.1], [0.6, 0.5], [0.7, 0.7]]), "volume": np.array([0.25, 0.25, 0.25, 0.25])}
    rho = np.array([1.0, 3.0, 5.0, 7.0])
    q_fine = np.column_stack([rho, 2 * rho]) * fine["volume"][:, None]
    owner = np.array([0, 0, 1, 1])  # coarse cell 2 holds no fine generator
    q = learned.restrict(q_fine, fine, coarse, owner)
    dens = q / coarse["volume"][:, None]
    assert np.allclose(dens[0], [2.0, 4.0]) and np.allclose(dens[1], [6.0, 12.0])
    # across the periodic x boundary (0.1, 0.2) is 0.18 from (0.95, 0.1) and (0.2, 0.1) 0.25: rho 1
    assert np.allclose(dens[2], [1.0, 2.0])
