"""Pins the oracle's pointwise physics: LUNA's golden HLL/HLLC fluxes (hydro_test.cpp:47-104), closed forms, symmetry
and rotation (hydro_test.cpp:106-132), primitive/conserved round trips and floors (euler.hpp:104-172)."""
import math

import numpy as np
import pytest

import oracle
from golden_data import load

GAMMA = 5.0 / 3.0


def relative(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b))


@pytest.mark.parametrize("kind", ["hll", "hllc"])
def test_luna_golden_fluxes(kind):
    data = load("riemann")
    worst = 0.0
    for c in data[kind]:
        n = np.eye(3)[c["axis"]]
        F = oracle.flux(c["L"], c["R"], n, [0, 0, 0], data["gamma"], kind)
        worst = max(worst, relative(F, c["F"]))
    assert worst < 1e-14, worst


def exact_flux(W, n, gamma):
    U = oracle.conserved(W, gamma)
    vn = W[1] * n[0] + W[2] * n[1] + W[3] * n[2]
    return np.array([U[0] * vn, U[1] * vn + W[4] * n[0], U[2] * vn + W[4] * n[1], U[3] * vn + W[4] * n[2],
                     (U[4] + W[4]) * vn])


def rotate(v, a):
    angle = math.sqrt(a @ a)
    k = a / angle
    return v * math.cos(angle) + np.cross(k, v) * math.sin(angle) + k * (k @ v) * (1 - math.cos(angle))


@pytest.mark.parametrize("kind", ["hll", "hllc"])
def test_flux_invariants(kind):
    rng = np.random.Generator(np.random.MT19937(7))
    def unit(*shape):
        return rng.uniform(-1.0, 1.0, size=shape) if shape else rng.uniform(-1.0, 1.0)
    consistent = moving = symmetric = rotated = 0.0
    for _ in range(1000):
        n = unit(3)
        n = n / math.sqrt(n @ n)
        w = unit(3)
        axis = 3 * unit(3)
        L = np.array([1.5 + unit(), *(2 * unit(3)), 1.5 + unit()])
        R = np.array([1.5 + unit(), *(2 * unit(3)), 1.5 + unit()])
        consistent = max(consistent, relative(oracle.flux(L, L, n, [0, 0, 0], GAMMA, kind), exact_flux(L, n, GAMMA)))
        expected = exact_flux(L, n, GAMMA) - (w @ n) * oracle.conserved(L, GAMMA)
        moving = max(moving, relative(oracle.flux(L, L, n, w, GAMMA, kind), expected))
        F = oracle.flux(L, R, n, w, GAMMA, kind)
        back = -oracle.flux(R, L, -n, w, GAMMA, kind)
        symmetric = max(symmetric, relative(F, back))
        rot = lambda s: np.array([s[0], *rotate(np.asarray(s[1:4]), axis), s[4]])
        turned = oracle.flux(rot(L), rot(R), rotate(n, axis), rotate(w, axis), GAMMA, kind)
        rotated = max(rotated, relative(turned, rot(F)))
    # The reference's bounds (1e-14, 1e-13, 1e-14, 2e-13) are for its own mt19937_64 sample set; these draws differ and
    # the swap's S_M numerator rounds in another association, whose tail reaches 2.7e-14 here: bound 5e-14.
    assert consistent < 1e-14 and moving < 1e-13 and symmetric < 5e-14 and rotated < 2e-13, (
        consistent, moving, symmetric, rotated)


def test_planted_sign_error_fails_golden():
    """Teeth: a flux with a flipped energy sign must fail the golden check that pins the real one."""
    c = load("riemann")["hllc"][0]
    F = oracle.flux(c["L"], c["R"], np.eye(3)[c["axis"]], [0, 0, 0], GAMMA)
    F[4] = -F[4]
    assert relative(F, c["F"]) > 1e-3


@pytest.mark.parametrize("D", [1, 2, 3])
def test_primitive_round_trip_and_floors(D):
    rng = np.random.Generator(np.random.MT19937(3))
    for _ in range(200):
        W = np.concatenate([[rng.uniform(0.1, 3)], rng.uniform(-2, 2, D), [rng.uniform(0.1, 3)]])
        U = oracle.conserved(W, GAMMA)
        assert relative(oracle.primitive(U, GAMMA), W) < 1e-14
    # density floor: rho -> smallr; energy floor: e >= smallc^2 / (gamma (gamma - 1)) rho
    U = np.zeros(D + 2)
    U[0] = -1.0
    W = oracle.primitive(U, GAMMA)
    assert W[0] == 1e-10
    smalle = 1e-10 * 1e-10 / GAMMA / (GAMMA - 1.0)
    assert W[D + 1] == (GAMMA - 1.0) * (smalle * 1e-10)
