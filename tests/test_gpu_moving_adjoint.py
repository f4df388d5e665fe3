"""The GPU's reverse of the moving step (luna_fv_moving_adjoint, order 1, fixed topology) against the oracle's (which central
differences pin, tests/test_oracle_moving_adjoint.py), on the same meshes with the mesh library's geometry derivative: 2D and
3D, HLLC and HLL, the grid steps held or free; and a uniform Q cotangent passes through unchanged (conservation)."""
import numpy as np
import pytest

import oracle
from luna_hydro import hydro, mesh, problems

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0


def _case(D, n, seed, riemann):
    rng = np.random.default_rng(seed)
    m0 = mesh.build(problems.jittered_lattice(n, D, 0.25, seed), 1.0)
    N, V = m0["cells"], D + 2
    rho = 1 + 0.2 * rng.uniform(-1, 1, N)
    P = 1 + 0.2 * rng.uniform(-1, 1, N)
    vel = 0.3 + 0.2 * rng.uniform(-1, 1, (N, D))
    vol = m0["volume"]
    q = np.zeros((N, V))
    q[:, 0] = rho * vol
    q[:, 1:1 + D] = (rho[:, None] * vel) * vol[:, None]
    q[:, V - 1] = (P / (GAMMA - 1) + 0.5 * rho * np.sum(vel ** 2, axis=1)) * vol
    fv = hydro.FiniteVolume(m0, GAMMA, riemann=riemann, motion="moving")
    fv.set_state(q)
    dt = fv.time_step(0.3)
    m1 = mesh.build(fv.moving_begin(dt), 1.0)
    fv.moving_finish(m1, dt)
    return rng, fv, m0, m1, q, dt


def _oracle_adjoint(m0, m1, q, dt, qbar, xbar, fixed, riemann):
    fvo = oracle.FiniteVolume(m0, GAMMA, moving=True, riemann=riemann)
    fvo.set_state(q)
    g0, g1 = mesh.GeometryDerivative(m0), mesh.GeometryDerivative(m1)
    return fvo.moving_adjoint(dt, qbar, xbar, m1, lambda *a: g0.adjoint(*a), lambda *a: g1.adjoint(*a), steps_fixed=fixed)


@pytest.mark.parametrize("D,n,riemann", [(2, 12, "hllc"), (2, 12, "hll"), (3, 6, "hllc"), (3, 8, "hll")])
@pytest.mark.parametrize("fixed", [True, False])
def test_moving_adjoint_matches_oracle(D, n, riemann, fixed):
    print("moving adjoint D=%d n=%d %s steps_fixed=%s" % (D, n, riemann, fixed), flush=True)
    rng, fv, m0, m1, q, dt = _case(D, n, 11, riemann)
    N, V = m0["cells"], D + 2
    qbar, xbar = rng.normal(size=(N, V)), rng.normal(size=(N, D))
    fv.set_state(q)
    qg, xg = fv.moving_adjoint(dt, m0, m1, qbar, xbar, steps_fixed=fixed)
    qo, xo = _oracle_adjoint(m0, m1, q, dt, qbar, xbar, fixed, riemann)
    for got, want, name in ((qg, qo, "qbar"), (xg, xo, "xbar")):
        scale = np.max(np.abs(want))
        err = np.max(np.abs(got - want)) / scale
        print("  %s: max |GPU - oracle| / max |oracle| = %.2e" % (name, err), flush=True)
        assert err < 1e-10, (name, err)
    # the handle is back on t^n with Q as set
    assert np.array_equal(fv.state(), q)
    fv.close()


def test_moving_adjoint_uniform_cotangent():
    rng, fv, m0, m1, q, dt = _case(3, 6, 5, "hllc")
    N, V = m0["cells"], 5
    for v in range(V):
        qbar = np.zeros((N, V))
        qbar[:, v] = 1.0
        fv.set_state(q)
        qb, xb = fv.moving_adjoint(dt, m0, m1, qbar, np.zeros((N, 3)))
        assert np.array_equal(qb, qbar) and not np.any(xb)
    fv.close()
