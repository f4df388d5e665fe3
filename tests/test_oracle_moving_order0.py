"""The oracle's order-0 step on a moving mesh (luna-cpp finite_volume.hpp:216-225: advance<0, true>() with the sweep's
point velocities, then move() off the grid, :1211-1250): a uniform flow on a lattice moving with it stays uniform to the
bit (Galilean invariance of the moving-face flux; the translated lattice has the same cells); conservation; the points
move by dt w, wrapped into the box."""
import math

import numpy as np
import pytest

import cases
import oracle
from luna_hydro import mesh, problems


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_uniform_flow_on_a_moving_lattice_stays_uniform(dim, n):
    m = cases.voronoi(n, dim, jitter=0.0)
    W = np.zeros((m["cells"], dim + 2))
    W[:, 0], W[:, dim + 1] = 1.0, 1.0
    W[:, 1:dim + 1] = [0.3, -0.2, 0.1][:dim]
    U0 = problems.conserved_rows(W, 5.0 / 3.0)
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, order=0, moving=True)
    fv.set_state(U0 * np.asarray(m["volume"])[:, None])
    for _ in range(3):
        dt = fv.time_step(0.3)
        x0 = fv.x.copy()
        x = fv.moving_begin(dt)
        moved = x0 + dt * fv.w
        assert np.array_equal(x, moved - np.floor(moved))  # unit box
        m = mesh.build(x, 1.0)
        fv.moving_finish(m, dt)
    assert np.max(np.abs(fv.q / np.asarray(m["volume"])[:, None] - U0)) == 0.0


def test_conserves():
    m = cases.voronoi(8, 3)
    q = cases.rough_state(m)
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, order=0, moving=True)
    fv.set_state(q)
    for _ in range(3):
        dt = fv.time_step(0.3)
        m = mesh.build(fv.moving_begin(dt), 1.0)
        fv.moving_finish(m, dt)
    for v in range(5):
        assert abs(math.fsum(fv.q[:, v]) - math.fsum(q[:, v])) <= 1e-14 * math.fsum(np.abs(q[:, v]))
