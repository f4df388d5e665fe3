"""luna_fv_step_adjoint_noalloc (the static adjoint with the library's scratch array for its work rows, no allocation)
against luna_fv_step_adjoint on the same state: equal to the bit, on two calls in a row, the second on a smaller mesh; a
mesh beyond the scratch is refused."""
import numpy as np
import pytest
import torch

import cases
from luna_hydro import hydro

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0


def _pair(dim, n, seed):
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m, seed=seed)
    fv = hydro.FiniteVolume(m, GAMMA)
    fv.set_state(q)
    rng = np.random.Generator(np.random.MT19937(seed))
    qbar = rng.normal(size=(m["cells"], dim + 2))
    return fv, q, qbar


def test_noalloc_adjoint_equals_the_allocating_one_twice():
    print("no-alloc adjoint: 3D 12^3, then 3D 8^3, then 2D 40^2", flush=True)
    for dim, n, seed in ((3, 12, 1), (3, 8, 2), (2, 40, 3)):
        fv, q, qbar = _pair(dim, n, seed)
        try:
            dt = fv.time_step(0.3)
            want = fv.step_adjoint(dt, torch.as_tensor(qbar).cuda()).cpu().numpy()
            fv.set_state(q)
            got = fv.step_adjoint_noalloc(dt, torch.as_tensor(qbar).cuda()).cpu().numpy()
            print("  dim=%d n=%d: bits equal %s" % (dim, n, np.array_equal(got, want)), flush=True)
            assert np.array_equal(got, want)
            assert np.array_equal(fv.state(), q)
        finally:
            fv.close()


def test_noalloc_adjoint_refuses_a_mesh_beyond_its_scratch():
    m = cases.voronoi(48, 3)  # 110,592 cells: beyond 64 MiB of work rows
    fv = hydro.FiniteVolume(m, GAMMA)
    try:
        fv.set_state(cases.rough_state(m))
        with pytest.raises(ValueError, match="scratch"):
            fv.step_adjoint_noalloc(1e-4, torch.zeros((m["cells"], 5), dtype=torch.float64).cuda())
    finally:
        fv.close()
