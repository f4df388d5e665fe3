"""Deterministic meshes and states shared by the tests (seeded; the same inputs for the oracle and the HIP path)."""
import numpy as np

from luna_hydro import mesh, problems

_CACHE = {}


def voronoi(n, dim, jitter=0.25, seed=1):
    key = (n, dim, jitter, seed)
    if key not in _CACHE:
        _CACHE[key] = mesh.build(problems.jittered_lattice(n, dim, jitter, seed), 1.0)
    return _CACHE[key]


def rough_state(m, gamma=5.0 / 3.0, seed=5, noise=0.3, blast=100.0):
    """Per cell W = (1 + noise u, 0.5 u, ..., 1 + noise u), pressure times `blast` near the box centre; Q extensive."""
    D, N = m["dim"], m["cells"]
    rng = np.random.Generator(np.random.MT19937(seed))
    W = np.empty((N, D + 2))
    W[:, 0] = 1 + noise * rng.uniform(-1, 1, N)
    W[:, 1:D + 1] = 0.5 * rng.uniform(-1, 1, (N, D))
    W[:, D + 1] = 1 + noise * rng.uniform(-1, 1, N)
    x = m["positions"]
    centre = np.linalg.norm(x - 0.5, axis=1) < 0.15
    W[centre, D + 1] *= blast
    return problems.conserved_rows(W, gamma) * m["volume"][:, None]


def uniform_state(m, W, gamma=5.0 / 3.0):
    W = np.asarray(W, dtype=np.float64)
    return problems.conserved_rows(np.tile(W, (m["cells"], 1)), gamma) * m["volume"][:, None]


def max_relative(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / scale))
