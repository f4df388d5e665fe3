"""Initial conditions and point sets of the benchmark problems (host side, numpy).

1D problems restate luna-cpp's src/hydro/problems_1d.hpp table (:209-284; LUNA's setups and the ones it lacks): the
Toro, Einfeldt, LA-14379, LeBlanc, pure-shock tubes, Noh, Shu-Osher, Titarev-Toro, Woodward-Colella, the linear waves,
the advected sound wave and the simple wave, as exact cell averages of the conserved state per unit length. 2D/3D setups are the benchmark configurations of this project:
jittered lattices (the convention of src/hydro/problems_3d.hpp:36-54, uniform jitter in [-j, j] spacings per axis),
the Sedov blast, the Kelvin-Helmholtz shear layer (sharp, and luna-cpp's smooth one) and the smooth acoustic wave.
"""
import math
from dataclasses import dataclass, field
from typing import Callable, Tuple

import numpy as np

GRID_BITS = 30  # positions snap to a grid of 2^-30 of the longest side


@dataclass
class Problem1D:
    L: float
    origin: float
    t_end: float
    gamma: float
    walls: Tuple[str, str]
    average: Callable  # (a, b) -> conserved state per unit length, averaged over [a, b]
    periodic: bool = False
    left: Tuple[float, float, float] = None   # a Riemann problem's primitives either side of its jump at x0
    right: Tuple[float, float, float] = None
    x0: float = 0.0


def _conserved1(W, gamma):
    # problems_1d.hpp:31-33 (its own operation order)
    return (W[0], W[0] * W[1], W[2] / (gamma - 1) + 0.5 * W[0] * W[1] * W[1])


def _fraction_below(x, a, b):
    return min(max((x - a) / (b - a), 0.0), 1.0)


def _tube(L, origin, t_end, gamma, left, right, x0, walls):
    l, r = _conserved1(left, gamma), _conserved1(right, gamma)

    def average(a, b):
        f = _fraction_below(x0, a, b)
        return tuple(f * l[k] + (1 - f) * r[k] for k in range(3))

    return Problem1D(L, origin, t_end, gamma, walls, average, False, tuple(left), tuple(right), x0)


def _shock_into_wave(a, b, left, x0, amp, k, gamma):
    h = b - a
    f = _fraction_below(x0, a, b)
    lo = max(a, x0)
    rho = f * left[0] + (1 - f)
    if b > lo:
        rho += amp * (math.cos(k * lo) - math.cos(k * b)) / k / h
    return (rho, f * left[0] * left[1],
            f * (left[2] / (gamma - 1) + 0.5 * left[0] * left[1] * left[1]) + (1 - f) / (gamma - 1))


def _woodward_colella(a, b):
    f1, f2 = _fraction_below(0.1, a, b), _fraction_below(0.9, a, b)
    p = f1 * 1000 + (f2 - f1) * 0.01 + (1 - f2) * 100
    return (1.0, 0.0, p / (1.4 - 1))


def _eigenmode(a, b, eps, v0, drho, dv, dp):
    k, G = 2 * math.pi, 5.0 / 3.0
    P0 = 1 / G
    h = b - a
    half = 0.5 * k * h
    s = math.sin(k * 0.5 * (a + b)) * math.sin(half) / half
    return (1 + eps * drho * s, v0 + eps * (dv + v0 * drho) * s,
            P0 / (G - 1) + 0.5 * v0 * v0 + eps * (dp / (G - 1) + 0.5 * v0 * v0 * drho + v0 * dv) * s)


PI = 3.141592653589793
LINEAR_AMPLITUDE = 1e-6  # LUNA's EPS of its linear-wave tests


def _noh(a, b):
    # problems_1d.hpp:63-67: cold inflow onto the wall at 0, brought to rest by a ramp towards the far wall, at the centre
    P0, U0, L, A = 1e-6, -1.0, 2.0, 1.0
    x = 0.5 * (a + b)
    u = U0 * min(max((L - x) / (L - A), 0.0), 1.0)
    return (1.0, u, P0 / (5.0 / 3.0 - 1) + 0.5 * u * u)


def _advected(a, b, eps):
    # problems_1d.hpp:79-83: the advected sound wave at the cell's centre
    s = eps * math.sin(2 * PI * 0.5 * (a + b))
    return _conserved1((1 + s, 1 + s, 1 + s), 5.0 / 3.0)


GAUSS4 = ((-0.8611363115940526, 0.3478548451374538), (-0.3399810435848563, 0.6521451548625461),
          (0.3399810435848563, 0.6521451548625461), (0.8611363115940526, 0.3478548451374538))


def gauss_average(a, b, f):
    """4-point Gauss-Legendre average over [a, b] of a pointwise conserved state f(x) (problems_1d.hpp:86-97)."""
    u = [0.0, 0.0, 0.0]
    for x, w in GAUSS4:
        at = f(0.5 * (a + b) + 0.5 * (b - a) * x)
        for v in range(3):
            u[v] += 0.5 * w * at[v]
    return tuple(u)


def _simple_wave(a, b, amp):
    # problems_1d.hpp:100-106: the isentropic simple wave, gamma 1.4: c = 1 + amp sin(2 pi x), u = 5 (c - 1), rho = c^5
    def at(x):
        g = 1.4
        c = 1 + amp * math.sin(2 * PI * x)
        rho = c ** (2 / (g - 1))
        return _conserved1((rho, 2 / (g - 1) * (c - 1), rho ** g / g), g)

    return gauss_average(a, b, at)


def _pure_shock(mach, walls):
    # problems_1d.hpp:109-115: a shock of Mach `mach` at x = 0.1 running right into (1, 0, 1)
    g = 1.4
    m2, s = mach * mach, mach * math.sqrt(g)
    rho2 = (g + 1) * m2 / ((g - 1) * m2 + 2)
    p2 = (2 * g * m2 - (g - 1)) / (g + 1)
    v2 = s * (1 - 1 / rho2)
    return _tube(1.0, 0.0, 0.5 / s, g, (rho2, v2, p2), (1, 0, 1), 0.1, walls)


def problem_1d(name, amplitude=0.0):
    """luna-cpp's 1D problem table (problems_1d.hpp:209-284): LUNA's setups and the ones it lacks; amplitude: of the
    linear waves (signed) and of the simple wave (0: their defaults, 1e-6 and 0.1)."""
    outflow, closed = ("outflow", "outflow"), ("reflective", "reflective")

    def toro(left, right, t):
        return _tube(1.0, 0.0, t, 1.4, left, right, 0.5, outflow)

    if name in ("sod", "toro-1"):
        return toro((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.25)
    if name == "toro-2":
        return toro((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.15)
    if name == "toro-3":
        return toro((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 0.012)
    if name == "toro-4":
        return toro((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), 0.035)
    if name == "toro-5":
        return toro((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 0.035)
    if name == "einfeldt-strong":
        return toro((1.0, -2.0, 0.4 * (3.0 - 0.5 * 4.0)), (1.0, 2.0, 0.4 * (3.0 - 0.5 * 4.0)), 0.15)
    if name == "einfeldt-shear":
        return toro((1.0, -1.0, 0.4 * (5.0 - 0.5 * 1.0)), (1.0, 1.0, 0.4 * (5.0 - 0.5 * 1.0)), 0.15)
    if name == "la-peak":
        return _tube(0.5, 0.1, 3.9e-3, 1.4, (0.1261192, 8.9047029, 782.92899), (6.591493, 2.2654207, 3.1544874), 0.5, outflow)
    if name == "la-rcvcr":
        return toro((1.0, -4.0, 0.4), (1.0, 4.0, 0.4), 0.15)
    if name == "la-stationary":
        return _tube(1.0, 0.0, 0.012, 1.4, (1.0, -19.59745, 1e3), (1.0, -19.59745, 1e-2), 0.8, outflow)
    if name == "la-slow":
        return toro((3.857143, -0.810631, 10.33333), (1.0, -3.44, 1.0), 1.0)
    if name == "la-streams":
        return toro((1.0, 2.0, 0.1), (1.0, -2.0, 0.1), 0.8)
    if name == "la-vacuum":
        return _tube(6.0, -5.0, 0.75, 1.4, (0.0, 0.0, 0.0), (1.0, 0.0, 1.0), 0.0, outflow)
    if name == "la-near-vacuum":
        return _tube(6.0, -5.0, 0.75, 1.4, (1e-6, 0.0, 1e-6), (1.0, 0.0, 1.0), 0.0, outflow)
    if name == "leblanc":
        return _tube(1.0, 0.0, 0.5, 5.0 / 3.0, (1.0, 0.0, (2.0 / 3.0) * 1e-1), (1e-3, 0.0, (2.0 / 3.0) * 1e-10), 0.3, outflow)
    if name == "noh":
        return Problem1D(2.0, 0.0, 0.6, 5.0 / 3.0, closed, _noh)
    if name == "shu-osher":
        left = (3.857143, 2.629369, 10.33333)
        return Problem1D(10.0, -5.0, 1.8, 1.4, outflow, lambda a, b: _shock_into_wave(a, b, left, -4, 0.2, 5, 1.4), False, left)
    if name == "titarev-toro":
        left = (1.515695, 0.523346, 1.805)
        return Problem1D(10.0, -5.0, 5.0, 1.4, outflow, lambda a, b: _shock_into_wave(a, b, left, -4.5, 0.1, 20 * PI, 1.4))
    if name == "woodward-colella":
        return Problem1D(1.0, 0.0, 0.038, 1.4, closed, _woodward_colella)
    eps = amplitude if amplitude != 0 else LINEAR_AMPLITUDE

    def wave(t, v0, drho, dv, dp):
        return Problem1D(1.0, 0.0, t, 5.0 / 3.0, outflow, lambda a, b: _eigenmode(a, b, eps, v0, drho, dv, dp), periodic=True)

    if name == "sound-right":
        return wave(1.0, 0.0, 1.0, 1.0, 1.0)
    if name == "sound-left":
        return wave(1.0, 0.0, 1.0, -1.0, 1.0)
    if name == "entropy":
        return wave(2.0, 0.5, 1.0, 0.0, 0.0)
    if name == "sound-advected":
        return Problem1D(1.0, 0.0, 1.0, 5.0 / 3.0, outflow, lambda a, b: _advected(a, b, eps), periodic=True)
    if name == "simple-wave":
        amp = amplitude if amplitude != 0 else 0.1
        return Problem1D(1.0, 0.0, 0.2, 1.4, outflow, lambda a, b: _simple_wave(a, b, amp), periodic=True)
    if name.startswith("shock-"):
        return _pure_shock(float(name[6:]), outflow)
    raise ValueError("unknown 1D problem " + name)


def moving_walls(pb):
    """The walls of a moving run (problems_1d.hpp:285-300): the static walls, except that an outflow end where the
    problem's state enters supersonically becomes an inflow wall holding that state. Returns (walls, (left, right)
    inflow primitives)."""
    walls, inflow = list(pb.walls), [(0.0, 0.0, 0.0), (0.0, 0.0, 0.0)]

    def entering(W, side):
        return W is not None and W[0] > 0 and side * W[1] > math.sqrt(pb.gamma * W[2] / W[0])

    if pb.walls[0] == "outflow" and entering(pb.left, 1):
        walls[0], inflow[0] = "inflow", tuple(pb.left)
    if pb.walls[1] == "outflow" and entering(pb.right, -1):
        walls[1], inflow[1] = "inflow", tuple(pb.right)
    return tuple(walls), tuple(inflow)


def line_points(n, L, jitter=0.0, seed=1):
    """n generators on [0, L]: L (i + 1/2 + jitter u_i) / n, u_i uniform in [-1/2, 1/2) from mt19937_64(seed)
    (hydro_1d.cpp:136-139); the jitter draws use numpy's MT19937 and are not the reference's bits."""
    i = np.arange(n, dtype=np.float64)
    if jitter == 0.0:
        return L * (i + 0.5) / n
    u = np.random.Generator(np.random.MT19937(seed)).uniform(-0.5, 0.5, n)
    return L * (i + 0.5 + jitter * u) / n


def line_grid(L):
    """The snapping grid of a line of length L: (step, period)."""
    period = 1 << GRID_BITS
    return L / period, period


def line_keys(x, L):
    step, period = line_grid(L)
    keys = np.rint(np.asarray(x) / step).astype(np.int64) % period
    return np.sort(keys), step, period


def line_cells(keys, step, period, periodic):
    """Each slot's cell [a, b] and volume V on the line (without the problem's origin), from the gaps of the sorted keys
    as the line computes V and the centroid (a wall's gap twice the distance to it): lists a, b, V."""
    n = len(keys)
    gaps = np.empty(n + 1, dtype=np.int64)  # gaps[s] = gap(s - 1)
    inner = np.diff(keys)
    if periodic:
        wrap = keys[0] + period - keys[-1]
        gaps[0], gaps[n] = wrap, wrap
    else:
        gaps[0], gaps[n] = 2 * keys[0], 2 * (period - keys[-1])
    gaps[1:n] = inner
    a, b, vol = [], [], []
    for s in range(n):
        V = (float(gaps[s]) + float(gaps[s + 1])) * (0.5 * step)
        c = (float(gaps[s + 1]) - float(gaps[s])) * (0.25 * step)
        x = float(keys[s]) * step
        lo = x + c - 0.5 * V
        a.append(lo)
        b.append(lo + V)
        vol.append(V)
    return a, b, vol


def line_initial_state(problem, keys, step, period):
    """Q (extensive, n x 3) from the problem's exact cell averages over each cell [c - V/2, c + V/2]."""
    a, b, vol = line_cells(keys, step, period, problem.periodic)
    q = np.empty((len(keys), 3))
    for s in range(len(keys)):
        V = vol[s]
        u = problem.average(a[s] + problem.origin, b[s] + problem.origin)
        q[s] = [u[k] * V for k in range(3)]
    return q


# ----------------------------------------------------------------------------------------------------------------- 2D / 3D


def morton_key(idx, bits):
    """Interleaved bits of integer lattice indices idx (N x D), axis 0 lowest."""
    idx = np.asarray(idx, dtype=np.uint64)
    D = idx.shape[1]
    key = np.zeros(idx.shape[0], dtype=np.uint64)
    for b in range(bits):
        for a in range(D):
            key |= ((idx[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * D + a)
    return key


def jittered_lattice(n, dim, jitter, seed, side=1.0, morton=True):
    """n^dim generators on the periodic box [0, side)^dim: lattice centres moved by up to `jitter` spacings per axis
    (uniform draws, numpy MT19937 seeded), snapped to the 2^-30 grid, in Morton order of their lattice cells."""
    h = side / n
    axes = np.meshgrid(*[np.arange(n)] * dim, indexing="ij")
    idx = np.stack([a.ravel() for a in axes], axis=1)  # lattice index, last axis fastest
    if morton:
        bits = max(1, int(math.ceil(math.log2(n))))
        order = np.argsort(morton_key(idx, bits), kind="stable")
        idx = idx[order]
    rng = np.random.Generator(np.random.MT19937(seed))
    u = rng.uniform(-1.0, 1.0, size=idx.shape)
    x = (idx + 0.5 + jitter * u) * h
    return snap_positions(x, side)


def grid_of(side):
    """The snapping grid of a box of sides `side`: step 2^-30 of the longest side, integer periods per axis."""
    side = np.asarray(side, dtype=np.float64)
    step = float(side.max()) / (1 << GRID_BITS)
    period = np.rint(side / step).astype(np.int64)
    return step, period


def snap_positions(x, side=1.0):
    x = np.asarray(x, dtype=np.float64)
    D = x.shape[1]
    sides = np.full(D, side, dtype=np.float64) if np.isscalar(side) else np.asarray(side, dtype=np.float64)
    step, period = grid_of(sides)
    k = np.rint(x * (1.0 / step)).astype(np.int64)
    k = np.mod(k, period[None, :])
    return k.astype(np.float64) * step


def conserved_rows(W, gamma):
    """W (N x V primitives) -> U per volume (euler.hpp:165-172 order: E = P/(gamma-1) + (0.5 rho) v.v)."""
    W = np.asarray(W, dtype=np.float64)
    D = W.shape[1] - 2
    rho = W[:, 0]
    vv = W[:, 1] * W[:, 1]
    for a in range(1, D):
        vv = vv + W[:, 1 + a] * W[:, 1 + a]
    U = np.empty_like(W)
    U[:, 0] = rho
    U[:, 1:D + 1] = rho[:, None] * W[:, 1:D + 1]
    U[:, D + 1] = W[:, D + 1] / (gamma - 1.0) + 0.5 * rho * vv
    return U


def sedov_state(mesh, gamma=5.0 / 3.0, rho=1.0, p=1e-5, energy=1.0, radius_cells=2.0, n=None):
    """The Sedov blast: uniform (rho, 0, p); energy E added as energy density E / sum V over the cells whose generator lies
    within radius_cells lattice spacings of the box centre (periodic distance), i.e. volume-weighted. Returns Q (extensive)."""
    D, N = mesh["dim"], mesh["cells"]
    side = np.asarray(mesh["side"][:D])
    n = n if n is not None else round(N ** (1.0 / D))
    h = side[0] / n
    x = np.asarray(mesh["positions"]).reshape(N, D)
    d = x - 0.5 * side[None, :]
    d = d - side[None, :] * np.round(d / side[None, :])
    inside = np.sqrt((d * d).sum(axis=1)) < radius_cells * h
    V = np.asarray(mesh["volume"])
    W = np.zeros((N, D + 2))
    W[:, 0] = rho
    W[:, D + 1] = p
    U = conserved_rows(W, gamma)
    deposit = V[inside].sum()
    U[inside, D + 1] += energy / deposit
    return U * V[:, None], int(inside.sum())


def kelvin_helmholtz_state(mesh, gamma=5.0 / 3.0):
    """2D shear layer on the periodic unit square: p = 2.5; rho = 2, vx = 0.5 for |y - 1/2| < 1/4, else rho = 1, vx = -0.5;
    vy = 0.01 sin(4 pi x) [exp(-(y - 1/4)^2 / (2 0.05^2)) + exp(-(y - 3/4)^2 / (2 0.05^2))], at the cell centroids."""
    N = mesh["cells"]
    x = np.asarray(mesh["positions"]).reshape(N, 2) + np.asarray(mesh["centroid"]).reshape(N, 2)
    X, Y = x[:, 0] % 1.0, x[:, 1] % 1.0
    inner = np.abs(Y - 0.5) < 0.25
    W = np.zeros((N, 4))
    W[:, 0] = np.where(inner, 2.0, 1.0)
    W[:, 1] = np.where(inner, 0.5, -0.5)
    s = 0.05
    W[:, 2] = 0.01 * np.sin(4 * np.pi * X) * (np.exp(-(Y - 0.25) ** 2 / (2 * s * s)) + np.exp(-(Y - 0.75) ** 2 / (2 * s * s)))
    W[:, 3] = 2.5
    return conserved_rows(W, gamma) * np.asarray(mesh["volume"])[:, None]


def _mcnally(y, outer, inner, layer=0.025):
    """McNally et al. (2012) eq. 1 for y in [0, 1), mirrored about 1/2: `outer` for |y - 1/2| > 1/4, `inner` inside,
    joined by exponential ramps of width `layer` (hydro_kh2d.cpp:57-60)."""
    h = np.where(y < 0.5, y, 1.0 - y)
    half = 0.5 * (outer - inner)
    return np.where(h < 0.25, outer - half * np.exp((h - 0.25) / layer), inner + half * np.exp((0.25 - h) / layer))


def kelvin_helmholtz_smooth_state(mesh, gamma=5.0 / 3.0):
    """The smooth shear layer of luna-cpp's KH benchmark (hydro_kh2d.cpp:27, 170-184): rho from 1 (outside) to 2 (inside)
    and vx from 0.5 to -0.5 along McNally's ramps, vy = 0.01 sin(4 pi x), p = 2.5, at the cell centroids. Unlike the
    sharp layer it converges under refinement, so errors against a finer run measure the method."""
    N = mesh["cells"]
    x = np.asarray(mesh["positions"]).reshape(N, 2) + np.asarray(mesh["centroid"]).reshape(N, 2)
    X, Y = x[:, 0], x[:, 1] - np.floor(x[:, 1])
    W = np.zeros((N, 4))
    W[:, 0] = _mcnally(Y, 1.0, 2.0)
    W[:, 1] = _mcnally(Y, 0.5, -0.5)
    W[:, 2] = 0.01 * np.sin(4 * np.pi * X)
    W[:, 3] = 2.5
    return conserved_rows(W, gamma) * np.asarray(mesh["volume"])[:, None]


def acoustic_state(mesh, gamma=5.0 / 3.0, amplitude=1e-3, k=(1, 1, 0)):
    """rho = 1 + amplitude sin(2 pi k.x) at the cell centroids, v = 0, p = 3/5 (c = 1 for gamma = 5/3)."""
    D, N = mesh["dim"], mesh["cells"]
    x = np.asarray(mesh["positions"]).reshape(N, D) + np.asarray(mesh["centroid"]).reshape(N, D)
    phase = 2 * np.pi * (x @ np.asarray(k[:D], dtype=np.float64))
    W = np.zeros((N, D + 2))
    W[:, 0] = 1 + amplitude * np.sin(phase)
    W[:, D + 1] = 3.0 / 5.0
    return conserved_rows(W, gamma) * np.asarray(mesh["volume"])[:, None]
