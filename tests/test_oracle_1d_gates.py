"""luna-cpp's 1D ladder gates on the oracle's static line (tests/gates_1d.py): the L1 errors of rho, v, P that LUNA's
runs reach (tests/golden/luna_1d_gates.json) on every problem at N = 256 with its limiter, and at N = 512 with minmod
and van Leer on the Toro tubes; the GPU test covers every static row. The exact solver's vacuum solution is pinned by
what fixes it: symmetry, empty space between the fans, and mass conservation."""
import math

import pytest

import exact_riemann
import gates_1d
import oracle

CASES = [(row, "luna") for row, series in gates_1d.static_cases(256) if series == "luna"]
CASES += [(row, s) for row, s in gates_1d.static_cases(512)
          if row["n"] == 512 and s in ("minmod", "vanleer") and row["problem"].startswith("toro-")]


@pytest.mark.parametrize("row,series", CASES, ids=["%s-%d-%s" % (r["problem"], r["n"], s) for r, s in CASES])
def test_static_gate(row, series):
    res = gates_1d.check_row(oracle.Line, row, series)
    assert res and all(ok for *_, ok in res), [(v, got, luna, ratio) for v, got, luna, ratio, ok in res if not ok]


def test_vacuum_generation_solution():
    """la-rcvcr's data (u = -4, 4) open a vacuum: the solution is symmetric, empty at s = 0, and the mass on [-X, X]
    beyond both heads is the initial mass less what the undisturbed states carry out through +-X, 2 X - 2 rho |u| t."""
    s = exact_riemann.Solver((1.0, -4.0, 0.4), (1.0, 4.0, 0.4), 1.4)
    assert s.vacuum and s.sample(0.0) == (0.0, 0.0, 0.0)
    for x in (0.3, 1.0, 3.0, 5.0):
        l, r = s.sample(-x), s.sample(x)
        assert math.isclose(l[0], r[0], rel_tol=1e-14) and math.isclose(l[1], -r[1], rel_tol=1e-14)
    c = math.sqrt(1.4 * 0.4)
    X, t = 4.0 + c + 1.0, 1.0  # beyond both heads at t = 1
    mass = s.average(-X, X, 0.0, t, samples=200000)[0] * 2 * X
    assert abs(mass - (2 * X - 2 * 1.0 * 4.0 * t)) < 1e-6
