"""luna-cpp's 1D ladder gates on the GPU's static line (hydro.Line through the C-ABI): for every static row of
tests/golden/luna_1d_gates.json (28 problems, N = 256 to 4096, LUNA's limiter and, where measured, minmod and van Leer),
the L1 errors of rho, v, P equal LUNA's to the row's tolerance (1e-6 relative; or LUNA's is under the row's floor).
The moving rows need the moving line, which this build does not have."""
import pytest

import gates_1d
from luna_hydro import hydro

pytestmark = pytest.mark.gpu


def test_every_static_gate():
    worst, linear, failed, runs = 0.0, 0.0, [], 0
    for row, series in gates_1d.static_cases():
        res = gates_1d.check_row(hydro.Line, row, series)
        runs += 1
        for v, got, luna, ratio, ok in res:
            if luna >= row["floor"]:
                if gates_1d.is_linear(row["problem"]):
                    linear = max(linear, ratio)  # the linear waves' gates: 1e-2 (L1 of a difference of two runs over 2 eps)
                else:
                    worst = max(worst, ratio)
            if not ok:
                failed.append((row["problem"], row["n"], series, v, got, luna, ratio))
        print("%s N=%d %s: %s" % (row["problem"], row["n"], series, " ".join("%.6e" % r[1] for r in res)), flush=True)
    print("1D gates on the GPU: %d runs, largest |L1 / LUNA's - 1| %.2e (linear waves %.2e)" % (runs, worst, linear), flush=True)
    assert not failed, failed
