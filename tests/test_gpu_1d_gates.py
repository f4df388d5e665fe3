"""luna-cpp's 1D ladder gates on the GPU's line (hydro.Line / hydro.MovingLine through the C-ABI): for every static row
of tests/golden/luna_1d_gates.json (28 problems, N = 256 to 4096, LUNA's limiter and, where measured, minmod and van
Leer), the L1 errors of rho, v, P equal LUNA's to the row's tolerance (1e-6 relative; or LUNA's is under the row's
floor); for every moving gate (19 problems, N = 256 and 512, Titarev-Toro at 1024 and 2048), the run reaches t_end
with each L1 at most 1.3 times the reference's own moving run's (its gate), and within 1 % of it."""
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


def test_every_moving_gate():
    worst, failed = 0.0, []
    for row in gates_1d.moving_cases():
        res, reached = gates_1d.check_moving_row(hydro.MovingLine, row)
        for v, got, ref, ok in res:
            if ref > 1e-6:  # the entropy wave's stored 1e-6 is a floor, not a measurement
                worst = max(worst, abs(got / ref - 1))
            if not (ok and reached):
                failed.append((row["problem"], row["n"], v, got, ref, reached))
        print("%s N=%d moving: %s" % (row["problem"], row["n"], " ".join("%.6e" % r[1] for r in res)), flush=True)
    print("1D moving gates on the GPU: largest |L1 / the reference's - 1| %.2e" % worst, flush=True)
    assert not failed, failed
    assert worst < 1e-2
