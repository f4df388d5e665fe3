"""Writes the golden vectors held in luna-cpp's own tests as JSON fixtures under tests/golden/.

Run once in a container that has the reference checkout (read as text only, nothing is compiled or imported):

    python tests/golden/make_fixtures.py /root/reference

Sources (paths relative to the reference checkout):
  src/test/hydro_test.cpp:47-80          LUNA's hll/hllc_flux_primitive values (gamma 5/3, jax x64)
  src/test/luna_1d_values.hpp:5-61       LUNA's final 1D states at N = 16 (MC, HLLC, CFL 0.8) and step counts
  src/test/luna_derivative_values.hpp    LUNA's 1D forward states after 5 fixed steps (and jvp/vjp vectors)
  src/test/hydro1d_test.cpp:234          LUNA's Sod L1 bounds at N = 128, 256
The fixtures are data only: inputs and expected outputs.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def numbers(text):
    return [float(x) for x in re.findall(NUM, text)]


def riemann_cases(src, name):
    start = src.index("const LunaCase %s[] = {" % name)
    end = src.index("};", start)
    body = src[start:end]
    cases = []
    for m in re.finditer(r"\{(\d), \{([^}]*)\}, \{([^}]*)\}, \{([^}]*)\}\}", body):
        cases.append({"axis": int(m.group(1)), "L": numbers(m.group(2)), "R": numbers(m.group(3)),
                      "F": numbers(m.group(4))})
    return cases


def table(src, name, rows, cols):
    start = src.index("const double %s[" % name)
    brace = src.index("= {", start)
    end = src.index("};", brace)
    vals = numbers(src[brace:end])
    assert len(vals) == rows * cols, (name, len(vals))
    return [vals[r * cols:(r + 1) * cols] for r in range(rows)]


def main(ref):
    out = {}
    hydro_test = open(os.path.join(ref, "src/test/hydro_test.cpp")).read()
    out["riemann"] = {"gamma": 5.0 / 3.0, "hll": riemann_cases(hydro_test, "luna_hll"),
                      "hllc": riemann_cases(hydro_test, "luna_hllc")}
    assert len(out["riemann"]["hll"]) == 12 and len(out["riemann"]["hllc"]) == 15

    v1d = open(os.path.join(ref, "src/test/luna_1d_values.hpp")).read()
    final = {}
    for key, prob in (("sod", "sod"), ("shu_osher", "shu-osher"), ("woodward_colella", "woodward-colella")):
        steps = int(re.search(r"const int luna_steps_%s = (\d+);" % key, v1d).group(1))
        final[prob] = {"n": 16, "steps": steps, "W": table(v1d, "luna_" + key, 16, 3)}
    out["luna_1d_final"] = final

    dv = open(os.path.join(ref, "src/test/luna_derivative_values.hpp")).read()
    cases = {}
    for m in re.finditer(r"// ([^\n]*)\nnamespace (luna_d_\w+) \{(.*?)\n\}", dv, re.S):
        label, ns, body = m.group(1), m.group(2), m.group(3)
        head = re.search(r"const int n = (\d+), limiter = (\d+), steps = (\d+);", body)
        n, limiter, steps = (int(head.group(k)) for k in (1, 2, 3))
        dt = float(re.search(r"const double dt = (" + NUM + r");", body).group(1))
        problem = "sod" if ns.startswith("luna_d_sod") else "sound-right"
        case = {"label": label, "problem": problem, "n": n, "limiter": limiter, "steps": steps, "dt": dt}
        for arr in ("U0", "dU", "w", "U", "JdU", "JTw"):
            case[arr] = table(body, arr, n, 3)
        cases[ns] = case
    assert len(cases) == 7, sorted(cases)
    out["luna_1d_forward"] = cases

    t1d = open(os.path.join(ref, "src/test/hydro1d_test.cpp")).read()
    b = re.search(r"const double bounds\[2\]\[3\] = \{\{([^}]*)\}, \{([^}]*)\}\};", t1d)
    out["sod_l1_bounds"] = {"128": numbers(b.group(1)), "256": numbers(b.group(2)), "t_end": 0.25}

    for key, value in out.items():
        with open(os.path.join(HERE, key + ".json"), "w") as f:
            json.dump(value, f, indent=1)
            f.write("\n")
    print("wrote", ", ".join(k + ".json" for k in out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
