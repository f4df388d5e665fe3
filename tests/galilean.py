"""The reference's Galilean test of the moving step (hydro_test.cpp:474-510), for any solver with the step's interface:
a blast at rest and the same blast moving with velocity (16, 0, 0), 40 fixed steps of 2^-12 on a 16^3 jittered mesh,
moving and static."""
import numpy as np

from luna_hydro import mesh, problems

GAMMA = 5.0 / 3.0


def blast(m, bulk, gamma=GAMMA):
    """hydro_test.cpp:294-305: pressure 100 within r < 0.15 of the box centre, else 1; rho 1; velocity (bulk, 0, 0)."""
    x = m["positions"]
    r2 = np.sum((x - 0.5) ** 2, axis=1)
    W = np.zeros((m["cells"], 5))
    W[:, 0] = 1.0
    W[:, 1] = bulk
    W[:, 4] = np.where(r2 < 0.15 * 0.15, 100.0, 1.0)
    return problems.conserved_rows(W, gamma) * m["volume"][:, None]


def primitive_rows(U, gamma=GAMMA):
    W = np.empty_like(U)
    W[:, 0] = U[:, 0]
    W[:, 1:4] = U[:, 1:4] / U[:, :1]
    W[:, 4] = (gamma - 1) * (U[:, 4] - 0.5 * np.sum(U[:, 1:4] * W[:, 1:4], axis=1))
    return W


def radial_profile(m, rho, cx):
    """hydro_test.cpp:308-324: the mean density in 8 radial bins (r < 0.4) around (cx, 0.5, 0.5), by generator."""
    x = m["positions"]
    d = x - np.array([cx, 0.5, 0.5])
    d -= np.round(d)
    b = (np.sqrt(np.sum(d * d, axis=1)) / 0.05).astype(int)
    total, count = np.zeros(8), np.zeros(8)
    sel = b < 8
    np.add.at(total, b[sel], rho[sel])
    np.add.at(count, b[sel], 1)
    return total / np.maximum(count, 1)


def runs(make, steps=40, bulk=16.0, dt=2.0 ** -12, n=16):
    """The reference's Galilean test (hydro_test.cpp:474-510) on the solver `make(mesh, moving)` builds: returns the
    largest per-cell difference of the moving blast from the one at rest (moving mesh), the two meshes' radial profile
    differences (moving, static) and the smallest CFL step seen."""
    points = problems.jittered_lattice(n, 3, 0.25, 1)
    cfl = 1.0
    out = {}
    for moving in (True, False):
        runs = []
        for b in (0.0, bulk):
            m = mesh.build(points, 1.0)
            fv = make(m, moving)
            fv.set_state(blast(m, b))
            runs.append([fv, m])
        for _ in range(steps):
            for r in runs:
                fv, m = r
                cfl = min(cfl, fv.time_step(0.3))
                if moving:
                    x = fv.moving_begin(dt)
                    r[1] = mesh.build(x, 1.0)
                    fv.moving_finish(r[1], dt)
                else:
                    fv.step(dt)
        (a, ma), (b, mb) = [(r[0].state(), r[1]) for r in runs]
        if moving:
            Wa = primitive_rows(a / ma["volume"][:, None])
            Wb = primitive_rows(b / mb["volume"][:, None])
            Wb[:, 1] -= bulk
            diff = np.abs(Wb[:, 0] - Wa[:, 0]) / Wa[:, 0] + np.abs(Wb[:, 4] - Wa[:, 4]) / Wa[:, 4]
            diff += np.sum(np.abs(Wb[:, 1:4] - Wa[:, 1:4]), axis=1) / bulk
            out["cells"] = float(diff.max())
        pa = radial_profile(ma, a[:, 0] / ma["volume"], 0.5)
        pb = radial_profile(mb, b[:, 0] / mb["volume"], 0.5 + bulk * dt * steps)
        out["moving" if moving else "static"] = float(np.sum(np.abs(pb - pa)) / np.sum(pa))
        out.setdefault("states", []).extend([a, b])
    out["cfl"] = cfl
    out["dt"] = dt
    return out
