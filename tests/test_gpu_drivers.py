"""The Sedov driver (hydro_sedov's loop and report lines) on the GPU: its checksum line equals the one computed from the
oracle's run of the same steps (the GPU reproduces the oracle's bits), and its snapshot holds the final state."""
import numpy as np
import pytest

import oracle
from luna_hydro import mesh, problems, sedov, snapshot

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["double", "default"])
def test_sedov_driver_checksum_matches_oracle(tmp_path, precision):
    """--precision default is the reference's 3D default, grad16."""
    from luna_hydro import hydro

    lines = []
    res = sedov.run(n=10, moving=False, t_end=0.05, courant=0.3, snapshots=10, out=str(tmp_path), max_steps=4, log=lines.append,
                    precision=precision)
    print("\n".join(lines), flush=True)
    # the reference's report lines, then this build's shock-radius line
    assert [l.split(":")[0] for l in lines] == ["run", "fallbacks", "phases_ms_median", "conservation", "checksum", "shock"]
    m = mesh.build(problems.jittered_lattice(10, 3, 0.25, 1), 1.0)
    q, _ = problems.sedov_state(m)
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, precision=hydro.policy(precision, 3))
    orc.set_state(q)
    t, nxt = 0.0, 1
    for _ in range(res["steps"]):
        dt = orc.time_step(0.3)
        t_next = 0.05 * nxt / 10
        if t + dt >= t_next:
            dt, t, nxt = t_next - t, t_next, nxt + 1
        else:
            t += dt
        orc.step(dt)
    want = "checksum: steps=%d Q=%x" % (res["steps"], sedov.bits_hash(orc.q))
    assert lines[4].startswith(want), (lines[4], want)
    snap = snapshot.read(str(tmp_path / "snapdir_000" / "snapshot_000.0.hdf5"))
    assert np.array_equal(snap["U"], (q / m["volume"][:, None]).T)


@pytest.mark.parametrize("ic,precision,order,riemann", [("smooth", "double", 1, "hllc"), ("sharp", "double", 1, "hllc"),
                                                        ("smooth", "default", 1, "hllc"), ("smooth", "double", 1, "hll"),
                                                        ("smooth", "double", 0, "hllc")])
def test_kh_driver_hash_and_modes_match_oracle(ic, precision, order, riemann):
    """The KH driver (hydro_kh2d's loop: outputs every 0.02, dt cut to land on them) on 24^2 to t = 0.06: its run line's
    hash is the FNV-1a of the oracle's final state after the same steps, and its mode series is the oracle's. --precision
    default is the reference's 2D default, grad32."""
    from luna_hydro import hydro, kh2d

    lines = []
    res = kh2d.run(n=24, ic=ic, t_end=0.06, log=lines.append, precision=precision, order=order, riemann=riemann)
    print("\n".join(lines), flush=True)
    assert [l.split(":")[0] for l in lines] == ["totals", "run", "ours"]
    m = mesh.build(problems.jittered_lattice(24, 2, 0.25, 1), 1.0)
    state = problems.kelvin_helmholtz_smooth_state if ic == "smooth" else problems.kelvin_helmholtz_state
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, step_rule="luna", order=order, riemann=riemann,
                              precision=hydro.policy(precision, 2) if order == 1 else "double")
    orc.set_state(state(m, 5.0 / 3.0))
    t, series = 0.0, []
    for k in range(1, 4):
        target = k * kh2d.OUTPUT_EVERY
        while t < target * (1 - 1e-13):
            dt = min(orc.time_step(0.8), target - t)
            orc.step(dt)
            t = target if dt == target - t else t + dt
        series.append(kh2d.mode_amplitude(orc.q, m["positions"], m["centroid"], m["volume"]))
    assert "hash=%016x" % sedov.bits_hash(orc.q, kh2d.FNV_BASIS) in lines[1]
    assert [s[1:] for s in res["series"][1:]] == series


@pytest.mark.parametrize("fused", [True, False])
def test_1d_driver_hash_matches_oracle(fused):
    """The 1D driver (hydro_1d's loop, fused LUNA-rule steps or time_step + step) on Sod with 256 cells: its hash is the
    FNV-1a of the oracle's final state after the same steps, and repeats give the same bits."""
    from luna_hydro import hydro1d

    lines = []
    res = hydro1d.run("sod", 256, fused=fused, repeats=2, log=lines.append)
    print("\n".join(lines), flush=True)
    assert [l.split(":")[0] for l in lines] == ["run", "run", "ours"] and res["same_bits"]
    pb = problems.problem_1d("sod")
    keys, step, period = problems.line_keys(problems.line_points(256, pb.L), pb.L)
    orc = oracle.Line(keys, period, step, pb.gamma, periodic=pb.periodic, walls=pb.walls)
    orc.set_state(problems.line_initial_state(pb, keys, step, period))
    t, steps = 0.0, 0
    nxt = orc.time_step(0.8) if fused else 0.0
    while t < pb.t_end * (1 - 1e-12):
        dt = min(nxt if fused else orc.time_step(0.8), pb.t_end - t)
        if fused:
            nxt = orc.step_and_time_step(dt, 0.8)
        else:
            orc.step(dt)
        t += dt
        steps += 1
    assert steps == res["steps"]
    assert res["hash"] == sedov.bits_hash(orc.state, hydro1d.FNV_BASIS)


def test_1d_driver_moving_hash_matches_oracle():
    """The 1D driver's moving mode (hydro_1d --moving: AREPO's rule at 0.4, cells from the inflow wall) on Shu-Osher with
    256 cells: its hash (Q, then the positions) is the oracle's after the same steps, with the same final cells."""
    from luna_hydro import hydro1d

    lines = []
    res = hydro1d.run("shu-osher", 256, moving=True, repeats=2, log=lines.append)
    print("\n".join(lines), flush=True)
    assert res["same_bits"] and "mesh=moving" in lines[-1]
    pb, orc = hydro1d.make_line("shu-osher", 256, moving=True, engine=oracle.MovingLine)
    t, steps = 0.0, 0
    while t < pb.t_end * (1 - 1e-12):
        dt = min(orc.time_step(0.4, "arepo"), pb.t_end - t)
        orc.step(dt)
        t += dt
        steps += 1
    assert steps == res["steps"] and orc.n == res["cells"] > 256
    h = sedov.bits_hash(orc.positions(), sedov.bits_hash(orc.state(), hydro1d.FNV_BASIS))
    assert res["hash"] == h
