"""The Sedov driver (hydro_sedov's loop and report lines) on the GPU: its checksum line equals the one computed from the
oracle's run of the same steps (the GPU reproduces the oracle's bits), and its snapshot holds the final state."""
import numpy as np
import pytest

import oracle
from luna_hydro import mesh, problems, sedov, snapshot

pytestmark = pytest.mark.gpu


def test_sedov_driver_checksum_matches_oracle(tmp_path):
    lines = []
    res = sedov.run(n=10, moving=False, t_end=0.05, courant=0.3, snapshots=10, out=str(tmp_path), max_steps=4, log=lines.append)
    print("\n".join(lines), flush=True)
    assert [l.split(":")[0] for l in lines] == ["run", "fallbacks", "phases_ms_median", "conservation", "checksum"]
    m = mesh.build(problems.jittered_lattice(10, 3, 0.25, 1), 1.0)
    q, _ = problems.sedov_state(m)
    orc = oracle.FiniteVolume(m, 5.0 / 3.0)
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
    assert lines[-1].startswith(want), (lines[-1], want)
    snap = snapshot.read(str(tmp_path / "snapdir_000" / "snapshot_000.0.hdf5"))
    assert np.array_equal(snap["U"], (q / m["volume"][:, None]).T)
