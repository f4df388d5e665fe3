"""The reference's Galilean test of the moving step (hydro_test.cpp:474-510) on the HIP path: a blast moving with the box at
velocity 16 equals the blast at rest cell by cell on a moving mesh (the reference's double-policy bounds, :683), its radial
profile follows the blast, a static mesh's does not; every final state equals the oracle's."""
import numpy as np
import pytest

import galilean
import oracle
from luna_hydro import hydro

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0


class _Oracle:
    def __init__(self, m, moving):
        self.fv = oracle.FiniteVolume(m, GAMMA, moving=moving)
        for name in ("set_state", "time_step", "step", "moving_begin", "moving_finish"):
            setattr(self, name, getattr(self.fv, name))

    def state(self):
        return self.fv.q.copy()


def test_galilean_blast_on_the_gpu():
    print("Galilean blast 16^3, 40 steps, moving and static, GPU", flush=True)
    made = []

    def make(m, moving):
        fv = hydro.FiniteVolume(m, GAMMA, motion="moving" if moving else "static")
        made.append(fv)
        return fv

    try:
        r = galilean.runs(make)
    finally:
        for fv in made:
            fv.close()
    print("  GPU:", {k: v for k, v in r.items() if k != "states"}, flush=True)
    assert r["dt"] <= r["cfl"]
    assert r["cells"] < 1e-6, r
    assert r["moving"] < 1e-10 and r["static"] > 1e3 * r["moving"], r
    print("Galilean blast, oracle", flush=True)
    o = galilean.runs(_Oracle)
    for a, b in zip(r["states"], o["states"]):
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) <= 1e-12
    print("  GPU states equal the oracle's to the bit: %s" % all(np.array_equal(a, b) for a, b in zip(r["states"], o["states"])),
          flush=True)
