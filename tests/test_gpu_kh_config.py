"""The 2D Kelvin-Helmholtz configuration (the sharp shear layer on a jittered periodic mesh, gamma 5/3, moving mesh, LUNA's
rule at 0.8) on the GPU: plain and learned limiter step for step against the oracle at 32^2 and 48^2 (the learned one with
the device's fractions applied to the oracle's face-centroid gradients), and at the configuration's 512^2 a few moving
steps that conserve mass, momentum and energy, with the learned limiter's gradients never steeper than the face-centroid
limiter's on the same state."""
import math

import numpy as np
import pytest

import oracle
from luna_hydro import hydro, learned, mesh, problems

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0
COURANT = 0.8


def say(msg):
    print(msg, flush=True)


def kh_mesh(n, seed=1):
    return mesh.build(problems.jittered_lattice(n, 2, 0.25, seed), 1.0)


@pytest.mark.parametrize("n", [32, 48])
@pytest.mark.parametrize("limiter", ["face", "learned"])
def test_kh_moving_matches_oracle(n, limiter):
    say("KH moving %d^2, %s limiter, against the oracle" % (n, limiter))
    m = kh_mesh(n)
    q = problems.kelvin_helmholtz_state(m, GAMMA)
    gpu = hydro.FiniteVolume(m, GAMMA, limiter=limiter, motion="moving", step_rule="luna")
    orc = oracle.FiniteVolume(m, GAMMA, limiter=limiter, moving=True, step_rule="luna")
    params = learned.default_params()
    trainer = learned.Trainer(params) if limiter == "learned" else None
    if trainer is not None:
        gpu.set_limiter_network(params)
    gpu.set_state(q)
    orc.set_state(q)
    try:
        for s in range(6):
            dt = orc.time_step(COURANT)
            assert gpu.time_step(COURANT) == dt, s
            if trainer is not None:
                x, t = gpu.limiter_features()
                assert np.array_equal(x, orc.features) and np.array_equal(t, orc.troubled), s
                orc.apply_fractions(trainer.forward(learned.device_rows(x)).cpu().numpy())
            assert np.array_equal(gpu.gradients(), orc.grad.reshape(gpu.gradients().shape)), s
            xg = gpu.moving_begin(dt)
            xo = orc.moving_begin(dt)
            assert np.array_equal(xg, xo), s
            m1 = mesh.build(xg, 1.0)
            gpu.moving_finish(m1, dt)
            orc.moving_finish(m1, dt)
            assert np.array_equal(gpu.state(), orc.q), s
            assert gpu.fallbacks() == orc.fallbacks()
        say("  6 steps: dt, gradients, positions and Q equal to the bit")
    finally:
        gpu.close()


def test_kh_512_moving_conserves_and_stays_within_the_face_limiter():
    n = 512
    say("KH moving 512^2: 3 steps, plain and learned")
    m = kh_mesh(n)
    q = problems.kelvin_helmholtz_state(m, GAMMA)
    total = [math.fsum(q[:, v]) for v in range(4)]
    scale = math.sqrt(2 * total[0] * total[3])
    for limiter in ("face", "learned"):
        fv = hydro.FiniteVolume(m, GAMMA, limiter=limiter, motion="moving", step_rule="luna")
        ref = hydro.FiniteVolume(m, GAMMA, limiter="face", step_rule="luna")
        if limiter == "learned":
            fv.set_limiter_network(learned.default_params())
        fv.set_state(q)
        mm = m
        try:
            for s in range(3):
                if s:
                    ref.set_geometry(mm)
                ref.set_state(fv.state())
                dt = fv.time_step(COURANT)
                ref.time_step(COURANT)
                g, g0 = fv.gradients().reshape(-1, 4, 2), ref.gradients().reshape(-1, 4, 2)
                ratio = np.linalg.norm(g, axis=2) / np.maximum(np.linalg.norm(g0, axis=2), 1e-300)
                assert np.all(ratio <= 1 + 1e-12), (limiter, s)
                mm = mesh.build(fv.moving_begin(dt), 1.0)
                fv.moving_finish(mm, dt)
                qs = fv.state()
                after = [math.fsum(qs[:, v]) for v in range(4)]
                assert abs(after[0] - total[0]) <= 1e-12 * total[0] and abs(after[3] - total[3]) <= 1e-12 * total[3]
                assert all(abs(after[v] - total[v]) <= 1e-12 * scale for v in (1, 2))
            say("  %s: 3 steps conserve to 1e-12; mean gradient ratio to the face limiter %.3f" % (limiter, float(ratio.mean())))
        finally:
            fv.close()
            ref.close()
