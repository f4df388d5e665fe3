"""The learned limiter on the GPU (through the C-ABI) against the oracle: the features and troubled flags the sweep
records (bit for bit), the face-centroid results without a network (bit for bit), the network's forward and backward
kernels (BF16 operands, FP32 accumulation) against oracle/mlp.py, a learned step against the oracle's step with the same
fractions (bit for bit), and training (the loss falls; the same batch gives the same bits)."""
import numpy as np
import pytest

import cases
import mlp
import oracle
from luna_hydro import hydro, problems

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0
# Network tolerances: the device accumulates in FP32 in the matrix cores' order, the oracle in float64 from the same
# BF16-rounded operands; a hidden unit within ~1e-7 of a BF16 rounding boundary can round the other way (2^-8 of it).
FRACTION_TOL = 1e-3  # absolute, on sigmoid(z)
GRAD_TOL = 1e-2      # relative, on the norm of the parameter gradient


def say(msg):
    print(msg, flush=True)


def state(m, dim):
    return problems.kelvin_helmholtz_state(m, GAMMA) if dim == 2 else cases.rough_state(m)


def same_features(x, y, what):
    """Bit-equal feature rows; on a difference, prints which features differ and a few bit pairs."""
    if np.array_equal(x, y):
        return True
    bad = x != y
    say("%s: %d of %d feature bits differ; per feature %s" % (what, int(bad.sum()), bad.size, bad.sum(axis=0).tolist()))
    r, k = np.nonzero(bad)
    for a, b in list(zip(r, k))[:8]:
        say("  row %d feature %d: device %#06x oracle %#06x" % (a, b, x[a, b], y[a, b]))
    return False


def pair(m, q, limiter="learned"):
    rule = "luna" if m["dim"] == 2 else "arepo"
    gpu = hydro.FiniteVolume(m, GAMMA, limiter=limiter, step_rule=rule)
    orc = oracle.FiniteVolume(m, GAMMA, limiter=limiter, step_rule=rule)
    gpu.set_state(q)
    orc.set_state(q)
    return gpu, orc


@pytest.mark.parametrize("dim,n", [(2, 32), (3, 12)])
def test_features_match_oracle(dim, n):
    say("features dim=%d n=%d" % (dim, n))
    m = cases.voronoi(n, dim)
    gpu, orc = pair(m, state(m, dim))
    for s in range(3):
        assert gpu.time_step(0.3) == orc.time_step(0.3)
        x, t = gpu.limiter_features()
        assert same_features(x, orc.features, "step %d" % s)
        assert np.array_equal(t, orc.troubled), s
        dt = orc.time_step(0.3)
        gpu.step(dt)
        orc.step(dt)
    say("  troubled %d of %d" % (int(t.sum()), m["cells"]))


@pytest.mark.parametrize("dim,n", [(2, 32), (3, 10)])
def test_without_network_is_the_face_limiter(dim, n):
    m = cases.voronoi(n, dim)
    q = state(m, dim)
    a = hydro.FiniteVolume(m, GAMMA, limiter="learned")
    b = hydro.FiniteVolume(m, GAMMA, limiter="face")
    a.set_state(q)
    b.set_state(q)
    for _ in range(3):
        dt = b.time_step(0.3)
        assert a.time_step(0.3) == dt
        a.step(dt)
        b.step(dt)
    assert np.array_equal(a.state(), b.state())


def trainer_rows(rows, seed):
    rng = np.random.default_rng(seed)
    x = mlp.bf16_bits(np.column_stack([rng.uniform(0, 1, rows), rng.normal(0, 1, (rows, 12)), rng.integers(0, 2, (rows, 3))]))
    return x


def test_forward_matches_oracle():
    from luna_hydro import learned

    rows = 5000 + 7  # not a multiple of 16: the last tile is partial
    x = trainer_rows(rows, 1)
    p = mlp.init_params(3)
    tr = learned.Trainer(p)
    frac = tr.forward(learned.device_rows(x)).cpu().numpy()
    ref = mlp.forward(p, x)["fraction"]
    err = np.abs(frac - ref).max()
    say("forward: max |fraction - oracle| %.2e over %d rows" % (err, rows))
    assert err < FRACTION_TOL
    assert np.array_equal(tr.params(), p)


def test_gradient_matches_oracle_and_is_deterministic():
    from luna_hydro import learned

    rows = 20000 + 3
    rng = np.random.default_rng(2)
    x = trainer_rows(rows, 2)
    t = rng.uniform(0, 1, rows).astype(np.float32)
    w = rng.uniform(0, 1, rows).astype(np.float32)
    p = mlp.init_params(4)
    tr = learned.Trainer(p)
    xd, td, wd = learned.device_rows(x, t, w)
    loss, g = tr.gradient(xd, td, wd, float(w.astype(np.float64).sum()))
    loss2, g2 = tr.gradient(xd, td, wd, float(w.astype(np.float64).sum()))
    assert loss == loss2 and np.array_equal(g, g2)
    ref_loss, ref = mlp.gradient(p, x, t, w, float(w.astype(np.float64).sum()))
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    blocks = {"W1": (mlp.P_W1, mlp.P_B1), "b1": (mlp.P_B1, mlp.P_W2), "W2": (mlp.P_W2, mlp.P_B2), "b2": (mlp.P_B2, mlp.P_W3),
              "w3": (mlp.P_W3, mlp.P_B3), "b3": (mlp.P_B3, mlp.PARAMS)}
    per = {k: float(np.linalg.norm(g[a:b] - ref[a:b]) / max(np.linalg.norm(ref[a:b]), 1e-30)) for k, (a, b) in blocks.items()}
    say("gradient: loss %.6g vs %.6g, relative error %.2e, per block %s" % (loss, ref_loss, rel, {k: "%.1e" % v for k, v in per.items()}))
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    assert rel < GRAD_TOL
    assert max(per.values()) < GRAD_TOL


def test_training_fits_a_teacher():
    """Targets from a teacher network: Adam steps on the device kernels bring the loss down at least five-fold."""
    from luna_hydro import learned

    rows = 32768
    x = trainer_rows(rows, 3)
    teacher = mlp.init_params(11)
    teacher[mlp.P_B3] = 0.0
    t = mlp.forward(teacher, x)["fraction"].astype(np.float32)
    w = np.ones(rows, dtype=np.float32)
    params, history = learned.train(x, t, w, steps=600, batch=8192, lr=3e-3, seed=5, every=200)
    say("teacher fit: loss %s" % ["%.3g" % l for _, l in history])
    assert history[-1][1] < 0.2 * history[0][1]
    again, history2 = learned.train(x, t, w, steps=600, batch=8192, lr=3e-3, seed=5, every=200)
    assert np.array_equal(params, again)


@pytest.mark.parametrize("dim,n", [(2, 32), (3, 10)])
def test_learned_step_matches_oracle_with_the_same_fractions(dim, n):
    """A learned step: the device's fractions (the kernel that scales the gradients, on the recorded features) applied
    to the oracle's face-centroid gradients give the device's state bit for bit, and every fraction is within
    FRACTION_TOL of the oracle network's."""
    from luna_hydro import learned

    m = cases.voronoi(n, dim)
    gpu, orc = pair(m, state(m, dim))
    p = mlp.init_params(7)
    p[mlp.P_B3] = 0.5  # fractions near 0.6: the scaling shows
    gpu.set_limiter_network(p)
    tr = learned.Trainer(p)
    for s in range(3):
        dt = orc.time_step(0.3)
        assert gpu.time_step(0.3) == dt
        x, t = gpu.limiter_features()
        assert same_features(x, orc.features, "step %d" % s)
        frac = tr.forward(learned.device_rows(x)).cpu().numpy()
        assert np.abs(frac - mlp.forward(p, x)["fraction"]).max() < FRACTION_TOL
        orc.apply_fractions(frac)
        assert np.array_equal(gpu.gradients(), orc.grad.reshape(gpu.gradients().shape)), s
        gpu.step(dt)
        orc.step(dt)
        assert np.array_equal(gpu.state(), orc.q), s
    say("learned step %dD: bits equal, mean fraction %.3f" % (dim, float(frac.mean())))


def test_shipped_network_steps_and_conserves():
    """The shipped network on the smooth KH layer (moving mesh): every step conserves mass, momentum and energy, and
    the gradients are never steeper than the face-centroid limiter's."""
    from luna_hydro import learned, mesh as meshlib

    m = learned.KH_SMOOTH.mesh(32, 3)
    q = learned.KH_SMOOTH.state(m, 32)
    fv = hydro.FiniteVolume(m, GAMMA, limiter="learned", motion="moving", step_rule="luna")
    ref = hydro.FiniteVolume(m, GAMMA, limiter="face", step_rule="luna")
    fv.set_limiter_network(learned.default_params())
    fv.set_state(q)
    total = q.sum(axis=0)
    for s in range(10):
        if s:
            ref.set_geometry(m)
        ref.set_state(fv.state())
        dt = fv.time_step(0.8)
        ref.time_step(0.8)
        g, g0 = fv.gradients().reshape(-1, 4, 2), ref.gradients().reshape(-1, 4, 2)
        ratio = np.linalg.norm(g, axis=2) / np.maximum(np.linalg.norm(g0, axis=2), 1e-300)
        assert np.all(ratio <= 1 + 1e-12), s
        x = fv.moving_begin(dt)
        m = meshlib.build(x, 1.0)
        fv.moving_finish(m, dt)
        assert np.allclose(fv.state().sum(axis=0), total, rtol=1e-12, atol=1e-12), s
    say("shipped network: 10 moving steps, mean gradient ratio %.3f" % float(ratio.mean()))
