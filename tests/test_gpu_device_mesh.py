"""The mesh in device memory (luna_fv_create_device / luna_fv_set_geometry_device): the arrays a GPU mesh producer leaves on
the device, read in place by the geometry pass. Steps on a device mesh equal the steps on the same mesh given as host
arrays and the oracle's, to the bit, static and moving, 2D and 3D; a malformed mesh is refused (its checks read nothing
out of range) and the handle refuses to step until a valid mesh is set; mesh_stats counts only the moving steps' meshes."""
import numpy as np
import pytest

import cases
import oracle
from luna_hydro import hydro, mesh

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("dim,n", [(3, 10), (2, 24)])
def test_static_device_mesh_equals_host_mesh_and_oracle(dim, n):
    say("device mesh, static, dim=%d n=%d" % (dim, n))
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    dev = hydro.FiniteVolume(hydro.DeviceMesh(m), GAMMA)
    host = hydro.FiniteVolume(m, GAMMA)
    orc = oracle.FiniteVolume(m, GAMMA)
    try:
        for fv in (dev, host, orc):
            fv.set_state(q)
        assert np.array_equal(dev.primitives()[:, -1], host.primitives()[:, -1])
        for s in range(3):
            dt = orc.time_step(0.3)
            assert dev.time_step(0.3) == dt and host.time_step(0.3) == dt
            dev.step(dt)
            host.step(dt)
            orc.step(dt)
            assert np.array_equal(dev.state(), host.state())
            assert np.array_equal(dev.state(), orc.q), s
    finally:
        dev.close()
        host.close()


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 20)])
def test_moving_device_mesh_equals_host_mesh_and_oracle(dim, n):
    say("device mesh, moving, dim=%d n=%d" % (dim, n))
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    dev = hydro.FiniteVolume(hydro.DeviceMesh(m), GAMMA, motion="moving")
    host = hydro.FiniteVolume(m, GAMMA, motion="moving")
    orc = oracle.FiniteVolume(m, GAMMA, moving=True)
    try:
        for fv in (dev, host, orc):
            fv.set_state(q)
        for s in range(3):
            dt = orc.time_step(0.3)
            assert dev.time_step(0.3) == dt and host.time_step(0.3) == dt
            x = orc.moving_begin(dt)
            assert np.array_equal(dev.moving_begin(dt), x) and np.array_equal(host.moving_begin(dt), x)
            nm = mesh.build(x, 1.0)
            dev.set_geometry(hydro.DeviceMesh(nm))
            dev.moving_finish(None, dt)
            host.moving_finish(nm, dt)
            orc.moving_finish(nm, dt)
            assert np.array_equal(dev.state(), host.state())
            assert np.array_equal(dev.state(), orc.q), s
        assert dev.mesh_updates() == 3 and host.mesh_updates() == 3
    finally:
        dev.close()
        host.close()


def malformed(m, what):
    bad = dict(m)
    if what == "list":
        lst = m["list"].copy()
        lst[7] = m["faces"] + 5
        bad["list"] = lst
    elif what == "owner":
        own = m["owner"].copy()
        own[3] = m["cells"] + 2
        bad["owner"] = own
    elif what == "start":
        st = m["start"].copy()
        st[5] = st[6] + 1  # decreasing
        bad["start"] = st
    return bad


@pytest.mark.parametrize("what", ["list", "owner", "start"])
def test_malformed_device_mesh_is_refused(what):
    say("malformed device mesh: %s" % what)
    m = cases.voronoi(8, 3)
    fv = hydro.FiniteVolume(hydro.DeviceMesh(m), GAMMA)
    try:
        fv.set_state(cases.rough_state(m))
        with pytest.raises(ValueError):
            fv.set_geometry(hydro.DeviceMesh(malformed(m, what)))
        with pytest.raises(RuntimeError):
            fv.time_step(0.3)
        fv.set_geometry(hydro.DeviceMesh(m))
        orc = oracle.FiniteVolume(m, GAMMA)
        orc.set_state(cases.rough_state(m))
        dt = orc.time_step(0.3)
        assert fv.time_step(0.3) == dt
        fv.step(dt)
        orc.step(dt)
        assert np.array_equal(fv.state(), orc.q)
    finally:
        fv.close()


def test_mesh_stats_counts_moving_updates_only():
    m = cases.voronoi(16, 2)
    a = hydro.FiniteVolume(m, GAMMA)
    c = hydro.FiniteVolume(m, GAMMA, motion="moving")
    try:
        a.set_geometry(m)
        a.set_geometry(hydro.DeviceMesh(m))
        assert a.mesh_updates() == 0
        c.set_state(cases.rough_state(m))
        c.set_geometry(m)  # mesh_moved(): not a mesh update of a step
        assert c.mesh_updates() == 0
        dt = c.time_step(0.3)
        c.moving_finish(mesh.build(c.moving_begin(dt), 1.0), dt)
        assert c.mesh_updates() == 1
    finally:
        a.close()
        c.close()


def test_lists_without_the_owner_bit_step_the_same():
    """The reference's owner_layout (finite_volume.hpp:78-87): a mesh whose lists carry plain face ids gets the owner bit
    from owner[], with the same bits as the mesh that carries it; a list entry naming a face of other cells is refused."""
    import cases

    m = cases.voronoi(8, 3)
    q = cases.rough_state(m)
    plain = dict(m)
    lst = np.asarray(m["list"]).copy()
    plain["list"] = np.where(lst < 0, ~lst, lst).astype(lst.dtype)
    assert np.any(lst < 0) and not np.any(plain["list"] < 0)
    runs = []
    for mm in (m, plain):
        fv = hydro.FiniteVolume(mm, GAMMA)
        fv.set_state(q)
        dt = fv.time_step(0.3)
        fv.step(dt)
        runs.append((dt, fv.gradients(), fv.state()))
        fv.close()
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])
    bad = dict(m)
    bl = lst.copy()
    start = np.asarray(m["start"])
    far = next(f for f in range(m["faces"]) if m["owner"][f] != 0 and m["neighbour"][f] != 0)
    bl[start[0]] = far
    bad["list"] = bl
    with pytest.raises(ValueError):
        hydro.FiniteVolume(bad, GAMMA)


def test_release_mesh_copies_a_device_meshs_face_arrays():
    """The step reads a device mesh's face arrays in place (as finite_volume.hpp:257-327 reads the mesh it holds);
    release_mesh() copies them into the handle, after which the caller's arrays are not read: overwritten with zeros
    (in range, so nothing is read out of bounds if they were), the steps still equal the oracle's to the bit."""
    say("release_mesh on a static device mesh")
    m = cases.voronoi(10, 3)
    q = cases.rough_state(m)
    dm = hydro.DeviceMesh(m)
    dev = hydro.FiniteVolume(dm, GAMMA)
    orc = oracle.FiniteVolume(m, GAMMA)
    try:
        dev.set_state(q)
        orc.set_state(q)
        dev.release_mesh()
        for t in dm.tensors.values():
            t.zero_()
        hydro.synchronize()
        del dm
        for s in range(2):
            dt = orc.time_step(0.3)
            assert dev.time_step(0.3) == dt
            dev.step(dt)
            orc.step(dt)
            assert np.array_equal(dev.state(), orc.q), s
    finally:
        dev.close()


def test_derivatives_on_a_device_mesh_equal_the_host_meshs():
    """On a device mesh the derivatives' face records are built on their first use (from the face arrays read in place):
    the tangent and the adjoint equal those of the same mesh given as host arrays, to the bit."""
    say("derivatives on a device mesh")
    m = cases.voronoi(8, 3)
    q = cases.rough_state(m)
    dev = hydro.FiniteVolume(hydro.DeviceMesh(m), GAMMA)
    host = hydro.FiniteVolume(m, GAMMA)
    try:
        for fv in (dev, host):
            fv.set_state(q)
        dt = host.time_step(0.3)
        assert dev.time_step(0.3) == dt
        rng = np.random.default_rng(5)
        qbar = rng.standard_normal(np.shape(q))
        assert np.array_equal(dev.step_adjoint(dt, qbar), host.step_adjoint(dt, qbar))
        dq = rng.standard_normal(np.shape(q))
        assert np.array_equal(dev.step_tangent(dt, dq), host.step_tangent(dt, dq))
    finally:
        dev.close()
        host.close()
