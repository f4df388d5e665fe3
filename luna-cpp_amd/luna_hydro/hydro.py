"""Python mirror of hydro::FiniteVolume<D> over the C-ABI of libluna_hydro.so (include/luna_hydro.h).

The names and argument meanings follow luna-cpp's src/hydro/finite_volume.hpp (D = 2, 3) and finite_volume_1d.hpp (D = 1):
state(), time_step(courant), step(dt), fallbacks(), positions(), point_velocities(); a moving step is
moving_begin(dt) -> (the caller builds the mesh of the returned positions) -> moving_finish(mesh, dt), the split at the
reference's call into the mesh library (move(), finite_volume.hpp:1212-1250). Errors raise ValueError for the
reference's std::invalid_argument, RuntimeError otherwise. There is no CPU fallback: without the HIP library or a GPU
every call fails.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_longlong)

OK, ERR_ARGUMENT, ERR_LOGIC, ERR_DEVICE, ERR_MEMORY = range(5)
LIMITERS = {"face": 0, "none": 1, "minmod": 2, "vanleer": 3, "learned": 4}
LIMITER_FEATURES, LIMITER_PARAMS = 16, 5313
RIEMANN = {"hllc": 0, "hll": 1}
MOTION = {"static": 0, "moving": 1}
RULES = {"arepo": 0, "luna": 1}
WALLS = {"reflective": 0, "outflow": 1, "inflow": 2}
PRECISIONS = {"double": 0, "grad32": 1, "grad16": 2}  # hydro::Precision's gradient slot (precision.hpp:41-54)
UPDATES = {"per_face": 0, "per_cell": 1}  # hydro::Update (finite_volume.hpp:25)


def policy(name, dim):
    """The drivers' --precision: default (DefaultPrecision<D>, precision.hpp:49-54: grad16 in 3D, grad32 in 2D), double,
    grad32 or grad16."""
    if name == "default":
        return "grad16" if dim == 3 else "grad32"
    if name not in PRECISIONS:
        raise ValueError("unknown --precision %s (default, double, grad32 or grad16)" % name)
    return name


class FVOptions(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("order", ctypes.c_int), ("limiter", ctypes.c_int), ("riemann", ctypes.c_int),
                ("motion", ctypes.c_int), ("step_rule", ctypes.c_int), ("gamma", ctypes.c_double),
                ("smallr", ctypes.c_double), ("smallc", ctypes.c_double), ("precision", ctypes.c_int),
                ("hll_compression", ctypes.c_double), ("update", ctypes.c_int), ("sweep_lanes", ctypes.c_int)]


class Geometry(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("cells", ctypes.c_int), ("faces", ctypes.c_int),
                ("owner", _ip), ("neighbour", _ip), ("start", _ip), ("list", _ip),
                ("area", _dp), ("face_centroid", _dp), ("centroid", _dp), ("volume", _dp), ("positions", _dp),
                ("side", ctypes.c_double * 3), ("step", ctypes.c_double), ("period", ctypes.c_longlong * 3)]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_longlong), ("total_ms", ctypes.c_double)]


class LineRecord(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_int), ("keys", _lp), ("period", ctypes.c_longlong), ("step", ctypes.c_double),
                ("periodic", ctypes.c_int), ("wall_left", ctypes.c_int), ("wall_right", ctypes.c_int)]


class MovingLineRecord(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_int), ("positions", _dp), ("side", ctypes.c_double), ("period", ctypes.c_longlong),
                ("periodic", ctypes.c_int), ("wall_left", ctypes.c_int), ("wall_right", ctypes.c_int),
                ("inflow_left", ctypes.c_double * 3), ("inflow_right", ctypes.c_double * 3), ("spacing", ctypes.c_double),
                ("regularise", ctypes.c_int)]


class FV1Options(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("limiter", ctypes.c_int), ("riemann", ctypes.c_int),
                ("gamma", ctypes.c_double), ("smallr", ctypes.c_double), ("smallc", ctypes.c_double)]


def library_path():
    """The built library; LUNA_HYDRO_LIBRARY names another build of the same sources (A/B timing of tuning knobs)."""
    return os.environ.get("LUNA_HYDRO_LIBRARY") or os.path.join(HERE, "libluna_hydro.so")


def lib():
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError("libluna_hydro.so is not built (run __graft_entry__.build()); there is no CPU fallback")
        # A process holds one libamdhip64.so.7, the first one loaded. torch ships its own and does not initialise on
        # /opt/rocm's ("No HIP GPUs are available"), while this library runs on either: load torch's first when torch
        # is installed, so that torch tensors and these entry points can share the process in any import order.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(path)
        L.luna_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc == OK:
        return
    msg = "%s: %s" % (what, lib().luna_last_error().decode(errors="replace"))
    if rc == ERR_ARGUMENT:
        raise ValueError(msg)
    raise RuntimeError(msg)


def synchronize():
    _check(lib().luna_synchronize(), "synchronize")


def exported_symbols():
    """The entry points include/luna_hydro.h declares, as the loaded library exports them."""
    names = ["luna_fv_create", "luna_fv_destroy", "luna_fv_set_state", "luna_fv_get_state", "luna_fv_state_device",
             "luna_fv_time_step", "luna_fv_step", "luna_fv_moving_begin", "luna_fv_moving_finish", "luna_fv_fallbacks",
             "luna_fv_positions", "luna_fv_point_velocities", "luna_fv_gradients", "luna_fv_primitives", "luna_fv_faces", "luna_fv_profile",
             "luna_fv_profile_read", "luna_fv1_create", "luna_fv1_destroy", "luna_fv1_set_state", "luna_fv1_get_state",
             "luna_fv1_time_step", "luna_fv1_step", "luna_fv1_step_and_time_step", "luna_fv1_step_tangent", "luna_fv1_step_adjoint",
             "luna_fv1_create_moving", "luna_fv1_cells", "luna_fv1_positions", "luna_fv1_line", "luna_fv1_set_inflow_spacing",
             "luna_fv1_mesh_stats",
             "luna_last_error",
             "luna_fv_set_geometry", "luna_fv_set_geometry_device", "luna_fv_create_device", "luna_synchronize", "luna_fv_step_tangent", "luna_fv_step_adjoint", "luna_fv_step_adjoint_noalloc", "luna_fv_moving_adjoint",
             "luna_fv_time_step_tangent", "luna_fv_moving_tangent", "luna_fv_set_step_rule", "luna_fv_release_mesh", "luna_fv_mesh_stats",
             "luna_fv_set_limiter_network", "luna_fv_limiter_features", "luna_mlp_create", "luna_mlp_destroy", "luna_mlp_set_params",
             "luna_mlp_get_params", "luna_mlp_forward", "luna_mlp_gradient", "luna_mlp_train_step", "luna_mlp_synchronize"]
    return {n: hasattr(lib(), n) for n in names}


def geometry_record(mesh):
    """A luna_geometry over the arrays of a mesh dict (kept alive by the returned holder)."""
    m = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in mesh.items()}
    for k in ("owner", "neighbour", "start", "list"):
        m[k] = np.ascontiguousarray(m[k], dtype=np.int32)
    for k in ("area", "face_centroid", "centroid", "volume", "positions"):
        m[k] = np.ascontiguousarray(m[k], dtype=np.float64)
    g = Geometry()
    g.dim, g.cells, g.faces = m["dim"], m["cells"], m["faces"]
    g.owner, g.neighbour = m["owner"].ctypes.data_as(_ip), m["neighbour"].ctypes.data_as(_ip)
    g.start, g.list = m["start"].ctypes.data_as(_ip), m["list"].ctypes.data_as(_ip)
    for k in ("area", "face_centroid", "centroid", "volume", "positions"):
        setattr(g, k, m[k].ctypes.data_as(_dp))
    for a in range(3):
        g.side[a] = m["side"][a] if a < len(m["side"]) else 0.0
        g.period[a] = int(m["period"][a]) if a < len(m["period"]) else 0
    g.step = m["step"]
    return g, m


class DeviceMesh:
    """A mesh's arrays in GPU memory (torch tensors), as a GPU mesh producer leaves them (voronoi-cpp's device views): the
    step reads them in place with set_geometry_device / FiniteVolume(DeviceMesh(...)). Built from a mesh dict by copying
    its arrays to the device once; `record` is the luna_geometry of device pointers."""

    def __init__(self, mesh, device="cuda"):
        import torch

        self.dim, self.cells, self.faces = mesh["dim"], mesh["cells"], mesh["faces"]
        self.tensors = {}
        for k in ("owner", "neighbour", "start", "list"):
            self.tensors[k] = torch.as_tensor(np.ascontiguousarray(mesh[k], dtype=np.int32)).to(device)
        for k in ("area", "face_centroid", "centroid", "volume", "positions"):
            self.tensors[k] = torch.as_tensor(np.ascontiguousarray(mesh[k], dtype=np.float64)).to(device)
        torch.cuda.synchronize()
        g = Geometry()
        g.dim, g.cells, g.faces = self.dim, self.cells, self.faces
        for k in ("owner", "neighbour", "start", "list"):
            setattr(g, k, ctypes.cast(ctypes.c_void_p(self.tensors[k].data_ptr()), _ip))
        for k in ("area", "face_centroid", "centroid", "volume", "positions"):
            setattr(g, k, ctypes.cast(ctypes.c_void_p(self.tensors[k].data_ptr()), _dp))
        for a in range(3):
            g.side[a] = mesh["side"][a] if a < len(mesh["side"]) else 0.0
            g.period[a] = int(mesh["period"][a]) if a < len(mesh["period"]) else 0
        g.step = mesh["step"]
        self.record = g

    def __getitem__(self, key):
        return {"dim": self.dim, "cells": self.cells, "faces": self.faces}[key]


class FiniteVolume:
    """hydro::FiniteVolume<D> (D = 2, 3) on the GPU: FiniteVolume(exec, mesh, gas, order, per_face, limiter, motion).
    mesh: a mesh dict (host arrays) or a DeviceMesh (read in place on the GPU). sweep_lanes: the reference's
    LUNA_SWEEP_LANES (LaneSweep with 2, 4 or 8 lanes per cell; 1 the one-lane sweep)."""

    def __init__(self, mesh, gamma, order=1, limiter="face", riemann="hllc", motion="static", step_rule="arepo",
                 smallr=1e-10, smallc=1e-10, precision="double", hll_compression=0.0, update="per_face", sweep_lanes=1):
        self.D, self.N = mesh["dim"], mesh["cells"]
        self.V = self.D + 2
        opt = FVOptions(self.D, order, LIMITERS[limiter], RIEMANN[riemann], MOTION[motion], RULES[step_rule],
                        gamma, smallr, smallc, PRECISIONS[precision], hll_compression, UPDATES[update], sweep_lanes)
        h = ctypes.c_void_p()
        if isinstance(mesh, DeviceMesh):
            self._mesh_arrays = mesh
            _check(lib().luna_fv_create_device(ctypes.byref(opt), ctypes.byref(mesh.record), ctypes.byref(h)), "FiniteVolume")
        else:
            g, self._mesh_arrays = geometry_record(mesh)
            _check(lib().luna_fv_create(ctypes.byref(opt), ctypes.byref(g), ctypes.byref(h)), "FiniteVolume")
        self._h = h
        self.motion = motion

    def close(self):
        if getattr(self, "_h", None):
            lib().luna_fv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.shape == (self.N, self.V)
        _check(lib().luna_fv_set_state(self._h, q.ctypes.data_as(_dp)), "set_state")

    def state(self):
        q = np.empty((self.N, self.V))
        _check(lib().luna_fv_get_state(self._h, q.ctypes.data_as(_dp)), "state")
        return q

    def time_step(self, courant):
        dt = ctypes.c_double()
        _check(lib().luna_fv_time_step(self._h, ctypes.c_double(courant), ctypes.byref(dt)), "time_step")
        return dt.value

    def step(self, dt):
        _check(lib().luna_fv_step(self._h, ctypes.c_double(dt)), "step")

    def set_step_rule(self, rule):
        """set_step_rule(StepRule) (finite_volume.hpp:181): "arepo", or "luna" in 2D."""
        _check(lib().luna_fv_set_step_rule(self._h, RULES[rule]), "set_step_rule")

    def release_mesh(self):
        """release_mesh() (finite_volume.hpp:134-139): the handle copies a device mesh's face arrays, which it read in place,
        and lets the mesh go; refused on a moving mesh."""
        _check(lib().luna_fv_release_mesh(self._h), "release_mesh")
        if isinstance(self._mesh_arrays, DeviceMesh):
            self._mesh_arrays = None

    def mesh_updates(self):
        """mesh_stats().updates: the meshes moving steps took (between moving_begin and moving_finish)."""
        n = ctypes.c_longlong()
        _check(lib().luna_fv_mesh_stats(self._h, ctypes.byref(n)), "mesh_stats")
        return n.value

    def time_step_tangent(self, courant, dq):
        """FiniteVolume::time_step_tangent(courant, dq): dt as time_step(), with the primitives' and the limited fit's
        tangents for dq, which the next step_tangent(dt, dq) with the same dq reuses. dq: N x V, host (numpy) or device
        (an object with data_ptr())."""
        dt = ctypes.c_double()
        if hasattr(dq, "data_ptr"):
            if not dq.is_contiguous() or dq.numel() != self.N * self.V:
                raise ValueError("time_step_tangent: a contiguous N x V device array")
            ptr, device = ctypes.c_void_p(dq.data_ptr()), 1
        else:
            rows = np.ascontiguousarray(np.array(dq, dtype=np.float64))
            if rows.size != self.N * self.V:
                raise ValueError("time_step_tangent: N x V rows")
            ptr, device = rows.ctypes.data_as(ctypes.c_void_p), 0
        _check(lib().luna_fv_time_step_tangent(self._h, ctypes.c_double(courant), ptr, device, ctypes.byref(dt)), "time_step_tangent")
        return dt.value

    def step_tangent(self, dt, dq):
        """FiniteVolume::step_tangent(): one static step of Q and of the tangent dq (N x V, extensive); returns J dq. dq: a
        numpy array (host) or an object with a device data_ptr() (e.g. a torch tensor on the GPU, updated in place)."""
        return self._derivative(lib().luna_fv_step_tangent, dt, dq, "step_tangent")

    def step_adjoint(self, dt, qbar):
        """FiniteVolume::step_adjoint(): J^T qbar of one static step from the current Q (which stays); qbar as step_tangent's dq."""
        return self._derivative(lib().luna_fv_step_adjoint, dt, qbar, "step_adjoint")

    def step_adjoint_noalloc(self, dt, qbar):
        """luna_fv_step_adjoint_noalloc: step_adjoint on a device array (an object with data_ptr(), e.g. a torch tensor on the
        GPU, N x V contiguous, overwritten with J^T qbar) with the library's scratch for its work rows: no allocation."""
        if not hasattr(qbar, "data_ptr") or not qbar.is_contiguous() or qbar.numel() != self.N * self.V:
            raise ValueError("step_adjoint_noalloc: a contiguous N x V device array")
        _check(lib().luna_fv_step_adjoint_noalloc(self._h, ctypes.c_double(dt), ctypes.c_void_p(qbar.data_ptr())), "step_adjoint_noalloc")
        return qbar

    def moving_adjoint(self, dt, mesh0, mesh1, qbar, xbar, steps_fixed=False):
        """FiniteVolume::step_moving_adjoint() (luna_fv_moving_adjoint; order 1, or order 0, where mesh1 is not read): the
        handle's Q is the state at t^n, mesh0
        the mesh of t^n, mesh1 the mesh the forward step reached; qbar (N x V) and xbar (N x D), the bars after the step,
        -> those before it, with the project's mesh library's geometry derivatives of both meshes. Leaves the handle on
        mesh0."""
        from . import mesh as mesh_lib

        g0, a0 = geometry_record(mesh0)
        g1, a1 = geometry_record(mesh1)
        d0, d1 = mesh_lib.GeometryDerivative(mesh0), mesh_lib.GeometryDerivative(mesh1)
        fn = ctypes.cast(mesh_lib.lib().luna_mesh_derivative_adjoint, ctypes.c_void_p)
        qb = np.ascontiguousarray(np.array(qbar, dtype=np.float64).reshape(self.N, self.V))
        xb = np.ascontiguousarray(np.array(xbar, dtype=np.float64).reshape(self.N, self.D))
        try:
            _check(lib().luna_fv_moving_adjoint(self._h, ctypes.c_double(dt), ctypes.byref(g0), ctypes.byref(g1), fn, d0._h, fn, d1._h,
                                                int(steps_fixed), qb.ctypes.data_as(_dp), xb.ctypes.data_as(_dp)), "moving_adjoint")
        finally:
            d0.close()
            d1.close()
        self._mesh_arrays = a0
        return qb, xb

    def moving_tangent(self, dt, mesh0, mesh1, dq, dx, steps_fixed=False):
        """FiniteVolume::step_moving_tangent() (luna_fv_moving_tangent; order 1, or order 0, where mesh1 is not read): the
        handle's Q is the state at t^n, mesh0 the mesh of t^n, mesh1 the mesh the forward step reached; dq (N x V) and dx
        (N x D), the tangents at t^n, -> those at t^n+1, with the project's mesh library's geometry tangents of both meshes.
        Leaves the handle on mesh0 with Q unchanged."""
        from . import mesh as mesh_lib

        g0, a0 = geometry_record(mesh0)
        g1, a1 = geometry_record(mesh1)
        d0, d1 = mesh_lib.GeometryDerivative(mesh0), mesh_lib.GeometryDerivative(mesh1)
        fn = ctypes.cast(mesh_lib.lib().luna_mesh_derivative_tangent, ctypes.c_void_p)
        tq = np.ascontiguousarray(np.array(dq, dtype=np.float64).reshape(self.N, self.V))
        tx = np.ascontiguousarray(np.array(dx, dtype=np.float64).reshape(self.N, self.D))
        try:
            _check(lib().luna_fv_moving_tangent(self._h, ctypes.c_double(dt), ctypes.byref(g0), ctypes.byref(g1), fn, d0._h, fn, d1._h,
                                                int(steps_fixed), tq.ctypes.data_as(_dp), tx.ctypes.data_as(_dp)), "moving_tangent")
        finally:
            d0.close()
            d1.close()
        self._mesh_arrays = a0
        return tq, tx

    def _derivative(self, fn, dt, rows, what):
        if hasattr(rows, "data_ptr"):
            if not rows.is_contiguous() or rows.numel() != self.N * self.V:
                raise ValueError("%s: a contiguous N x V device array" % what)
            _check(fn(self._h, ctypes.c_double(dt), ctypes.c_void_p(rows.data_ptr()), 1), what)
            return rows
        out = np.ascontiguousarray(np.array(rows, dtype=np.float64))
        if out.size != self.N * self.V:
            raise ValueError("%s: N x V rows" % what)
        _check(fn(self._h, ctypes.c_double(dt), out.ctypes.data_as(_dp), 0), what)
        return out

    def set_limiter_network(self, params):
        """The learned limiter's network (LIMITER_PARAMS FP32 values), or None to remove it (then the learned mode limits
        as the face-centroid limiter and still records features)."""
        if params is None:
            _check(lib().luna_fv_set_limiter_network(self._h, None, 0), "set_limiter_network")
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib().luna_fv_set_limiter_network(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size),
               "set_limiter_network")

    def limiter_features(self):
        """The last sweep's learned-limiter features, (N V) x 16 BF16 bit patterns (uint16), and troubled flags (N, uint8)."""
        x = np.empty((self.N * self.V, LIMITER_FEATURES), dtype=np.uint16)
        t = np.empty(self.N, dtype=np.uint8)
        _check(lib().luna_fv_limiter_features(self._h, x.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort)),
                                              t.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))), "limiter_features")
        return x, t

    def moving_begin(self, dt, positions=True):
        """The first half of a moving step; returns the new positions (None with positions=False)."""
        x = np.empty((self.N, self.D)) if positions else None
        ptr = x.ctypes.data_as(_dp) if positions else None
        _check(lib().luna_fv_moving_begin(self._h, ctypes.c_double(dt), ptr), "moving_begin")
        return x

    def set_geometry(self, mesh):
        """set_geometry with a mesh dict (host arrays) or a DeviceMesh (luna_fv_set_geometry_device; the step reads its face
        arrays in place until the next mesh, so the handle keeps a reference to it)."""
        if isinstance(mesh, DeviceMesh):
            _check(lib().luna_fv_set_geometry_device(self._h, ctypes.byref(mesh.record)), "set_geometry_device")
            self._mesh_arrays = mesh
            return
        g, arrays = geometry_record(mesh)
        _check(lib().luna_fv_set_geometry(self._h, ctypes.byref(g)), "set_geometry")
        self._mesh_arrays = arrays

    def moving_finish(self, mesh, dt):
        """The second half: mesh is the mesh of the new positions, or None after set_geometry()."""
        if mesh is None:
            _check(lib().luna_fv_moving_finish(self._h, None, ctypes.c_double(dt)), "moving_finish")
            return
        g, arrays = geometry_record(mesh)
        _check(lib().luna_fv_moving_finish(self._h, ctypes.byref(g), ctypes.c_double(dt)), "moving_finish")
        self._mesh_arrays = arrays

    def fallbacks(self):
        c = ctypes.c_longlong()
        _check(lib().luna_fv_fallbacks(self._h, ctypes.byref(c)), "fallbacks")
        return c.value

    def positions(self):
        x = np.empty((self.N, self.D))
        _check(lib().luna_fv_positions(self._h, x.ctypes.data_as(_dp)), "positions")
        return x

    def point_velocities(self):
        w = np.empty((self.N, self.D))
        _check(lib().luna_fv_point_velocities(self._h, w.ctypes.data_as(_dp)), "point_velocities")
        return w

    def gradients(self):
        g = np.empty((self.N, self.V * self.D))
        _check(lib().luna_fv_gradients(self._h, g.ctypes.data_as(_dp)), "gradients")
        return g

    def primitives(self):
        p = np.empty((self.N, self.V + 2))
        _check(lib().luna_fv_primitives(self._h, p.ctypes.data_as(_dp)), "primitives")
        return p

    def faces(self):
        f = ctypes.c_int()
        _check(lib().luna_fv_faces(self._h, ctypes.byref(f)), "faces")
        return f.value

    def profile(self, on=True):
        _check(lib().luna_fv_profile(self._h, int(on)), "profile")

    def kernel_times(self):
        out = (KernelTime * 32)()
        n = ctypes.c_int()
        _check(lib().luna_fv_profile_read(self._h, out, 32, ctypes.byref(n)), "profile_read")
        return {out[k].name.decode(): (out[k].launches, out[k].total_ms) for k in range(min(n.value, 32))}


class Line:
    """hydro::FiniteVolume<1> on a static sorted line of grid keys (walls or periodic)."""

    def __init__(self, keys, period, step, gamma, periodic=False, walls=("outflow", "outflow"), order=1,
                 limiter="face", riemann="hllc", smallr=1e-10, smallc=1e-10):
        self.keys = np.ascontiguousarray(np.asarray(keys, dtype=np.int64))
        self.n = len(self.keys)
        rec = LineRecord(self.n, self.keys.ctypes.data_as(_lp), int(period), float(step), int(periodic),
                         WALLS[walls[0]], WALLS[walls[1]])
        opt = FV1Options(order, LIMITERS[limiter], RIEMANN[riemann], gamma, smallr, smallc)
        h = ctypes.c_void_p()
        _check(lib().luna_fv1_create(ctypes.byref(opt), ctypes.byref(rec), ctypes.byref(h)), "FiniteVolume<1>")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().luna_fv1_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.shape == (self.n, 3)
        _check(lib().luna_fv1_set_state(self._h, q.ctypes.data_as(_dp)), "set_state")

    def state(self):
        q = np.empty((self.n, 3))
        _check(lib().luna_fv1_get_state(self._h, q.ctypes.data_as(_dp)), "state")
        return q

    def time_step(self, courant, rule="luna"):
        dt = ctypes.c_double()
        _check(lib().luna_fv1_time_step(self._h, ctypes.c_double(courant), RULES[rule], ctypes.byref(dt)), "time_step")
        return dt.value

    def step(self, dt):
        _check(lib().luna_fv1_step(self._h, ctypes.c_double(dt)), "step")

    def step_tangent(self, dt, dq):
        """FiniteVolume<1>::step_tangent(): one step of Q and of the tangent dq (cells x 3, extensive); returns J dq."""
        dq = np.ascontiguousarray(np.array(dq, dtype=np.float64))
        _check(lib().luna_fv1_step_tangent(self._h, ctypes.c_double(dt), dq.ctypes.data_as(_dp)), "step_tangent")
        return dq

    def step_adjoint(self, dt, qbar):
        """FiniteVolume<1>::step_adjoint(): J^T qbar of one step from the current Q (which stays)."""
        qbar = np.ascontiguousarray(np.array(qbar, dtype=np.float64))
        _check(lib().luna_fv1_step_adjoint(self._h, ctypes.c_double(dt), qbar.ctypes.data_as(_dp)), "step_adjoint")
        return qbar

    def step_and_time_step(self, dt, courant):
        nxt = ctypes.c_double()
        _check(lib().luna_fv1_step_and_time_step(self._h, ctypes.c_double(dt), ctypes.c_double(courant),
                                                  ctypes.byref(nxt)), "step_and_time_step")
        return nxt.value


class MovingLine(Line):
    """hydro::FiniteVolume<1> on a moving line (Motion::moving): n generators at positions x on [0, side] (walls) or a
    periodic line with `period` grid keys; walls as ("outflow", "inflow", ...) with the inflow primitives in `inflow`
    (left, right); open walls supply cells `spacing` apart. state() and positions() are by cell (new cells after the
    last), line() is (keys, cells) by slot."""

    def __init__(self, x, side, period, gamma, periodic=False, walls=("outflow", "outflow"), inflow=((0, 0, 0), (0, 0, 0)),
                 order=1, limiter="face", riemann="hllc", spacing=0.0, regularise=True, smallr=1e-10, smallc=1e-10):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        rec = MovingLineRecord(len(x), x.ctypes.data_as(_dp), float(side), int(period), int(periodic), WALLS[walls[0]],
                               WALLS[walls[1]], (ctypes.c_double * 3)(*inflow[0]), (ctypes.c_double * 3)(*inflow[1]),
                               float(spacing), int(regularise))
        opt = FV1Options(order, LIMITERS[limiter], RIEMANN[riemann], gamma, smallr, smallc)
        h = ctypes.c_void_p()
        _check(lib().luna_fv1_create_moving(ctypes.byref(opt), ctypes.byref(rec), ctypes.byref(h)), "FiniteVolume<1> (moving)")
        self._h = h
        self.period, self.side, self.periodic = int(period), float(side), bool(periodic)

    @property
    def n(self):
        cells = ctypes.c_int()
        _check(lib().luna_fv1_cells(self._h, ctypes.byref(cells), None), "cells")
        return cells.value

    def cells_added(self):
        added = ctypes.c_longlong()
        _check(lib().luna_fv1_cells(self._h, None, ctypes.byref(added)), "cells_added")
        return added.value

    def positions(self):
        x = np.empty(self.n)
        _check(lib().luna_fv1_positions(self._h, x.ctypes.data_as(_dp)), "positions")
        return x

    def line(self):
        n = self.n
        keys, cells = np.empty(n, dtype=np.int64), np.empty(n, dtype=np.int32)
        _check(lib().luna_fv1_line(self._h, keys.ctypes.data_as(_lp), cells.ctypes.data_as(_ip)), "line")
        return keys, cells

    def set_inflow_spacing(self, spacing):
        _check(lib().luna_fv1_set_inflow_spacing(self._h, ctypes.c_double(spacing)), "set_inflow_spacing")

    def mesh_stats(self):
        updates, sorts = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().luna_fv1_mesh_stats(self._h, ctypes.byref(updates), ctypes.byref(sorts)), "mesh_stats")
        return updates.value, sorts.value

    def sorts(self):
        return self.mesh_stats()[1]

    def time_step(self, courant, rule="arepo"):
        return super().time_step(courant, rule)
