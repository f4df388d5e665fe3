"""ctypes front of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (built by oracle/Makefile) and mirrors the solver's calls on numpy arrays. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the product never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_longlong)


class Gas(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("smallr", ctypes.c_double), ("smallc", ctypes.c_double),
                ("inverse_gm1", ctypes.c_double)]


class _Geometry(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int), ("N", ctypes.c_int), ("F", ctypes.c_int),
                ("owner", _ip), ("neighbour", _ip), ("start", _ip), ("list", _ip),
                ("area", _dp), ("face_centroid", _dp), ("centroid", _dp), ("volume", _dp), ("positions", _dp),
                ("side", ctypes.c_double * 3), ("step", ctypes.c_double), ("period", ctypes.c_longlong * 3)]


class _FV(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int), ("N", ctypes.c_int), ("order", ctypes.c_int), ("limiter", ctypes.c_int),
                ("riemann", ctypes.c_int), ("moving", ctypes.c_int), ("step_rule", ctypes.c_int), ("gas", Gas),
                ("q", _dp), ("prims", _dp), ("grad", _dp), ("w", _dp), ("x", _dp), ("old_centroid", _dp),
                ("steps", _ip), ("flux", _dp), ("fallbacks", ctypes.c_longlong * 2), ("prims_ready", ctypes.c_int),
                ("features", ctypes.POINTER(ctypes.c_ushort)), ("troubled", ctypes.POINTER(ctypes.c_ubyte)),
                ("hold_steps", ctypes.c_int), ("precision", ctypes.c_int), ("hll_compression", ctypes.c_double),
                ("hll", ctypes.POINTER(ctypes.c_ubyte)), ("sweep_lanes", ctypes.c_int)]


_GEO_ADJOINT = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp)
_GEO_TANGENT = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp)


class _Line(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("periodic", ctypes.c_int), ("wall_left", ctypes.c_int),
                ("wall_right", ctypes.c_int), ("order", ctypes.c_int), ("limiter", ctypes.c_int),
                ("riemann", ctypes.c_int), ("gas", Gas), ("key", _lp), ("period", ctypes.c_longlong),
                ("step", ctypes.c_double), ("q", _dp), ("next", _dp), ("moving", ctypes.c_int), ("regularise", ctypes.c_int),
                ("id", _ip), ("side", ctypes.c_double), ("inflow", (ctypes.c_double * 3) * 2)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_time_step.restype = ctypes.c_double
        L.orc_line_time_step.restype = ctypes.c_double
        L.orc_line_step_fused.restype = ctypes.c_double
        L.orc_line_volume.restype = ctypes.c_double
        L.orc_line_centroid.restype = ctypes.c_double
        L.orc_sound_speed.restype = ctypes.c_double
        L.orc_snap.restype = ctypes.c_longlong
        L.orc_half_from_float.restype = ctypes.c_ushort
        L.orc_half_to_float.restype = ctypes.c_float
        L.orc_power_of_two.restype = ctypes.c_float
        L.orc_snap.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_longlong]
        L.orc_delta.argtypes = [ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong]
        L.orc_moving_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                         _GEO_ADJOINT, ctypes.c_void_p, _GEO_ADJOINT, ctypes.c_void_p]
        L.orc_mline_create.restype = ctypes.c_void_p
        L.orc_mline_create.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_mline_destroy.argtypes = [ctypes.c_void_p]
        L.orc_mline_cells.argtypes = [ctypes.c_void_p]
        L.orc_mline_set_state.argtypes = [ctypes.c_void_p, _dp]
        L.orc_mline_get_state.argtypes = [ctypes.c_void_p, _dp]
        L.orc_mline_get_positions.argtypes = [ctypes.c_void_p, _dp]
        L.orc_mline_get_line.argtypes = [ctypes.c_void_p, _lp, _ip]
        L.orc_mline_stat.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_mline_stat.restype = ctypes.c_longlong
        L.orc_mline_time_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.orc_mline_time_step.restype = ctypes.c_double
        L.orc_mline_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_mline_step_fused.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_mline_step_fused.restype = ctypes.c_double
        _LIB = L
    return _LIB


def set_threads(n):
    """The OpenMP threads of the oracle's loops."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def _ptr(a, kind=_dp):
    return a.ctypes.data_as(kind) if a is not None else None


def gas(gamma, smallr=1e-10, smallc=1e-10):
    g = Gas()
    lib().orc_gas_init(ctypes.byref(g), ctypes.c_double(gamma))
    g.smallr, g.smallc = smallr, smallc
    return g


def _vec(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def primitive(U, gamma):
    U = _vec(U)
    W = np.empty_like(U)
    g = gas(gamma)
    lib().orc_primitive(len(U) - 2, ctypes.byref(g), _ptr(U), _ptr(W))
    return W


def conserved(W, gamma):
    W = _vec(W)
    U = np.empty_like(W)
    g = gas(gamma)
    lib().orc_conserved(len(W) - 2, ctypes.byref(g), _ptr(W), _ptr(U))
    return U


def flux(L, R, n, w, gamma, riemann="hllc"):
    L, R, n, w = _vec(L), _vec(R), _vec(n), _vec(w)
    F = np.empty_like(L)
    g = gas(gamma)
    fn = lib().orc_hll if riemann == "hll" else lib().orc_hllc
    fn(len(L) - 2, ctypes.byref(g), _ptr(L), _ptr(R), _ptr(n), _ptr(w), _ptr(F))
    return F


def primitive_tangent(U, dU, gamma, smallr=1e-10, smallc=1e-10):
    U, dU = _vec(U), _vec(dU)
    dW = np.empty_like(U)
    g = gas(gamma, smallr, smallc)
    lib().orc_primitive_tangent(len(U) - 2, ctypes.byref(g), _ptr(U), _ptr(dU), _ptr(dW))
    return dW


def primitive_adjoint(U, Wbar, gamma, smallr=1e-10, smallc=1e-10):
    U, Wbar = _vec(U), _vec(Wbar)
    Ubar = np.empty_like(U)
    g = gas(gamma, smallr, smallc)
    lib().orc_primitive_adjoint(len(U) - 2, ctypes.byref(g), _ptr(U), _ptr(Wbar), _ptr(Ubar))
    return Ubar


def half_from_float(f):
    """precision.hpp half::from_float: the binary16 bits of a float (round to nearest even)."""
    return lib().orc_half_from_float(ctypes.c_float(f))


def half_to_float(h):
    return lib().orc_half_to_float(ctypes.c_ushort(h))


def group_exponent(m):
    return lib().orc_group_exponent(ctypes.c_float(m))


def half_step(W, grad, dt, gamma):
    """W: [V]; grad: [V][D]."""
    W, grad = _vec(W), _vec(grad)
    out = np.empty_like(W)
    g = gas(gamma)
    lib().orc_half_step(len(W) - 2, ctypes.byref(g), _ptr(W), _ptr(grad), ctypes.c_double(dt), _ptr(out))
    return out


def half_step_tangent(W, dW, grad, dgrad, dt, gamma):
    W, dW, grad, dgrad = _vec(W), _vec(dW), _vec(grad), _vec(dgrad)
    out = np.empty_like(W)
    g = gas(gamma)
    lib().orc_half_step_tangent(len(W) - 2, ctypes.byref(g), _ptr(W), _ptr(dW), _ptr(grad), _ptr(dgrad), ctypes.c_double(dt), _ptr(out))
    return out


def half_step_adjoint(W, grad, dt, hbar, gamma):
    """Returns (Wbar, gbar) = J^T hbar."""
    W, grad, hbar = _vec(W), _vec(grad), _vec(hbar)
    Wbar, gbar = np.zeros_like(W), np.zeros_like(grad)
    g = gas(gamma)
    lib().orc_half_step_adjoint(len(W) - 2, ctypes.byref(g), _ptr(W), _ptr(grad), ctypes.c_double(dt), _ptr(hbar), _ptr(Wbar), _ptr(gbar))
    return Wbar, gbar


def hllc_tangent(L, R, dL, dR, n, w, gamma, riemann="hllc"):
    """Returns (F, dF): the flux (HLLC or HLL) from the tangent's own intermediates and J (dL, dR)."""
    L, R, dL, dR, n, w = _vec(L), _vec(R), _vec(dL), _vec(dR), _vec(n), _vec(w)
    F, dF = np.empty_like(L), np.empty_like(L)
    g = gas(gamma)
    (lib().orc_hll_tangent if riemann == "hll" else lib().orc_hllc_tangent)(len(L) - 2, ctypes.byref(g), _ptr(L), _ptr(R), _ptr(dL), _ptr(dR), _ptr(n), _ptr(w), _ptr(F), _ptr(dF))
    return F, dF


def hllc_adjoint(L, R, n, w, Fbar, gamma, riemann="hllc"):
    """Returns (Lbar, Rbar) = J^T Fbar of the flux (HLLC or HLL)."""
    L, R, n, w, Fbar = _vec(L), _vec(R), _vec(n), _vec(w), _vec(Fbar)
    Lbar, Rbar = np.zeros_like(L), np.zeros_like(R)
    g = gas(gamma)
    (lib().orc_hll_adjoint if riemann == "hll" else lib().orc_hllc_adjoint)(len(L) - 2, ctypes.byref(g), _ptr(L), _ptr(R), _ptr(n), _ptr(w), _ptr(Fbar), _ptr(Lbar), _ptr(Rbar))
    return Lbar, Rbar


def face_velocity(d, xf, wL, wR):
    """FaceFlux::face_velocity (finite_volume.hpp:1380-1397): d owner -> neighbour, xf relative to the owner's generator."""
    d, xf, wL, wR = _vec(d), _vec(xf), _vec(wL), _vec(wR)
    wf = np.empty_like(d)
    lib().orc_face_velocity(len(d), _ptr(d), _ptr(xf), _ptr(wL), _ptr(wR), _ptr(wf))
    return wf


def regularised_point_velocity(W_row, centroid, max_angle):
    """The sweep's regularised point velocity (finite_volume.hpp:526-537): W_row = (W, c_s), centroid relative to the
    generator, max_angle the cell's largest face angle."""
    W_row, centroid = _vec(W_row), _vec(centroid)
    w = np.empty_like(centroid)
    lib().orc_regularised_point_velocity(len(centroid), _ptr(W_row), _ptr(centroid), ctypes.c_double(max_angle), _ptr(w))
    return w


def snap(x, step, period):
    return lib().orc_snap(ctypes.c_double(x), ctypes.c_double(step), ctypes.c_longlong(period))


class Geometry:
    """Holds a mesh record (dict of numpy arrays, the voronoi-cpp output contract) as the oracle reads it."""

    def __init__(self, mesh):
        self.mesh = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in mesh.items()}
        m = self.mesh
        g = _Geometry()
        g.D, g.N, g.F = m["dim"], m["cells"], m["faces"]
        g.owner, g.neighbour = _ptr(m["owner"], _ip), _ptr(m["neighbour"], _ip)
        g.start, g.list = _ptr(m["start"], _ip), _ptr(m["list"], _ip)
        g.area, g.face_centroid = _ptr(m["area"]), _ptr(m["face_centroid"])
        g.centroid, g.volume, g.positions = _ptr(m["centroid"]), _ptr(m["volume"]), _ptr(m["positions"])
        for a in range(3):
            g.side[a] = m["side"][a] if a < len(m["side"]) else 0.0
            g.period[a] = int(m["period"][a]) if a < len(m["period"]) else 0
        g.step = m["step"]
        self.c = g


class FiniteVolume:
    """The oracle's FiniteVolume<D> (D = 2, 3): state, time_step, step, and the moving step split at the mesh update."""

    def __init__(self, mesh, gamma, order=1, limiter="face", riemann="hllc", moving=False, step_rule="arepo", precision="double",
                 hll_compression=0.0, sweep_lanes=1):
        D, N = mesh["dim"], mesh["cells"]
        if precision != "double" and limiter == "learned":
            raise ValueError("the learned limiter's features need the double policy")
        if sweep_lanes not in (0, 1, 2, 4, 8) or (sweep_lanes > 1 and (order != 1 or precision != "double" or limiter == "learned")):
            raise ValueError("LaneSweep: 2, 4 or 8 lanes, order 1, the double policy, the face-centroid limiter or none")
        V = D + 2
        self.D, self.N, self.V = D, N, V
        self.q = np.zeros((N, V))
        self.prims = np.zeros((N, V + 2))
        self.grad = np.zeros((N, V * D))
        self.w = np.zeros((N, D))
        self.x = np.array(mesh["positions"], dtype=np.float64).reshape(N, D).copy()
        self.old_centroid = np.zeros((N, D))
        self.steps = np.zeros((N, D), dtype=np.int32)
        self.flux = np.zeros((max(1, mesh["faces"]), V))
        c = _FV()
        c.D, c.N, c.order = D, N, order
        c.limiter = {"face": 0, "none": 1, "learned": 2}[limiter]
        c.riemann = {"hllc": 0, "hll": 1}[riemann]
        c.moving = int(moving)
        c.step_rule = {"arepo": 0, "luna": 1}[step_rule]
        c.precision = {"double": 0, "grad32": 1, "grad16": 2}[precision]
        # the HLL fallback's flags (written by each sweep when hll_compression > 0)
        self.hll = np.zeros(N, dtype=np.uint8)
        c.hll_compression = hll_compression
        c.hll = self.hll.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
        c.sweep_lanes = sweep_lanes
        c.gas = gas(gamma)
        c.q, c.prims, c.grad, c.w = _ptr(self.q), _ptr(self.prims), _ptr(self.grad), _ptr(self.w)
        c.x, c.old_centroid, c.steps, c.flux = _ptr(self.x), _ptr(self.old_centroid), _ptr(self.steps, _ip), _ptr(self.flux)
        c.prims_ready = 0
        # the learned limiter's record of each sweep (limiter "learned"; its gradients are the face limiter's until
        # apply_fractions scales them)
        self.features = np.zeros((N * V, 16), dtype=np.uint16)
        self.troubled = np.zeros(N, dtype=np.uint8)
        c.features = self.features.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort))
        c.troubled = self.troubled.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
        self.c = c
        self.set_mesh(mesh)

    def set_mesh(self, mesh):
        self.geo = Geometry(mesh)
        if mesh["faces"] > self.flux.shape[0]:
            self.flux = np.zeros((mesh["faces"], self.V))
            self.c.flux = _ptr(self.flux)

    def set_state(self, q):
        self.q[...] = q
        self.c.prims_ready = 0

    def time_step(self, courant):
        return lib().orc_time_step(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.c_double(courant))

    def step(self, dt):
        lib().orc_step_static(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.c_double(dt))

    def apply_fractions(self, fraction):
        """The learned limiter's scaling of the last sweep's gradients: row i V + v of fraction (float32, N V) scales
        variable v of cell i, in double, unless the cell is troubled (the device's k_limiter_apply)."""
        frac = np.asarray(fraction, dtype=np.float32).reshape(self.N, self.V).astype(np.float64)
        keep = self.troubled.astype(bool)
        frac[keep] = 1.0
        g = self.grad.reshape(self.N, self.V, self.D)
        g *= frac[:, :, None]

    def step_tangent(self, dt, dq):
        """The static order-1 step of q and of its tangent dq (N x V, extensive): returns J dq; q advances."""
        dq = np.ascontiguousarray(np.array(dq, dtype=np.float64))
        if lib().orc_step_tangent(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.c_double(dt), _ptr(dq)) != 0:
            raise ValueError("step_tangent: a static mesh, order 1")
        return dq

    def step_adjoint(self, dt, qbar):
        """J^T qbar of the static order-1 step from the current q (which stays)."""
        qbar = np.ascontiguousarray(np.array(qbar, dtype=np.float64))
        if lib().orc_step_adjoint(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.c_double(dt), _ptr(qbar)) != 0:
            raise ValueError("step_adjoint: a static mesh, order 1")
        return qbar

    def face_sides(self, f, dt, pass_=-1, unit=0.0):
        """Face f's two states as the flux solves them: (L, R, n, wf, |A|, fallbacks). pass_: -1 static (dt the step's),
        0 / 1 a moving step's first / second pass (dt 0 / the step's, unit = grid step / dt), 2 order 0 moving."""
        D, V = self.D, self.V
        L, R, n, wf = np.zeros(V), np.zeros(V), np.zeros(D), np.zeros(D)
        fb = ctypes.c_int(0)
        L_ = lib()
        L_.orc_face_sides.restype = ctypes.c_double
        mag = L_.orc_face_sides(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.c_int(f), ctypes.c_double(dt), ctypes.c_int(pass_),
                                ctypes.c_double(unit), _ptr(L), _ptr(R), _ptr(n), _ptr(wf), ctypes.byref(fb))
        return L, R, n, wf, mag, fb.value

    def moving_begin(self, dt):
        lib().orc_moving_begin(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.c_double(dt))
        return self.x.copy()

    def moving_finish(self, mesh, dt):
        self.set_mesh(mesh)
        lib().orc_moving_finish(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.c_double(dt))

    def fallbacks(self):
        return int(self.c.fallbacks[0] + self.c.fallbacks[1])

    def moving_adjoint(self, dt, qbar, xbar, mesh1, adjoint0, adjoint1, steps_fixed=True):
        """The moving step's reverse (order 1, fixed topology) from the state held (q, x at t^n on the current mesh):
        qbar, xbar (bars of Q, x after the step) -> before it. mesh1: the mesh the forward step reached. adjoint0/1:
        the geometry adjoints of the two meshes, f(Vbar, cbar, Abar, xfbar, xbar) adding into xbar (numpy views)."""
        D, N, V = self.D, self.N, self.V
        geo1 = Geometry(mesh1)
        F0, F1 = self.geo.c.F, geo1.c.F

        def wrap(fn, F):
            def cb(_ctx, Vb, cb_, Ab, xfb, xb):
                as_ = np.ctypeslib.as_array
                fn(as_(Vb, (N,)), as_(cb_, (N, D)), as_(Ab, (max(F, 1), D))[:F], as_(xfb, (max(F, 1), D))[:F], as_(xb, (N, D)))
            return _GEO_ADJOINT(cb)

        c0, c1 = wrap(adjoint0, F0), wrap(adjoint1, F1)
        qb = np.ascontiguousarray(np.array(qbar, dtype=np.float64).reshape(N, V))
        xb = np.ascontiguousarray(np.array(xbar, dtype=np.float64).reshape(N, D))
        rc = lib().orc_moving_adjoint(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.byref(geo1.c), ctypes.c_double(dt),
                                      int(steps_fixed), _ptr(qb), _ptr(xb), c0, None, c1, None)
        if rc != 0:
            raise ValueError("moving_adjoint: a moving mesh, order 1")
        return qb, xb


    def moving_tangent(self, dt, dq, dx, mesh1, tangent0, tangent1, steps_fixed=True):
        """The moving step's tangent (orders 0 and 1, fixed topology) at the state held (q, x at t^n on the current mesh):
        (dq, dx) at t^n -> at t^n+1; q, x stay. mesh1: the mesh the forward step reached (order 1). tangent0/1: the two
        meshes' geometry tangents, f(dx) -> (dV, dc, dA, dxf) (mesh.GeometryDerivative.tangent)."""
        D, N, V = self.D, self.N, self.V
        geo1 = Geometry(mesh1)
        F0, F1 = self.geo.c.F, geo1.c.F

        def wrap(fn, F):
            def cb(_ctx, dxp, dVp, dcp, dAp, dxfp):
                as_ = np.ctypeslib.as_array
                dV, dc, dA, dxf = fn(as_(dxp, (N, D)).copy())
                as_(dVp, (N,))[...] = dV
                as_(dcp, (N, D))[...] = dc
                if F:
                    as_(dAp, (F, D))[...] = dA
                    as_(dxfp, (F, D))[...] = dxf
            return _GEO_TANGENT(cb)

        c0, c1 = wrap(tangent0, F0), wrap(tangent1, F1)
        tq = np.ascontiguousarray(np.array(dq, dtype=np.float64).reshape(N, V))
        tx = np.ascontiguousarray(np.array(dx, dtype=np.float64).reshape(N, D))
        L = lib()
        L.orc_moving_tangent.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                         _GEO_TANGENT, ctypes.c_void_p, _GEO_TANGENT, ctypes.c_void_p]
        rc = L.orc_moving_tangent(ctypes.byref(self.c), ctypes.byref(self.geo.c), ctypes.byref(geo1.c), ctypes.c_double(dt),
                                  int(steps_fixed), _ptr(tq), _ptr(tx), c0, None, c1, None)
        if rc != 0:
            raise ValueError("moving_tangent: a moving mesh, order 0 or 1, the double policy")
        return tq, tx


def face_velocity_tangent(d, xf, wL, wR, dd, dxf, dwL, dwR):
    """FaceVelocity::tangent (finite_volume.hpp:2322-2330)."""
    args = [_vec(a) for a in (d, xf, wL, wR, dd, dxf, dwL, dwR)]
    out = np.empty_like(args[0])
    lib().orc_face_velocity_tangent(len(args[0]), *(_ptr(a) for a in args), _ptr(out))
    return out


def moving_face_tangent(L, R, A, w, dL, dR, dA, dw, gamma, riemann="hllc"):
    """MovingFace::tangent (euler.hpp:861-891): dG of G = |A| F(L, R, A / |A|, w)."""
    args = [_vec(a) for a in (L, R, A, w, dL, dR, dA, dw)]
    out = np.empty_like(args[0])
    g = gas(gamma)
    lib().orc_moving_face_tangent({"hllc": 0, "hll": 1}[riemann], len(args[2]), ctypes.byref(g), *(_ptr(a) for a in args), _ptr(out))
    return out


WALLS = {"reflective": 0, "outflow": 1, "inflow": 2}
LIMITERS = {"face": 0, "none": 1, "minmod": 2, "vanleer": 3}


class Line:
    """The oracle's static FiniteVolume<1> on a sorted line of grid keys (walls or periodic)."""

    def __init__(self, keys, period, step, gamma, periodic=False, walls=("outflow", "outflow"), order=1,
                 limiter="face", riemann="hllc"):
        self.keys = np.ascontiguousarray(np.asarray(keys, dtype=np.int64))
        n = len(self.keys)
        self.n = n
        self.q = np.zeros((n, 3))
        self.next = np.zeros((n, 3))
        c = _Line()
        c.n, c.periodic = n, int(periodic)
        c.wall_left, c.wall_right = WALLS[walls[0]], WALLS[walls[1]]
        c.order, c.limiter, c.riemann = order, LIMITERS[limiter], {"hllc": 0, "hll": 1}[riemann]
        c.gas = gas(gamma)
        c.key, c.period, c.step = _ptr(self.keys, _lp), period, step
        self._bufs = [self.q, self.next]
        c.q, c.next = _ptr(self.q), _ptr(self.next)
        self.c = c

    def _sync(self):
        # The C side swaps q and next after each step; follow its pointers.
        addr = ctypes.cast(self.c.q, ctypes.c_void_p).value
        if addr != self._bufs[0].ctypes.data:
            self._bufs.reverse()

    @property
    def state(self):
        self._sync()
        return self._bufs[0]

    def set_state(self, q):
        self.state[...] = q

    def volume(self, s):
        return lib().orc_line_volume(ctypes.byref(self.c), s)

    def centroid(self, s):
        return lib().orc_line_centroid(ctypes.byref(self.c), s)

    def time_step(self, courant, rule="luna"):
        return lib().orc_line_time_step(ctypes.byref(self.c), ctypes.c_double(courant), int(rule == "luna"))

    def step(self, dt):
        lib().orc_line_step(ctypes.byref(self.c), ctypes.c_double(dt))

    def step_tangent(self, dt, dq):
        """One step of q and of its tangent dq (extensive, N x 3): returns J dq; q advances."""
        dq = np.ascontiguousarray(np.array(dq, dtype=np.float64))
        self._sync()
        lib().orc_line_step_tangent(ctypes.byref(self.c), ctypes.c_double(dt), _ptr(dq))
        return dq

    def step_adjoint(self, dt, qbar):
        """J^T qbar of one step from the current q (which stays): qbar after the step -> before it."""
        qbar = np.ascontiguousarray(np.array(qbar, dtype=np.float64))
        self._sync()
        lib().orc_line_step_adjoint(ctypes.byref(self.c), ctypes.c_double(dt), _ptr(qbar))
        return qbar

    def step_and_time_step(self, dt, courant):
        return lib().orc_line_step_fused(ctypes.byref(self.c), ctypes.c_double(dt), ctypes.c_double(courant))


class MovingLine:
    """The oracle's FiniteVolume<1> on a moving line: n generators at positions x on [0, side] (walls) or a periodic
    line; walls as ("outflow", "inflow", ...) with the inflow primitives in `inflow` (left, right); cells come from open
    walls every `spacing`. state() and positions() are by cell (in the order cells were made), line() is (keys, ids) by
    slot."""

    def __init__(self, x, side, period, gamma, periodic=False, walls=("outflow", "outflow"), inflow=((0, 0, 0), (0, 0, 0)),
                 order=1, limiter="face", riemann="hllc", spacing=0.0, regularise=True):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        il = np.ascontiguousarray(np.asarray(inflow[0], dtype=np.float64))
        ir = np.ascontiguousarray(np.asarray(inflow[1], dtype=np.float64))
        self._h = lib().orc_mline_create(len(x), _ptr(x), float(side), int(period), int(periodic), WALLS[walls[0]], WALLS[walls[1]],
                                         _ptr(il), _ptr(ir), order, LIMITERS[limiter], {"hllc": 0, "hll": 1}[riemann], float(gamma),
                                         float(spacing), int(regularise))
        self.period, self.side, self.grid_step, self.periodic = int(period), float(side), float(side) / int(period), bool(periodic)

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_mline_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def n(self):
        return lib().orc_mline_cells(self._h)

    def set_state(self, q):
        q = np.ascontiguousarray(np.asarray(q, dtype=np.float64))
        assert q.shape == (self.n, 3)
        lib().orc_mline_set_state(self._h, _ptr(q))

    def state(self):
        q = np.empty((self.n, 3))
        lib().orc_mline_get_state(self._h, _ptr(q))
        return q

    def positions(self):
        x = np.empty(self.n)
        lib().orc_mline_get_positions(self._h, _ptr(x))
        return x

    def line(self):
        keys = np.empty(self.n, dtype=np.int64)
        ids = np.empty(self.n, dtype=np.int32)
        lib().orc_mline_get_line(self._h, _ptr(keys, _lp), _ptr(ids, _ip))
        return keys, ids

    def cells_added(self):
        return lib().orc_mline_stat(self._h, 0)

    def sorts(self):
        return lib().orc_mline_stat(self._h, 2)

    def time_step(self, courant, rule="arepo"):
        return lib().orc_mline_time_step(self._h, courant, int(rule == "luna"))

    def step(self, dt):
        lib().orc_mline_step(self._h, dt)

    def step_and_time_step(self, dt, courant):
        return lib().orc_mline_step_fused(self._h, dt, courant)
