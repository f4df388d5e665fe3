"""Python front of the periodic Voronoi geometry producer (libluna_mesh.so, include/luna_mesh.h).

build(points, side) returns the mesh record as a dict of numpy arrays: the voronoi-cpp output contract the hydro step
reads (owner, neighbour, area, face_centroid, centroid, volume, positions, start, list, side, step, period).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Record(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("cells", ctypes.c_int), ("faces", ctypes.c_int),
                ("owner", _ip), ("neighbour", _ip), ("start", _ip), ("list", _ip),
                ("area", _dp), ("face_centroid", _dp), ("centroid", _dp), ("volume", _dp), ("positions", _dp),
                ("side", ctypes.c_double * 3), ("step", ctypes.c_double), ("period", ctypes.c_longlong * 3),
                ("build_seconds", ctypes.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libluna_mesh.so")
        if not os.path.exists(path):
            raise RuntimeError("libluna_mesh.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.luna_mesh_build.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.POINTER(Record)]
        L.luna_mesh_release.argtypes = [ctypes.POINTER(Record)]
        L.luna_mesh_last_error.restype = ctypes.c_char_p
        L.luna_mesh_derivative_create.argtypes = [ctypes.POINTER(Record), ctypes.POINTER(ctypes.c_void_p)]
        L.luna_mesh_derivative_destroy.argtypes = [ctypes.c_void_p]
        L.luna_mesh_derivative_tangent.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.luna_mesh_derivative_adjoint.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        _LIB = L
    return _LIB


def _copy(ptr, count, dtype, shape=None):
    if count == 0:
        return np.zeros(shape or (0,), dtype=dtype)
    arr = np.ctypeslib.as_array(ptr, shape=(count,)).copy().astype(dtype, copy=False)
    return arr.reshape(shape) if shape else arr


def build(points, side=1.0):
    """The periodic Voronoi mesh of `points` (N x D) in the box [0, side)^D (side: scalar or D lengths)."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
    N, D = pts.shape
    sides = np.ascontiguousarray(np.full(D, side, dtype=np.float64) if np.isscalar(side) else np.asarray(side, np.float64))
    rec = Record()
    rc = lib().luna_mesh_build(D, N, pts.ctypes.data_as(_dp), sides.ctypes.data_as(_dp), ctypes.byref(rec))
    if rc != 0:
        raise RuntimeError(lib().luna_mesh_last_error().decode())
    try:
        F, E = rec.faces, None
        start = _copy(rec.start, N + 1, np.int32)
        E = int(start[-1])
        mesh = {
            "dim": D, "cells": N, "faces": F,
            "owner": _copy(rec.owner, F, np.int32), "neighbour": _copy(rec.neighbour, F, np.int32),
            "start": start, "list": _copy(rec.list, E, np.int32),
            "area": _copy(rec.area, F * D, np.float64, (F, D)),
            "face_centroid": _copy(rec.face_centroid, F * D, np.float64, (F, D)),
            "centroid": _copy(rec.centroid, N * D, np.float64, (N, D)),
            "volume": _copy(rec.volume, N, np.float64),
            "positions": _copy(rec.positions, N * D, np.float64, (N, D)),
            "side": [rec.side[a] for a in range(D)], "step": rec.step,
            "period": [int(rec.period[a]) for a in range(D)],
            "build_seconds": rec.build_seconds,
        }
    finally:
        lib().luna_mesh_release(ctypes.byref(rec))
    return mesh


def record_of(mesh):
    """A luna_mesh_record over the arrays of a mesh dict (kept alive by the returned holder)."""
    m = dict(mesh)
    for k in ("owner", "neighbour", "start", "list"):
        m[k] = np.ascontiguousarray(m[k], dtype=np.int32)
    for k in ("area", "face_centroid", "centroid", "volume", "positions"):
        m[k] = np.ascontiguousarray(m[k], dtype=np.float64)
    rec = Record()
    rec.dim, rec.cells, rec.faces = m["dim"], m["cells"], m["faces"]
    for k in ("owner", "neighbour", "start", "list"):
        setattr(rec, k, m[k].ctypes.data_as(_ip))
    for k in ("area", "face_centroid", "centroid", "volume", "positions"):
        setattr(rec, k, m[k].ctypes.data_as(_dp))
    for a in range(3):
        rec.side[a] = m["side"][a] if a < len(m["side"]) else 0.0
        rec.period[a] = int(m["period"][a]) if a < len(m["period"]) else 0
    rec.step = m["step"]
    return rec, m


class GeometryDerivative:
    """The derivative of a mesh's geometry (V, c per cell; A, x_f per face) in the generator positions at fixed topology
    (luna_mesh_derivative_*; the role of the reference's GeometryDerivative, mesh_derivative.hpp:21-46)."""

    def __init__(self, mesh):
        self.D, self.N, self.F = mesh["dim"], mesh["cells"], mesh["faces"]
        rec, self._keep = record_of(mesh)
        h = ctypes.c_void_p()
        if lib().luna_mesh_derivative_create(ctypes.byref(rec), ctypes.byref(h)) != 0:
            raise RuntimeError(lib().luna_mesh_last_error().decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().luna_mesh_derivative_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def tangent(self, dx):
        """dx (N x D) -> dV (N), dc (N x D), dA (F x D), dxf (F x D)."""
        dx = np.ascontiguousarray(dx, dtype=np.float64).reshape(self.N, self.D)
        dV, dc = np.zeros(self.N), np.zeros((self.N, self.D))
        dA, dxf = np.zeros((self.F, self.D)), np.zeros((self.F, self.D))
        lib().luna_mesh_derivative_tangent(self._h, *(a.ctypes.data_as(_dp) for a in (dx, dV, dc, dA, dxf)))
        return dV, dc, dA, dxf

    def adjoint(self, Vbar, cbar, Abar, xfbar, xbar=None):
        """The transpose: adds to xbar (N x D, zeros if None) and returns it."""
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (Vbar, cbar, Abar, xfbar)]
        xbar = np.zeros((self.N, self.D)) if xbar is None else np.ascontiguousarray(xbar, dtype=np.float64)
        lib().luna_mesh_derivative_adjoint(self._h, *(a.ctypes.data_as(_dp) for a in args), xbar.ctypes.data_as(_dp))
        return xbar
