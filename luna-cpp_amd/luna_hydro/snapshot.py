"""LUNA-layout HDF5 snapshots through libluna_snapshot.so (include/luna_snapshot.h; the reference's writer is
src/hydro/snapshot.hpp:68-135): <dir>/snapdir_NNN/snapshot_NNN.0.hdf5."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)


class Snapshot(ctypes.Structure):
    _fields_ = [("index", ctypes.c_int), ("time", ctypes.c_double), ("sides", ctypes.c_double * 3), ("order", ctypes.c_int),
                ("gamma", ctypes.c_double), ("smallr", ctypes.c_double), ("smallc", ctypes.c_double),
                ("length_cm", ctypes.c_double), ("mass_g", ctypes.c_double), ("velocity_cm_s", ctypes.c_double),
                ("cells", ctypes.c_longlong), ("positions", _dp), ("volume", _dp), ("q", _dp), ("commit", ctypes.c_char_p),
                ("dirty", ctypes.c_int), ("mesh_commit", ctypes.c_char_p), ("mesh_dirty", ctypes.c_int),
                ("mesh_origin", ctypes.c_char_p)]

# The mesh producer whose meshes the snapshots hold (written as voronoi_cpp_origin, the attribute the reference fills
# with its mesh library's origin).
MESH_ORIGIN = "luna-cpp_amd/csrc/mesh/voronoi.cpp (periodic Voronoi producer of this project, voronoi-cpp's output contract)"


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libluna_snapshot.so")
        if not os.path.exists(path):
            raise RuntimeError("libluna_snapshot.so is not built (python luna-cpp_amd/luna_hydro/build.py)")
        _LIB = ctypes.CDLL(path)
        _LIB.luna_snapshot_error.restype = ctypes.c_char_p
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s: %s" % (what, lib().luna_snapshot_error().decode(errors="replace")))


def write(directory, index, time, positions, volume, q, gamma, sides=(1.0, 1.0, 1.0), order=1, units=None, commit="", dirty=False,
          smallr=1e-10, smallc=1e-10, mesh_commit="", mesh_dirty=False, mesh_origin=MESH_ORIGIN):
    """Writes one snapshot of a 3D state (Q extensive, N x 5); returns the file's path. units: (length_cm, mass_g,
    velocity_cm_s) or None."""
    x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    vol = np.ascontiguousarray(volume, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 5)
    s = Snapshot()
    s.index, s.time, s.order, s.gamma, s.smallr, s.smallc = index, time, order, gamma, smallr, smallc
    for a in range(3):
        s.sides[a] = sides[a]
    s.length_cm, s.mass_g, s.velocity_cm_s = units if units else (1.0, 1.0, 1.0)
    s.cells = x.shape[0]
    s.positions, s.volume, s.q = x.ctypes.data_as(_dp), vol.ctypes.data_as(_dp), q.ctypes.data_as(_dp)
    s.commit, s.dirty = commit.encode(), int(dirty)
    s.mesh_commit, s.mesh_dirty, s.mesh_origin = mesh_commit.encode(), int(mesh_dirty), mesh_origin.encode()
    name = ctypes.create_string_buffer(4096)
    _check(lib().luna_snapshot_write(os.fsencode(directory), ctypes.byref(s), name, 4096), "write")
    return name.value.decode()


def read(path, names=("x", "y", "z", "volume", "density", "velocity_x", "velocity_y", "velocity_z", "internal_energy", "pressure")):
    """The datasets of group `cells` (U as 5 x N) and its numeric attributes."""
    out = {}
    n = ctypes.c_double()
    _check(lib().luna_snapshot_read_attribute(os.fsencode(path), b"cells", b"N_cells", ctypes.byref(n)), "N_cells")
    N = int(n.value)
    for name in list(names) + ["U"]:
        a = np.empty(5 * N if name == "U" else N)
        _check(lib().luna_snapshot_read_dataset(os.fsencode(path), name.encode(), a.ctypes.data_as(_dp), a.size), name)
        out[name] = a.reshape(5, N) if name == "U" else a
    attrs = {}
    for name in ("time", "step", "gamma", "spatial_dimension", "order", "Lx", "Ly", "Lz", "N_cells", "luna_cpp_dirty",
                 "voronoi_cpp_dirty"):
        v = ctypes.c_double()
        _check(lib().luna_snapshot_read_attribute(os.fsencode(path), b"cells", name.encode(), ctypes.byref(v)), name)
        attrs[name] = v.value
    buf = ctypes.create_string_buffer(256)
    _check(lib().luna_snapshot_read_string(os.fsencode(path), b"cells", b"mesh", buf, 256), "mesh")
    attrs["mesh"] = buf.value.decode()
    for name in ("luna_cpp_commit", "voronoi_cpp_commit", "voronoi_cpp_origin"):
        _check(lib().luna_snapshot_read_string(os.fsencode(path), b"cells", name.encode(), buf, 256), name)
        attrs[name] = buf.value.decode()
    out["attributes"] = attrs
    return out
