"""Build recipe of the native libraries (in-tree, so they travel with the source tree):

  libluna_hydro.so  hand-written HIP for gfx950 (csrc/hydro/*.hip), hipcc, FP64 with -ffp-contract=off
  libluna_mesh.so   the periodic Voronoi geometry producer (csrc/mesh/voronoi.cpp), g++ with OpenMP
  libluna_snapshot.so  LUNA-layout HDF5 snapshots (csrc/snapshot/snapshot.cpp), g++ with the image's HDF5

Run `python luna-cpp_amd/luna_hydro/build.py`, or __graft_entry__.build().
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["hydro/mesh_kernels.hip", "hydro/geometry_kernels.hip", "hydro/line_kernels.hip", "hydro/adjoint_kernels.hip",
               "hydro/limiter_kernels.hip",
               "hydro/api.hip"]
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that line is harmless.
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out)
    if res.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_hydro():
    """Each HIP source compiled on its own (in parallel, objects under _build/), then linked into libluna_hydro.so."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = os.path.join(HERE, "libluna_hydro.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.basename(src).replace(".hip", ".o"))
        jobs.append((obj, [hipcc] + compile_flags + ["-I" + INCLUDE, "-c", "-o", obj, os.path.join(CSRC, src)]))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, cmd) for _, cmd in jobs]:
            f.result()
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + [obj for obj, _ in jobs])
    os.replace(out + ".tmp", out)
    return out


def build_mesh():
    out = os.path.join(HERE, "libluna_mesh.so")
    _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra", "-I" + INCLUDE,
          "-o", out + ".tmp", os.path.join(CSRC, "mesh", "voronoi.cpp")])
    os.replace(out + ".tmp", out)
    return out


HDF5 = os.environ.get("LUNA_HDF5", "/opt/conda")  # the image's HDF5 C library (include/, lib/)


def build_snapshot():
    out = os.path.join(HERE, "libluna_snapshot.so")
    lib = os.path.join(HDF5, "lib")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-I" + INCLUDE, "-I" + os.path.join(HDF5, "include"),
          "-o", out + ".tmp", os.path.join(CSRC, "snapshot", "snapshot.cpp"), "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    return [build_mesh(), build_snapshot(), build_hydro()]


if __name__ == "__main__":
    build_all()
    sys.exit(0)
