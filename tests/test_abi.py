"""The C-ABI boundary on the CPU: every function include/*.h declares is exported by its library, the C++ facade
compiles and links against libluna_hydro.so, and the product path fails loudly (no CPU fallback) where there is no GPU."""
import os
import re
import subprocess

import pytest

from luna_hydro import hydro, mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "luna-cpp_amd", "luna_hydro")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(luna_\w+)\s*\(", text)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIBDIR, lib)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("header,lib", [("luna_hydro.h", "libluna_hydro.so"), ("luna_mesh.h", "libluna_mesh.so")])
def test_every_declared_symbol_is_exported(header, lib):
    names = declared(header)
    assert names
    missing = [n for n in names if n not in exported(lib)]
    assert not missing, missing


def test_python_mirror_binds_every_entry_point():
    syms = hydro.exported_symbols()
    assert all(syms.values()), [k for k, v in syms.items() if not v]
    assert set(syms) == set(declared("luna_hydro.h"))


def test_cpp_facade_compiles_and_links(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('''
#include "luna/finite_volume.hpp"
#include "luna_mesh.h"
int main() {
    luna_mesh_record m{};
    double x[2 * 64];
    for (int i = 0; i < 64; ++i) { x[2 * i] = (i % 8 + 0.5) / 8; x[2 * i + 1] = (i / 8 + 0.5) / 8 + 0.01 * (i % 3); }
    const double side[2] = {1, 1};
    if (luna_mesh_build(2, 64, x, side, &m) != 0) return 2;
    try {
        luna::FiniteVolume<2> fv(reinterpret_cast<const luna_geometry&>(m), luna::Gas{}, 1);
        (void)fv;
    } catch (const std::runtime_error&) {
        luna_mesh_release(&m);
        return 3;  // no GPU here: the facade reports the device error
    }
    luna_mesh_release(&m);
    return 0;
}
''')
    exe = tmp_path / "use"
    subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + LIBDIR, "-lluna_hydro", "-lluna_mesh", "-Wl,-rpath," + LIBDIR], check=True)
    rc = subprocess.run([str(exe)]).returncode
    assert rc in (0, 3)


def test_layout_of_mesh_record_matches_geometry():
    """luna_geometry is the leading part of luna_mesh_record (the facade passes one as the other)."""
    from luna_hydro.hydro import Geometry
    from luna_hydro.mesh import Record
    for (a, ta), (b, tb) in zip(Geometry._fields_, Record._fields_):
        assert a == b and getattr(Geometry, a).offset == getattr(Record, b).offset


def test_no_cpu_fallback_without_gpu():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU device node is present")
    m = mesh.build([[(i % 4 + 0.5) / 4, (i // 4 % 4 + 0.5) / 4, (i // 16 + 0.5) / 4 + 0.01 * (i % 3)] for i in range(64)], 1.0)
    with pytest.raises(RuntimeError):
        hydro.FiniteVolume(m, 5.0 / 3.0)
