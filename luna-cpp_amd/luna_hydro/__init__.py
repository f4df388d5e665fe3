"""luna_hydro: luna-cpp's finite-volume hydro step on MI355X (gfx950).

  hydro     FiniteVolume (D = 2, 3) and Line (D = 1) over the C-ABI of libluna_hydro.so (HIP kernels)
  mesh      the periodic Voronoi geometry producer (libluna_mesh.so) that feeds the step its mesh
  problems  initial conditions and point sets of the benchmark problems
"""
from . import mesh, problems  # noqa: F401
from .hydro import FiniteVolume, Line  # noqa: F401
