"""MI355X-native implicit P1-FEM Richards time step (drop-in for JAX-WSPM's path).

Layers (reference files in parentheses):
  config     parameter dataclasses            (config/settings.py)
  mesh       loader, generators, boundaries   (src/mesh/loader.py, src/models/boundary.py)
  problems   Test1-3 / Test3D / unit-domain van Genuchten cases
  device     C-ABI context over libwspm.so    (assembly, Krylov, Picard in HIP)
  linear     solve_system / soil-model / assembly entry points with the
             reference's names                (src/solvers/solver_types.py, ...)
  richards   RichardsSolver(config).solve()   (src/solvers/richards_solver_v2.py, _3D_v2.py)
  cli        the reference's command line     (src/cli.py)
"""
__version__ = "0.1.0"
