"""MI355X-native (gfx950) SIMPLE / pressure-correction path for 2-D curvilinear
mixed-convection flow, with the operators of mhsuhail00/CFD_Solver.

  cfd_solver_amd.grid      grids, metrics, INP.DAT format
  cfd_solver_amd.stages    RefStages: the reference's stage functions on the GPU
  cfd_solver_amd.cavity    CavityCase / CavityBatch: device-resident steady cavity solves,
                           one case or many equally sized cases per launch
  cfd_solver_amd.build     compiles libcfd_gpu.so for gfx950

The HIP library (libcfd_gpu.so) is required; nothing falls back to the CPU.
"""
from .grid import Grid, cavity_grid, ogrid_ellipse, metrics, read_inp_dat, write_inp_dat  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    if name in ("RefStages",):
        from .stages import RefStages
        return RefStages
    if name in ("CavityCase", "CavityBatch", "default_params"):
        from . import cavity
        return getattr(cavity, name)
    raise AttributeError(name)
