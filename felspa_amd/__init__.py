"""felspa_amd: MI355X (gfx950) replacement for FELSPA's Q2-Q1 Stokes
saddle-point solve (StokesLinearSystem::solve), behind a C-ABI
(include/fs_stokes.h) with hand-written HIP kernels."""
from . import _lib
from .configs import CONFIGS, make_problem, solver_settings
from .stokes import (FC, SCR, CompareTest, NoConvergence, SparseILU, SparseMatrix,
                     StokesLinearSystem, StokesSolverControl, block_color_ordering,
                     ilu_symbolic)
from .system import DeviceAssembler, StokesSystem, load_blocks, problem, save_blocks

__all__ = ["StokesLinearSystem", "StokesSolverControl", "SparseILU", "SparseMatrix",
           "NoConvergence", "StokesSystem", "DeviceAssembler", "problem", "CONFIGS", "make_problem",
           "solver_settings", "ilu_symbolic", "block_color_ordering", "save_blocks", "load_blocks", "FC", "SCR",
           "CompareTest", "_lib"]
