"""The five Stokes workloads (matrix/field/BC parameters and solver settings).

Solver settings per workload: ILU levels, lattice nodes per ILU block edge,
and the flexible method's Krylov basis (restart) and inner ILU-PCG steps on A
per outer iteration. The last two are the measured optimum on one MI355X
(mixed precision, FGMRES restart {60, 250} x inner steps 0..60; 0 = one ILU
application per outer iteration, the reference's A preconditioner): 2D needs
15-30 inner steps (C2 stagnates with 8), 3D is fastest with none, where the
outer iterations grow less than the inner work shrinks.
The inner CG on M_p (the Schur complement's approximate inverse) stops at a
relative residual of 0.1 after at most one step, in both precisions: one
ILU(0)-PCG step reaches it on all five workloads (a 1e-2 tolerance takes two
steps and 4-10 % more solve time), and a one-step CG runs as ordinary
launches instead of the cooperative kernel.

C1  2D inclusion, 64x64 Q2-Q1, shear top/bottom + periodic x, ILU(0) (CPU-runnable)
C2  2D Rayleigh-Taylor, 256x256, free slip, gravity, ILU(1)
C3  3D porphyroclast, 32^3, linear simple shear, ILU(1)
C4  2D 128x128 seeded ellipses (learned warm start), simple shear
C5  3D 48^3 seeded ellipsoids, simple shear, ILU(2)
Seeds follow seed = 1000 * config_id + sample_id.
"""
from . import _lib as L
from .system import problem

CONFIGS = {
    "C1": dict(dim=2, n_cells=64, field=L.FS_FIELD_INCLUSION, bc=L.FS_BC_TOPBOTTOM_SHEAR_PERIODIC,
               eta_matrix=1.0, eta_incl=1e3, radius=0.1, ilu_k_A=0, ilu_k_S=0, hint_nodes=5, restart=250,
               inner_a=30, inner_cg_tol=0.1, inner_cg_max=1),
    "C2": dict(dim=2, n_cells=256, field=L.FS_FIELD_LAYERS, bc=L.FS_BC_FREE_SLIP,
               eta_matrix=1.0, eta_incl=100.0, rho_matrix=1.0, rho_incl=1.1, amp=0.02,
               gravity=1.0, ilu_k_A=1, ilu_k_S=0, hint_nodes=5, restart=250, inner_a=30,
               inner_cg_tol=0.1, inner_cg_max=1),
    "C3": dict(dim=3, n_cells=32, field=L.FS_FIELD_INCLUSION, bc=L.FS_BC_LINEAR_SHEAR,
               eta_matrix=1.0, eta_incl=1e3, radius=0.2, ilu_k_A=1, ilu_k_S=0, hint_nodes=3, restart=250,
               inner_a=0, inner_cg_tol=0.1, inner_cg_max=1),
    "C4": dict(dim=2, n_cells=128, field=L.FS_FIELD_ELLIPSES, bc=L.FS_BC_LINEAR_SHEAR,
               eta_matrix=1.0, n_ellipses=(1, 8), axes=(0.05, 0.2), logc=(0.0, 3.0),
               ilu_k_A=1, ilu_k_S=0, hint_nodes=5, restart=250, inner_a=20,
               inner_cg_tol=0.1, inner_cg_max=1),
    "C5": dict(dim=3, n_cells=48, field=L.FS_FIELD_ELLIPSES, bc=L.FS_BC_LINEAR_SHEAR,
               eta_matrix=1.0, n_ellipses=(1, 4), axes=(0.08, 0.25), logc=(1.0, 3.0),
               ilu_k_A=2, ilu_k_S=0, hint_nodes=3, restart=250, inner_a=0,
               inner_cg_tol=0.1, inner_cg_max=1),
}

_SOLVER_KEYS = ("ilu_k_A", "ilu_k_S", "hint_nodes", "restart", "inner_a", "inner_cg_tol", "inner_cg_max")


def make_problem(name, n_cells=None, sample=0):
    c = dict(CONFIGS[name])
    for k in _SOLVER_KEYS:
        c.pop(k, None)
    if n_cells is not None:
        c["n_cells"] = n_cells
    cid = int(name[1:])
    c["seed"] = 1000 * cid + sample
    return problem(**c)


def solver_settings(name):
    c = CONFIGS[name]
    return {k: c[k] for k in _SOLVER_KEYS}
