"""Device assembly (StokesAssembler::local_assembly, stokes_ilu.cc:739-821,
with copy_local_to_global, stokes_common.implement.h:658-669) against the
host generator: the quadrature samples and every value of A, B^T, B, K11, M_p
and the rhs bit for bit, over both dimensions, the four fields and the three
boundary conditions (periodic seams, Dirichlet elimination, the pinned
pressure dof); then the device pipeline sample -> assemble ->
fs_stokes_refactor_device -> fs_stokes_solve_device against the host
pipeline on the same handle settings."""
import time

import numpy as np
import pytest

import felspa_amd as fa
from felspa_amd import _lib as L
from felspa_amd.system import BLOCK_NAMES, DeviceAssembler, StokesSystem, problem

pytestmark = pytest.mark.gpu

CASES = {
    "2d-inclusion-periodic-gravity": dict(dim=2, n_cells=8, field=L.FS_FIELD_INCLUSION,
                                          bc=L.FS_BC_TOPBOTTOM_SHEAR_PERIODIC, radius=0.2, gravity=1.0,
                                          rho_incl=2.0),
    "2d-layers-freeslip-gravity": dict(dim=2, n_cells=10, field=L.FS_FIELD_LAYERS, bc=L.FS_BC_FREE_SLIP,
                                       eta_incl=100.0, rho_incl=1.1, amp=0.05, gravity=1.0),
    "2d-ellipses-shear-odd": dict(dim=2, n_cells=9, field=L.FS_FIELD_ELLIPSES, bc=L.FS_BC_LINEAR_SHEAR, seed=7),
    "2d-constant-periodic": dict(dim=2, n_cells=6, field=L.FS_FIELD_CONSTANT, bc=L.FS_BC_TOPBOTTOM_SHEAR_PERIODIC),
    "3d-inclusion-periodic-gravity": dict(dim=3, n_cells=4, field=L.FS_FIELD_INCLUSION,
                                          bc=L.FS_BC_TOPBOTTOM_SHEAR_PERIODIC, radius=0.3, gravity=1.0),
    "3d-ellipsoids-freeslip-gravity": dict(dim=3, n_cells=5, field=L.FS_FIELD_ELLIPSES, bc=L.FS_BC_FREE_SLIP,
                                           seed=11, gravity=2.0, n_ellipses=(2, 4), logc=(1.0, 3.0),
                                           axes=(0.1, 0.3)),
    "3d-layers-shear-gravity": dict(dim=3, n_cells=3, field=L.FS_FIELD_LAYERS, bc=L.FS_BC_LINEAR_SHEAR,
                                    amp=0.1, gravity=1.0),
}


def _bits_equal(got, want, what):
    got = np.ascontiguousarray(got, dtype=np.float64)
    want = np.ascontiguousarray(want, dtype=np.float64)
    assert got.shape == want.shape, what
    bad = got.view(np.uint64) != want.view(np.uint64)
    if bad.any():
        d = np.abs(got - want)[bad].max()
        raise AssertionError("%s: %d of %d values differ (max |diff| %.3e)" % (what, bad.sum(), got.size, d))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("name", list(CASES))
def test_device_assembly_is_bit_identical_to_host(name):
    S = StokesSystem(problem(**CASES[name]))
    asm = DeviceAssembler(S)
    eta, rho = asm.sample_field()
    he, hr = S.quadrature_fields()
    _bits_equal(eta.cpu().numpy(), he.ravel(), name + " eta_q")
    _bits_equal(rho.cpu().numpy(), hr.ravel(), name + " rho_q")
    for rep in range(2):  # a second call on the same buffers starts from zero again
        asm.assemble(eta, rho)
        for b in BLOCK_NAMES:
            _bits_equal(asm.values[b].cpu().numpy(), S.blocks[b].data, "%s %s (call %d)" % (name, b, rep))
        _bits_equal(asm.rhs.cpu().numpy(), S.rhs, "%s rhs (call %d)" % (name, rep))
    asm.close()


def _pair(dim, n, field, bc, **kw):
    """Two systems on one pattern, different viscosity / density."""
    a = StokesSystem(problem(dim, n, field, bc, eta_incl=1e3, rho_incl=2.0, gravity=1.0, **kw))
    b = StokesSystem(problem(dim, n, field, bc, eta_incl=7.5, rho_incl=0.5, gravity=1.0, **kw))
    for k in BLOCK_NAMES:
        assert np.array_equal(a.blocks[k].indptr, b.blocks[k].indptr)
        assert np.array_equal(a.blocks[k].indices, b.blocks[k].indices)
    return a, b


def test_caller_samples_assemble_another_field():
    """The assembler of one system, fed another field's samples on the same
    mesh, produces that other system's blocks bit for bit."""
    S1, S2 = _pair(3, 4, L.FS_FIELD_INCLUSION, L.FS_BC_LINEAR_SHEAR, radius=0.3)
    asm = DeviceAssembler(S1)
    e2, r2 = S2.quadrature_fields()
    asm.assemble(_dev(e2.ravel()), _dev(r2.ravel()))
    for b in BLOCK_NAMES:
        _bits_equal(asm.values[b].cpu().numpy(), S2.blocks[b].data, "field 2 " + b)
    _bits_equal(asm.rhs.cpu().numpy(), S2.rhs, "field 2 rhs")
    asm.close()


@pytest.mark.parametrize("dim,n,bc", [(2, 24, L.FS_BC_TOPBOTTOM_SHEAR_PERIODIC), (3, 6, L.FS_BC_LINEAR_SHEAR)])
def test_device_pipeline_matches_host_pipeline(dim, n, bc):
    """A time step without the host: a handle made on field 1 receives field
    2's device-assembled values (fs_stokes_refactor_device) and solves field
    2's device rhs; a handle made directly on field 2's host blocks solves the
    host rhs. Same pattern, same settings: the same solution."""
    import torch
    S1, S2 = _pair(dim, n, L.FS_FIELD_INCLUSION, bc, radius=0.25)
    hint = S1.block_hint(3 if dim == 3 else 5)
    kw = dict(A_lof=1, constraints=S1.constrained, block_hint=hint, rtol=1e-10)
    ref = fa.StokesLinearSystem(S2.blocks, **kw)
    x_ref = np.zeros(S2.n)
    st_ref = ref.solve(x_ref, S2.rhs)
    lin = fa.StokesLinearSystem(S1.blocks, **kw)
    asm = DeviceAssembler(S1)
    e2, r2 = S2.quadrature_fields()
    asm.assemble(_dev(e2.ravel()), _dev(r2.ravel()))
    lin.refactor_device([asm.values[b].data_ptr() for b in BLOCK_NAMES])
    d_x = torch.zeros(S2.n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc, st = lin.solve_device(asm.rhs.data_ptr(), d_x.data_ptr())
    assert rc == L.FS_OK
    x = d_x.cpu().numpy()
    assert st["outer_iters"] == st_ref["outer_iters"]
    scale = np.abs(x_ref).max()
    assert np.abs(x - x_ref).max() <= 1e-13 * scale, np.abs(x - x_ref).max() / scale
    # the host refactor of the same values agrees with the device one
    lin.refactor(S2.blocks)
    x_h = np.zeros(S2.n)
    lin.solve(x_h, S2.rhs)
    assert np.abs(x_h - x).max() <= 1e-13 * scale
    for o in (lin, ref, asm):
        o.close()


def test_assembly_time_3d():
    """The 3D ellipsoid field at 24^3 cells: bit-identical values and the GPU
    assembly time beside the host generator's (printed)."""
    t0 = time.perf_counter()
    S = StokesSystem(fa.make_problem("C5", n_cells=24, sample=3))
    t_host = time.perf_counter() - t0
    asm = DeviceAssembler(S)
    eta, rho = asm.sample_field()
    asm.assemble(eta, rho)
    ms = min(asm.assemble(eta, rho) for _ in range(3))
    for b in BLOCK_NAMES:
        _bits_equal(asm.values[b].cpu().numpy(), S.blocks[b].data, "24^3 " + b)
    _bits_equal(asm.rhs.cpu().numpy(), S.rhs, "24^3 rhs")
    print("\n24^3 Q2-Q1: device assembly %.2f ms, host generator (patterns + values) %.2f s, nnz(A) %d"
          % (ms, t_host, S.blocks["A"].nnz))
    asm.close()
