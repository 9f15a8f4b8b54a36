"""Host-side parts of the library (no GPU): the symbolic ILU(k) phase
against the oracle, input validation, defaults, and that the C-ABI library
loads and exports every symbol its headers declare."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import felspa_amd as fa
from felspa_amd import _lib as L

from conftest import ROOT, system


@pytest.mark.parametrize("k", [0, 1, 2, 3])
@pytest.mark.parametrize("cfg,n", [("C1", 8), ("C2", 8), ("C3", 3)])
def test_symbolic_matches_oracle(oracle, cfg, n, k):
    A = system(cfg, n).blocks["A"]
    P = fa.ilu_symbolic(A, k)
    F = oracle.SparseILU(A, k).factors()
    assert (P.indptr == F.indptr).all() and (P.indices == F.indices).all()


def test_symbolic_accepts_diagonal_first_rows(oracle):
    A = system("C1", 6).blocks["A"]
    # deal.II layout: diagonal first in each row
    rows = []
    for r in range(A.shape[0]):
        cols = list(A.indices[A.indptr[r]:A.indptr[r + 1]])
        cols.remove(r)
        rows.append([r] + cols)
    rp = np.zeros(A.shape[0] + 1, np.int64)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.int32)
    D = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=A.shape)
    D.has_sorted_indices = False
    P = fa.ilu_symbolic(D, 2)
    F = oracle.SparseILU(A, 2).factors()
    assert (P.indptr == F.indptr).all() and (P.indices == F.indices).all()


def test_symbolic_rejects_bad_input():
    with pytest.raises(L.FsException) as e:
        fa.ilu_symbolic(sp.csr_matrix(np.ones((2, 3))), 0)
    assert e.value.code == L.FS_ERR_BAD_INPUT
    bad = sp.csr_matrix((np.ones(1), np.array([5], np.int32), np.array([0, 1, 1])), shape=(2, 2),
                        )
    with pytest.raises(Exception):
        fa.ilu_symbolic(bad, 0)


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(fs_[A-Za-z_0-9]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    declared = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            declared |= _declared(h)
    assert "fs_stokes_solve" in declared and "fs_system_create" in declared
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(L.EXPORTS) <= declared


def test_default_options_follow_reference():
    o = fa.stokes.default_opts()
    # restart 0: each method's own basis, 23 for the reference GMRES
    # (max_n_tmp_vectors = 25, stokes_ilu.cc:538), 60 for FGMRES, the default
    # method, with 60 inner ILU-PCG steps on A (the converging configuration)
    assert o.restart == 0 and o.method == L.FS_METHOD_FGMRES and o.inner_a_iters == 60
    assert o.rtol == 1e-8            # tol = 1e-8 ||rhs|| (stokes_ilu.cc:533)
    assert o.reduce == 1e-6          # StokesSolverControl reduce (stokes_common.h:343)
    assert o.inner_cg_tol == 1e-12   # InverseMatrixCG (linear_operators.h:44-45)
    assert o.ilu_k_A == 1 and o.ilu_k_S == 0  # set_level_of_fill defaults


def test_setters_reject_a_null_handle():
    lib = L.lib()
    assert lib.fs_stokes_set_krylov(None, 60, 30, -1.0, -1, -1.0, -1) == L.FS_ERR_BAD_INPUT
    assert lib.fs_stokes_set_method(None, L.FS_METHOD_FGMRES) == L.FS_ERR_BAD_INPUT
    assert lib.fs_stokes_set_precision(None, L.FS_PREC_MIXED) == L.FS_ERR_BAD_INPUT


def test_create_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    S = system("C1", 4)
    with pytest.raises(L.FsException) as e:
        fa.StokesLinearSystem(S.blocks)
    assert e.value.code == L.FS_ERR_NO_DEVICE


def test_solver_control_statistics(tmp_path):
    c = fa.StokesSolverControl()
    c.n_gmres_iter.append(5)
    c.gmres_error.append(1e-9)
    c.gmres_timer.append(0.1)
    c.log_gmres = True
    c.write_statistics(str(tmp_path / "stats"))
    lines = open(str(tmp_path / "stats.csv")).read().splitlines()
    assert lines[0] == "gmres_error,gmres_timer,n_gmres_iter,"
    assert lines[1].startswith("1e-09,0.1,5")


@pytest.mark.parametrize("cfg,n,k,nodes", [("C1", 8, 0, None), ("C2", 16, 1, 5), ("C3", 4, 1, 3)])
def test_block_color_schedule_is_valid(cfg, n, k, nodes):
    """The schedule the GPU ILU runs (fs_block_color_ordering): a permutation,
    blocks of at most 256 rows, and no ILU(k) factor entry joining two blocks
    of one colour stage in the permuted matrix (the blocks of a stage are
    solved at once)."""
    S = system(cfg, n)
    A = S.blocks["A"]
    hint = S.block_hint(nodes) if nodes else None
    perm, blk, stg = fa.block_color_ordering(A, k, 0 if nodes else 64, hint)
    m = A.shape[0]
    assert np.array_equal(np.sort(perm), np.arange(m))
    sizes = np.diff(blk)
    assert blk[0] == 0 and blk[-1] == m and sizes.min() >= 1 and sizes.max() <= 256
    assert stg[0] == 0 and stg[-1] == len(blk) - 1 and np.diff(stg).min() >= 1
    B = A[perm][:, perm].tocsr()
    B.sort_indices()
    F = fa.ilu_symbolic(B, k).tocoo()
    block_of = np.repeat(np.arange(len(blk) - 1), sizes)
    stage_of = np.repeat(np.arange(len(stg) - 1), np.diff(stg))
    bi, bj = block_of[F.row], block_of[F.col]
    assert not np.any((bi != bj) & (stage_of[bi] == stage_of[bj]))
    # every row's entries in other blocks lie in earlier stages (lower part)
    # or later stages (upper part): the forward sweep needs only finished stages
    lower = F.col < F.row
    assert np.all(stage_of[bj[lower & (bi != bj)]] < stage_of[bi[lower & (bi != bj)]])


@pytest.mark.parametrize("seed", range(6))
def test_block_color_schedule_nonsymmetric_k2(seed):
    """A structurally nonsymmetric pattern at level of fill 2: a block can be
    reached from blocks it does not reach itself, so the colouring must treat
    the coupling in both directions. No ILU(2) entry of the permuted matrix
    may join two blocks of one stage, and the stage-parallel symbolic phase
    that runs on that schedule must not meet such a pair."""
    import scipy.sparse as sp
    rng = np.random.default_rng(100 + seed)
    m = 600
    rows, cols = [np.arange(m)], [np.arange(m)]
    for _ in range(3):  # one-directional couplings only
        r = rng.integers(0, m, m)
        c = (r + rng.integers(1, 40, m)) % m
        rows.append(r)
        cols.append(c)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    A = sp.csr_matrix((np.ones(len(r)), (r, c)), shape=(m, m))
    A.sum_duplicates()
    A.sort_indices()
    assert (A != A.T).nnz > 0
    perm, blk, stg = fa.block_color_ordering(A, 2, 8, None)
    B = A[perm][:, perm].tocsr()
    B.sort_indices()
    F = fa.ilu_symbolic(B, 2).tocoo()
    block_of = np.repeat(np.arange(len(blk) - 1), np.diff(blk))
    stage_of = np.repeat(np.arange(len(stg) - 1), np.diff(stg))
    bi, bj = block_of[F.row], block_of[F.col]
    assert not np.any((bi != bj) & (stage_of[bi] == stage_of[bj]))
