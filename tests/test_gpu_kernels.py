"""GPU parity of the per-function kernels against the CPU oracle, through
the C-ABI: SparseMatrix::vmult, SparseILU initialize (numeric factors),
vmult and Tvmult, in the reference order and in the block-colour order."""
import numpy as np
import pytest

import felspa_amd as fa
from felspa_amd import _lib as L
from conftest import rel, system

pytestmark = pytest.mark.gpu


def _perm_matrix(A, perm):
    return A[perm][:, perm].tocsr()


@pytest.mark.parametrize("cfg,n", [("C1", 8), ("C2", 8), ("C3", 3)])
def test_spmv_matches_oracle(oracle, cfg, n):
    S = system(cfg, n)
    K = S.K()
    rng = np.random.default_rng(1)
    x = rng.standard_normal(K.shape[1])
    M = fa.SparseMatrix(K)
    y = M.vmult(x)
    y_ref = oracle.spmv(K, x)
    assert rel(y, y_ref) < 1e-14
    y32 = M.vmult(x, precision=L.FS_PREC_MIXED)
    assert rel(y32, y_ref) < 2e-6


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("cfg,n", [("C1", 8), ("C3", 3)])
def test_ilu_factors_reference_order(oracle, cfg, n, k):
    A = system(cfg, n).blocks["A"]
    ilu = fa.SparseILU(ordering=L.FS_ORDER_REFERENCE)
    ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=k))
    F, perm = ilu.factors()
    assert (perm == np.arange(A.shape[0])).all()
    R = oracle.SparseILU(A, k).factors()
    assert (F.indptr == R.indptr).all() and (F.indices == R.indices).all()
    assert np.abs(F.data - R.data).max() <= 1e-12 * np.abs(R.data).max()
    rng = np.random.default_rng(2)
    b = rng.standard_normal(A.shape[0])
    ref = oracle.SparseILU(A, k)
    assert rel(ilu.vmult(b), ref.vmult(b)) < 1e-12
    assert rel(ilu.Tvmult(b), ref.Tvmult(b)) < 1e-12
    assert rel(ilu.vmult(b, precision=L.FS_PREC_MIXED), ref.vmult(b)) < 1e-4


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("cfg,n", [("C1", 16), ("C2", 16), ("C3", 4)])
def test_ilu_block_color(oracle, cfg, n, k):
    """The block-colour ILU is the reference SparseILU of P A P^T."""
    S = system(cfg, n)
    A = S.blocks["A"]
    ilu = fa.SparseILU(ordering=L.FS_ORDER_BLOCK_COLOR, block_hint=S.block_hint(2 * k + 3))
    ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=k))
    F, perm = ilu.factors()
    assert sorted(perm) == list(range(A.shape[0]))
    PA = _perm_matrix(A, perm)
    R = oracle.SparseILU(PA, k).factors()
    assert (F.indptr == R.indptr).all() and (F.indices == R.indices).all()
    assert np.abs(F.data - R.data).max() <= 1e-12 * np.abs(R.data).max()
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.shape[0])
    ref = oracle.SparseILU(PA, k)
    # vmult in the caller's order equals the oracle on the permuted system
    want = np.empty_like(b)
    want[perm] = ref.vmult(b[perm])
    assert rel(ilu.vmult(b), want) < 1e-12
    wantT = np.empty_like(b)
    wantT[perm] = ref.Tvmult(b[perm])
    assert rel(ilu.Tvmult(b), wantT) < 1e-12


@pytest.mark.parametrize("cfg,n,hint", [("C3", 6, 3), ("C2", 24, 5)])
def test_ilu_dense_stages_large_blocks(oracle, cfg, n, hint):
    """ILU(1) in lattice blocks as the solver uses them (3D: 81-row blocks
    whose thousands of external chunks are summed in pieces by several
    workgroups); FP64 and FP32 applications against the oracle."""
    S = system(cfg, n)
    A = S.blocks["A"]
    ilu = fa.SparseILU(ordering=L.FS_ORDER_BLOCK_COLOR, block_hint=S.block_hint(hint))
    ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=1))
    F, perm = ilu.factors()
    PA = _perm_matrix(A, perm)
    ref = oracle.SparseILU(PA, 1)
    R = ref.factors()
    assert (F.indptr == R.indptr).all() and (F.indices == R.indices).all()
    assert np.abs(F.data - R.data).max() <= 1e-12 * np.abs(R.data).max()
    # largest block's count of 4-entry chunks of external L entries: above
    # 1536 its sums are split into pieces
    h = np.asarray(S.block_hint(hint))[perm]
    starts = np.flatnonzero(np.r_[True, h[1:] != h[:-1]])
    chunks = []
    for a, e in zip(starts, np.r_[starts[1:], len(h)]):
        c = 0
        for r in range(a, e):
            cols = F.indices[F.indptr[r]:F.indptr[r + 1]]
            c += (int((cols < a).sum()) + 3) // 4
        chunks.append(c)
    if cfg == "C3":
        assert max(chunks) > 1536
    rng = np.random.default_rng(4)
    for trial in range(2):  # the second application reuses the pieces' buffers
        b = rng.standard_normal(A.shape[0])
        want = np.empty_like(b)
        want[perm] = ref.vmult(b[perm])
        assert rel(ilu.vmult(b), want) < 1e-12
        assert rel(ilu.vmult(b, precision=L.FS_PREC_MIXED), want) < 1e-4


def test_ilu_zero_pivot_reported():
    import scipy.sparse as sp
    A = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 2.0]]))
    ilu = fa.SparseILU(ordering=L.FS_ORDER_REFERENCE)
    with pytest.raises(L.FsException) as e:
        ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=0))
    assert e.value.code == L.FS_ERR_ZERO_PIVOT


@pytest.mark.parametrize("cfg,n,hint", [("C5", 4, 3), ("C5", 6, 3), ("C3", 4, 3), ("C2", 16, 5)])
def test_ilu2_block_color(oracle, cfg, n, hint):
    """Level of fill 2 in the product path (block-colour order, lattice
    blocks, dense stage kernels with split pieces, the numeric kernel's
    windowed L columns): the factor is the reference SparseILU of P A P^T,
    pattern bit for bit and values to 1e-12; FP64 applications to 1e-12 and
    FP32 ones to 1e-4 of the oracle."""
    S = system(cfg, n)
    A = S.blocks["A"]
    ilu = fa.SparseILU(ordering=L.FS_ORDER_BLOCK_COLOR, block_hint=S.block_hint(hint))
    ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=2))
    F, perm = ilu.factors()
    assert sorted(perm) == list(range(A.shape[0]))
    PA = _perm_matrix(A, perm)
    ref = oracle.SparseILU(PA, 2)
    R = ref.factors()
    assert (F.indptr == R.indptr).all() and (F.indices == R.indices).all()
    assert np.abs(F.data - R.data).max() <= 1e-12 * np.abs(R.data).max()
    rng = np.random.default_rng(5)
    for trial in range(2):
        b = rng.standard_normal(A.shape[0])
        want = np.empty_like(b)
        want[perm] = ref.vmult(b[perm])
        assert rel(ilu.vmult(b), want) < 1e-12
        assert rel(ilu.vmult(b, precision=L.FS_PREC_MIXED), want) < 1e-4
    # refactorisation with new values on the same pattern
    A2 = A.copy()
    A2.data = A.data * (1.0 + 0.1 * rng.random(A.nnz))
    A2 = (A2 + A2.T) * 0.5
    A2 = A2.tocsr()
    A2.sort_indices()
    if A2.nnz == A.nnz:
        ilu.initialize(A2, fa.SparseILU.AdditionalData(use_previous_sparsity=True, level_of_fill=2))
        F2, _ = ilu.factors()
        R2 = oracle.SparseILU(_perm_matrix(A2, perm), 2).factors()
        assert np.abs(F2.data - R2.data).max() <= 1e-12 * np.abs(R2.data).max()


@pytest.mark.parametrize("cfg,n,hint", [("C5", 6, 3), ("C3", 4, 3)])
def test_ilu2_rowwise_in_block_stages(oracle, monkeypatch, cfg, n, hint):
    """The in-block stages one row at a time (the form long-row stages take
    when three rows do not fit in LDS, forced by FS_ILU_ROWWISE=1): the
    factor is the reference SparseILU of P A P^T to 1e-12."""
    monkeypatch.setenv("FS_ILU_ROWWISE", "1")
    S = system(cfg, n)
    A = S.blocks["A"]
    ilu = fa.SparseILU(ordering=L.FS_ORDER_BLOCK_COLOR, block_hint=S.block_hint(hint))
    ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=2))
    F, perm = ilu.factors()
    R = oracle.SparseILU(_perm_matrix(A, perm), 2).factors()
    assert (F.indptr == R.indptr).all() and (F.indices == R.indices).all()
    assert np.abs(F.data - R.data).max() <= 1e-12 * np.abs(R.data).max()


def test_ilu_oversize_row_unsupported():
    """A factor row longer than the numeric kernel's LDS plan (an arrow
    matrix: the last row and column full) is refused at the symbolic phase
    with FS_ERR_UNSUPPORTED, before any kernel runs."""
    import scipy.sparse as sp
    m = 9000
    r = np.concatenate([np.arange(m), np.full(m - 1, m - 1), np.arange(m - 1)])
    c = np.concatenate([np.arange(m), np.arange(m - 1), np.full(m - 1, m - 1)])
    v = np.concatenate([np.full(m, 4.0), np.full(2 * (m - 1), 1e-3)])
    A = sp.csr_matrix((v, (r, c)), shape=(m, m))
    A.sort_indices()
    hint = np.arange(m, dtype=np.int32) // 64
    hint[-1] = m  # the arrow row in a block of its own, coloured last
    ilu = fa.SparseILU(ordering=L.FS_ORDER_BLOCK_COLOR, block_hint=hint)
    with pytest.raises(L.FsException) as e:
        ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=2))
    assert e.value.code == L.FS_ERR_UNSUPPORTED
