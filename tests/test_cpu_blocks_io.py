"""The binary block-CSR interchange (include/fs_blocks_io.h): a round trip
keeps every array bit for bit (deal.II's diagonal-first rows included), and
malformed or corrupted files are refused."""
import numpy as np
import pytest
import scipy.sparse as sp

import felspa_amd as fa
from felspa_amd import _lib as L
from conftest import system


def _same(a, b):
    return (a.shape == b.shape and np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
            and np.array_equal(a.data, b.data))


@pytest.mark.parametrize("cfg,n", [("C2", 8), ("C3", 3)])
def test_round_trip(tmp_path, cfg, n):
    S = system(cfg, n)
    path = str(tmp_path / "k.fsbk")
    fa.save_blocks(path, S.blocks, S.rhs)
    blocks, rhs = fa.load_blocks(path)
    for k in fa.system.BLOCK_NAMES:
        assert _same(blocks[k], S.blocks[k]), k
    assert np.array_equal(rhs, S.rhs)
    fa.save_blocks(path, S.blocks)  # without a right-hand side
    _, rhs2 = fa.load_blocks(path)
    assert rhs2 is None


def test_diagonal_first_rows_survive(tmp_path):
    """deal.II stores square blocks with the diagonal first in every row; the
    file keeps the writer's order (the solver sorts on input)."""
    S = system("C1", 4)
    blocks = dict(S.blocks)
    A = S.blocks["A"].tocsr()
    rows = []
    for r in range(A.shape[0]):
        c = A.indices[A.indptr[r]:A.indptr[r + 1]]
        v = A.data[A.indptr[r]:A.indptr[r + 1]]
        o = np.argsort(c != r, kind="stable")  # diagonal first, then ascending
        rows.append((c[o], v[o]))
    A2 = sp.csr_matrix((np.concatenate([x[1] for x in rows]), np.concatenate([x[0] for x in rows]), A.indptr),
                       shape=A.shape)
    A2.has_sorted_indices = False
    blocks["A"] = A2
    path = str(tmp_path / "d.fsbk")
    fa.save_blocks(path, blocks, S.rhs)
    got, _ = fa.load_blocks(path)
    assert np.array_equal(got["A"].indices, A2.indices) and np.array_equal(got["A"].data, A2.data)


def test_corrupted_and_malformed_files_are_refused(tmp_path):
    S = system("C2", 6)
    path = str(tmp_path / "c.fsbk")
    fa.save_blocks(path, S.blocks, S.rhs)
    raw = bytearray(open(path, "rb").read())
    raw[len(raw) // 2] ^= 0x40
    open(path, "wb").write(bytes(raw))
    with pytest.raises(L.FsException) as e:
        fa.load_blocks(path)
    assert e.value.code == L.FS_ERR_BAD_INPUT
    open(path, "wb").write(bytes(raw[: len(raw) // 3]))
    with pytest.raises(L.FsException):
        fa.load_blocks(path)
    open(path, "wb").write(b"not a block file at all")
    with pytest.raises(L.FsException):
        fa.load_blocks(path)
    bad = dict(S.blocks)
    bad["B"] = S.blocks["B"][:-1]  # wrong shape
    with pytest.raises(L.FsException):
        fa.save_blocks(str(tmp_path / "x.fsbk"), bad, S.rhs)
