"""fs_ilu_vmult_noalloc: the ILU application on device vectors whose work
vectors sit in the library's static device area (no allocation), against
fs_ilu_vmult (host vectors, library-allocated work vectors) on the same
factor: the same kernels, so the same values; two calls in a row, the second
on a smaller system, and a factor beyond the static area refused."""
import numpy as np
import pytest

import felspa_amd as fa
from felspa_amd import _lib as L
from conftest import system

pytestmark = pytest.mark.gpu


def _apply_noalloc(ilu, b, precision):
    import torch
    d_src = torch.from_numpy(b).cuda()
    d_dst = torch.full_like(d_src, float("nan"))
    torch.cuda.synchronize()
    ilu.vmult_device_noalloc(d_src.data_ptr(), d_dst.data_ptr(), precision)
    return d_dst.cpu().numpy()


def test_noalloc_matches_vmult_twice_smaller_second():
    rng = np.random.default_rng(21)
    for cfg, n, k, hint in (("C2", 24, 1, 5), ("C1", 8, 0, 5)):  # the second call is the smaller input
        S = system(cfg, n)
        A = S.blocks["A"]
        ilu = fa.SparseILU(ordering=L.FS_ORDER_BLOCK_COLOR, block_hint=S.block_hint(hint))
        ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=k))
        b = rng.standard_normal(A.shape[0])
        for prec in (L.FS_PREC_FP64, L.FS_PREC_MIXED):
            want = ilu.vmult(b, precision=prec)
            got = _apply_noalloc(ilu, b, prec)
            assert np.all(np.isfinite(got))
            assert np.array_equal(got, want), (cfg, prec, np.abs(got - want).max())


def test_noalloc_refuses_more_rows_than_the_static_area():
    import scipy.sparse as sp
    import torch
    m = (1 << 21) + 1
    A = sp.identity(m, format="csr") * 2.0
    A.sort_indices()
    ilu = fa.SparseILU(ordering=L.FS_ORDER_REFERENCE)
    ilu.initialize(A, fa.SparseILU.AdditionalData(level_of_fill=0))
    d = torch.zeros(m, dtype=torch.float64, device="cuda")
    with pytest.raises(L.FsException) as e:
        ilu.vmult_device_noalloc(d.data_ptr(), d.data_ptr(), L.FS_PREC_FP64)
    assert e.value.code == L.FS_ERR_UNSUPPORTED
