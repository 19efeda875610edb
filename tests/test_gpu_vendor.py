"""The vendor comparison (rocSPARSE generic SpGEMM in binary64, the counterpart of the
reference's efficient/mkl_*.cpp programs) on the config-1 operands of A-P-V-P-A's final product:
the same structure as this library's product, row by row, and counts equal to the uint64 counts
(every count is below 2^53 here, so binary64 holds it exactly)."""
import numpy as np
import pytest

from resmcm_ari_amd import configs, smcm, vendor

pytestmark = pytest.mark.gpu


def test_rocsparse_same_structure_config1(ctx):
    hg = configs.CFG1.build()
    G = smcm.Graph(ctx, hg)
    try:
        X, _ = G.run(configs.metapath_from_letters("A-P-V", "bib"))
        Y, _ = G.run(configs.metapath_from_letters("V-P-A", "bib"))
        xp, xc, xv = X.to_host()
        yp, yc, yv = Y.to_host()
        X.free()
        Y.free()
    finally:
        G.close()
    off = hg.type_offset
    xc = (xc - off[configs.V]).astype(np.int32)
    yc = (yc - off[configs.A]).astype(np.int32)
    m, k, n = len(xp) - 1, len(yp) - 1, int(hg.dimensions[configs.A])
    r = ctx.spgemm_csr(m, k, n, xp, xc, xv, yp, yc, yv)
    gp, gc, gv = r.to_host()
    r.free()
    _, nnz, (vp, vc, vv) = vendor.spgemm_f64(m, k, n, xp, xc, xv.astype(np.float64), yp, yc, yv.astype(np.float64),
                                             reps=1)
    assert nnz == len(gc)
    np.testing.assert_array_equal(vp, gp)
    rows = np.repeat(np.arange(m), np.diff(vp))
    order = np.lexsort((vc, rows))  # rocSPARSE may leave a row's columns unsorted
    np.testing.assert_array_equal(vc[order], gc)
    assert gv.max() < 2**53
    np.testing.assert_array_equal(vv[order], gv.astype(np.float64))


def test_rocsparse_rejects_decreasing_row_pointers():
    ptr = np.array([0, 2, 1, 3], np.int64)
    col = np.array([0, 1, 0], np.int32)
    with pytest.raises(ValueError):
        vendor.spgemm_f64(3, 2, 2, ptr, col, None, np.array([0, 1, 2], np.int64), np.array([0, 1], np.int32), None)
