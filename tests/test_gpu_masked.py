"""GPU parity of the non0tab-masked GEMM against the CPU oracle (relative Frobenius <= 1e-12)."""
import numpy as np
import pytest

import cfgmm
import oracle
from cfgmm import workloads as W
from helpers import op_view, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("blksize", [16, 48, 64, 128])
@pytest.mark.parametrize("ta", [False, True])
def test_masked_gemm_vs_oracle(blksize, ta):
    rng = np.random.default_rng(blksize * 2 + ta)
    m, n, k = 300, 1000, 530
    lda = m if ta else k
    a = rng.uniform(-1, 1, size=(k if ta else m) * lda)
    b = rng.uniform(-1, 1, size=k * n)  # nonzero under off bits: the mask must exclude them
    nblk = (n + blksize - 1) // blksize
    tab = W.periodic_mask(k, nblk, W.MASK_30)
    c = rng.uniform(-1, 1, size=m * n)
    c_ref = c.copy()
    assert oracle.matmul_non0tab(c_ref, a, b, m, n, k, lda, n, n, ta, False, tab, nblk, blksize) == 0
    cfgmm.MatmulLoops().matmul_loop_parallel_mnk_pack_a_non0tab(blksize, c, a, b, m, n, k, lda, n, n, ta, False, tab,
                                                                nblk)
    assert rel_fro(c, c_ref) <= TOL


def test_masked_gemm_reference_test_shape():
    """src/impl_matmul_non0tab.rs:329-377: m = k = 2228, n = 16384, sin/cos, 30 % periodic mask, BLKSIZE 48."""
    m, n, k = 2228, 16384, 2228
    a, b = W.sin_cos_operands(m * k, k * n)
    nblk = (n + 47) // 48
    tab = W.periodic_mask(k, nblk, W.MASK_30)
    W.zero_under_mask(b, tab, 48, n)
    c = np.zeros(m * n)
    cfgmm.matmul_anyway_full_non0tab(c, a, b, m, n, k, k, n, n, False, False, tab, nblk)
    c_ref = np.zeros(m * n)
    assert oracle.matmul_non0tab(c_ref, a, b, m, n, k, k, n, n, False, False, tab, nblk, 48) == 0
    assert rel_fro(c, c_ref) <= TOL


def test_masked_gemm_empty_mask_leaves_c():
    rng = np.random.default_rng(1)
    m, n, k = 50, 100, 70
    a = rng.uniform(-1, 1, m * k)
    b = rng.uniform(-1, 1, k * n)
    c = rng.uniform(-1, 1, m * n)
    c0 = c.copy()
    tab = np.zeros((k, 2), np.uint8)
    cfgmm.MatmulLoops().matmul_loop_parallel_mnk_pack_a_non0tab(64, c, a, b, m, n, k, k, n, n, False, False, tab, 2)
    assert np.array_equal(c, c0)
