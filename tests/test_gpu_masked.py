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


@pytest.mark.parametrize("blksize", [16, 48, 64])
@pytest.mark.parametrize("ta", [False, True])
def test_masked_gemm_transb_ex(blksize, ta):
    """cfgmm_dgemm_non0tab_ex_dev with transb: checked against the oracle on an explicitly transposed
    copy of B (op(B)(p, j) = b[j*ldb + p]); padded leading dimensions, a partial last column block,
    nonzero B under off bits."""
    rng = np.random.default_rng(100 + blksize + ta)
    m, n, k = 211, 700, 389
    lda = (m if ta else k) + 3
    ldb = k + 5
    ldc = n + 2
    a = rng.uniform(-1, 1, size=(k if ta else m) * lda)
    b = rng.uniform(-1, 1, size=n * ldb)
    nblk = (n + blksize - 1) // blksize
    tab = W.periodic_mask(k, nblk + 1, W.MASK_30)
    c = rng.uniform(-1, 1, size=m * ldc)
    bt = np.ascontiguousarray(b.reshape(n, ldb)[:, :k].T).ravel()  # k x n, row-major
    c_ref = c.copy()
    assert oracle.matmul_non0tab(c_ref, a, bt, m, n, k, lda, n, ldc, ta, False, tab, nblk + 1, blksize) == 0
    s = cfgmm.Stream()
    dA, dB, dC, dT = (cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(x)) for x in (a, b, c, tab))
    cfgmm.dgemm_non0tab_dev(s, dC, dA, dB, m, n, k, lda, ldb, ldc, ta, True, dT, nblk + 1, blksize,
                            allow_transb=True)
    s.synchronize()
    got = dC.to_array(np.float64, (m * ldc,))
    with pytest.raises(cfgmm.CfgmmError) as e:  # the reference-shaped entry point still rejects transb
        cfgmm.dgemm_non0tab_dev(s, dC, dA, dB, m, n, k, lda, ldb, ldc, ta, True, dT, nblk + 1, blksize)
    assert e.value.status == cfgmm.EINVAL_TRANSB_MASK
    for x in (dA, dB, dC, dT):
        x.free()
    s.destroy()
    G, R = got.reshape(m, ldc), c_ref.reshape(m, ldc)
    assert rel_fro(G[:, :n], R[:, :n]) <= TOL
    assert np.array_equal(G[:, n:], c.reshape(m, ldc)[:, n:])


@pytest.mark.parametrize("blksize", [16, 64])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_masked_sgemm_f32(blksize, ta, tb):
    """cfgmm_sgemm_non0tab_dev (T = f32, not implemented in the reference): against the f64 oracle on
    the same float32 values (relative Frobenius <= 1e-5, f32 accumulation over ~160 listed terms),
    padded leading dimensions, a partial last column block, nonzero B under off bits; bitwise
    repeatable."""
    rng = np.random.default_rng(300 + blksize + 2 * ta + tb)
    m, n, k = 211, 700, 530
    lda = (m if ta else k) + 3
    ldb = (k if tb else n) + 5
    ldc = n + 2
    a = rng.uniform(-1, 1, size=(k if ta else m) * lda).astype(np.float32)
    b = rng.uniform(-1, 1, size=(n if tb else k) * ldb).astype(np.float32)
    nblk = (n + blksize - 1) // blksize
    tab = W.periodic_mask(k, nblk + 1, W.MASK_30)
    c = rng.uniform(-1, 1, size=m * ldc).astype(np.float32)
    bk = b.reshape(n, ldb)[:, :k].T if tb else b.reshape(k, ldb)[:, :n]  # op(B), k x n
    c_ref = c.astype(np.float64)
    assert oracle.matmul_non0tab(c_ref, a.astype(np.float64), np.ascontiguousarray(bk, dtype=np.float64).ravel(), m,
                                 n, k, lda, n, ldc, ta, False, tab, nblk + 1, blksize) == 0
    s = cfgmm.Stream()
    dA, dB, dT = (cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(x)) for x in (a, b, tab))
    outs = []
    for _ in range(2):
        dC = cfgmm.DeviceBuffer.from_array(c)
        cfgmm.sgemm_non0tab_dev(s, dC, dA, dB, m, n, k, lda, ldb, ldc, ta, tb, dT, nblk + 1, blksize)
        s.synchronize()
        outs.append(dC.to_array(np.float32, (m, ldc)))
        dC.free()
    for x in (dA, dB, dT):
        x.free()
    s.destroy()
    assert np.array_equal(outs[0], outs[1])
    G, R = outs[0].astype(np.float64), c_ref.reshape(m, ldc)
    assert rel_fro(G[:, :n], R[:, :n]) <= 1e-5
    assert np.array_equal(outs[0][:, n:], c.reshape(m, ldc)[:, n:])


@pytest.mark.parametrize("blksize", [16, 32, 48, 64])
@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("n", [998, 999])
def test_masked_gemm_long_lists_both_paths(blksize, ta, n):
    """B not transposed: even n takes the LDS-copy kernel (op(A) transposed into the workspace;
    128-row tiles, and 256-row tiles under CFGMM_TUNE_MASKED_LDS = 2), odd n the register-staged
    one.  A 90 % mask over k = 2600 gives lists of ~2340 entries (longer than the register kernel's
    2048-entry LDS segments, ending in a partial slice) beside one empty list; odd ldb, padded lda
    and ldc, a partial last column block.  Each path against the oracle, the default bitwise
    repeatable."""
    rng = np.random.default_rng(7000 + blksize + 2 * ta + n)
    m, k = 200, 2600
    lda = (m if ta else k) + 3
    ldb, ldc = n + 3, n + 1
    a = rng.uniform(-1, 1, size=(k if ta else m) * lda)
    b = rng.uniform(-1, 1, size=k * ldb)
    nblk = (n + blksize - 1) // blksize
    tab = (rng.random((k, nblk)) < 0.9).astype(np.uint8)
    tab[:, nblk // 2] = 0  # one empty list
    c = rng.uniform(-1, 1, size=m * ldc)
    c_ref = c.copy()
    bk = np.ascontiguousarray(b.reshape(k, ldb)[:, :n]).ravel()
    assert oracle.matmul_non0tab(c_ref, a, bk, m, n, k, lda, n, ldc, ta, False, tab, nblk, blksize) == 0
    s = cfgmm.Stream()
    dA, dB, dT = (cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(x)) for x in (a, b, tab))
    outs = []
    old = cfgmm.tune_get(cfgmm.TUNE_MASKED_LDS)
    try:
        for mode in (0, 0, 1, 2):
            cfgmm.tune_set(cfgmm.TUNE_MASKED_LDS, mode)
            dC = cfgmm.DeviceBuffer.from_array(c)
            cfgmm.dgemm_non0tab_dev(s, dC, dA, dB, m, n, k, lda, ldb, ldc, ta, False, dT, nblk, blksize)
            s.synchronize()
            outs.append(dC.to_array(np.float64, (m, ldc)))
            dC.free()
    finally:
        cfgmm.tune_set(cfgmm.TUNE_MASKED_LDS, old)
    for x in (dA, dB, dT):
        x.free()
    s.destroy()
    assert np.array_equal(outs[0], outs[1])
    R = c_ref.reshape(m, ldc)
    for G in outs[1:]:
        assert rel_fro(G[:, :n], R[:, :n]) <= TOL
        assert np.array_equal(G[:, n:], c.reshape(m, ldc)[:, n:])
