"""GPU parity of the FP64-MFMA GEMM (through the C ABI) against the CPU oracle.

Tolerance: relative Frobenius error <= 1e-12 (plain FP64)."""
import itertools

import numpy as np
import pytest

import cfgmm
import oracle
from cfgmm import workloads as W
from helpers import op_view, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _run_both(m, n, k, ta, tb, seed, pad=(0, 0, 0), c_init=True, api="host"):
    rng = np.random.default_rng(seed)
    lda = (m if ta else k) + pad[0]
    ldb = (k if tb else n) + pad[1]
    ldc = n + pad[2]
    a = rng.uniform(-1, 1, size=(k if ta else m) * lda)
    b = rng.uniform(-1, 1, size=(n if tb else k) * ldb)
    c = rng.uniform(-1, 1, size=m * ldc) if c_init else np.zeros(m * ldc)
    c_ref = c.copy()
    oracle.matmul(c_ref, a, b, m, n, k, lda, ldb, ldc, ta, tb)
    c_gpu = c.copy()
    if api == "host":
        cfgmm.MatmulLoops(252, 512, 240, 14, 2, 8).matmul_loop_parallel_mnk_pack_a(c_gpu, a, b, m, n, k, lda, ldb,
                                                                                   ldc, ta, tb)
    else:
        s = cfgmm.Stream()
        dA, dB, dC = (cfgmm.DeviceBuffer.from_array(x) for x in (a, b, c_gpu))
        cfgmm.dgemm_dev(s, dC, dA, dB, m, n, k, lda, ldb, ldc, ta, tb)
        s.synchronize()
        dC.download(c_gpu)
        s.destroy()
    return c, c_ref, c_gpu, ldc


def test_dgemm_small_nt_workload():
    """C(1031x977) += A(1031x1543) . B(977x1543)^T, uniform[-1,1) seed 0, C = 0."""
    g = W.dgemm_small_nt(seed=0)
    c_ref = np.zeros(g.m * g.ldc)
    oracle.matmul(c_ref, g.a, g.b, g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.transa, g.transb)
    c = np.zeros(g.m * g.ldc)
    cfgmm.MatmulLoops().matmul_loop_macro_mb(c, g.a, g.b, g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.transa, g.transb)
    err = rel_fro(c, c_ref)
    print(f"dgemm 1031x977x1543 NT rel_fro = {err:.3e}")
    assert err <= TOL
    # numpy (OpenBLAS) as a second, independent anchor
    A, B = g.dense()
    assert rel_fro(c.reshape(g.m, g.n), A @ B) <= TOL


@pytest.mark.parametrize("ta,tb", list(itertools.product([False, True], repeat=2)))
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 7), (130, 129, 100), (300, 520, 700), (1100, 1200, 300)])
def test_dgemm_trans_ragged(m, n, k, ta, tb):
    c0, c_ref, c_gpu, ldc = _run_both(m, n, k, ta, tb, seed=m + n + k, pad=(1, 2, 3))
    assert rel_fro(c_gpu.reshape(m, ldc)[:, :n], c_ref.reshape(m, ldc)[:, :n]) <= TOL
    # columns of C beyond n are never written
    assert np.array_equal(c_gpu.reshape(m, ldc)[:, n:], c0.reshape(m, ldc)[:, n:])


@pytest.mark.parametrize("api", ["host", "dev"])
def test_dgemm_many_small_tiles(api):
    """Enough 64x64 tiles to fill the chip without splitting K."""
    m, n, k = 2100, 2300, 333
    _, c_ref, c_gpu, _ = _run_both(m, n, k, False, False, seed=1, api=api)
    assert rel_fro(c_gpu, c_ref) <= TOL


@pytest.mark.parametrize("ta,tb", [(False, False), (True, True)])
def test_dgemm_whole_tiles_plus_split_tail(ta, tb):
    """529 tiles of 128x128: 512 whole tiles, the last 17 split along K and reduced in a fixed order."""
    m, n, k = 2900, 2900, 1000
    c0, c_ref, c_gpu, ldc = _run_both(m, n, k, ta, tb, seed=4, pad=(0, 0, 5))
    assert rel_fro(c_gpu.reshape(m, ldc)[:, :n], c_ref.reshape(m, ldc)[:, :n]) <= TOL
    assert np.array_equal(c_gpu.reshape(m, ldc)[:, n:], c0.reshape(m, ldc)[:, n:])


def test_dgemm_split_k_path_deterministic():
    """Few tiles + deep K takes the split-K path; the fixed-order reduction is bitwise repeatable."""
    m, n, k = 200, 190, 4100
    _, c_ref, c1, _ = _run_both(m, n, k, True, True, seed=2)
    assert rel_fro(c1, c_ref) <= TOL
    _, _, c2, _ = _run_both(m, n, k, True, True, seed=2)
    assert np.array_equal(c1, c2)


def test_dgemm_integer_exact():
    """Small-integer operands: every partial sum is an exact f64 integer -> bitwise exact result."""
    rng = np.random.default_rng(7)
    m, n, k = 700, 650, 900
    a = rng.integers(-3, 4, size=m * k).astype(np.float64)
    b = rng.integers(-3, 4, size=n * k).astype(np.float64)
    c = rng.integers(-5, 6, size=m * n).astype(np.float64)
    ref = c.reshape(m, n).astype(np.int64) + a.reshape(m, k).astype(np.int64) @ b.reshape(n, k).astype(np.int64).T
    cfgmm.matmul_anyway_full(c, a, b, m, n, k, k, k, n, False, True)
    assert np.array_equal(c.reshape(m, n).astype(np.int64), ref)


def test_dgemm_accumulates_twice():
    """Two calls add twice (C += op(A) op(B) has no beta)."""
    rng = np.random.default_rng(8)
    m, n, k = 64, 80, 96
    a = rng.uniform(-1, 1, m * k)
    b = rng.uniform(-1, 1, k * n)
    c = np.zeros(m * n)
    cfgmm.matmul_anyway_full(c, a, b, m, n, k, k, n, n, False, False)
    cfgmm.matmul_anyway_full(c, a, b, m, n, k, k, n, n, False, False)
    assert rel_fro(c.reshape(m, n), 2 * (a.reshape(m, k) @ b.reshape(k, n))) <= TOL


@pytest.mark.parametrize("tile,splits", [(64, 1), (64, 3), (64, 5), (128, 2), (128, 7)])
@pytest.mark.parametrize("ta,tb", [(False, True), (True, False)])
def test_dgemm_small_plans(tile, splits, ta, tb):
    """Every tile edge x K-range plan the autotuner may pick for a small GEMM meets the FP64 bar."""
    old = (cfgmm.tune_get(cfgmm.TUNE_DGEMM_SMALL_TILE), cfgmm.tune_get(cfgmm.TUNE_DGEMM_SMALL_SPLITS))
    try:
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_SMALL_TILE, tile)
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_SMALL_SPLITS, splits)
        c, c_ref, c_gpu, ldc = _run_both(301, 277, 899, ta, tb, seed=tile + splits, pad=(3, 1, 2), api="dev")
    finally:
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_SMALL_TILE, old[0])
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_SMALL_SPLITS, old[1])
    G = c_gpu.reshape(301, ldc)
    R = c_ref.reshape(301, ldc)
    assert rel_fro(G[:, :277], R[:, :277]) <= TOL
    assert np.array_equal(G[:, 277:], c.reshape(301, ldc)[:, 277:])


def test_dgemm_large_nn_full_size():
    """C(3527x9583) += A(3527x6581) B(6581x9583), NN, uniform[-1, 1) seed 0 (README.md:11-17): the
    plan takes whole 128x128 tiles plus a K-split tail reduced in a fixed order."""
    g = W.dgemm_large_nn(seed=0)
    c_ref = np.zeros(g.m * g.ldc)
    oracle.matmul(c_ref, g.a, g.b, g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.transa, g.transb)
    s = cfgmm.Stream()
    dA, dB = cfgmm.DeviceBuffer.from_array(g.a), cfgmm.DeviceBuffer.from_array(g.b)
    dC = cfgmm.DeviceBuffer(8 * g.m * g.ldc)
    dC.zero()
    cfgmm.dgemm_dev(s, dC, dA, dB, g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.transa, g.transb)
    s.synchronize()
    c = dC.to_array(np.float64, (g.m * g.ldc,))
    s.destroy()
    err = rel_fro(c, c_ref)
    print(f"dgemm 3527x9583x6581 NN rel_fro = {err:.3e}", flush=True)
    assert err <= TOL
    A, B = g.dense()
    assert rel_fro(c.reshape(g.m, g.n), A @ B) <= TOL


# The reference's own GEMM tests (src/impl_matmul.rs:250-350): a[x] = sin(x), b[x] = cos(x) over
# the flat index, m = 3527, n = 9583, k = 6581, through matmul_anyway_full.
@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False)], ids=["NN", "NT", "TN"])
def test_dgemm_reference_harness_sincos(ta, tb):
    m, n, k = 3527, 9583, 6581
    lda = m if ta else k
    ldb = k if tb else n
    a, b = W.sin_cos_operands((k if ta else m) * lda, (n if tb else k) * ldb)
    c_ref = np.zeros(m * n)
    oracle.matmul(c_ref, a, b, m, n, k, lda, ldb, n, ta, tb)
    c = np.zeros(m * n)
    cfgmm.matmul_anyway_full(c, a, b, m, n, k, lda, ldb, n, ta, tb)
    err = rel_fro(c, c_ref)
    print(f"dgemm sin/cos {m}x{n}x{k} trans={ta}{tb} rel_fro = {err:.3e}", flush=True)
    assert err <= TOL


# Shapes whose 64-row tiles of width 64, 80 or 96 give every tile a CU of its own (256 CUs) with
# 3/4 of the CUs or more covered: the in-workgroup K split (dgemm_kgroup_kernel).  k covers whole
# rounds of the four groups (64, 1543 = 24 rounds + a tail), leftover full steps only (224 = 3
# rounds + 2 steps), and a tail round alone (37).
@pytest.mark.parametrize("m,n", [(1000, 961), (1031, 1039), (1031, 1300)], ids=["bn64", "bn80", "bn96"])
@pytest.mark.parametrize("k", [64, 1543, 224, 37])
def test_dgemm_kgroup_plan(m, n, k):
    def run(kgroups):
        old = cfgmm.tune_get(cfgmm.TUNE_DGEMM_KGROUPS)
        try:
            cfgmm.tune_set(cfgmm.TUNE_DGEMM_KGROUPS, kgroups)
            return _run_both(m, n, k, False, True, seed=m + n + k, pad=(2, 3, 1), api="dev")
        finally:
            cfgmm.tune_set(cfgmm.TUNE_DGEMM_KGROUPS, old)
    c0, c_ref, c_kg, ldc = run(2)
    _, _, c_split, _ = run(1)
    G, R, S = (x.reshape(m, ldc) for x in (c_kg, c_ref, c_split))
    assert rel_fro(G[:, :n], R[:, :n]) <= TOL
    assert rel_fro(S[:, :n], R[:, :n]) <= TOL
    assert np.array_equal(G[:, n:], c0.reshape(m, ldc)[:, n:])
    # the fixed group order makes the result bitwise repeatable
    _, _, c_kg2, _ = run(2)
    assert np.array_equal(c_kg, c_kg2)


@pytest.mark.parametrize("kgroups", [2, 0], ids=["kgroup", "default"])
def test_dgemm_kgroup_integer_exact(kgroups):
    """Small-integer operands through the in-workgroup K split, and through the default plan of
    this shape (the K split over workgroups): every partial sum and the reduction of the partial
    tiles are exact, so the result is bitwise the integer product."""
    old = cfgmm.tune_get(cfgmm.TUNE_DGEMM_KGROUPS)
    cfgmm.tune_set(cfgmm.TUNE_DGEMM_KGROUPS, kgroups)
    try:
        _kgroup_integer_exact()
    finally:
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_KGROUPS, old)


def _kgroup_integer_exact():
    rng = np.random.default_rng(9)
    m, n, k = 1031, 977, 1543
    a = rng.integers(-2, 3, size=m * k).astype(np.float64)
    b = rng.integers(-2, 3, size=n * k).astype(np.float64)
    c = rng.integers(-5, 6, size=m * n).astype(np.float64)
    ref = c.reshape(m, n).astype(np.int64) + a.reshape(m, k).astype(np.int64) @ b.reshape(n, k).astype(np.int64).T
    cfgmm.matmul_anyway_full(c, a, b, m, n, k, k, k, n, False, True)
    assert np.array_equal(c.reshape(m, n).astype(np.int64), ref)
