"""Pins of the CPU oracle (oracle/cfgmm_oracle.c), which restates the reference crate's loops.

The reference's own tests only print errors (src/impl_matmul.rs:276-278), so the
oracle is pinned by: numpy float64 (a library routine), exact integer-valued
closed forms, brute-force loops on tiny shapes, the masked == zero-filled-dense
invariant, mask authority, linearity/positivity of rho, and the reference
tests' own sin/cos/periodic-mask generators.  Each check is chosen so that a
dropped term, a wrong index, a swapped transpose or a mis-mapped mask column
fails at least one of them.
"""
import itertools

import numpy as np
import pytest

import oracle
from cfgmm import workloads as W
from helpers import op_view, rel_fro

# blocking boundaries of the restated loop nest: MC=252, KC=512, NC=240, MR=14, NR=16
SHAPES = [(1, 1, 1), (13, 17, 5), (14, 16, 512), (15, 15, 513), (253, 241, 40), (300, 260, 530)]


def _gemm_case(m, n, k, ta, tb, rng, pad=(0, 0, 0), integer=False):
    lda = (m if ta else k) + pad[0]
    ldb = (k if tb else n) + pad[1]
    ldc = n + pad[2]
    a_rows, b_rows = (k if ta else m), (n if tb else k)
    if integer:
        a = rng.integers(-3, 4, size=a_rows * lda).astype(np.float64)
        b = rng.integers(-3, 4, size=b_rows * ldb).astype(np.float64)
        c = rng.integers(-5, 6, size=m * ldc).astype(np.float64)
    else:
        a = rng.uniform(-1, 1, size=a_rows * lda)
        b = rng.uniform(-1, 1, size=b_rows * ldb)
        c = rng.uniform(-1, 1, size=m * ldc)
    return a, b, c, lda, ldb, ldc


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("ta,tb", list(itertools.product([False, True], repeat=2)))
def test_gemm_matches_numpy(m, n, k, ta, tb):
    rng = np.random.default_rng(m * 1000 + n * 10 + k + 7 * ta + 3 * tb)
    a, b, c, lda, ldb, ldc = _gemm_case(m, n, k, ta, tb, rng, pad=(3, 1, 2))
    c0 = c.copy()
    oracle.matmul(c, a, b, m, n, k, lda, ldb, ldc, ta, tb)
    A, B = op_view(a, m, k, lda, ta), op_view(b, k, n, ldb, tb)
    ref = c0.reshape(m, ldc)[:, :n] + A @ B
    got = c.reshape(m, ldc)
    assert rel_fro(got[:, :n], ref) <= 1e-14
    # padding columns of C are untouched (src/impl_matmul.rs:52-70 writes only j < nr)
    assert np.array_equal(got[:, n:], c0.reshape(m, ldc)[:, n:])


@pytest.mark.parametrize("ta,tb", list(itertools.product([False, True], repeat=2)))
def test_gemm_integer_exact(ta, tb):
    """Small-integer operands: every partial sum is an exact f64 integer, so the result is exact."""
    rng = np.random.default_rng(11)
    m, n, k = 260, 250, 600
    a, b, c, lda, ldb, ldc = _gemm_case(m, n, k, ta, tb, rng, integer=True)
    c0 = c.copy()
    oracle.matmul(c, a, b, m, n, k, lda, ldb, ldc, ta, tb)
    A = op_view(a, m, k, lda, ta).astype(np.int64)
    B = op_view(b, k, n, ldb, tb).astype(np.int64)
    ref = c0.reshape(m, ldc).astype(np.int64) + A @ B
    assert np.array_equal(c.reshape(m, ldc).astype(np.int64), ref)


def test_gemm_bruteforce_tiny():
    rng = np.random.default_rng(5)
    for m, n, k in [(1, 1, 1), (2, 3, 4), (14, 16, 3), (15, 17, 2)]:
        for ta, tb in itertools.product([False, True], repeat=2):
            a, b, c, lda, ldb, ldc = _gemm_case(m, n, k, ta, tb, rng)
            c0 = c.copy()
            oracle.matmul(c, a, b, m, n, k, lda, ldb, ldc, ta, tb)
            for i in range(m):
                for j in range(n):
                    s = c0[i * ldc + j]
                    for p in range(k):
                        av = a[p * lda + i] if ta else a[i * lda + p]
                        bv = b[j * ldb + p] if tb else b[p * ldb + j]
                        s += av * bv
                    assert abs(c[i * ldc + j] - s) <= 1e-14 * (1 + abs(s))


def test_gemm_zero_dims_noop():
    c = np.ones(12)
    a = np.ones(12)
    b = np.ones(12)
    oracle.matmul(c, a, b, 3, 4, 0, 1, 4, 4, False, False)
    assert np.all(c == 1.0)


def _masked_case(m, n, k, blksize, ta, rng, density_pattern=W.MASK_30):
    lda = m if ta else k
    a = rng.uniform(-1, 1, size=(k if ta else m) * lda)
    b = rng.uniform(-1, 1, size=k * n)
    nblk = (n + blksize - 1) // blksize
    tab = W.periodic_mask(k, nblk, density_pattern)
    return a, b, tab, lda, nblk


@pytest.mark.parametrize("blksize", [16, 48, 64])
@pytest.mark.parametrize("ta", [False, True])
def test_masked_gemm_equals_zero_filled_dense(blksize, ta):
    """Masked GEMM == dense GEMM with B zeroed under off bits; B under off bits is never read."""
    rng = np.random.default_rng(blksize + ta)
    m, n, k = 270, 500, 530
    a, b, tab, lda, nblk = _masked_case(m, n, k, blksize, ta, rng)
    bz = b.copy()
    W.zero_under_mask(bz, tab, blksize, n)
    c = rng.uniform(-1, 1, size=m * n)
    c0 = c.copy()
    # b (not zeroed) must give the same result as bz: the mask is authoritative
    assert oracle.matmul_non0tab(c, a, b, m, n, k, lda, n, n, ta, False, tab, nblk, blksize) == 0
    ref = c0.reshape(m, n) + op_view(a, m, k, lda, ta) @ bz.reshape(k, n)
    assert rel_fro(c.reshape(m, n), ref) <= 1e-14
    # and the zero-filled dense GEMM through the plain oracle agrees
    c2 = c0.copy()
    oracle.matmul(c2, a, bz, m, n, k, lda, n, n, ta, False)
    assert rel_fro(c, c2) <= 1e-14


def test_masked_gemm_column_block_map():
    """One on-bit per row at a distinct column block pins the p -> j/blksize mapping exactly."""
    m, n, k, blksize = 3, 96, 2, 48
    a = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])  # 3x2
    b = np.ones(k * n)
    tab = np.array([[1, 0], [0, 1]], dtype=np.uint8)  # row p=0 on in block 0, p=1 on in block 1
    c = np.zeros(m * n)
    assert oracle.matmul_non0tab(c, a, b, m, n, k, k, n, n, False, False, tab, 2, blksize) == 0
    C = c.reshape(m, n)
    assert np.all(C[:, :48] == np.array([[1.0], [3.0], [5.0]]))
    assert np.all(C[:, 48:] == np.array([[2.0], [4.0], [6.0]]))


def test_masked_gemm_transb_rejected():
    c = np.zeros(4)
    a = np.zeros(4)
    b = np.zeros(4)
    tab = np.ones(4, dtype=np.uint8)
    assert oracle.matmul_non0tab(c, a, b, 2, 2, 2, 2, 2, 2, False, True, tab, 1, 16) == oracle.EINVAL_TRANSB_MASK


def test_masked_gemm_reference_generators():
    """src/impl_matmul_non0tab.rs:329-377 at reduced size: sin/cos operands, 30 % periodic mask, BLKSIZE 48."""
    m, n, k = 300, 1000, 320
    a, b = W.sin_cos_operands(m * k, k * n)
    nblk = (n + 47) // 48
    tab = W.periodic_mask(k, nblk, W.MASK_30)
    W.zero_under_mask(b, tab, 48, n)
    c = np.zeros(m * n)
    assert oracle.matmul_non0tab(c, a, b, m, n, k, k, n, n, False, False, tab, nblk, 48) == 0
    ref = a.reshape(m, k) @ b.reshape(k, n)
    assert rel_fro(c.reshape(m, n), ref) <= 1e-14


def _rho_case(nao, ngrid, nset, blksize, rng, pattern=W.MASK_25, sym=False):
    nblk = (ngrid + blksize - 1) // blksize
    lhs = rng.standard_normal(nao * ngrid)
    rhs = rng.standard_normal(nao * ngrid)
    dms = rng.standard_normal(nset * nao * nao)
    if sym:
        D = dms.reshape(nset, nao, nao)
        dms = (D + D.transpose(0, 2, 1)).reshape(-1) * 0.5
    tab = W.periodic_mask(nao, nblk, pattern)
    return lhs, rhs, dms, tab, nblk


def _rho_numpy(rhos0, dms, lhs, rhs, nao, ngrid, nset, tab, blksize):
    on = np.repeat(tab.astype(bool), blksize, axis=1)[:, :ngrid]
    L = np.where(on, lhs.reshape(nao, ngrid), 0.0)
    R = np.where(on, rhs.reshape(nao, ngrid), 0.0)
    D = dms.reshape(nset, nao, nao)
    return rhos0.reshape(nset, ngrid) + np.stack([np.einsum("ug,ug->g", L, D[s] @ R) for s in range(nset)])


@pytest.mark.parametrize("blksize", [16, 48, 64])
@pytest.mark.parametrize("nao,ngrid,nset", [(30, 200, 1), (260, 777, 2), (530, 500, 3)])
def test_rho_matches_numpy_einsum(blksize, nao, ngrid, nset):
    rng = np.random.default_rng(nao + ngrid + blksize)
    lhs, rhs, dms, tab, nblk = _rho_case(nao, ngrid, nset, blksize, rng)
    rhos = rng.standard_normal(nset * ngrid)
    rhos0 = rhos.copy()
    # lhs/rhs NOT zeroed under off bits: the mask must still exclude them
    assert oracle.aodmao2rho(rhos, dms, lhs, rhs, nao, ngrid, nao, ngrid, nset, tab, nblk, blksize) == 0
    ref = _rho_numpy(rhos0, dms, lhs, rhs, nao, ngrid, nset, tab, blksize)
    assert rel_fro(rhos.reshape(nset, ngrid), ref) <= 1e-13


def test_rho_strided_and_untouched_padding():
    rng = np.random.default_rng(3)
    nao, ngrid, nset, blksize = 40, 100, 2, 16
    sa, sg = nao + 3, ngrid + 5
    nblk = (ngrid + blksize - 1) // blksize
    lhs = rng.standard_normal(nao * sg)
    dms = rng.standard_normal(nset * sa * sa)
    tab = W.periodic_mask(nao, nblk + 1, W.MASK_30)  # ldtab = nblk + 1
    rhos = rng.standard_normal(nset * sg)
    rhos0 = rhos.copy()
    assert oracle.aodmao2rho(rhos, dms, lhs, lhs, sa, sg, nao, ngrid, nset, tab, nblk + 1, blksize) == 0
    L = lhs.reshape(nao, sg)[:, :ngrid]
    D = dms.reshape(nset, sa, sa)[:, :nao, :nao]
    on = np.repeat(tab[:, :nblk].astype(bool), blksize, axis=1)[:, :ngrid]
    Lm = np.where(on, L, 0.0)
    ref = rhos0.reshape(nset, sg)[:, :ngrid] + np.stack([np.einsum("ug,ug->g", Lm, D[s] @ Lm) for s in range(nset)])
    got = rhos.reshape(nset, sg)
    assert rel_fro(got[:, :ngrid], ref) <= 1e-13
    assert np.array_equal(got[:, ngrid:], rhos0.reshape(nset, sg)[:, ngrid:])


def test_rho_psd_nonnegative_and_linear():
    """D PSD and lhs == rhs => rho >= 0; rho is linear in D (sets add)."""
    rng = np.random.default_rng(9)
    nao, ngrid, blksize = 120, 320, 64
    nblk = (ngrid + blksize - 1) // blksize
    phi = rng.standard_normal(nao * ngrid)
    C = rng.standard_normal((nao, 20))
    D1 = C @ C.T / 20
    D2 = rng.standard_normal((nao, nao))
    tab = W.periodic_mask(nao, nblk, W.MASK_30)
    dms = np.concatenate([D1.ravel(), D2.ravel(), (D1 + D2).ravel()])
    rhos = np.zeros(3 * ngrid)
    assert oracle.aodmao2rho(rhos, dms, phi, phi, nao, ngrid, nao, ngrid, 3, tab, nblk, blksize) == 0
    r = rhos.reshape(3, ngrid)
    assert np.all(r[0] >= -1e-12 * np.max(np.abs(r[0])))
    assert rel_fro(r[2], r[0] + r[1]) <= 1e-13


def test_rho_empty_and_full_masks():
    rng = np.random.default_rng(4)
    nao, ngrid, nset, blksize = 50, 130, 1, 48
    nblk = (ngrid + blksize - 1) // blksize
    lhs = rng.standard_normal(nao * ngrid)
    dms = rng.standard_normal(nao * nao)
    rhos = rng.standard_normal(ngrid)
    rhos0 = rhos.copy()
    assert oracle.aodmao2rho(rhos, dms, lhs, lhs, nao, ngrid, nao, ngrid, nset,
                             np.zeros((nao, nblk), np.uint8), nblk, blksize) == 0
    assert np.array_equal(rhos, rhos0)
    assert oracle.aodmao2rho(rhos, dms, lhs, lhs, nao, ngrid, nao, ngrid, nset,
                             np.ones((nao, nblk), np.uint8), nblk, blksize) == 0
    L = lhs.reshape(nao, ngrid)
    ref = rhos0 + np.einsum("ug,ug->g", L, dms.reshape(nao, nao) @ L)
    assert rel_fro(rhos, ref) <= 1e-13


def test_rho_reference_generators():
    """src/impl_aodmao2rho.rs:264-345 at reduced size: sin/cos phi, D = sin(0.38x+1.2),
    25 % periodic mask, BLKSIZE 48, odd ngrid (edge blocks)."""
    nao, ngrid, nset = 300, 1201, 1
    lhs, rhs = W.sin_cos_operands(nao * ngrid, nao * ngrid)
    dms = np.sin(np.arange(nset * nao * nao, dtype=np.float64) * 0.38 + 1.2)
    nblk = (ngrid + 47) // 48
    tab = W.periodic_mask(nao, nblk, W.MASK_25)
    W.zero_under_mask(lhs, tab, 48, ngrid)
    W.zero_under_mask(rhs, tab, 48, ngrid)
    rhos = np.zeros(nset * ngrid)
    assert oracle.aodmao2rho(rhos, dms, lhs, rhs, nao, ngrid, nao, ngrid, nset, tab, nblk, 48) == 0
    # the reference test's dense check: (dm % ao_rhs) * ao_lhs summed over AOs
    ref = (lhs.reshape(nao, ngrid) * (dms.reshape(nao, nao) @ rhs.reshape(nao, ngrid))).sum(axis=0)
    assert rel_fro(rhos, ref) <= 1e-13


def test_rho_generated_workload_small():
    p = W.rho_problem(seed=1, nshell=60, nao=190, ngrid=640, nset=2, blksize=64, density=0.3)
    assert p.nao == 190 and p.tab.shape == (190, 10)
    assert abs(p.shell_tab.mean() - 0.3) < 0.02
    rhos = np.zeros(p.nset * p.ngrid)
    assert oracle.aodmao2rho(rhos, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(), p.nao, p.ngrid, p.nao, p.ngrid,
                             p.nset, np.ascontiguousarray(p.tab), p.nblk, p.blksize) == 0
    ref = W.rho_reference_dense(p)
    assert rel_fro(rhos.reshape(p.nset, p.ngrid), ref) <= 1e-13
    assert np.all(rhos >= -1e-12 * np.max(np.abs(rhos)))  # PSD D, lhs == rhs


def test_symmetric_form_flop_count():
    """The flop count bench.py prices the symmetric rho form at: brute force over each block's
    64-row chunks (rows of chunk c times the list positions from its own first row on), equal to
    the general count when every block fits one chunk, and within (1/2, 1] of it otherwise."""
    p = W.rho_problem(seed=2, nshell=120, nao=400, ngrid=64 * 9 + 20, nset=2, density=0.4)
    brute = 0.0
    for b in range(p.nblk):
        n = int(p.tab[:, b].sum())
        npts = min(p.blksize, p.ngrid - b * p.blksize)
        for c in range(0, n, 64):
            rows = min(64, n - c)
            brute += npts * (2.0 * rows * (n - c) + 2.0 * rows)
    brute *= p.nset
    assert p.flops_symmetric == pytest.approx(brute, rel=1e-15)
    assert 0.5 * p.flops < p.flops_symmetric < p.flops
    q = W.rho_problem(seed=2, nshell=30, nao=60, ngrid=64 * 4, nset=1, density=0.5)
    assert q.flops_symmetric == pytest.approx(q.flops, rel=1e-15)
