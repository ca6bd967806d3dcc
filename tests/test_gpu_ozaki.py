"""Split-precision (Ozaki, balanced int8 digits of radix 254) GEMM through the C ABI, against
the CPU oracle.

Tolerances (relative Frobenius vs the FP64 oracle, uniform[-1, 1) data): 2e-9 / 1e-11 /
1e-13 for 4 / 5 / 6 slices.  An exact-integer emulation of the digit scheme in numpy gives
5.3e-10 / 2.3e-12 / 9.9e-15 at 301x263x1100 and at k = 6581 alike.

Element-wise, every scaled product term carries at most (S + 1.01) 254^-S of digit residual
and dropped levels, so |C - AB|_ij <= k 2^(ea_i + eb_j) ((S + 1.01) 254^-S + (S + 2) eps)
+ k eps (|A||B|)_ij; this is checked on data whose rows and columns span ten decades.
Operands that are integers over 127 with a row / column maximum in [1/2, 1) are their own
first digit, so two slices multiply them exactly in int32."""
import itertools

import numpy as np
import pytest

import cfgmm
import oracle
from helpers import rel_fro

pytestmark = pytest.mark.gpu
TOL = {4: 2e-9, 5: 1e-11, 6: 1e-13}


def _ozaki(a, b, c, m, n, k, lda, ldb, ldc, ta, tb, ns):
    s = cfgmm.Stream()
    dA, dB, dC = (cfgmm.DeviceBuffer.from_array(x) for x in (a, b, c))
    cfgmm.dgemm_dev(s, dC, dA, dB, m, n, k, lda, ldb, ldc, ta, tb, precision=cfgmm.PREC_OZAKI_INT8, nslices=ns)
    s.synchronize()
    out = dC.to_array(np.float64, c.shape)
    s.destroy()
    return out


@pytest.mark.parametrize("ns", [4, 5, 6])
@pytest.mark.parametrize("ta,tb", list(itertools.product([False, True], repeat=2)))
def test_ozaki_vs_oracle(ns, ta, tb):
    rng = np.random.default_rng(ns * 10 + 2 * ta + tb)
    m, n, k = 301, 263, 1100
    lda, ldb, ldc = (m if ta else k) + 1, (k if tb else n) + 2, n + 3
    a = rng.uniform(-1, 1, (k if ta else m) * lda)
    b = rng.uniform(-1, 1, (n if tb else k) * ldb)
    c = rng.uniform(-1, 1, m * ldc)
    ref = c.copy()
    oracle.matmul(ref, a, b, m, n, k, lda, ldb, ldc, ta, tb)
    got = _ozaki(a, b, c, m, n, k, lda, ldb, ldc, ta, tb, ns)
    G, R = got.reshape(m, ldc), ref.reshape(m, ldc)
    err = rel_fro(G[:, :n] - c.reshape(m, ldc)[:, :n], R[:, :n] - c.reshape(m, ldc)[:, :n])
    print(f"ozaki ns={ns} trans={ta}{tb} rel_fro(AB) = {err:.3e}")
    assert err <= TOL[ns]
    assert np.array_equal(G[:, n:], c.reshape(m, ldc)[:, n:])  # padding untouched


def test_ozaki_first_digit_exact():
    """a = qa / 127 and b = qb / 127 with integer |q| <= 126 and a maximum of at least 64 in every
    row / column: the first balanced digit is q itself and the second is zero, so the int32 sums
    are the exact integer products and 127^2 C rounds to qa @ qb."""
    rng = np.random.default_rng(3)
    m, n, k = 200, 150, 700
    qa = rng.integers(-126, 127, (m, k))
    qb = rng.integers(-126, 127, (k, n))
    qa[:, 0] = 100
    qb[0, :] = -90
    a, b = (qa / 127.0).ravel(), (qb / 127.0).ravel()
    got = _ozaki(a, b, np.zeros(m * n), m, n, k, k, n, n, False, False, 2).reshape(m, n) * 127.0 ** 2
    ref = qa.astype(np.int64) @ qb.astype(np.int64)
    assert np.array_equal(np.rint(got).astype(np.int64), ref)
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_ozaki_row_scales():
    """Rows and columns spanning 1e-8 .. 1e2: per-row/column exponents keep the error relative."""
    rng = np.random.default_rng(5)
    m, n, k = 257, 190, 640
    a = rng.uniform(-1, 1, (m, k)) * (10.0 ** rng.uniform(-8, 2, (m, 1)))
    b = rng.uniform(-1, 1, (k, n)) * (10.0 ** rng.uniform(-8, 2, (1, n)))
    a, b = a.ravel(), b.ravel()
    c = np.zeros(m * n)
    eps = np.finfo(np.float64).eps
    A, B = a.reshape(m, k), b.reshape(k, n)
    ref = A @ B
    absab = np.abs(A) @ np.abs(B)
    ea = np.frexp(np.max(np.abs(A), axis=1))[1].astype(np.float64)
    eb = np.frexp(np.max(np.abs(B), axis=0))[1].astype(np.float64)
    scale = np.exp2(ea[:, None] + eb[None, :])
    for S in (4, 5, 6):
        got = _ozaki(a, b, c, m, n, k, k, n, n, False, False, S).reshape(m, n)
        bound = k * scale * ((S + 1.01) * 254.0 ** (-S) + (S + 2) * eps) + 2 * k * eps * absab
        assert np.all(np.abs(got - ref) <= bound), S


def test_ozaki_nonfinite_rows_give_nan():
    m, n, k = 64, 64, 64
    a = np.ones(m * k)
    a[5 * k + 3] = np.inf
    b = np.ones(k * n)
    got = _ozaki(a, b, np.zeros(m * n), m, n, k, k, n, n, False, False, 4).reshape(m, n)
    assert np.all(np.isnan(got[5])) and np.all(np.isfinite(np.delete(got, 5, axis=0)))


def test_ozaki_rejects_bad_slice_count():
    s = cfgmm.Stream()
    buf = cfgmm.DeviceBuffer(8 * 16)
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.dgemm_dev(s, buf, buf, buf, 4, 4, 4, 4, 4, 4, False, False, precision=cfgmm.PREC_OZAKI_INT8, nslices=7)
    assert e.value.status == cfgmm.EINVAL


@pytest.mark.parametrize("ns", [5, 6])
def test_ozaki_large_nn_full_size(ns):
    """C(3527x9583) += A(3527x6581) B(6581x9583), uniform[-1, 1) seed 0 (the README's benchmark
    shape, README.md:11-17): k = 6581 puts every level's int32 sum near its largest, and the
    128x64 tile grid has ragged edges in m and n."""
    from cfgmm import workloads as W
    g = W.dgemm_large_nn(seed=0)
    ref = np.zeros(g.m * g.ldc)
    oracle.matmul(ref, g.a, g.b, g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.transa, g.transb)
    got = _ozaki(g.a, g.b, np.zeros(g.m * g.ldc), g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.transa, g.transb, ns)
    err = rel_fro(got, ref)
    print(f"ozaki 3527x9583x6581 NN ns={ns} rel_fro = {err:.3e}", flush=True)
    assert err <= min(TOL[ns], 1e-10)
