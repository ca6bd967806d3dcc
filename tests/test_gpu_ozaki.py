"""Split-precision (Ozaki, int8 slices) GEMM through the C ABI, against the CPU oracle.

Tolerances (relative Frobenius vs the FP64 oracle, uniform[-1, 1) data):
5e-8 / 5e-10 / 5e-12 for 4 / 5 / 6 slices (measured 1.5e-8 / 1.4e-10 / 1.3e-12).
Element-wise, the method's bound |C - AB| <= (S+2) 2^(-7S) (|A||B|) + k eps (|A||B|)
is checked on data whose rows and columns span ten decades.  Small-integer
inputs are exact (every entry fits in the first slice; int32 sums are exact)."""
import itertools

import numpy as np
import pytest

import cfgmm
import oracle
from helpers import rel_fro

pytestmark = pytest.mark.gpu
TOL = {4: 5e-8, 5: 5e-10, 6: 5e-12}


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


def test_ozaki_small_integers_exact():
    rng = np.random.default_rng(3)
    m, n, k = 200, 150, 700
    a = rng.integers(-100, 101, m * k).astype(np.float64)
    b = rng.integers(-100, 101, k * n).astype(np.float64)
    c = np.zeros(m * n)
    got = _ozaki(a, b, c, m, n, k, k, n, n, False, False, 2)
    ref = a.reshape(m, k).astype(np.int64) @ b.reshape(k, n).astype(np.int64)
    assert np.array_equal(got.reshape(m, n), ref.astype(np.float64))


def test_ozaki_row_scales():
    """Rows and columns spanning 1e-8 .. 1e2: per-row/column exponents keep the error relative."""
    rng = np.random.default_rng(5)
    m, n, k = 257, 190, 640
    a = rng.uniform(-1, 1, (m, k)) * (10.0 ** rng.uniform(-8, 2, (m, 1)))
    b = rng.uniform(-1, 1, (k, n)) * (10.0 ** rng.uniform(-8, 2, (1, n)))
    a, b = a.ravel(), b.ravel()
    c = np.zeros(m * n)
    S = 6
    got = _ozaki(a, b, c, m, n, k, k, n, n, False, False, S).reshape(m, n)
    A, B = a.reshape(m, k), b.reshape(k, n)
    ref = A @ B
    absab = np.abs(A) @ np.abs(B)
    bound = ((S + 2) * 2.0 ** (-7 * S) + k * np.finfo(np.float64).eps) * absab
    assert np.all(np.abs(got - ref) <= bound)


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
