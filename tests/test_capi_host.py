"""Host-side checks of the C ABI that need no GPU: the library loads, exports every
symbol include/cfgmm.h declares, and rejects bad arguments the way the reference
panics (before any device work)."""
import ctypes
import os
import re

import numpy as np
import pytest

import cfgmm

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "cfgmm.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfgmm_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for name in ["cfgmm_matmul_loop_parallel_mnk_pack_a", "cfgmm_matmul_loop_macro_mb", "cfgmm_matmul_anyway_full",
                 "cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab", "cfgmm_matmul_anyway_full_non0tab",
                 "cfgmm_aodmao2rho_loop_parallel", "cfgmm_aodmao2rho_anyway", "cfgmm_dgemm_dev",
                 "cfgmm_dgemm_non0tab_dev", "cfgmm_aodmao2rho_dev"]:
        assert name in syms


def test_library_exports_every_declared_symbol():
    L = cfgmm.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # the binding declares a signature for every header symbol
    assert set(declared_symbols()) == set(cfgmm._SIGNATURES)


def _gemm_args(m=4, n=5, k=6):
    a = np.zeros(m * k)
    b = np.zeros(k * n)
    c = np.zeros(m * n)
    return c, a, b


def test_masked_transb_is_rejected_first():
    c, a, b = _gemm_args()
    tab = np.ones((6, 1), np.uint8)
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.matmul_anyway_full_non0tab(c, a, b, 4, 5, 6, 6, 6, 5, False, True, tab, 1)
    assert e.value.status == cfgmm.EINVAL_TRANSB_MASK
    assert "transb=true is not implemented" in e.value.message


def test_short_slices_are_out_of_bounds():
    c, a, b = _gemm_args()
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.matmul_anyway_full(c, a[:-1], b, 4, 5, 6, 6, 5, 5, False, False)
    assert e.value.status == cfgmm.EBOUNDS
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.matmul_anyway_full(c[:-1], a, b, 4, 5, 6, 6, 5, 5, False, False)
    assert e.value.status == cfgmm.EBOUNDS
    # transposed A: extent (k-1)*lda + m
    a_t = np.zeros(5 * 4 + 4 - 1)
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.matmul_anyway_full(c, a_t, b, 4, 5, 6, 4, 5, 5, True, False)
    assert e.value.status == cfgmm.EBOUNDS


def test_leading_dimension_too_small():
    c, a, b = _gemm_args()
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.matmul_anyway_full(c, a, b, 4, 5, 6, 5, 5, 5, False, False)  # lda < k
    assert e.value.status == cfgmm.EINVAL_DIMS


def test_bad_blksize():
    rhos = np.zeros(10)
    dms = np.zeros(4)
    phi = np.zeros(20)
    tab = np.ones((2, 1), np.uint8)
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.MatmulLoops.aodmao2rho_loop_parallel(40, rhos, dms, phi, phi, [2, 10], [2, 10], 1, tab, 1)
    assert e.value.status == cfgmm.EINVAL


def test_rho_short_mask():
    nao, ngrid = 3, 100
    rhos = np.zeros(ngrid)
    dms = np.zeros(nao * nao)
    phi = np.zeros(nao * ngrid)
    tab = np.ones(nao * 3 - 1, np.uint8)  # needs nao x ceil(100/48)=3
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.aodmao2rho_anyway(rhos, dms, phi, phi, [nao, ngrid], [nao, ngrid], 1, tab, 3)
    assert e.value.status == cfgmm.EBOUNDS


def test_empty_problem_is_a_noop_without_device():
    c = np.ones(20)
    a = np.zeros(0)
    b = np.zeros(0)
    cfgmm.matmul_anyway_full(c, a, b, 4, 5, 0, 1, 5, 5, False, False)  # k = 0: C untouched, no device needed
    assert np.all(c == 1.0)


def test_last_error_is_thread_local_and_cleared():
    c, a, b = _gemm_args()
    with pytest.raises(cfgmm.CfgmmError):
        cfgmm.matmul_anyway_full(c, a[:-1], b, 4, 5, 6, 6, 5, 5, False, False)
    assert "index out of bounds" in cfgmm.last_error()
    cfgmm.matmul_anyway_full(c, a, b, 4, 5, 0, 6, 5, 5, False, False)
    assert cfgmm.last_error() == ""
