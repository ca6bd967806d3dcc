"""CPU oracle for the two hot paths -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers over ``oracle/cfgmm_oracle.c`` (a C restatement of the
reference crate's CPU loop nest; see that file's header for the function map).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this package; the product library never calls it.

Every function here is pinned in ``tests/test_oracle.py`` against numpy
float64 (a library routine), exact integer-valued closed forms, brute-force
loops on tiny shapes, mask invariants, and the reference tests' own generators.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_lib = None

OK, EINVAL, EINVAL_TRANSB_MASK, EOOM = 0, 1, 2, 4


class OracleError(RuntimeError):
    pass


def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in (line + " ")
    except OSError:
        pass
    return False


def build() -> None:
    """Compile the oracle with the committed Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    name = "libcfgmm_oracle_v4.so" if _cpu_has_avx512() else "libcfgmm_oracle_v3.so"
    path = os.path.join(_BUILD, name)
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    sz, dp, cdp, u8p, i = (ctypes.c_size_t, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                           ctypes.POINTER(ctypes.c_uint8), ctypes.c_int)
    L.oracle_matmul_loop_parallel_mnk_pack_a.argtypes = [dp, cdp, cdp, sz, sz, sz, sz, sz, sz, i, i, i]
    L.oracle_matmul_loop_parallel_mnk_pack_a_non0tab.argtypes = [dp, cdp, cdp, sz, sz, sz, sz, sz, sz, i, i, u8p,
                                                                 sz, sz, i]
    L.oracle_aodmao2rho_loop_parallel.argtypes = [dp, cdp, cdp, cdp, sz, sz, sz, sz, sz, u8p, sz, sz, i]
    L.oracle_default_threads.argtypes = []
    for f in (L.oracle_matmul_loop_parallel_mnk_pack_a, L.oracle_matmul_loop_parallel_mnk_pack_a_non0tab,
              L.oracle_aodmao2rho_loop_parallel, L.oracle_default_threads):
        f.restype = ctypes.c_int
    _lib = L
    return L


def default_threads() -> int:
    return lib().oracle_default_threads()


def _dptr(x: np.ndarray):
    assert x.dtype == np.float64 and x.flags.c_contiguous
    return x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u8ptr(x: np.ndarray):
    assert x.dtype == np.uint8 and x.flags.c_contiguous
    return x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _need(name: str, arr: np.ndarray, rows: int, cols: int, ld: int) -> None:
    """The C restatement indexes like the reference; refuse buffers shorter than that extent."""
    need = 0 if rows == 0 or cols == 0 else (rows - 1) * ld + cols
    if arr.size < need:
        raise ValueError(f"oracle: {name} has {arr.size} elements, the arguments index {need}")


def _check(st: int, what: str) -> None:
    if st != OK:
        raise OracleError(f"{what} failed with status {st}")


def matmul(c, a, b, m, n, k, lda, ldb, ldc, transa, transb, nthreads=0):
    """C += op(A) op(B) in place on flat float64 arrays (src/impl_matmul.rs:155-231)."""
    if m and n and k:
        _need("a", a, k if transa else m, m if transa else k, lda)
        _need("b", b, n if transb else k, k if transb else n, ldb)
    _need("c", c, m, n, ldc)
    _check(lib().oracle_matmul_loop_parallel_mnk_pack_a(_dptr(c), _dptr(a), _dptr(b), m, n, k, lda, ldb, ldc,
                                                        int(transa), int(transb), nthreads), "oracle matmul")
    return c


def matmul_non0tab(c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab, ldtab, blksize=48, nthreads=0):
    """Masked C += op(A)(B masked by tab) (src/impl_matmul_non0tab.rs:210-305); returns the status."""
    if m and n and k and not transb and blksize:
        _need("a", a, k if transa else m, m if transa else k, lda)
        _need("b", b, k, n, ldb)
        _need("tab", tab.reshape(-1), k, -(-n // blksize), ldtab)
    _need("c", c, m, n, ldc)
    return lib().oracle_matmul_loop_parallel_mnk_pack_a_non0tab(_dptr(c), _dptr(a), _dptr(b), m, n, k, lda, ldb, ldc,
                                                                int(transa), int(transb), _u8ptr(tab), ldtab,
                                                                blksize, nthreads)


def aodmao2rho(rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize=48,
               nthreads=0):
    """rho += sum phi_lhs D phi_rhs under the mask (src/impl_aodmao2rho.rs:171-244); returns the status.

    rho rows are strided by stride_grid, like the phi rows (src/impl_aodmao2rho.rs:80,88)."""
    if nao and ngrid and nset and blksize:
        _need("rhos", rhos, nset, ngrid, stride_grid)
        _need("ao_lhs", ao_lhs, nao, ngrid, stride_grid)
        _need("ao_rhs", ao_rhs, nao, ngrid, stride_grid)
        _need("dms", dms, (nset - 1) * stride_ao + nao, nao, stride_ao)
        _need("tab", tab.reshape(-1), nao, -(-ngrid // blksize), ldtab)
    return lib().oracle_aodmao2rho_loop_parallel(_dptr(rhos), _dptr(dms), _dptr(ao_lhs), _dptr(ao_rhs), stride_ao,
                                                 stride_grid, nao, ngrid, nset, _u8ptr(tab), ldtab, blksize, nthreads)
