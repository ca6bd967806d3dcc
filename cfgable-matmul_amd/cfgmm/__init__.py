"""Python binding of libcfgmm (the C ABI in include/cfgmm.h) for tests, bench and serving.

The reference crate's user-facing surface is a Rust associated-function API on
``MatmulLoops<T, MC, KC, NC, MR, NR_LANE, LANE>`` plus three f64 wrappers
(src/structs.rs:15-17, src/impl_matmul.rs:155-248,
src/impl_matmul_non0tab.rs:210-326, src/impl_aodmao2rho.rs:171-261).  This
module mirrors it with the same names and argument meaning on flat numpy
float64 buffers: outputs are accumulated into in place, and where the reference
panics the binding raises :class:`CfgmmError` carrying the C status code.

There is no CPU fallback: if ``lib/libcfgmm.so`` is missing or the device is
not gfx950, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libcfgmm.so")

OK = 0
EINVAL = 1
EINVAL_TRANSB_MASK = 2
EINVAL_DIMS = 3
EOOM = 4
EHIP = 5
ENODEV = 6
EBOUNDS = 7

PREC_FP64 = 0
PREC_OZAKI_INT8 = 1

TUNE_BATCH_WGS_PER_CU = 0
TUNE_BATCH_MAX_SETS = 1
TUNE_DGEMM_GROUP_M = 2
TUNE_RHO_MAX_SETS = 3
TUNE_DGEMM_SMALL_TILE = 4
TUNE_DGEMM_SMALL_SPLITS = 5
TUNE_RESIDENT_DEPTH = 6
TUNE_RESIDENT_COLS = 7
TUNE_RHO_ORDER = 8
TUNE_RHO_SYMMETRIC = 9
TUNE_DGEMM_KGROUPS = 10
TUNE_MASKED_LDS = 11
TUNE_RESIDENT_PERSISTENT = 12

_lib = None

_SZ = ctypes.c_size_t
_I = ctypes.c_int
_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_U8P = ctypes.POINTER(ctypes.c_uint8)

# Signatures of every symbol declared in include/cfgmm.h.
_SIGNATURES = {
    "cfgmm_last_error": ([], ctypes.c_char_p),
    "cfgmm_version": ([], _I),
    "cfgmm_matmul_loop_parallel_mnk_pack_a": ([_VP, _SZ, _VP, _SZ, _VP, _SZ] + [_SZ] * 6 + [_I, _I], _I),
    "cfgmm_matmul_loop_macro_mb": ([_VP, _SZ, _VP, _SZ, _VP, _SZ] + [_SZ] * 6 + [_I, _I], _I),
    "cfgmm_matmul_anyway_full": ([_VP, _SZ, _VP, _SZ, _VP, _SZ] + [_SZ] * 6 + [_I, _I], _I),
    "cfgmm_dgemm_dev": ([_VP, _VP, _VP, _VP] + [_SZ] * 6 + [_I, _I], _I),
    "cfgmm_dgemm_ex_dev": ([_VP, _VP, _VP, _VP] + [_SZ] * 6 + [_I, _I, _I, _I], _I),
    "cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab": (
        [_VP, _SZ, _VP, _SZ, _VP, _SZ] + [_SZ] * 6 + [_I, _I, _VP, _SZ, _SZ, _SZ], _I),
    "cfgmm_matmul_anyway_full_non0tab": ([_VP, _SZ, _VP, _SZ, _VP, _SZ] + [_SZ] * 6 + [_I, _I, _VP, _SZ, _SZ], _I),
    "cfgmm_dgemm_non0tab_dev": ([_VP, _VP, _VP, _VP] + [_SZ] * 6 + [_I, _I, _VP, _SZ, _SZ], _I),
    "cfgmm_dgemm_non0tab_ex_dev": ([_VP, _VP, _VP, _VP] + [_SZ] * 6 + [_I, _I, _VP, _SZ, _SZ], _I),
    "cfgmm_sgemm_non0tab_dev": ([_VP, _VP, _VP, _VP] + [_SZ] * 6 + [_I, _I, _VP, _SZ, _SZ], _I),
    "cfgmm_aodmao2rho_loop_parallel": ([_VP, _SZ, _VP, _SZ, _VP, _SZ, _VP, _SZ] + [_SZ] * 5 + [_VP, _SZ, _SZ, _SZ],
                                       _I),
    "cfgmm_aodmao2rho_anyway": ([_VP, _SZ, _VP, _SZ, _VP, _SZ, _VP, _SZ] + [_SZ] * 5 + [_VP, _SZ, _SZ], _I),
    "cfgmm_aodmao2rho_dev": ([_VP, _VP, _VP, _VP, _VP] + [_SZ] * 5 + [_VP, _SZ, _SZ], _I),
    "cfgmm_aodmao2rho_dev_noalloc": ([_VP, _VP, _VP, _VP, _VP] + [_SZ] * 5 + [_VP, _SZ, _SZ], _I),
    "cfgmm_aodmao2rho_ex_dev": ([_VP, _VP, _VP, _VP, _VP] + [_SZ] * 5 + [_VP, _SZ, _SZ, _I, _I], _I),
    "cfgmm_aodmao2rho_multi_dev": ([_VP, _VP, _SZ, _VP, _VP, _SZ, _SZ, _VP] + [_SZ] * 5 + [_VP, _SZ, _SZ], _I),
    "cfgmm_aogridao2mat_dev": ([_VP, _VP, _SZ, _VP, _VP, _VP] + [_SZ] * 3 + [_VP, _SZ, _SZ], _I),
    "cfgmm_eval_gto_dev": ([_VP, _VP, _SZ, _VP, _SZ, _VP, _SZ, _VP, _VP, _SZ, _VP, _SZ, _SZ, _SZ,
                            ctypes.c_double], _I),
    "cfgmm_aodmao2rho_batched_dev": ([_VP, _VP, _SZ, _I, _I], _I),
    "cfgmm_rho_resident_create": ([_VP, _VP, _VP, _VP] + [_SZ] * 5 + [_VP, _SZ, _SZ, _I, _I, ctypes.POINTER(_VP)], _I),
    "cfgmm_rho_resident_eval": ([_VP, _VP, _VP, _SZ, _SZ, _SZ], _I),
    "cfgmm_rho_resident_eval_batched": ([_VP, _VP, _SZ], _I),
    "cfgmm_rho_resident_bytes": ([_VP, ctypes.POINTER(_SZ)], _I),
    "cfgmm_rho_resident_destroy": ([_VP], _I),
    "cfgmm_server_create": ([_VP, ctypes.POINTER(_VP)], _I),
    "cfgmm_server_add_molecule": ([_VP, _I, _VP, _VP, _VP, _SZ, _SZ, _SZ, _SZ], _I),
    "cfgmm_server_warm": ([_VP, _VP, _SZ], _I),
    "cfgmm_server_run": ([_VP, _VP, _SZ, _VP], _I),
    "cfgmm_server_get_stats": ([_VP, _VP], _I),
    "cfgmm_server_destroy": ([_VP], _I),
    "cfgmm_tune_set": ([_I, _I], _I),
    "cfgmm_tune_get": ([_I, ctypes.POINTER(_I)], _I),
    "cfgmm_device_count": ([ctypes.POINTER(_I)], _I),
    "cfgmm_device_malloc": ([ctypes.POINTER(_VP), _SZ], _I),
    "cfgmm_device_free": ([_VP], _I),
    "cfgmm_memcpy_htod": ([_VP, _VP, _SZ], _I),
    "cfgmm_memcpy_dtoh": ([_VP, _VP, _SZ], _I),
    "cfgmm_memcpy_htod_async": ([_VP, _VP, _SZ, _VP], _I),
    "cfgmm_memcpy_dtoh_async": ([_VP, _VP, _SZ, _VP], _I),
    "cfgmm_memcpy_dtod_async": ([_VP, _VP, _SZ, _VP], _I),
    "cfgmm_memset_dev": ([_VP, _I, _SZ], _I),
    "cfgmm_mem_get_info": ([ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], _I),
    "cfgmm_stream_create": ([ctypes.POINTER(_VP)], _I),
    "cfgmm_stream_destroy": ([_VP], _I),
    "cfgmm_stream_synchronize": ([_VP], _I),
    "cfgmm_device_synchronize": ([], _I),
    "cfgmm_event_create": ([ctypes.POINTER(_VP)], _I),
    "cfgmm_event_destroy": ([_VP], _I),
    "cfgmm_event_record": ([_VP, _VP], _I),
    "cfgmm_event_synchronize": ([_VP], _I),
    "cfgmm_event_query": ([_VP], _I),
    "cfgmm_event_elapsed_ms": ([ctypes.POINTER(ctypes.c_float), _VP, _VP], _I),
    "cfgmm_trace_enable": ([_I], _I),
    "cfgmm_trace_reset": ([], _I),
    "cfgmm_trace_read": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)], _I),
}


class CfgmmError(RuntimeError):
    """A libcfgmm call returned a non-zero status (where the reference would panic)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"cfgmm status {status}: {message}")
        self.status = status
        self.message = message


def lib() -> ctypes.CDLL:
    """Load lib/libcfgmm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `make -C {PKG_ROOT}` "
                                "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, (argtypes, restype) in _SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = argtypes
        f.restype = restype
    _lib = L
    return L


def last_error() -> str:
    return lib().cfgmm_last_error().decode()


def check(status: int) -> None:
    if status != OK:
        raise CfgmmError(status, last_error())


def _ptr(x: np.ndarray | None):
    if x is None:
        return None
    if not x.flags.c_contiguous:
        raise ValueError("arrays passed to libcfgmm must be C-contiguous")
    return x.ctypes.data


def _f64(x: np.ndarray) -> np.ndarray:
    if x.dtype != np.float64:
        raise TypeError(f"expected float64, got {x.dtype}")
    return x


def _u8(tab: np.ndarray) -> np.ndarray:
    if tab.dtype == np.bool_:
        tab = tab.view(np.uint8)
    if tab.dtype != np.uint8:
        raise TypeError(f"mask must be bool or uint8, got {tab.dtype}")
    return tab


# ---------------------------------------------------------------------------
# reference-shaped host API (host pointers, synchronous)
# ---------------------------------------------------------------------------
class MatmulLoops:
    """Mirror of ``MatmulLoops<f64, MC, KC, NC, MR, NR_LANE, LANE>`` (src/structs.rs:15-17).

    The CPU blocking const generics carry no meaning on the GPU; they are
    accepted and ignored, so ``MatmulLoops(252, 512, 240, 14, 2, 8)`` and
    ``MatmulLoops()`` behave identically.
    """

    def __init__(self, *const_generics: int):
        self.const_generics = const_generics

    @staticmethod
    def matmul_loop_parallel_mnk_pack_a(c, a, b, m, n, k, lda, ldb, ldc, transa, transb):
        """src/impl_matmul.rs:155-231."""
        c, a, b = _f64(c), _f64(a), _f64(b)
        check(lib().cfgmm_matmul_loop_parallel_mnk_pack_a(_ptr(c), c.size, _ptr(a), a.size, _ptr(b), b.size, m, n, k,
                                                          lda, ldb, ldc, int(transa), int(transb)))

    @staticmethod
    def matmul_loop_macro_mb(c, a, b, m, n, k, lda, ldb, ldc, transa, transb):
        """README.md:22-31 name of the same function."""
        c, a, b = _f64(c), _f64(a), _f64(b)
        check(lib().cfgmm_matmul_loop_macro_mb(_ptr(c), c.size, _ptr(a), a.size, _ptr(b), b.size, m, n, k, lda, ldb,
                                               ldc, int(transa), int(transb)))

    @staticmethod
    def matmul_loop_parallel_mnk_pack_a_non0tab(blksize, c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab, ldtab):
        """src/impl_matmul_non0tab.rs:210-305 (BLKSIZE passed first, as the turbofish)."""
        c, a, b, tab = _f64(c), _f64(a), _f64(b), _u8(tab)
        check(lib().cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab(_ptr(c), c.size, _ptr(a), a.size, _ptr(b), b.size,
                                                                  m, n, k, lda, ldb, ldc, int(transa), int(transb),
                                                                  _ptr(tab), tab.size, ldtab, blksize))

    @staticmethod
    def aodmao2rho_loop_parallel(blksize, rhos, dms, ao_lhs, ao_rhs, strides, shape, nset, tab, ldtab):
        """src/impl_aodmao2rho.rs:171-244; strides = [stride_ao, stride_grid], shape = [nao, ngrid]."""
        rhos, dms, ao_lhs, ao_rhs, tab = _f64(rhos), _f64(dms), _f64(ao_lhs), _f64(ao_rhs), _u8(tab)
        stride_ao, stride_grid = strides
        nao, ngrid = shape
        check(lib().cfgmm_aodmao2rho_loop_parallel(_ptr(rhos), rhos.size, _ptr(dms), dms.size, _ptr(ao_lhs),
                                                   ao_lhs.size, _ptr(ao_rhs), ao_rhs.size, stride_ao, stride_grid,
                                                   nao, ngrid, nset, _ptr(tab), tab.size, ldtab, blksize))


def matmul_anyway_full(c, a, b, m, n, k, lda, ldb, ldc, transa, transb):
    """src/impl_matmul.rs:234-248."""
    c, a, b = _f64(c), _f64(a), _f64(b)
    check(lib().cfgmm_matmul_anyway_full(_ptr(c), c.size, _ptr(a), a.size, _ptr(b), b.size, m, n, k, lda, ldb, ldc,
                                         int(transa), int(transb)))


def matmul_anyway_full_non0tab(c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab, ldtab):
    """src/impl_matmul_non0tab.rs:308-326 (BLKSIZE = 48)."""
    c, a, b, tab = _f64(c), _f64(a), _f64(b), _u8(tab)
    check(lib().cfgmm_matmul_anyway_full_non0tab(_ptr(c), c.size, _ptr(a), a.size, _ptr(b), b.size, m, n, k, lda,
                                                 ldb, ldc, int(transa), int(transb), _ptr(tab), tab.size, ldtab))


def aodmao2rho_anyway(rhos, dms, ao_lhs, ao_rhs, strides, shape, nset, tab, ldtab):
    """src/impl_aodmao2rho.rs:247-261 (BLKSIZE = 48)."""
    rhos, dms, ao_lhs, ao_rhs, tab = _f64(rhos), _f64(dms), _f64(ao_lhs), _f64(ao_rhs), _u8(tab)
    check(lib().cfgmm_aodmao2rho_anyway(_ptr(rhos), rhos.size, _ptr(dms), dms.size, _ptr(ao_lhs), ao_lhs.size,
                                        _ptr(ao_rhs), ao_rhs.size, strides[0], strides[1], shape[0], shape[1], nset,
                                        _ptr(tab), tab.size, ldtab))


# ---------------------------------------------------------------------------
# device plumbing
# ---------------------------------------------------------------------------
def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().cfgmm_device_count(ctypes.byref(n)))
    return n.value


def mem_get_info() -> tuple[int, int]:
    f, t = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(lib().cfgmm_mem_get_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


class DeviceBuffer:
    """An HBM allocation owned by Python (freed on .free() or garbage collection)."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        check(lib().cfgmm_device_malloc(ctypes.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, x: np.ndarray, stream=None) -> "DeviceBuffer":
        x = np.ascontiguousarray(x)
        buf = cls(x.nbytes)
        buf.upload(x, stream)
        return buf

    def upload(self, x: np.ndarray, stream=None) -> None:
        x = np.ascontiguousarray(x)
        assert x.nbytes <= self.nbytes
        if stream is None:
            check(lib().cfgmm_memcpy_htod(self.ptr, x.ctypes.data, x.nbytes))
        else:
            check(lib().cfgmm_memcpy_htod_async(self.ptr, x.ctypes.data, x.nbytes, stream.handle))

    def download(self, out: np.ndarray, stream=None) -> np.ndarray:
        assert out.flags.c_contiguous and out.nbytes <= self.nbytes
        if stream is None:
            check(lib().cfgmm_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        else:
            check(lib().cfgmm_memcpy_dtoh_async(out.ctypes.data, self.ptr, out.nbytes, stream.handle))
        return out

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        return self.download(out)

    def zero(self) -> None:
        check(lib().cfgmm_memset_dev(self.ptr, 0, self.nbytes))

    def free(self) -> None:
        if self.ptr is not None and _lib is not None:
            lib().cfgmm_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().cfgmm_stream_create(ctypes.byref(p)))
        self.handle = p.value

    def synchronize(self) -> None:
        check(lib().cfgmm_stream_synchronize(self.handle))

    def destroy(self) -> None:
        if self.handle is not None:
            lib().cfgmm_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().cfgmm_event_create(ctypes.byref(p)))
        self.handle = p.value

    def record(self, stream: Stream | None) -> None:
        check(lib().cfgmm_event_record(self.handle, stream.handle if stream else None))

    def synchronize(self) -> None:
        check(lib().cfgmm_event_synchronize(self.handle))

    def done(self) -> bool:
        st = lib().cfgmm_event_query(self.handle)
        if st == 1:
            return False
        check(st)
        return True

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        check(lib().cfgmm_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle))
        return ms.value

    def destroy(self) -> None:
        if self.handle is not None:
            lib().cfgmm_event_destroy(self.handle)
            self.handle = None


def trace_enable(on: bool = True) -> None:
    check(lib().cfgmm_trace_enable(int(on)))


def trace_reset() -> None:
    check(lib().cfgmm_trace_reset())


def trace_read(name: str) -> tuple[float, int]:
    """(summed milliseconds, launch count) of kernel `name` since the last reset."""
    ms, n = ctypes.c_double(0.0), ctypes.c_longlong(0)
    check(lib().cfgmm_trace_read(name.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def aodmao2rho_multi_dev(stream, rhos, rho_comp_stride, dms, ao_lhs, lhs_comp_stride, ncomp, ao_rhs, stride_ao,
                         stride_grid, nao, ngrid, nset, tab, ldtab, blksize):
    """ncomp left-hand components (ao_lhs + c*lhs_comp_stride) sharing D, rhs and the mask."""
    check(lib().cfgmm_aodmao2rho_multi_dev(_h(stream), _dev(rhos), rho_comp_stride, _dev(dms), _dev(ao_lhs),
                                           lhs_comp_stride, ncomp, _dev(ao_rhs), stride_ao, stride_grid, nao, ngrid,
                                           nset, _dev(tab), ldtab, blksize))


def aogridao2mat_dev(stream, V, ldv, ao_lhs, ao_rhs, w, stride_grid, nao, ngrid, tab, ldtab, blksize):
    """V += sum over grid points of lhs[u][g] w[g] rhs[v][g] for (u, v) active in g's block (w may be None)."""
    check(lib().cfgmm_aogridao2mat_dev(_h(stream), _dev(V), ldv, _dev(ao_lhs), _dev(ao_rhs),
                                       None if w is None else _dev(w), stride_grid, nao, ngrid, _dev(tab), ldtab,
                                       blksize))


class GtoShell(ctypes.Structure):
    """cfgmm_gto_shell: a Cartesian GTO shell (centre, l <= 4, primitive range, first AO)."""
    _fields_ = [("center", ctypes.c_double * 3), ("l", _I), ("nprim", _I), ("prim_off", _I), ("ao_off", _I)]


def gto_shells(centers, ls, nprims, prim_offs, ao_offs):
    """A ctypes array of GtoShell from per-shell sequences."""
    arr = (GtoShell * len(ls))()
    for i, (c, l, n, po, ao) in enumerate(zip(centers, ls, nprims, prim_offs, ao_offs)):
        arr[i].center[:] = [float(x) for x in c]
        arr[i].l, arr[i].nprim, arr[i].prim_off, arr[i].ao_off = int(l), int(n), int(po), int(ao)
    return arr


def eval_gto_dev(stream, phi, stride_grid, tab, ldtab, shells, exps, coefs, coords, ngrid, nao, blksize, cutoff):
    """Tabulates the Cartesian GTO shells (a GtoShell array; exps/coefs: host float64 arrays) at the
    device grid coords (ngrid x 3) into phi (nao x stride_grid) and the per-AO block mask tab."""
    exps = np.ascontiguousarray(exps, dtype=np.float64)
    coefs = np.ascontiguousarray(coefs, dtype=np.float64)
    if exps.shape != coefs.shape:
        raise ValueError("exps and coefs differ in length")
    check(lib().cfgmm_eval_gto_dev(_h(stream), _dev(phi), stride_grid, _dev(tab), ldtab, ctypes.addressof(shells), len(shells),
                                   exps.ctypes.data, coefs.ctypes.data, exps.size, _dev(coords), ngrid, nao,
                                   blksize, float(cutoff)))


class RhoProblem(ctypes.Structure):
    """cfgmm_rho_problem: one problem of a batched rho call (device pointers)."""
    _fields_ = [("rhos", _VP), ("dms", _VP), ("ao_lhs", _VP), ("ao_rhs", _VP), ("stride_ao", _SZ),
                ("stride_grid", _SZ), ("nao", _SZ), ("ngrid", _SZ), ("nset", _SZ), ("tab", _VP), ("ldtab", _SZ),
                ("blksize", _SZ)]


def aodmao2rho_batched_dev(stream, problems, precision=PREC_FP64, nslices=0):
    """problems: a sequence of RhoProblem (or a ctypes array of them)."""
    arr = (RhoProblem * len(problems))(*problems) if not isinstance(problems, ctypes.Array) else problems
    check(lib().cfgmm_aodmao2rho_batched_dev(_h(stream), ctypes.cast(arr, _VP), len(arr), precision, nslices))


class RhoResident:
    """Resident operands of one masked rho problem (cfgmm_rho_resident_*), split precision.

    Built once from device buffers (dms, phi, mask); every ``eval`` adds
    rho_s(g) for the leading ``ngrid`` points and ``nset`` matrices into ``rhos``.
    """

    def __init__(self, stream, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize,
                 precision=PREC_OZAKI_INT8, nslices=6):
        h = _VP()
        check(lib().cfgmm_rho_resident_create(_h(stream), _dev(dms), _dev(ao_lhs), _dev(ao_rhs), stride_ao,
                                              stride_grid, nao, ngrid, nset, _dev(tab), ldtab, blksize, precision,
                                              nslices, ctypes.byref(h)))
        self.handle = h.value
        self.ngrid, self.nset, self.blksize = ngrid, nset, blksize

    @property
    def bytes(self) -> int:
        v = _SZ(0)
        check(lib().cfgmm_rho_resident_bytes(self.handle, ctypes.byref(v)))
        return v.value

    def eval(self, stream, rhos, stride_rho, ngrid=None, nset=None) -> None:
        check(lib().cfgmm_rho_resident_eval(_h(stream), self.handle, _dev(rhos), stride_rho,
                                            self.ngrid if ngrid is None else ngrid,
                                            self.nset if nset is None else nset))

    def destroy(self) -> None:
        if self.handle:
            check(lib().cfgmm_rho_resident_destroy(self.handle))
            self.handle = None


REQ_RHO = 0
REQ_GEMM = 1


class ServerConfig(ctypes.Structure):
    """cfgmm_server_config."""
    _fields_ = [("hbm_budget_bytes", _SZ), ("slots", _I), ("rho_precision", _I), ("rho_nslices", _I),
                ("max_batch", _I), ("split_tile_cap_bytes", _SZ)]


class Request(ctypes.Structure):
    """cfgmm_request."""
    _fields_ = [("kind", _I), ("molecule", _I), ("ngrid", _SZ), ("nset", _SZ), ("m", _SZ), ("n", _SZ), ("k", _SZ),
                ("seed", ctypes.c_ulonglong), ("arrival_s", ctypes.c_double), ("out", _VP)]


class RequestResult(ctypes.Structure):
    """cfgmm_request_result."""
    _fields_ = [("admit_s", ctypes.c_double), ("done_s", ctypes.c_double), ("checksum", ctypes.c_ulonglong),
                ("status", _I), ("batch", _I), ("batch_size", _I)]


class ServerStats(ctypes.Structure):
    """cfgmm_server_stats."""
    _fields_ = [(k, ctypes.c_ulonglong) for k in ("rho_requests", "gemm_requests", "batches", "hits", "misses",
                                                   "uploads", "evictions")] + [("peak_bytes", _SZ),
                                                                               ("held_bytes", _SZ),
                                                                               ("rho_fp64_requests",
                                                                                ctypes.c_ulonglong)]


class Server:
    """The serving loop (cfgmm_server_*): molecules are registered as host numpy arrays, which
    this object keeps alive; ``run`` replays a request list in real time."""

    def __init__(self, hbm_budget_bytes, slots=4, rho_precision=PREC_FP64, rho_nslices=6, max_batch=16,
                 split_tile_cap_bytes=0):
        cfg = ServerConfig(hbm_budget_bytes, slots, rho_precision, rho_nslices, max_batch, split_tile_cap_bytes)
        h = _VP()
        check(lib().cfgmm_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        self._arrays = []

    def add_molecule(self, mid, dms, phi, tab, blksize):
        dms, phi, tab = _f64(dms), _f64(phi), _u8(tab)
        nset, nao = dms.shape[0], dms.shape[1]
        ngrid = phi.shape[1]
        self._arrays += [dms, phi, tab]
        check(lib().cfgmm_server_add_molecule(self.handle, mid, _ptr(dms), _ptr(phi), _ptr(tab), nao, ngrid, nset,
                                              blksize))

    def warm(self, ids):
        arr = (ctypes.c_int * len(ids))(*ids)
        check(lib().cfgmm_server_warm(self.handle, ctypes.cast(arr, _VP), len(ids)))

    def run(self, requests):
        """requests: sequence of Request; returns a ctypes array of RequestResult."""
        reqs = (Request * len(requests))(*requests)
        res = (RequestResult * len(requests))()
        check(lib().cfgmm_server_run(self.handle, ctypes.cast(reqs, _VP), len(reqs), ctypes.cast(res, _VP)))
        return res

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().cfgmm_server_get_stats(self.handle, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def destroy(self):
        if self.handle:
            check(lib().cfgmm_server_destroy(self.handle))
            self.handle = None


class ResidentEval(ctypes.Structure):
    """cfgmm_rho_resident_job: one entry of a batched resident evaluation."""
    _fields_ = [("h", _VP), ("rhos", _VP), ("stride_rho", _SZ), ("ngrid_used", _SZ), ("nset_used", _SZ)]


def rho_resident_eval_batched(stream, evals):
    """evals: sequence of (RhoResident, rhos buffer, stride_rho, ngrid, nset)."""
    arr = (ResidentEval * len(evals))(*[ResidentEval(h.handle, _dev(r), st, ng, ns) for h, r, st, ng, ns in evals])
    check(lib().cfgmm_rho_resident_eval_batched(_h(stream), ctypes.cast(arr, _VP), len(arr)))


def _dev(x):
    """Device address of a DeviceBuffer or a raw integer pointer."""
    return x.ptr if hasattr(x, "ptr") else x


def tune_set(param: int, value: int) -> None:
    check(lib().cfgmm_tune_set(param, value))


def tune_get(param: int) -> int:
    v = ctypes.c_int(0)
    check(lib().cfgmm_tune_get(param, ctypes.byref(v)))
    return v.value


def _h(stream):
    return stream.handle if stream is not None else None


def dgemm_dev(stream, c: DeviceBuffer, a: DeviceBuffer, b: DeviceBuffer, m, n, k, lda, ldb, ldc, transa, transb,
              precision=PREC_FP64, nslices=0):
    if precision == PREC_FP64:
        check(lib().cfgmm_dgemm_dev(_h(stream), c.ptr, a.ptr, b.ptr, m, n, k, lda, ldb, ldc, int(transa),
                                    int(transb)))
    else:
        check(lib().cfgmm_dgemm_ex_dev(_h(stream), c.ptr, a.ptr, b.ptr, m, n, k, lda, ldb, ldc, int(transa),
                                       int(transb), precision, nslices))


def dgemm_non0tab_dev(stream, c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab: DeviceBuffer, ldtab, blksize,
                      allow_transb=False):
    """Masked GEMM on device buffers; allow_transb selects cfgmm_dgemm_non0tab_ex_dev, which takes transb
    (the reference-shaped entry point rejects it, as the reference does)."""
    fn = lib().cfgmm_dgemm_non0tab_ex_dev if allow_transb else lib().cfgmm_dgemm_non0tab_dev
    check(fn(_h(stream), c.ptr, a.ptr, b.ptr, m, n, k, lda, ldb, ldc, int(transa), int(transb), tab.ptr, ldtab,
             blksize))


def sgemm_non0tab_dev(stream, c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab: DeviceBuffer, ldtab, blksize):
    """The masked GEMM with float32 buffers (cfgmm_sgemm_non0tab_dev; transb allowed, f32 accumulation)."""
    check(lib().cfgmm_sgemm_non0tab_dev(_h(stream), c.ptr, a.ptr, b.ptr, m, n, k, lda, ldb, ldc, int(transa),
                                        int(transb), tab.ptr, ldtab, blksize))


def aodmao2rho_dev(stream, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize,
                   precision=PREC_FP64, nslices=0):
    if precision == PREC_FP64:
        check(lib().cfgmm_aodmao2rho_dev(_h(stream), rhos.ptr, dms.ptr, ao_lhs.ptr, ao_rhs.ptr, stride_ao,
                                         stride_grid, nao, ngrid, nset, tab.ptr, ldtab, blksize))
    else:
        check(lib().cfgmm_aodmao2rho_ex_dev(_h(stream), rhos.ptr, dms.ptr, ao_lhs.ptr, ao_rhs.ptr, stride_ao,
                                            stride_grid, nao, ngrid, nset, tab.ptr, ldtab, blksize, precision,
                                            nslices))
