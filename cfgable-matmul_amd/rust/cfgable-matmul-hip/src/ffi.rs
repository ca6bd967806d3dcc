//! Raw declarations of libcfgmm's C ABI (`include/cfgmm.h`), one `extern "C"` item per C
//! declaration with the same parameter list: `size_t` is `usize`, `int` is `c_int`,
//! `double *` is `*mut f64`, `uint8_t *` is `*const u8`, `void *` stream / handle arguments are
//! `*mut c_void`, structures are passed by pointer.  A test of the repository compares this
//! list with the header declaration by declaration.
#![allow(non_camel_case_types)]
use std::os::raw::{c_char, c_int, c_void};

/// `cfgmm_rho_problem`: one problem of a batched rho call (device pointers).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct cfgmm_rho_problem {
    pub rhos: *mut f64,
    pub dms: *const f64,
    pub ao_lhs: *const f64,
    pub ao_rhs: *const f64,
    pub stride_ao: usize,
    pub stride_grid: usize,
    pub nao: usize,
    pub ngrid: usize,
    pub nset: usize,
    pub tab: *const u8,
    pub ldtab: usize,
    pub blksize: usize,
}

pub const CFGMM_OK: c_int = 0;
pub const CFGMM_PREC_FP64: c_int = 0;
pub const CFGMM_PREC_OZAKI_INT8: c_int = 1;

extern "C" {
    pub fn cfgmm_last_error() -> *const c_char;
    pub fn cfgmm_version() -> c_int;

    // host-slice drop-ins (src/impl_matmul.rs, src/impl_matmul_non0tab.rs, src/impl_aodmao2rho.rs)
    pub fn cfgmm_matmul_loop_parallel_mnk_pack_a(c: *mut f64, c_len: usize, a: *const f64, a_len: usize,
        b: *const f64, b_len: usize, m: usize, n: usize, k: usize, lda: usize, ldb: usize, ldc: usize,
        transa: c_int, transb: c_int) -> c_int;
    pub fn cfgmm_matmul_loop_macro_mb(c: *mut f64, c_len: usize, a: *const f64, a_len: usize, b: *const f64,
        b_len: usize, m: usize, n: usize, k: usize, lda: usize, ldb: usize, ldc: usize, transa: c_int,
        transb: c_int) -> c_int;
    pub fn cfgmm_matmul_anyway_full(c: *mut f64, c_len: usize, a: *const f64, a_len: usize, b: *const f64,
        b_len: usize, m: usize, n: usize, k: usize, lda: usize, ldb: usize, ldc: usize, transa: c_int,
        transb: c_int) -> c_int;
    pub fn cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab(c: *mut f64, c_len: usize, a: *const f64, a_len: usize,
        b: *const f64, b_len: usize, m: usize, n: usize, k: usize, lda: usize, ldb: usize, ldc: usize,
        transa: c_int, transb: c_int, tab: *const u8, tab_len: usize, ldtab: usize, blksize: usize) -> c_int;
    pub fn cfgmm_matmul_anyway_full_non0tab(c: *mut f64, c_len: usize, a: *const f64, a_len: usize,
        b: *const f64, b_len: usize, m: usize, n: usize, k: usize, lda: usize, ldb: usize, ldc: usize,
        transa: c_int, transb: c_int, tab: *const u8, tab_len: usize, ldtab: usize) -> c_int;
    pub fn cfgmm_aodmao2rho_loop_parallel(rhos: *mut f64, rhos_len: usize, dms: *const f64, dms_len: usize,
        ao_lhs: *const f64, ao_lhs_len: usize, ao_rhs: *const f64, ao_rhs_len: usize, stride_ao: usize,
        stride_grid: usize, nao: usize, ngrid: usize, nset: usize, tab: *const u8, tab_len: usize, ldtab: usize,
        blksize: usize) -> c_int;
    pub fn cfgmm_aodmao2rho_anyway(rhos: *mut f64, rhos_len: usize, dms: *const f64, dms_len: usize,
        ao_lhs: *const f64, ao_lhs_len: usize, ao_rhs: *const f64, ao_rhs_len: usize, stride_ao: usize,
        stride_grid: usize, nao: usize, ngrid: usize, nset: usize, tab: *const u8, tab_len: usize,
        ldtab: usize) -> c_int;

    // device pointers and a stream (additive)
    pub fn cfgmm_dgemm_dev(stream: *mut c_void, c: *mut f64, a: *const f64, b: *const f64, m: usize, n: usize,
        k: usize, lda: usize, ldb: usize, ldc: usize, transa: c_int, transb: c_int) -> c_int;
    pub fn cfgmm_dgemm_ex_dev(stream: *mut c_void, c: *mut f64, a: *const f64, b: *const f64, m: usize,
        n: usize, k: usize, lda: usize, ldb: usize, ldc: usize, transa: c_int, transb: c_int, precision: c_int,
        nslices: c_int) -> c_int;
    pub fn cfgmm_dgemm_non0tab_dev(stream: *mut c_void, c: *mut f64, a: *const f64, b: *const f64, m: usize,
        n: usize, k: usize, lda: usize, ldb: usize, ldc: usize, transa: c_int, transb: c_int, tab: *const u8,
        ldtab: usize, blksize: usize) -> c_int;
    pub fn cfgmm_aodmao2rho_dev(stream: *mut c_void, rhos: *mut f64, dms: *const f64, ao_lhs: *const f64,
        ao_rhs: *const f64, stride_ao: usize, stride_grid: usize, nao: usize, ngrid: usize, nset: usize,
        tab: *const u8, ldtab: usize, blksize: usize) -> c_int;
    pub fn cfgmm_aodmao2rho_dev_noalloc(stream: *mut c_void, rhos: *mut f64, dms: *const f64,
        ao_lhs: *const f64, ao_rhs: *const f64, stride_ao: usize, stride_grid: usize, nao: usize, ngrid: usize,
        nset: usize, tab: *const u8, ldtab: usize, blksize: usize) -> c_int;
    pub fn cfgmm_aodmao2rho_ex_dev(stream: *mut c_void, rhos: *mut f64, dms: *const f64, ao_lhs: *const f64,
        ao_rhs: *const f64, stride_ao: usize, stride_grid: usize, nao: usize, ngrid: usize, nset: usize,
        tab: *const u8, ldtab: usize, blksize: usize, precision: c_int, nslices: c_int) -> c_int;
    pub fn cfgmm_aodmao2rho_batched_dev(stream: *mut c_void, problems: *const cfgmm_rho_problem,
        nproblems: usize, precision: c_int, nslices: c_int) -> c_int;

    // plumbing
    pub fn cfgmm_device_malloc(ptr: *mut *mut c_void, bytes: usize) -> c_int;
    pub fn cfgmm_device_free(ptr: *mut c_void) -> c_int;
    pub fn cfgmm_memcpy_htod(dst: *mut c_void, src: *const c_void, bytes: usize) -> c_int;
    pub fn cfgmm_memcpy_dtoh(dst: *mut c_void, src: *const c_void, bytes: usize) -> c_int;
    pub fn cfgmm_stream_create(stream: *mut *mut c_void) -> c_int;
    pub fn cfgmm_stream_destroy(stream: *mut c_void) -> c_int;
    pub fn cfgmm_stream_synchronize(stream: *mut c_void) -> c_int;
}
