//! Row-major f64 GEMM `C += op(A) op(B)` (`src/impl_matmul.rs`), on FP64 MFMA.
use crate::ffi;
use crate::structs::{check, MatmulLoops};
use std::os::raw::c_int;

impl<
        const MC: usize,
        const KC: usize,
        const NC: usize,
        const MR: usize,
        const NR_LANE: usize,
        const LANE: usize,
        const MB: usize,
    > MatmulLoops<f64, MC, KC, NC, MR, NR_LANE, LANE, MB>
{
    /// `src/impl_matmul.rs:155-231`: `op(A)(i, p) = transa ? a[p*lda + i] : a[i*lda + p]`,
    /// `op(B)(p, j) = transb ? b[j*ldb + p] : b[p*ldb + j]`, `c[i*ldc + j] += sum_p op(A) op(B)`
    /// for `i < m, j < n` (columns of C past n untouched).  Synchronous.
    pub fn matmul_loop_parallel_mnk_pack_a(
        c: &mut [f64],
        a: &[f64],
        b: &[f64],
        m: usize,
        n: usize,
        k: usize,
        lda: usize,
        ldb: usize,
        ldc: usize,
        transa: bool,
        transb: bool,
    ) {
        check(unsafe {
            ffi::cfgmm_matmul_loop_parallel_mnk_pack_a(
                c.as_mut_ptr(),
                c.len(),
                a.as_ptr(),
                a.len(),
                b.as_ptr(),
                b.len(),
                m,
                n,
                k,
                lda,
                ldb,
                ldc,
                transa as c_int,
                transb as c_int,
            )
        });
    }

    /// README's name for the same call (`README.md:22-31`).
    pub fn matmul_loop_macro_mb(
        c: &mut [f64],
        a: &[f64],
        b: &[f64],
        m: usize,
        n: usize,
        k: usize,
        lda: usize,
        ldb: usize,
        ldc: usize,
        transa: bool,
        transb: bool,
    ) {
        check(unsafe {
            ffi::cfgmm_matmul_loop_macro_mb(
                c.as_mut_ptr(),
                c.len(),
                a.as_ptr(),
                a.len(),
                b.as_ptr(),
                b.len(),
                m,
                n,
                k,
                lda,
                ldb,
                ldc,
                transa as c_int,
                transb as c_int,
            )
        });
    }
}

/// `src/impl_matmul.rs:234-248`.
pub fn matmul_anyway_full(
    c: &mut [f64],
    a: &[f64],
    b: &[f64],
    m: usize,
    n: usize,
    k: usize,
    lda: usize,
    ldb: usize,
    ldc: usize,
    transa: bool,
    transb: bool,
) {
    check(unsafe {
        ffi::cfgmm_matmul_anyway_full(
            c.as_mut_ptr(),
            c.len(),
            a.as_ptr(),
            a.len(),
            b.as_ptr(),
            b.len(),
            m,
            n,
            k,
            lda,
            ldb,
            ldc,
            transa as c_int,
            transb as c_int,
        )
    });
}
