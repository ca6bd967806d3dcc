//! The GEMM with B rows masked per (k row, column block) (`src/impl_matmul_non0tab.rs`).
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
    /// `src/impl_matmul_non0tab.rs:210-305`:
    /// `c[i*ldc + j] += sum_{p : tab[p*ldtab + j/BLKSIZE]} op(A)(i, p) b[p*ldb + j]`; B values
    /// under a false bit are never read.  `transb = true` panics, as `unimplemented!` does in
    /// the reference (`src/impl_matmul_non0tab.rs:228-230`).
    pub fn matmul_loop_parallel_mnk_pack_a_non0tab<const BLKSIZE: usize>(
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
        tab: &[bool],
        ldtab: usize,
    ) {
        check(unsafe {
            ffi::cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab(
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
                tab.as_ptr() as *const u8, // bool: one byte, 0 or 1
                tab.len(),
                ldtab,
                BLKSIZE,
            )
        });
    }
}

/// `src/impl_matmul_non0tab.rs:308-326` (BLKSIZE 48).
pub fn matmul_anyway_full_non0tab(
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
    tab: &[bool],
    ldtab: usize,
) {
    check(unsafe {
        ffi::cfgmm_matmul_anyway_full_non0tab(
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
            tab.as_ptr() as *const u8,
            tab.len(),
            ldtab,
        )
    });
}
