//! The fused masked AO-DM-AO density contraction (`src/impl_aodmao2rho.rs`).
use crate::ffi;
use crate::structs::{check, MatmulLoops};

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
    /// `src/impl_aodmao2rho.rs:171-244`:
    /// `rhos[s*stride_grid + g] += sum_{u, v in L_b} ao_lhs[u*stride_grid + g]
    ///  dms[s*stride_ao^2 + u*stride_ao + v] ao_rhs[v*stride_grid + g]`, `b = g / BLKSIZE`,
    /// `L_b = { x : tab[x*ldtab + b] }`, with `strides = [stride_ao, stride_grid]` and
    /// `shape = [nao, ngrid]`.  BLKSIZE must be a positive multiple of 16 (the reference asserts
    /// `NC % BLKSIZE == 0`, `src/impl_aodmao2rho.rs:188`; here any such BLKSIZE works).
    pub fn aodmao2rho_loop_parallel<const BLKSIZE: usize>(
        rhos: &mut [f64],
        dms: &[f64],
        ao_lhs: &[f64],
        ao_rhs: &[f64],
        strides: [usize; 2],
        shape: [usize; 2],
        nset: usize,
        tab: &[bool],
        ldtab: usize,
    ) {
        let [stride_ao, stride_grid] = strides;
        let [nao, ngrid] = shape;
        check(unsafe {
            ffi::cfgmm_aodmao2rho_loop_parallel(
                rhos.as_mut_ptr(),
                rhos.len(),
                dms.as_ptr(),
                dms.len(),
                ao_lhs.as_ptr(),
                ao_lhs.len(),
                ao_rhs.as_ptr(),
                ao_rhs.len(),
                stride_ao,
                stride_grid,
                nao,
                ngrid,
                nset,
                tab.as_ptr() as *const u8,
                tab.len(),
                ldtab,
                BLKSIZE,
            )
        });
    }
}

/// `src/impl_aodmao2rho.rs:247-261` (BLKSIZE 48).
pub fn aodmao2rho_anyway(
    rhos: &mut [f64],
    dms: &[f64],
    ao_lhs: &[f64],
    ao_rhs: &[f64],
    strides: [usize; 2],
    shape: [usize; 2],
    nset: usize,
    tab: &[bool],
    ldtab: usize,
) {
    let [stride_ao, stride_grid] = strides;
    let [nao, ngrid] = shape;
    check(unsafe {
        ffi::cfgmm_aodmao2rho_anyway(
            rhos.as_mut_ptr(),
            rhos.len(),
            dms.as_ptr(),
            dms.len(),
            ao_lhs.as_ptr(),
            ao_lhs.len(),
            ao_rhs.as_ptr(),
            ao_rhs.len(),
            stride_ao,
            stride_grid,
            nao,
            ngrid,
            nset,
            tab.as_ptr() as *const u8,
            tab.len(),
            ldtab,
        )
    });
}
