//! `cfgable-matmul` on MI355X: the reference crate's public calls with the same module paths,
//! type and argument lists, running on the GPU through libcfgmm's C ABI (`include/cfgmm.h`).
//!
//! | Reference (file:line) | Here |
//! |---|---|
//! | `structs::MatmulLoops<T, MC, KC, NC, MR, NR_LANE, LANE>` `src/structs.rs:15-17` | [`structs::MatmulLoops`] (plus README's `MB`, default 0) |
//! | `MatmulLoops::matmul_loop_parallel_mnk_pack_a` `src/impl_matmul.rs:155-231` | [`impl_matmul`] |
//! | README `matmul_loop_macro_mb` `README.md:22-31` | [`impl_matmul`] |
//! | `impl_matmul::matmul_anyway_full` `src/impl_matmul.rs:234-248` | [`impl_matmul::matmul_anyway_full`] |
//! | `MatmulLoops::matmul_loop_parallel_mnk_pack_a_non0tab::<BLKSIZE>` `src/impl_matmul_non0tab.rs:210-305` | [`impl_matmul_non0tab`] |
//! | `impl_matmul_non0tab::matmul_anyway_full_non0tab` `src/impl_matmul_non0tab.rs:308-326` | [`impl_matmul_non0tab::matmul_anyway_full_non0tab`] |
//! | `MatmulLoops::aodmao2rho_loop_parallel::<BLKSIZE>` `src/impl_aodmao2rho.rs:171-244` | [`impl_aodmao2rho`] |
//! | `impl_aodmao2rho::aodmao2rho_anyway` `src/impl_aodmao2rho.rs:247-261` | [`impl_aodmao2rho::aodmao2rho_anyway`] |
//!
//! The CPU blocking constants (`MC`, `KC`, `NC`, `MR`, `NR_LANE`, `LANE`, `MB`) are accepted and
//! have no effect: the GPU kernels pick their own tiles.  The calls keep the reference's
//! semantics: outputs are accumulated into, rho rows are strided by `stride_grid`, a `bool`
//! mask is one byte per (row, block).  Where the reference panics (`unimplemented!` for transb
//! with a mask, `assert!` on BLKSIZE, a slice index out of bounds) these calls panic too, with
//! the library's message; no out-of-bounds access happens first, since every slice goes to the
//! library with its length.
//!
//! The microkernel traits of the reference (`structs::MatmulMicroKernelAPI`,
//! `impl_matmul_non0tab::MatmulMicroKernelMaskKForBAPI`) are a CPU register-tile plug-in point
//! with no GPU counterpart and are not provided.  [`device`] adds the device-pointer and
//! split-precision entry points, which the reference does not have.
#![allow(clippy::too_many_arguments)]
#![allow(non_snake_case)]

pub mod device;
pub mod ffi;
pub mod impl_aodmao2rho;
pub mod impl_matmul;
pub mod impl_matmul_non0tab;
pub mod prelude;
pub mod structs;
