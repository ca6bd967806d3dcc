//! Re-exports, as the reference's `prelude` does for its own modules (`src/prelude.rs:1-13`).
pub use crate::impl_aodmao2rho::aodmao2rho_anyway;
pub use crate::impl_matmul::matmul_anyway_full;
pub use crate::impl_matmul_non0tab::matmul_anyway_full_non0tab;
pub use crate::structs::MatmulLoops;
