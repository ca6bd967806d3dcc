//! `MatmulLoops`, the reference's API anchor (`src/structs.rs:15-17`).
use core::marker::PhantomData;
use std::ffi::CStr;
use std::os::raw::c_int;

/// `MatmulLoops<T, MC, KC, NC, MR, NR_LANE, LANE>` of the reference, with README's `MB`
/// (`README.md:22-31`) as a defaulted eighth constant so that both turbofish forms,
/// `MatmulLoops::<f64, 252, 512, 240, 14, 2, 8>` and `MatmulLoops::<f64, 252, 512, 240, 14, 2, 8, 2360>`,
/// name this type.  The methods exist for `T = f64`, as the reference's microkernel does
/// (`src/structs.rs:19-47`).  The constants are CPU blocking parameters and are ignored here.
pub struct MatmulLoops<
    T,
    const MC: usize,
    const KC: usize,
    const NC: usize,
    const MR: usize,
    const NR_LANE: usize,
    const LANE: usize,
    const MB: usize = 0,
> {
    _phantom: PhantomData<T>,
}

/// Panics with libcfgmm's message on a non-zero status: the reference panics in the same
/// situations (argument errors), and a GPU error has no other way out of a `()` return.
pub(crate) fn check(status: c_int) {
    if status != 0 {
        let msg = unsafe { CStr::from_ptr(crate::ffi::cfgmm_last_error()) }.to_string_lossy().into_owned();
        panic!("cfgmm status {status}: {msg}");
    }
}
