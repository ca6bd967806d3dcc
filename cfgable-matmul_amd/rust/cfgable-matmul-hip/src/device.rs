//! Device-resident operands (no counterpart in the reference): buffers that stay in HBM across
//! calls, a stream, and the device-pointer forms of the GEMM and rho, in FP64 or split
//! precision.  Errors panic with the library's message, as the host-slice calls do.
use crate::ffi;
use crate::structs::check;
use std::os::raw::{c_int, c_void};

/// Precision of a device call: FP64 MFMA, or split precision with `n` balanced int8 digits per
/// operand (Ozaki scheme; 5 keep the GEMM within about 1e-11 of FP64 on well-scaled data).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Precision {
    Fp64,
    Split(u8),
}

impl Precision {
    fn raw(self) -> (c_int, c_int) {
        match self {
            Precision::Fp64 => (ffi::CFGMM_PREC_FP64, 0),
            Precision::Split(n) => (ffi::CFGMM_PREC_OZAKI_INT8, n as c_int),
        }
    }
}

/// A HIP stream created by the library (non-blocking, default priority).
pub struct Stream(*mut c_void);

impl Stream {
    pub fn new() -> Self {
        let mut s = core::ptr::null_mut();
        check(unsafe { ffi::cfgmm_stream_create(&mut s) });
        Stream(s)
    }
    pub fn synchronize(&self) {
        check(unsafe { ffi::cfgmm_stream_synchronize(self.0) });
    }
    pub fn raw(&self) -> *mut c_void {
        self.0
    }
}

impl Default for Stream {
    fn default() -> Self {
        Self::new()
    }
}

impl Drop for Stream {
    fn drop(&mut self) {
        unsafe { ffi::cfgmm_stream_destroy(self.0) };
    }
}

/// `len` elements of `T` in device memory.
pub struct DeviceBuffer<T: Copy> {
    ptr: *mut T,
    len: usize,
}

impl<T: Copy> DeviceBuffer<T> {
    pub fn new(len: usize) -> Self {
        let mut p = core::ptr::null_mut();
        check(unsafe { ffi::cfgmm_device_malloc(&mut p, len * core::mem::size_of::<T>()) });
        DeviceBuffer { ptr: p as *mut T, len }
    }
    pub fn from_slice(x: &[T]) -> Self {
        let b = Self::new(x.len());
        b.upload(x);
        b
    }
    pub fn upload(&self, x: &[T]) {
        assert!(x.len() <= self.len);
        check(unsafe {
            ffi::cfgmm_memcpy_htod(self.ptr as *mut c_void, x.as_ptr() as *const c_void, core::mem::size_of_val(x))
        });
    }
    pub fn download(&self, out: &mut [T]) {
        assert!(out.len() <= self.len);
        check(unsafe {
            ffi::cfgmm_memcpy_dtoh(out.as_mut_ptr() as *mut c_void, self.ptr as *const c_void, core::mem::size_of_val(out))
        });
    }
    pub fn len(&self) -> usize {
        self.len
    }
    pub fn is_empty(&self) -> bool {
        self.len == 0
    }
    pub fn as_ptr(&self) -> *const T {
        self.ptr
    }
    pub fn as_mut_ptr(&mut self) -> *mut T {
        self.ptr
    }
}

impl<T: Copy> Drop for DeviceBuffer<T> {
    fn drop(&mut self) {
        unsafe { ffi::cfgmm_device_free(self.ptr as *mut c_void) };
    }
}

/// `C += op(A) op(B)` on device buffers (`cfgmm_dgemm_ex_dev`), arguments as
/// `MatmulLoops::matmul_loop_parallel_mnk_pack_a`; asynchronous on `stream`.
pub fn dgemm(stream: &Stream, c: &mut DeviceBuffer<f64>, a: &DeviceBuffer<f64>, b: &DeviceBuffer<f64>, m: usize,
             n: usize, k: usize, lda: usize, ldb: usize, ldc: usize, transa: bool, transb: bool, prec: Precision) {
    let (p, ns) = prec.raw();
    check(unsafe {
        ffi::cfgmm_dgemm_ex_dev(stream.raw(), c.as_mut_ptr(), a.as_ptr(), b.as_ptr(), m, n, k, lda, ldb, ldc,
                                transa as c_int, transb as c_int, p, ns)
    });
}

/// Masked rho on device buffers (`cfgmm_aodmao2rho_ex_dev`), arguments as
/// `MatmulLoops::aodmao2rho_loop_parallel` with BLKSIZE at run time; asynchronous on `stream`.
pub fn aodmao2rho(stream: &Stream, rhos: &mut DeviceBuffer<f64>, dms: &DeviceBuffer<f64>,
                  ao_lhs: &DeviceBuffer<f64>, ao_rhs: &DeviceBuffer<f64>, strides: [usize; 2], shape: [usize; 2],
                  nset: usize, tab: &DeviceBuffer<u8>, ldtab: usize, blksize: usize, prec: Precision) {
    let (p, ns) = prec.raw();
    check(unsafe {
        ffi::cfgmm_aodmao2rho_ex_dev(stream.raw(), rhos.as_mut_ptr(), dms.as_ptr(), ao_lhs.as_ptr(),
                                     ao_rhs.as_ptr(), strides[0], strides[1], shape[0], shape[1], nset,
                                     tab.as_ptr(), ldtab, blksize, p, ns)
    });
}
