// Links libcfgmm.so (built by `make -C cfgable-matmul_amd`).  CFGMM_LIB_DIR overrides the
// default, the in-tree cfgable-matmul_amd/lib two levels above this crate.
use std::path::PathBuf;

fn main() {
    let dir = match std::env::var("CFGMM_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(std::env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../lib"),
    };
    let dir = dir.canonicalize().unwrap_or(dir);
    println!("cargo:rerun-if-env-changed=CFGMM_LIB_DIR");
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=cfgmm");
    // the shared library is found at run time without LD_LIBRARY_PATH
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
