//! Links libscatrs_hip.so: the directory comes from SCATRS_HIP_LIB_DIR (default: the library's
//! in-tree build location, scatrs_amd/ beside this crate's rust/ directory). The library itself
//! is built with hipcc for gfx950 (`python -m scatrs_amd.build`); it carries its own HIP runtime
//! dependency (libamdhip64), which the dynamic loader resolves from the ROCm installation.
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = env::var("SCATRS_HIP_LIB_DIR").map(PathBuf::from).unwrap_or_else(|_| {
        PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../scatrs_amd")
    });
    println!("cargo:rerun-if-env-changed=SCATRS_HIP_LIB_DIR");
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=scatrs_hip");
    // Let binaries find the library next to where it was linked from.
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
