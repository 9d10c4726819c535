// Links libnewt_gpu.so (built by `python -m tooling_amd.build` into
// tooling_amd/_build).  NEWT_GPU_LIB_DIR overrides the directory.
fn main() {
    let dir = std::env::var("NEWT_GPU_LIB_DIR").unwrap_or_else(|_| {
        let here = std::path::PathBuf::from(std::env::var("CARGO_MANIFEST_DIR").unwrap());
        here.join("../../_build").canonicalize().map(|p| p.display().to_string()).unwrap_or_default()
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=newt_gpu");
    println!("cargo:rerun-if-env-changed=NEWT_GPU_LIB_DIR");
}
