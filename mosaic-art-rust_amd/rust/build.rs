// Links libmosaic_hip.so (built by mosaic-art-rust_amd/build.py: hipcc
// --offload-arch=gfx950) and the ROCm runtime it needs.
fn main() {
    let dir = std::env::var("MOSAIC_HIP_DIR")
        .expect("set MOSAIC_HIP_DIR to the directory holding libmosaic_hip.so");
    let rocm = std::env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string());
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=mosaic_hip");
    println!("cargo:rustc-link-search=native={rocm}/lib");
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    println!("cargo:rerun-if-env-changed=MOSAIC_HIP_DIR");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");
}
