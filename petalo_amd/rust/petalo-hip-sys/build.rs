// Links libpetalo_hip.so.  PETALO_HIP_LIB_DIR names the directory that holds
// it (default: the package directory of this repository, where
// `python -m petalo_amd.build` puts it); the ROCm runtime it depends on must
// be on the loader path.
fn main() {
    let dir = std::env::var("PETALO_HIP_LIB_DIR").unwrap_or_else(|_| {
        let here = std::path::PathBuf::from(std::env::var("CARGO_MANIFEST_DIR").unwrap());
        here.join("../..").to_string_lossy().into_owned()
    });
    println!("cargo:rerun-if-env-changed=PETALO_HIP_LIB_DIR");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=petalo_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
}
