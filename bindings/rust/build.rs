// Links libhyperprec_hip.so.  HYPERPREC_HIP_DIR names the directory holding it (the build
// writes it to quadprec_amd/ of this repository: `python -m quadprec_amd.build`).
fn main() {
    let dir = std::env::var("HYPERPREC_HIP_DIR").unwrap_or_else(|_| "../../quadprec_amd".to_string());
    println!("cargo:rerun-if-env-changed=HYPERPREC_HIP_DIR");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=hyperprec_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
}
