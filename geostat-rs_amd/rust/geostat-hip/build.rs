// Links libgeostat_hip.so (built by geostat-rs_amd/build_lib.py for gfx950).
// GEOSTAT_HIP_LIB_DIR overrides the default location next to this crate.
fn main() {
    let dir = std::env::var("GEOSTAT_HIP_LIB_DIR").unwrap_or_else(|_| {
        let manifest = std::env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR");
        format!("{manifest}/../../build")
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=geostat_hip");
    println!("cargo:rerun-if-env-changed=GEOSTAT_HIP_LIB_DIR");
}
