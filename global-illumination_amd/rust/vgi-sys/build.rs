//! Links `libvgi.so`. The library is built by global-illumination_amd/build.py
//! (hipcc for gfx950) into global-illumination_amd/; `VGI_LIB_DIR` overrides
//! where it is looked for.
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = match env::var("VGI_LIB_DIR") {
        Ok(dir) => PathBuf::from(dir),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("..").join(".."),
    };
    println!("cargo:rerun-if-env-changed=VGI_LIB_DIR");
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=vgi");
}
