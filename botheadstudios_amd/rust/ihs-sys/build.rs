// Links libihs_sph.so, built in-tree by `python -m botheadstudios_amd.build` (hipcc, gfx950). The library sits two
// folders above this crate (botheadstudios_amd/); IHS_SPH_LIB_DIR overrides the folder.
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = match env::var("IHS_SPH_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("cargo sets CARGO_MANIFEST_DIR"))
            .join("..")
            .join(".."),
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=ihs_sph");
    println!("cargo:rerun-if-env-changed=IHS_SPH_LIB_DIR");
    println!("cargo:rerun-if-changed=build.rs");
}
