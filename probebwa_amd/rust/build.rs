// Builds libprobebwa_amd.so with the library's own Makefile (hipcc for gfx950)
// and links it.  PROBEBWA_AMD_LIB_DIR points at a prebuilt library instead.
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    println!("cargo:rerun-if-env-changed=PROBEBWA_AMD_LIB_DIR");
    let dir = match env::var("PROBEBWA_AMD_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => {
            let pkg = PathBuf::from(env!("CARGO_MANIFEST_DIR")).join("..");
            let csrc = pkg.join("csrc");
            for f in ["Makefile", "pbw_kernels.hip", "pbw_mlp.hip", "pbw_index.hip", "pbw_host.cpp", "pbw_io.cpp",
                      "pbw_internal.h", "pbw_pair_kernels.inc"] {
                println!("cargo:rerun-if-changed={}", csrc.join(f).display());
            }
            let ok = Command::new("make")
                .arg("-C")
                .arg(&csrc)
                .status()
                .expect("make not found")
                .success();
            assert!(ok, "hipcc build of libprobebwa_amd.so failed");
            pkg
        }
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=probebwa_amd");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
