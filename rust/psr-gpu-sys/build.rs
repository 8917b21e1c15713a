//! Builds libpsrgpu.so for gfx950 with the repository's own recipe
//! (pulsarsr_amd/build.py: hipcc --offload-arch=gfx950 -ffp-contract=off) and
//! links it.  PSRGPU_SRC names the checkout of this repository; PSRGPU_LIB may
//! instead name a directory holding an already built libpsrgpu.so.
use std::{env, path::PathBuf, process::Command};

fn main() {
    println!("cargo:rerun-if-env-changed=PSRGPU_SRC");
    println!("cargo:rerun-if-env-changed=PSRGPU_LIB");
    let lib_dir = if let Ok(dir) = env::var("PSRGPU_LIB") {
        PathBuf::from(dir)
    } else {
        let root = PathBuf::from(env::var("PSRGPU_SRC").expect(
            "set PSRGPU_SRC to the pulsarsr_amd checkout (or PSRGPU_LIB to a directory with libpsrgpu.so)",
        ));
        let status = Command::new("python3")
            .arg(root.join("pulsarsr_amd").join("build.py"))
            .status()
            .expect("python3 and hipcc (ROCm) are needed to build libpsrgpu");
        assert!(status.success(), "building libpsrgpu failed");
        println!("cargo:rerun-if-changed={}", root.join("pulsarsr_amd").join("csrc").display());
        println!("cargo:rerun-if-changed={}", root.join("include").join("psr_gpu.h").display());
        root.join("pulsarsr_amd")
    };
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=psrgpu");
    // the shared library is found at run time through the same directory
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
}
