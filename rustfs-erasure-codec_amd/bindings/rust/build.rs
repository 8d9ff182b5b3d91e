//! `build.rs` additions of the reference crate (feature `gpu-rocm`): compiles the
//! HIP sources for gfx950 with hipcc and links `librsgpu.so`.  `RSGPU_CSRC` points
//! at the `csrc/` directory of this package.
//!
//! `apply_inst.hip` holds one instantiation set of the tiled matrix-apply kernel;
//! it is compiled once per (tile shape, element type) pair, as the package's
//! `build.py` does.

use std::path::{Path, PathBuf};
use std::process::Command;

const HIPCC: &str = "/opt/rocm/bin/hipcc";
const FLAGS: [&str; 9] = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-c",
];

fn compile(src: &Path, obj: &Path, defs: &[String]) {
    let st = Command::new(HIPCC).args(FLAGS).args(defs).arg(src).arg("-o").arg(obj).status()
        .expect("run hipcc");
    assert!(st.success(), "hipcc failed for {}", src.display());
}

pub fn build_rsgpu() {
    let csrc = PathBuf::from(std::env::var("RSGPU_CSRC").unwrap_or_else(|_| "rsgpu/csrc".into()));
    let out = PathBuf::from(std::env::var("OUT_DIR").unwrap());
    let mut objs = Vec::new();
    let mut units: Vec<String> = std::fs::read_dir(&csrc)
        .unwrap()
        .filter_map(|e| e.ok()?.file_name().into_string().ok())
        .filter(|f| f.ends_with(".hip") && f != "apply_inst.hip")
        .collect();
    units.sort();
    for f in &units {
        let obj = out.join(f.replace(".hip", ".o"));
        compile(&csrc.join(f), &obj, &[]);
        objs.push(obj);
        println!("cargo:rerun-if-changed={}", csrc.join(f).display());
    }
    for tower in 0..2 {
        for tile in 0..5 {
            let obj = out.join(format!("apply_t{tile}_w{tower}.o"));
            compile(&csrc.join("apply_inst.hip"), &obj,
                    &[format!("-DAPPLY_TILE={tile}"), format!("-DAPPLY_TOWER={tower}")]);
            objs.push(obj);
        }
    }
    let so = out.join("librsgpu.so");
    let st = Command::new(HIPCC)
        .args(["--offload-arch=gfx950", "-shared", "-fPIC", "-o"])
        .arg(&so)
        .args(&objs)
        .arg("-lpthread")
        .status()
        .expect("link librsgpu.so");
    assert!(st.success());
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=rsgpu");
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
    println!("cargo:rustc-link-lib=dylib=amdhip64");
}

fn main() {
    if std::env::var_os("CARGO_FEATURE_GPU_ROCM").is_some() {
        build_rsgpu();
    }
}
