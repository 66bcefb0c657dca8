//! Link libgpurify.so.  GPURIFY_LIB_DIR names the directory holding it (the
//! in-tree build puts it in gpurify_amd/); an rpath keeps the binary
//! runnable without LD_LIBRARY_PATH.
fn main() {
    println!("cargo:rerun-if-env-changed=GPURIFY_LIB_DIR");
    if let Ok(dir) = std::env::var("GPURIFY_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    }
    println!("cargo:rustc-link-lib=dylib=gpurify");
}
