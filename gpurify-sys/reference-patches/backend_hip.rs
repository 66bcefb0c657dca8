// Additions to crates/backend/src/lib.rs for the HIP field-solver backend.
// The enum there is documented as deliberately closed (lib.rs:17); adding a
// variant is an explicit API change of that crate.

// (1) In `pub enum Backend` (lib.rs:19-25), after `Gpu`:
//
//     /// MI355X (gfx950) field solver through libgpurify (feature `hip`);
//     /// callers fall back to [`Backend::Cpu`] when unavailable.
//     Hip,

// (2) In `Backend::as_str` (lib.rs:27-36), the new arm:
//
//     Backend::Hip => "hip",

// (3) Beside `gpu_ready` (lib.rs:50-59):

/// Is a gfx950 device usable by libgpurify right now?  `false` without the
/// `hip` feature.  The probe creates no context and cannot abort: the library
/// reports a missing or foreign device as "not ready".
#[must_use]
pub fn hip_ready() -> bool {
    #[cfg(feature = "hip")]
    {
        gpurify_sys::ready()
    }
    #[cfg(not(feature = "hip"))]
    {
        false
    }
}

// (4) In `available_backends()`, after the Gpu entry:
//
//     if hip_ready() {
//         out.push(Backend::Hip);
//     }
