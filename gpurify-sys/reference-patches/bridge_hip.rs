// The device solve inside crates/pex/src/bridge.rs::solve_capacitance
// (feature `hip`).  It replaces the homogeneous branch's method selection and
// solve (bridge.rs:397-408) and keeps everything else of the function: the
// mesher, the plate, the eps scaling, the cap_af loop and its errors.
//
// In the homogeneous branch, before building `geometry`:
//
//     let mut solved = match hip::cap_solve(&panels, names.len()) {
//         hip::Outcome::Solved(c) => CapResult::from_row_major(names.len(), c),
//         hip::Outcome::Failed(message) => return Err(BridgeError::Capacitance(message)),
//         hip::Outcome::Unavailable => { /* the existing CPU code, unchanged */ }
//     };
//
// NO_DEVICE, OOM and INTERNAL statuses never reach the caller as errors: the
// CPU path runs instead (checked and unchecked callers alike), so a missing
// or exhausted device cannot turn a solvable layout into an error.  Every
// other status is an answer the reference itself gives (invalid input,
// capacities, singular or empty system, non-finite result) and maps onto
// BridgeError::Capacitance with the library's message.

#[cfg(feature = "hip")]
pub(crate) mod hip {
    use std::cell::RefCell;

    use gpurify_sys::{default_opts, Context, Panels};

    use crate::quasistatic::integrals::panel::Panel;

    pub(crate) enum Outcome {
        /// C, n_conductors^2 row-major, farads, vacuum.
        Solved(Vec<f64>),
        /// An answer of the reference's own error classes.
        Failed(String),
        /// No usable device (or the device path failed): run the CPU path.
        Unavailable,
    }

    thread_local! {
        // One library context per thread, created on first use and destroyed
        // when the thread ends (Context implements Drop).  A failed creation
        // is remembered so the probe is not repeated on every call.
        static CTX: RefCell<Option<Result<Context, ()>>> = const { RefCell::new(None) };
    }

    pub(crate) fn cap_solve(panels: &[Panel], ncond: usize) -> Outcome {
        if !gdsverify_backend::hip_ready() {
            return Outcome::Unavailable;
        }
        // 4 vertices per panel: a triangle repeats its last vertex, which
        // leaves every integral of the panel unchanged
        let mut vertices = Vec::with_capacity(panels.len() * 12);
        let mut conductor = Vec::with_capacity(panels.len());
        for p in panels {
            for k in 0..4 {
                let q = p.vertices[k.min(p.vertices.len() - 1)];
                vertices.extend_from_slice(&[q.x, q.y, q.z]);
            }
            conductor.push(p.conductor as u32);
        }
        let set = Panels { vertices: &vertices, conductor: &conductor, n_conductors: ncond as u32 };
        CTX.with(|cell| {
            let mut slot = cell.borrow_mut();
            if slot.is_none() {
                *slot = Some(Context::new(0).map_err(|_| ()));
            }
            let Some(Ok(ctx)) = slot.as_ref() else {
                return Outcome::Unavailable;
            };
            // the reference's rule and constants: dense below 500 panels,
            // else FMM order 4, depth 4, tol 1e-8, CPQR-truncated M2L
            match ctx.cap_solve(&set, &default_opts()) {
                Ok((c, _report)) => Outcome::Solved(c),
                Err(e) if e.is_fallback() => Outcome::Unavailable,
                Err(e) => Outcome::Failed(e.message),
            }
        })
    }
}
