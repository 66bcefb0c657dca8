//! FFI of libgpurify (include/gpurify.h): the MI355X field-solver capacitance
//! engine behind the reference's `solve_capacitance` (crates/pex/src/bridge.rs:307-443).
//!
//! The raw items mirror the C header one for one (`#[repr(C)]` structs in
//! field order, `extern "C"` functions).  Above them a small safe layer:
//! [`Context`] owns a library context and destroys it on drop, every call
//! returns `Result<_, Error>`, and [`Error::is_fallback`] says whether the
//! caller should run its CPU path instead (no usable device, device memory
//! exhausted, HIP failure) or report the error as the reference would.
//!
//! Source only in this repository: the image that builds the library has no
//! Rust toolchain.  `tests/test_abi.py` checks that the structs and the
//! function list here match the header's.
#![allow(non_camel_case_types)]

use std::ffi::CStr;
use std::os::raw::{c_char, c_int};

// ------------------------------------------------------------ status codes
pub const GP_OK: c_int = 0;
pub const GP_ERR_INVALID_INPUT: c_int = 1;
pub const GP_ERR_NO_DEVICE: c_int = 2;
pub const GP_ERR_OOM: c_int = 3;
pub const GP_ERR_NOT_CONVERGED: c_int = 4;
pub const GP_ERR_NON_FINITE: c_int = 5;
pub const GP_ERR_INTERNAL: c_int = 6;
pub const GP_ERR_GRID_CAPACITY: c_int = 7;
pub const GP_ERR_PANEL_CAPACITY: c_int = 8;
pub const GP_ERR_SINGULAR: c_int = 9;
pub const GP_ERR_EMPTY: c_int = 10;

pub const GP_PREC_FP64: u32 = 0;
pub const GP_PREC_MIXED: u32 = 1;
pub const GP_METHOD_AUTO: u32 = 0;
pub const GP_METHOD_DENSE: u32 = 1;
pub const GP_METHOD_FMM: u32 = 2;
pub const GP_M2L_FP64: u32 = 0;
pub const GP_M2L_FP32: u32 = 1;
pub const GP_M2L_BF16: u32 = 2;
pub const GP_M2L_FP64_VALU: u32 = 3;
pub const GP_M2L_BF16_VALU: u32 = 4;
pub const GP_BASIS_FP32: u32 = 0;
pub const GP_BASIS_FP64: u32 = 1;
pub const GP_NEAR_FP32: u32 = 0;
pub const GP_NEAR_FP16: u32 = 1;
pub const GP_GROUND_NONE: u32 = 0;
pub const GP_GROUND_IMAGE: u32 = 1;
pub const GP_ROLE_CONDUCTOR: u32 = 0;
pub const GP_ROLE_DIELECTRIC: u32 = 1;
pub const GP_NOALLOC_MAX_PANELS: u32 = 512;
pub const GP_NOALLOC_MAX_CONDUCTORS: u32 = 64;
pub const GP_NOALLOC_SCRATCH_BYTES: usize = 2726656;

// ------------------------------------------------------------ structs
#[repr(C)]
pub struct gp_ctx {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct gp_panels {
    pub n_panels: u32,
    pub n_conductors: u32,
    pub vertices: *const f64,
    pub conductor: *const u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct gp_cap_opts {
    pub precision: u32,
    pub method: u32,
    pub fmm_order: u32,
    pub max_level: u32,
    pub leaf_target: u32,
    pub restart: u32,
    pub max_restarts: u32,
    pub rhs_block: u32,
    pub tol: f64,
    pub inner_tol: f64,
    pub max_refine: u32,
    pub m2l_inner: u32,
    pub m2l_outer: u32,
    pub strict: u32,
    pub dense_limit: u32,
    pub rhs_first: u32,
    pub rhs_count: u32,
    pub inner_passes: u32,
    pub inner_basis: u32,
    pub m2l_rank: u32,
    pub step_graphs: u32,
    pub m2l_compress: u32,
    pub fuse_leaf: u32,
    pub skip_converged: u32,
    pub batch_groups: u32,
    pub near_inner: u32,
    pub inner_order: u32,
    pub ground: u32,
    pub ground_z: f64,
    pub m2l_tile_order: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct gp_cap_report {
    pub n_panels: u32,
    pub n_conductors: u32,
    pub method_used: u32,
    pub precision_used: u32,
    pub iterations: u32,
    pub matvecs: u32,
    pub refine_steps: u32,
    pub converged: u32,
    pub max_level: u32,
    pub n_leaves: u32,
    pub n_boxes: u32,
    pub fmm_order: u32,
    pub near_nnz: u64,
    pub m2l_pairs: u64,
    pub final_rel_residual: f64,
    pub seconds_total: f64,
    pub seconds_setup: f64,
    pub seconds_solve: f64,
    pub seconds_near: f64,
    pub seconds_far: f64,
    pub seconds_orth: f64,
    pub seconds_upload: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct gp_param_map {
    pub n_params: u32,
    pub param_off: *const u32,
    pub panel: *const u32,
    pub dverts: *const f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct gp_volume_group {
    pub net: u32,
    pub n_rects: u32,
    pub height_nm: f64,
    pub thickness_nm: f64,
    pub dielectric_k: f64,
    pub rects: *const i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct gp_panel_role {
    pub kind: u32,
    pub conductor: u32,
    pub eps_surrounding: f64,
    pub eps_out: f64,
    pub eps_in: f64,
    pub reference: [f64; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct gp_res_polygon {
    pub net: u32,
    pub n_vertices: u32,
    pub xy: *const i32,
    pub has_pex: u32,
    pub volume_conductor: u32,
    pub sheet_res_ohm_sq: f64,
    pub via_res_ohm: f64,
    pub thickness_nm: f64,
    pub height_nm: f64,
}

// ------------------------------------------------------------ functions
#[link(name = "gpurify")]
extern "C" {
    pub fn gp_ctx_create(out: *mut *mut gp_ctx, device_ordinal: u32) -> c_int;
    pub fn gp_ctx_destroy(ctx: *mut gp_ctx);
    pub fn gp_last_error(ctx: *const gp_ctx) -> *const c_char;
    pub fn gp_device_ready() -> c_int;
    pub fn gp_device_synchronize(ctx: *mut gp_ctx) -> c_int;
    pub fn gp_device_info(ctx: *mut gp_ctx, buf: *mut c_char, len: usize) -> c_int;
    pub fn gp_version() -> *const c_char;
    pub fn gp_cap_opts_default(opts: *mut gp_cap_opts);
    pub fn gp_cap_solve(ctx: *mut gp_ctx, panels: *const gp_panels, opts: *const gp_cap_opts,
                        c: *mut f64, report: *mut gp_cap_report) -> c_int;
    pub fn gp_cap_solve_batched(ctx: *mut gp_ctx, problems: *const gp_panels, n_problems: u32,
                                opts: *const gp_cap_opts, c: *mut f64,
                                reports: *mut gp_cap_report) -> c_int;
    pub fn gp_cap_jacobian(ctx: *mut gp_ctx, panels: *const gp_panels, opts: *const gp_cap_opts,
                           params: *const gp_param_map, c: *mut f64, j: *mut f64,
                           report: *mut gp_cap_report) -> c_int;
    pub fn gp_fmm_apply(ctx: *mut gp_ctx, panels: *const gp_panels, opts: *const gp_cap_opts,
                        nr: u32, precision: u32, x: *const f64, y: *mut f64) -> c_int;
    pub fn gp_m2l_operator(order: u32, o0: i32, o1: i32, o2: i32, compressed: u32, g: *mut f64,
                           rank: *mut u32) -> c_int;
    pub fn gp_m2l_factors(order: u32, o0: i32, o1: i32, o2: i32, u: *mut f64, vt: *mut f64,
                          rank: *mut u32) -> c_int;
    pub fn gp_cap_solve_dense_noalloc(ctx: *mut gp_ctx, panels: *const gp_panels, c: *mut f64,
                                      report: *mut gp_cap_report) -> c_int;
    pub fn gp_profile_enable(ctx: *mut gp_ctx, enable: c_int) -> c_int;
    pub fn gp_profile_report(ctx: *mut gp_ctx, json: *mut c_char, len: usize) -> c_int;
    pub fn gp_tune_force(ctx: *mut gp_ctx, variant: c_int) -> c_int;
    pub fn gp_tune_report(ctx: *mut gp_ctx, json: *mut c_char, len: usize) -> c_int;
    pub fn gp_mesh_groups(groups: *const gp_volume_group, n_groups: u32, xy_scale: f64,
                          refine: u32, max_panels: u32, n_panels: *mut u32, vertices: *mut f64,
                          conductor: *mut u32, panel_eps: *mut f64) -> c_int;
    pub fn gp_solve_capacitance(ctx: *mut gp_ctx, groups: *const gp_volume_group, n_groups: u32,
                                xy_scale: f64, refine: u32, max_panels: u32,
                                opts: *const gp_cap_opts, n_nets: *mut u32, nets_out: *mut u32,
                                cap_af_out: *mut f64, c_out: *mut f64,
                                report: *mut gp_cap_report) -> c_int;
    pub fn gp_cap_solve_dielectric(ctx: *mut gp_ctx, panels: *const gp_panels,
                                   roles: *const gp_panel_role, c: *mut f64,
                                   report: *mut gp_cap_report) -> c_int;
    pub fn gp_cap_solve_dielectric_fmm(ctx: *mut gp_ctx, panels: *const gp_panels,
                                       roles: *const gp_panel_role, opts: *const gp_cap_opts,
                                       c: *mut f64, report: *mut gp_cap_report) -> c_int;
    pub fn gp_qui_parse(text: *const c_char, text_len: usize, scale: f64, n_panels: *mut u32,
                        n_conductors: *mut u32, names_len: *mut usize, vertices: *mut f64,
                        conductor: *mut u32, n_vertices: *mut u8, names: *mut c_char) -> c_int;
    pub fn gp_solve_resistance(polys: *const gp_res_polygon, n_polys: u32, dbu_nm: f64,
                               n_nets: *mut u32, nets_out: *mut u32, r_ohm_out: *mut f64) -> c_int;
}

// ------------------------------------------------------------ safe layer

/// A failed call: the status code and the library's message.
#[derive(Clone, Debug, PartialEq, Eq)]
pub struct Error {
    pub code: c_int,
    pub message: String,
}

impl Error {
    /// No usable device, device memory exhausted or a HIP failure: nothing
    /// is wrong with the layout, so a checked caller runs its CPU path.
    /// Every other code is an answer the reference gives itself (bad input,
    /// grid or panel capacity, non-finite result, singular or empty system,
    /// non-convergence with `strict`).
    pub fn is_fallback(&self) -> bool {
        matches!(self.code, GP_ERR_NO_DEVICE | GP_ERR_OOM | GP_ERR_INTERNAL)
    }
}

impl std::fmt::Display for Error {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "libgpurify status {}: {}", self.code, self.message)
    }
}

impl std::error::Error for Error {}

fn message(ctx: *const gp_ctx) -> String {
    // SAFETY: gp_last_error never returns NULL; the text lives until the
    // next call on ctx (or on this thread for a null ctx).
    unsafe { CStr::from_ptr(gp_last_error(ctx)) }.to_string_lossy().into_owned()
}

fn check(ctx: *const gp_ctx, rc: c_int) -> Result<(), Error> {
    if rc == GP_OK {
        Ok(())
    } else {
        Err(Error { code: rc, message: message(ctx) })
    }
}

/// gpu_ready() of the backend crate for this backend: a gfx950 device is usable now.
pub fn ready() -> bool {
    unsafe { gp_device_ready() == 1 }
}

/// The reference's own solver constants (bridge.rs:397-404, cap/fmm_solver.rs:164).
pub fn default_opts() -> gp_cap_opts {
    let mut o = gp_cap_opts::default();
    unsafe { gp_cap_opts_default(&mut o) };
    o
}

/// One panel set borrowed for a call: 12 vertex coordinates per panel (a
/// triangle repeats its last vertex), metres, and the conductor of each panel.
#[derive(Clone, Copy, Debug)]
pub struct Panels<'a> {
    pub vertices: &'a [f64],
    pub conductor: &'a [u32],
    pub n_conductors: u32,
}

impl Panels<'_> {
    fn raw(&self) -> gp_panels {
        assert_eq!(self.vertices.len(), 12 * self.conductor.len(), "12 coordinates per panel");
        gp_panels {
            n_panels: self.conductor.len() as u32,
            n_conductors: self.n_conductors,
            vertices: self.vertices.as_ptr(),
            conductor: self.conductor.as_ptr(),
        }
    }
}

/// Per-net result of [`Context::solve_capacitance`] (ascending net order).
#[derive(Clone, Debug, Default)]
pub struct NetCapacitance {
    pub nets: Vec<u32>,
    pub cap_af: Vec<f64>,
    /// (n_nets + 1)^2 farads with permittivity applied, substrate last.
    pub c: Vec<f64>,
    pub report: gp_cap_report,
}

/// A library context on one device, destroyed on drop.  Calls on one
/// context are serialized by the library; use one context per thread for
/// concurrency.
pub struct Context {
    raw: *mut gp_ctx,
}

// SAFETY: the library locks a context for the duration of every call.
unsafe impl Send for Context {}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { gp_ctx_destroy(self.raw) };
    }
}

impl Context {
    pub fn new(device: u32) -> Result<Context, Error> {
        let mut raw = std::ptr::null_mut();
        let rc = unsafe { gp_ctx_create(&mut raw, device) };
        check(std::ptr::null(), rc)?;
        Ok(Context { raw })
    }

    /// cap::solve / cap::fmm_solver::solve: C (n_conductors^2, F, vacuum).
    pub fn cap_solve(&self, p: &Panels<'_>, opts: &gp_cap_opts)
                     -> Result<(Vec<f64>, gp_cap_report), Error> {
        let raw = p.raw();
        let nc = p.n_conductors as usize;
        let mut c = vec![0.0; nc * nc];
        let mut rep = gp_cap_report::default();
        let rc = unsafe { gp_cap_solve(self.raw, &raw, opts, c.as_mut_ptr(), &mut rep) };
        check(self.raw, rc)?;
        Ok((c, rep))
    }

    /// The per-net solves of one layout in one call: C of each problem.
    pub fn cap_solve_batched(&self, problems: &[Panels<'_>], opts: &gp_cap_opts)
                             -> Result<(Vec<Vec<f64>>, Vec<gp_cap_report>), Error> {
        let raws: Vec<gp_panels> = problems.iter().map(|p| p.raw()).collect();
        let sizes: Vec<usize> = problems.iter().map(|p| (p.n_conductors as usize).pow(2)).collect();
        let mut flat = vec![0.0; sizes.iter().sum()];
        let mut reps = vec![gp_cap_report::default(); problems.len()];
        let rc = unsafe {
            gp_cap_solve_batched(self.raw, raws.as_ptr(), raws.len() as u32, opts,
                                 flat.as_mut_ptr(), reps.as_mut_ptr())
        };
        check(self.raw, rc)?;
        let mut out = Vec::with_capacity(sizes.len());
        let mut off = 0;
        for s in sizes {
            out.push(flat[off..off + s].to_vec());
            off += s;
        }
        Ok((out, reps))
    }

    /// solve_capacitance (bridge.rs:307-443) from the (net, layer) volume
    /// groups in the reference's BTreeMap order.
    pub fn solve_capacitance(&self, groups: &[gp_volume_group], xy_scale: f64, refine: u32,
                             max_panels: u32, opts: &gp_cap_opts)
                             -> Result<NetCapacitance, Error> {
        let mut distinct: Vec<u32> = groups.iter().map(|g| g.net).collect();
        distinct.sort_unstable();
        distinct.dedup();
        let nn = distinct.len();
        let mut out = NetCapacitance {
            nets: vec![0; nn],
            cap_af: vec![0.0; nn],
            c: vec![0.0; (nn + 1) * (nn + 1)],
            report: gp_cap_report::default(),
        };
        let mut n_nets = 0u32;
        let rc = unsafe {
            gp_solve_capacitance(self.raw, groups.as_ptr(), groups.len() as u32, xy_scale, refine,
                                 max_panels, opts, &mut n_nets, out.nets.as_mut_ptr(),
                                 out.cap_af.as_mut_ptr(), out.c.as_mut_ptr(), &mut out.report)
        };
        check(self.raw, rc)?;
        out.nets.truncate(n_nets as usize);
        out.cap_af.truncate(n_nets as usize);
        Ok(out)
    }

    /// cap::solve_dielectric (cap/dielectric.rs:47-151): C with permittivities.
    pub fn cap_solve_dielectric(&self, p: &Panels<'_>, roles: &[gp_panel_role])
                                -> Result<Vec<f64>, Error> {
        let raw = p.raw();
        assert_eq!(roles.len(), p.conductor.len(), "one role per panel");
        let nc = p.n_conductors as usize;
        let mut c = vec![0.0; nc * nc];
        let rc = unsafe {
            gp_cap_solve_dielectric(self.raw, &raw, roles.as_ptr(), c.as_mut_ptr(),
                                    std::ptr::null_mut())
        };
        check(self.raw, rc)?;
        Ok(c)
    }

    /// C and J = dC/dtheta (n_params x n_conductors^2) for geometric parameters.
    pub fn cap_jacobian(&self, p: &Panels<'_>, params: &Params, opts: &gp_cap_opts)
                        -> Result<(Vec<f64>, Vec<f64>, gp_cap_report), Error> {
        let raw = p.raw();
        let nc = p.n_conductors as usize;
        let map = params.raw();
        let mut c = vec![0.0; nc * nc];
        let mut j = vec![0.0; params.len() * nc * nc];
        let mut rep = gp_cap_report::default();
        let rc = unsafe {
            gp_cap_jacobian(self.raw, &raw, opts, &map, c.as_mut_ptr(), j.as_mut_ptr(), &mut rep)
        };
        check(self.raw, rc)?;
        Ok((c, j, rep))
    }
}

/// Geometric parameters in the CSR layout of gp_param_map: parameter p moves
/// panels panel[off[p]..off[p+1]] with 12 vertex velocities each.
#[derive(Clone, Debug, Default)]
pub struct Params {
    pub off: Vec<u32>,
    pub panel: Vec<u32>,
    pub dverts: Vec<f64>,
}

impl Params {
    pub fn len(&self) -> usize {
        self.off.len().saturating_sub(1)
    }
    pub fn is_empty(&self) -> bool {
        self.len() == 0
    }
    fn raw(&self) -> gp_param_map {
        assert_eq!(self.dverts.len(), 12 * self.panel.len(), "12 velocities per moved panel");
        gp_param_map {
            n_params: self.len() as u32,
            param_off: self.off.as_ptr(),
            panel: self.panel.as_ptr(),
            dverts: self.dverts.as_ptr(),
        }
    }
}

/// The forward/backward pair of capacitance sensitivities, modelled on
/// torch.autograd.Function: forward returns C and keeps J; backward maps
/// dL/dC to dL/dtheta_p = sum_ij dL/dC_ij J[p][i][j].
pub struct CapFunction {
    jac: Vec<f64>,
    n_params: usize,
    nc2: usize,
}

impl CapFunction {
    pub fn forward(ctx: &Context, p: &Panels<'_>, params: &Params, opts: &gp_cap_opts)
                   -> Result<(Vec<f64>, CapFunction), Error> {
        let (c, jac, _) = ctx.cap_jacobian(p, params, opts)?;
        let nc2 = (p.n_conductors as usize).pow(2);
        Ok((c, CapFunction { jac, n_params: params.len(), nc2 }))
    }

    pub fn backward(&self, grad_c: &[f64]) -> Vec<f64> {
        assert_eq!(grad_c.len(), self.nc2);
        (0..self.n_params)
            .map(|p| {
                self.jac[p * self.nc2..(p + 1) * self.nc2]
                    .iter()
                    .zip(grad_c)
                    .map(|(a, b)| a * b)
                    .sum()
            })
            .collect()
    }
}

/// r_ohm per net of extract_quasistatic's resistance half (host only).
pub fn solve_resistance(polys: &[gp_res_polygon], dbu_nm: f64) -> Result<Vec<(u32, f64)>, Error> {
    let mut n = 0u32;
    let rc = unsafe {
        gp_solve_resistance(polys.as_ptr(), polys.len() as u32, dbu_nm, &mut n,
                            std::ptr::null_mut(), std::ptr::null_mut())
    };
    check(std::ptr::null(), rc)?;
    let mut nets = vec![0u32; n as usize];
    let mut r = vec![0.0f64; n as usize];
    let rc = unsafe {
        gp_solve_resistance(polys.as_ptr(), polys.len() as u32, dbu_nm, &mut n, nets.as_mut_ptr(),
                            r.as_mut_ptr())
    };
    check(std::ptr::null(), rc)?;
    Ok(nets.into_iter().zip(r).collect())
}
