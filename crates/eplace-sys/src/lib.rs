//! FFI to `libeplace.so` (the C ABI declared in `include/eplace.h`) and a
//! small safe layer over it.
//!
//! The library is the device half of ElectroPlace
//! (`crates/nextpnr/src/placer/electro_place/mod.rs:65-204`); the glue that
//! flattens a `Context` and binds the result lives in
//! `glue/electro_place_gpu.rs`, to be added to the nextpnr crate as
//! `placer/electro_place/gpu.rs`.  Layouts here mirror the header field for
//! field (the repository's tests compile a C program that checks every
//! offset of the same structs).
#![allow(non_camel_case_types)]

use std::ffi::CStr;
use std::os::raw::{c_char, c_int};
use std::sync::{Mutex, OnceLock};

pub const EP_OK: c_int = 0;
pub const EP_PENDING: c_int = 1;
pub const EP_ERR_INVALID_ARG: c_int = -1;
pub const EP_ERR_NO_BELS: c_int = -2;
pub const EP_ERR_PLACEMENT_FAILED: c_int = -3;
pub const EP_ERR_OOM: c_int = -4;
pub const EP_ERR_ADMISSION_REFUSED: c_int = -5;
pub const EP_ERR_DEVICE: c_int = -6;
pub const EP_ERR_NONFINITE: c_int = -7;
pub const EP_ERR_NOT_FOUND: c_int = -8;
pub const EP_ERR_CANCELLED: c_int = -9;

pub const EP_PREC_FP64: i32 = 0;
pub const EP_PREC_FAST: i32 = 1;
pub const EP_FIELD_REFERENCE: i32 = 0;
pub const EP_FIELD_IDXST: i32 = 1;

#[repr(C)]
pub struct ep_ctx {
    _opaque: [u8; 0],
}

#[repr(C)]
pub struct ep_design {
    pub n_mov: i32,
    pub n_fixed: i32,
    pub n_nets: i32,
    pub w: i32,
    pub h: i32,
    pub bx: i32,
    pub by: i32,
    pub reserved: i32,
    pub n_pins: i64,
    pub x0: *const f64,
    pub y0: *const f64,
    pub area: *const f32,
    pub fixed_x: *const f64,
    pub fixed_y: *const f64,
    pub net_off: *const u32,
    pub pin_cell: *const u32,
    pub pinw: *const f64,
    pub net_weight: *const f64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct ep_config {
    pub seed: u64,
    pub wl_coeff: f64,
    pub target_util: f64,
    pub timing_driven: i32,
    pub max_iters: i32,
    pub target_density: f64,
    pub timing_weight: f64,
    pub nesterov_step_size: f64,
    pub legalize_interval: i32,
    pub min_iters: i32,
    pub overflow_target: f64,
    pub density_eta: f64,
    pub grow_mult: f64,
    pub grow_threshold: f64,
    pub grow_add: f64,
    pub stop_every_iter: i32,
    pub log_hpwl: i32,
    pub check_nonfinite: i32,
    pub field_mode: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct ep_result {
    pub status: i32,
    pub iters: i32,
    pub converged: i32,
    pub precision: i32,
    pub overflow: f64,
    pub hpwl: f64,
    pub lambda: f64,
    pub t_submit_ms: f64,
    pub t_admit_ms: f64,
    pub t_done_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct ep_iter_rec {
    pub iter: i32,
    pub checked: i32,
    pub overflow: f64,
    pub lambda: f64,
    pub step_size_x: f64,
    pub step_size_y: f64,
    pub step_norm_x: f64,
    pub step_norm_y: f64,
    pub hpwl: f64,
    pub wl_norm: f64,
    pub den_norm: f64,
}

#[repr(C)]
pub struct ep_legal_problem {
    pub n_bels: i32,
    pub n_children: i32,
    pub bel_x: *const i32,
    pub bel_y: *const i32,
    pub bel_bucket: *const i32,
    pub bel_taken: *const u8,
    pub cell_bucket: *const i32,
    pub child_root: *const i32,
    pub child_dx: *const i32,
    pub child_dy: *const i32,
    pub child_bucket: *const i32,
    pub child_fixed: *const i32,
    pub net_has_driver: *const u8,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct ep_legal_result {
    pub best_iter: i32,
    pub legalizations: i32,
    pub best_hpwl: f64,
    pub final_displacement: f64,
}

#[link(name = "eplace")]
extern "C" {
    pub fn ep_create(out: *mut *mut ep_ctx, device: c_int, hbm_budget_bytes: u64) -> c_int;
    pub fn ep_destroy(ctx: *mut ep_ctx);
    pub fn ep_last_error(ctx: *mut ep_ctx) -> *const c_char;
    pub fn ep_config_default(cfg: *mut ep_config);
    pub fn ep_version() -> c_int;
    pub fn ep_design_upload(ctx: *mut ep_ctx, d: *const ep_design, precision: i32, id: *mut u32) -> c_int;
    pub fn ep_design_free(ctx: *mut ep_ctx, id: u32) -> c_int;
    pub fn ep_design_bytes(d: *const ep_design, precision: i32) -> u64;
    pub fn ep_place(ctx: *mut ep_ctx, id: u32, cfg: *const ep_config, res: *mut ep_result) -> c_int;
    pub fn ep_place_noalloc(ctx: *mut ep_ctx, id: u32, cfg: *const ep_config, res: *mut ep_result) -> c_int;
    pub fn ep_place_batch(ctx: *mut ep_ctx, ids: *const u32, n: u32, cfg: *const ep_config,
                          res: *mut ep_result) -> c_int;
    pub fn ep_positions(ctx: *mut ep_ctx, id: u32, x: *mut f64, y: *mut f64) -> c_int;
    pub fn ep_iter_log(ctx: *mut ep_ctx, id: u32, recs: *mut ep_iter_rec, cap: u32, n: *mut u32) -> c_int;
    pub fn ep_place_legalized(ctx: *mut ep_ctx, id: u32, cfg: *const ep_config,
                              lp: *const ep_legal_problem, res: *mut ep_result,
                              lres: *mut ep_legal_result, cell_bel: *mut i32, child_bel: *mut i32) -> c_int;
    pub fn ep_legalize_problem(lp: *const ep_legal_problem, n_cells: i32, x: *const f64, y: *const f64,
                               cell_bel: *mut i32, child_bel: *mut i32, displacement: *mut f64) -> c_int;
    pub fn ep_legal_hpwl(d: *const ep_design, lp: *const ep_legal_problem, cell_bel: *const i32,
                         child_bel: *const i32, hpwl: *mut f64) -> c_int;
}

/// A failed library call: its status and the context's message.
#[derive(Debug, Clone)]
pub struct Error {
    pub status: c_int,
    pub message: String,
}

/// The flattened problem, owning its arrays (`ep_design` + `ep_legal_problem`).
#[derive(Default, Clone)]
pub struct Flat {
    pub w: i32,
    pub h: i32,
    pub x0: Vec<f64>,
    pub y0: Vec<f64>,
    pub fixed_x: Vec<f64>,
    pub fixed_y: Vec<f64>,
    pub net_off: Vec<u32>,
    pub pin_cell: Vec<u32>,
    pub pinw: Vec<f64>,
    pub bel_x: Vec<i32>,
    pub bel_y: Vec<i32>,
    pub bel_bucket: Vec<i32>,
    pub bel_taken: Vec<u8>,
    pub cell_bucket: Vec<i32>,
    pub child_root: Vec<i32>,
    pub child_dx: Vec<i32>,
    pub child_dy: Vec<i32>,
    pub child_bucket: Vec<i32>,
    pub child_fixed: Vec<i32>,
    pub net_has_driver: Vec<u8>,
}

impl Flat {
    fn design(&self) -> ep_design {
        ep_design {
            n_mov: self.x0.len() as i32,
            n_fixed: self.fixed_x.len() as i32,
            n_nets: (self.net_off.len().max(1) - 1) as i32,
            w: self.w,
            h: self.h,
            bx: 0, // bins = tiles, as the reference (mod.rs:75-76)
            by: 0,
            reserved: 0,
            n_pins: self.pin_cell.len() as i64,
            x0: self.x0.as_ptr(),
            y0: self.y0.as_ptr(),
            area: std::ptr::null(),
            fixed_x: self.fixed_x.as_ptr(),
            fixed_y: self.fixed_y.as_ptr(),
            net_off: self.net_off.as_ptr(),
            pin_cell: self.pin_cell.as_ptr(),
            pinw: self.pinw.as_ptr(),
            net_weight: std::ptr::null(),
        }
    }

    fn legal(&self) -> ep_legal_problem {
        ep_legal_problem {
            n_bels: self.bel_x.len() as i32,
            n_children: self.child_root.len() as i32,
            bel_x: self.bel_x.as_ptr(),
            bel_y: self.bel_y.as_ptr(),
            bel_bucket: self.bel_bucket.as_ptr(),
            bel_taken: self.bel_taken.as_ptr(),
            cell_bucket: self.cell_bucket.as_ptr(),
            child_root: self.child_root.as_ptr(),
            child_dx: self.child_dx.as_ptr(),
            child_dy: self.child_dy.as_ptr(),
            child_bucket: self.child_bucket.as_ptr(),
            child_fixed: self.child_fixed.as_ptr(),
            net_has_driver: self.net_has_driver.as_ptr(),
        }
    }
}

/// What `place_legalized` returns: the BEL (index into `Flat::bel_*`) of
/// every movable cell and cluster child, and the run's figures.
#[derive(Debug, Clone, Default)]
pub struct Placed {
    pub cell_bel: Vec<i32>,
    pub child_bel: Vec<i32>,
    pub result: ep_result,
    pub legal: ep_legal_result,
}

/// One library context on one GPU, kept for the life of the process so its
/// transform matrices and pooled device memory are reused across calls.
pub struct Gpu {
    ctx: *mut ep_ctx,
}

// The library serialises every call on a context with its own lock.
unsafe impl Send for Gpu {}

impl Gpu {
    /// The process-wide context on device 0, created on first use.
    pub fn shared() -> Result<&'static Mutex<Gpu>, Error> {
        static GPU: OnceLock<Result<Mutex<Gpu>, Error>> = OnceLock::new();
        GPU.get_or_init(|| {
            let mut c = std::ptr::null_mut();
            let rc = unsafe { ep_create(&mut c, 0, 0) };
            if rc != EP_OK {
                return Err(Error { status: rc, message: last_error(std::ptr::null_mut()) });
            }
            Ok(Mutex::new(Gpu { ctx: c }))
        })
        .as_ref()
        .map_err(|e| e.clone())
    }

    fn check(&self, rc: c_int) -> Result<(), Error> {
        if rc == EP_OK {
            Ok(())
        } else {
            Err(Error { status: rc, message: last_error(self.ctx) })
        }
    }

    /// The reference defaults (`ElectroPlaceCfg::default` and mod.rs:31-43).
    pub fn default_config() -> ep_config {
        let mut c = std::mem::MaybeUninit::<ep_config>::zeroed();
        unsafe {
            ep_config_default(c.as_mut_ptr());
            c.assume_init()
        }
    }

    /// place_electro end to end on the device (ep_place_legalized): the
    /// design is uploaded, placed with the best legalized snapshot kept, and
    /// freed again.
    pub fn place_legalized(&mut self, flat: &Flat, cfg: &ep_config, precision: i32) -> Result<Placed, Error> {
        let d = flat.design();
        let lp = flat.legal();
        let mut id = 0u32;
        self.check(unsafe { ep_design_upload(self.ctx, &d, precision, &mut id) })?;
        let mut out = Placed {
            cell_bel: vec![-1; flat.x0.len()],
            child_bel: vec![-1; flat.child_root.len()],
            ..Default::default()
        };
        let rc = unsafe {
            ep_place_legalized(self.ctx, id, cfg, &lp, &mut out.result, &mut out.legal,
                               out.cell_bel.as_mut_ptr(), out.child_bel.as_mut_ptr())
        };
        let msg = if rc != EP_OK { last_error(self.ctx) } else { String::new() };
        unsafe { ep_design_free(self.ctx, id) };
        if rc != EP_OK {
            return Err(Error { status: rc, message: msg });
        }
        Ok(out)
    }
}

impl Drop for Gpu {
    fn drop(&mut self) {
        unsafe { ep_destroy(self.ctx) };
    }
}

fn last_error(ctx: *mut ep_ctx) -> String {
    unsafe {
        let p = ep_last_error(ctx);
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}
