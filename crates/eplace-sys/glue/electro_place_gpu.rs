//! GPU body of `place_electro`, to be added to the nextpnr crate as
//! `crates/nextpnr/src/placer/electro_place/gpu.rs` (with `pub mod gpu;` in
//! `electro_place/mod.rs` and `eplace-sys` as a dependency), then called
//! from `PlacerElectro::place` in place of `place_electro`:
//!
//! ```ignore
//! fn place(&self, ctx: &mut Context, cfg: &Self::Config) -> Result<(), PlacerError> {
//!     gpu::place_electro_gpu(ctx, cfg)
//! }
//! ```
//!
//! The selection string stays "electro" (`crates/npnr-python/src/lib.rs:293-302`),
//! so there is one code path.  What stays on the Rust side is what the
//! reference's own host code does: initial placement, flattening in
//! reference order (below), binding the result and `validate_all_placed`.
//! The device runs the loop of `mod.rs:104-197`; at every check iteration
//! the library legalizes with the reference's rule on the host, scores the
//! binding with `total_hpwl` and keeps the best snapshot, and the final
//! binding is legalize(best) (`mod.rs:176-199`, `common.rs:485-495`).

use eplace_sys as ffi;
use rustc_hash::FxHashMap;

use crate::common::PlaceStrength;
use crate::context::Context;
use crate::netlist::CellId;

use super::super::common::{collect_movable_cells, initial_placement, unbind_movable_cells, validate_all_placed};
use super::super::PlacerError;
use super::ElectroPlaceCfg;

/// The field mode: the reference's (default, bit-compatible with
/// `density.rs:155-163`) or the index-shifted field that spreads cells
/// (`NEXTPNR_ELECTRO_FIELD=idxst`).
fn field_mode() -> i32 {
    match std::env::var("NEXTPNR_ELECTRO_FIELD").as_deref() {
        Ok("idxst") => ffi::EP_FIELD_IDXST,
        _ => ffi::EP_FIELD_REFERENCE,
    }
}

/// Flattens the context in reference order:
/// * movable cells by solver index (`collect_movable_cells`, common.rs:155-174),
///   initial positions from their BELs or the grid centre (common.rs:179-199);
/// * every alive net in slot order, driver first then users
///   (common.rs:228-240); a pin on a movable cell names its solver index, a
///   pin on a placed non-movable cell a fixed cell at its BEL location (first
///   appearance order), a pin on an unplaced one is dropped (common.rs:268-291);
/// * pin weights as `compute_pin_weights` counts them, dropped pins in the
///   degree (common.rs:381-407);
/// * the BELs of every bucket the movable cells and cluster children use, in
///   bucket order (`bels_for_bucket`, context/buckets.rs:25-34), those bound
///   to cells the legalization never unbinds marked taken;
/// * each movable root's unlocked constraint children (common.rs:294-374);
///   a child with pins is the fixed cell of its pins.
/// Returns the problem, the BEL id of every problem BEL and the child cells.
pub(crate) fn flatten_for_gpu(
    ctx: &Context,
    cell_to_idx: &FxHashMap<CellId, usize>,
    idx_to_cell: &[CellId],
) -> Result<(ffi::Flat, Vec<crate::chipdb::BelId>, Vec<CellId>), PlacerError> {
    let w = ctx.chipdb().width();
    let h = ctx.chipdb().height();
    let mut f = ffi::Flat { w, h, ..Default::default() };
    for &cell_id in idx_to_cell {
        let cell = ctx.design.cell(cell_id);
        match cell.bel {
            Some(bel) => {
                let loc = ctx.bel(bel).loc();
                f.x0.push(loc.x as f64);
                f.y0.push(loc.y as f64);
            }
            None => {
                f.x0.push(w as f64 / 2.0);
                f.y0.push(h as f64 / 2.0);
            }
        }
    }
    // cluster children of the movable roots, in solver order
    let mut children: Vec<CellId> = Vec::new();
    for (i, &root) in idx_to_cell.iter().enumerate() {
        if let Some(cluster) = ctx.design.clusters.get(&root) {
            for &child_id in &cluster.constr_children {
                let child = ctx.design.cell(child_id);
                if !child.alive || child.bel_strength.is_locked() {
                    continue;
                }
                children.push(child_id);
                f.child_root.push(i as i32);
                f.child_dx.push(child.constr_x);
                f.child_dy.push(child.constr_y);
            }
        }
    }
    let child_set: FxHashMap<CellId, usize> = children.iter().enumerate().map(|(k, &c)| (c, k)).collect();
    // nets, fixed cells and pin weights
    let n = idx_to_cell.len();
    let mut fixed_of: FxHashMap<CellId, u32> = FxHashMap::default();
    f.pinw = vec![0.0; n];
    f.net_off.push(0);
    for (_, net) in ctx.design.iter_alive_nets() {
        let driver = net.driver();
        let users = net.users();
        let pins = driver.iter().map(|p| p.cell).chain(users.iter().map(|u| u.cell));
        for cell_id in pins {
            if let Some(&idx) = cell_to_idx.get(&cell_id) {
                f.pin_cell.push(idx as u32);
                continue;
            }
            if !cell_id.is_some() {
                continue; // an invalid user names no cell
            }
            match ctx.design.cell(cell_id).bel {
                Some(bel) => {
                    let next = fixed_of.len() as u32;
                    let k = *fixed_of.entry(cell_id).or_insert_with(|| {
                        let loc = ctx.bel(bel).loc();
                        f.fixed_x.push(loc.x as f64);
                        f.fixed_y.push(loc.y as f64);
                        next
                    });
                    f.pin_cell.push(n as u32 + k);
                }
                None if child_set.contains_key(&cell_id) => {
                    return Err(PlacerError::PlacementFailed(
                        "cluster child with pins is unplaced after initial placement".to_owned(),
                    ));
                }
                None => {} // unplaced non-movable pin: dropped
            }
        }
        f.net_off.push(f.pin_cell.len() as u32);
        f.net_has_driver.push(driver.is_some() as u8);
        let degree = driver.is_some() as usize + users.len();
        if degree >= 2 {
            let wgt = 1.0 / degree as f64;
            if let Some(dp) = driver {
                if let Some(&idx) = cell_to_idx.get(&dp.cell) {
                    f.pinw[idx] += wgt;
                }
            }
            for u in users {
                if let Some(&idx) = cell_to_idx.get(&u.cell) {
                    f.pinw[idx] += wgt;
                }
            }
        }
    }
    f.child_fixed = children.iter().map(|c| fixed_of.get(c).map_or(-1, |&k| k as i32)).collect();
    // BELs of the buckets in use, each bucket in its own order
    let mut bucket_of_type: FxHashMap<crate::common::IdString, i32> = FxHashMap::default();
    let mut bucket_by_name: FxHashMap<String, i32> = FxHashMap::default();
    let mut bels: Vec<crate::chipdb::BelId> = Vec::new();
    let mut unbind: FxHashMap<CellId, ()> = idx_to_cell.iter().map(|&c| (c, ())).collect();
    unbind.extend(children.iter().map(|&c| (c, ())));
    let types = idx_to_cell.iter().chain(children.iter()).map(|&c| ctx.design.cell(c).cell_type);
    let mut type_list: Vec<crate::common::IdString> = Vec::new();
    for t in types {
        if bucket_of_type.contains_key(&t) {
            continue;
        }
        let first = ctx.bels_for_bucket(t).next();
        let name = first.map(|b| b.bucket().to_owned()).unwrap_or_default();
        let next = bucket_by_name.len() as i32;
        let b = *bucket_by_name.entry(name).or_insert_with(|| {
            type_list.push(t);
            next
        });
        bucket_of_type.insert(t, b);
    }
    for (b, &t) in type_list.iter().enumerate() {
        for bel in ctx.bels_for_bucket(t) {
            let loc = bel.loc();
            f.bel_x.push(loc.x);
            f.bel_y.push(loc.y);
            f.bel_bucket.push(b as i32);
            let held = bel.bound_cell().is_some_and(|c| !unbind.contains_key(&c.id()));
            f.bel_taken.push(held as u8);
            bels.push(bel.id());
        }
    }
    f.cell_bucket = idx_to_cell.iter().map(|&c| bucket_of_type[&ctx.design.cell(c).cell_type]).collect();
    f.child_bucket = children.iter().map(|&c| bucket_of_type[&ctx.design.cell(c).cell_type]).collect();
    Ok((f, bels, children))
}

fn gpu_err(e: ffi::Error) -> PlacerError {
    match e.status {
        ffi::EP_ERR_NO_BELS => PlacerError::NoBelsAvailable(e.message),
        _ => PlacerError::PlacementFailed(format!("GPU placer status {}: {}", e.status, e.message)),
    }
}

/// `place_electro` (mod.rs:65-204) with the loop on the GPU.
pub fn place_electro_gpu(ctx: &mut Context, cfg: &ElectroPlaceCfg) -> Result<(), PlacerError> {
    ctx.reseed_rng(cfg.seed);
    initial_placement(ctx)?;
    ctx.populate_bel_buckets();
    let (cell_to_idx, idx_to_cell) = collect_movable_cells(ctx);
    if idx_to_cell.is_empty() {
        return Ok(());
    }
    let (flat, bels, children) = flatten_for_gpu(ctx, &cell_to_idx, &idx_to_cell)?;
    let mut ec = ffi::Gpu::default_config();
    ec.seed = cfg.seed;
    ec.wl_coeff = cfg.wl_coeff;
    ec.target_util = cfg.target_util;
    ec.timing_driven = cfg.timing_driven as i32;
    ec.max_iters = cfg.max_iters as i32;
    ec.target_density = cfg.target_density;
    ec.timing_weight = cfg.timing_weight;
    ec.nesterov_step_size = cfg.nesterov_step_size;
    ec.legalize_interval = cfg.legalize_interval as i32;
    ec.field_mode = field_mode();
    let placed = {
        let gpu = ffi::Gpu::shared().map_err(gpu_err)?;
        let mut gpu = gpu.lock().map_err(|_| PlacerError::PlacementFailed("GPU context poisoned".into()))?;
        gpu.place_legalized(&flat, &ec, ffi::EP_PREC_FAST).map_err(gpu_err)?
    };
    // bind the legalized best snapshot (the bind_bel calls of legalize_electro)
    unbind_movable_cells(ctx, &idx_to_cell);
    let mut ch = 0;
    for (i, &cell_id) in idx_to_cell.iter().enumerate() {
        let bel = bels[placed.cell_bel[i] as usize];
        if !ctx.bind_bel(bel, cell_id, PlaceStrength::Placer) {
            return Err(PlacerError::PlacementFailed("failed to bind a cell to its legalized BEL".into()));
        }
        while ch < children.len() && flat.child_root[ch] as usize == i {
            let cbel = bels[placed.child_bel[ch] as usize];
            if !ctx.bind_bel(cbel, children[ch], PlaceStrength::Placer) {
                return Err(PlacerError::PlacementFailed("failed to bind a cluster child".into()));
            }
            ch += 1;
        }
    }
    log::info!(
        "ElectroPlace (GPU): {} iterations, best legalized HPWL {:.0} at iteration {}",
        placed.result.iters, placed.legal.best_hpwl, placed.legal.best_iter
    );
    validate_all_placed(ctx)
}
