"""Host-side mirror of the reference's ElectroPlace interface.

Names and meanings follow the reference:
  * ``ElectroPlaceCfg``   crates/nextpnr/src/placer/electro_place/config.rs:5-40
  * ``PlacerError``       crates/nextpnr/src/placer/mod.rs:20-28
  * ``PlacerElectro``     crates/nextpnr/src/placer/electro_place/mod.rs:45-63
  * ``place_electro``     crates/nextpnr/src/placer/electro_place/mod.rs:65-204
  * ``Context.place``     crates/npnr-python/src/lib.rs:224-307 ("electro" arm :293-302)
  * ``Context.total_hpwl`` crates/nextpnr/src/metrics/wirelength.rs:46-51

The reference's ``Context`` is its whole chip database and netlist; here a
``Context`` holds the flattened SoA snapshot the GPU path consumes
(``FlatDesign``) plus the BEL grid its legalizer binds to.  All compute runs in
libeplace.so on a HIP device; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, fields
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


# ------------------------------------------------------------------ errors --

class PlacerError(RuntimeError):
    """crates/nextpnr/src/placer/mod.rs:20-28."""


class NoBelsAvailable(PlacerError):
    pass


class PlacementFailed(PlacerError):
    pass


class InitialPlacementFailed(PlacerError):
    pass


class DeviceError(PlacerError):
    pass


class NonFiniteError(PlacerError):
    pass


class AdmissionRefused(PlacerError):
    pass


def _check(rc: int, ctx=None, what: str = "") -> int:
    if rc >= 0:
        return rc
    msg = L.lib().ep_last_error(ctx).decode() if ctx is not None else ""
    text = f"{what}: {msg}" if what else msg
    if rc == L.EP_ERR_NO_BELS:
        raise NoBelsAvailable(text)
    if rc == L.EP_ERR_PLACEMENT_FAILED:
        raise PlacementFailed(text)
    if rc == L.EP_ERR_NONFINITE:
        raise NonFiniteError(text)
    if rc == L.EP_ERR_ADMISSION_REFUSED:
        raise AdmissionRefused(text)
    if rc == L.EP_ERR_INVALID_ARG:
        raise ValueError(text)
    if rc == L.EP_ERR_NOT_FOUND:
        raise KeyError(text)
    if rc in (L.EP_ERR_DEVICE, L.EP_ERR_OOM):
        raise DeviceError(text)
    raise PlacerError(f"status {rc}: {text}")


# ------------------------------------------------------------------ config --

@dataclass
class ElectroPlaceCfg:
    """ElectroPlaceCfg with the reference defaults (config.rs:26-39), the
    constants of mod.rs:31-43, and the GPU path's knobs."""

    seed: int = 1
    wl_coeff: float = 0.5
    target_util: float = 0.7
    timing_driven: bool = False
    target_density: float = 1.0
    timing_weight: float = 0.0
    nesterov_step_size: float = 0.1
    max_iters: int = 500
    legalize_interval: int = 5
    # mod.rs:31-43
    min_iters: int = 20
    overflow_target: float = 0.1
    density_eta: float = 0.1
    grow_mult: float = 1.025
    grow_threshold: float = 50.0
    grow_add: float = 1.0
    # GPU path
    stop_every_iter: bool = False
    log_hpwl: bool = True
    check_nonfinite: bool = True
    field_mode: str = "reference"   # or "idxst"

    def to_c(self) -> L.EpConfig:
        c = L.EpConfig()
        L.lib().ep_config_default(C.byref(c))
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "field_mode":
                v = {"reference": L.EP_FIELD_REFERENCE, "idxst": L.EP_FIELD_IDXST}[v]
            elif isinstance(v, bool):
                v = int(v)
            setattr(c, f.name, v)
        return c


# ------------------------------------------------------------------ design --

@dataclass
class FlatDesign:
    """Flattened placement problem (include/eplace.h ep_design)."""

    W: int
    H: int
    x0: np.ndarray
    y0: np.ndarray
    fixed_x: np.ndarray
    fixed_y: np.ndarray
    net_off: np.ndarray
    pin_cell: np.ndarray
    area: Optional[np.ndarray] = None
    pinw: Optional[np.ndarray] = None
    Bx: int = 0
    By: int = 0
    # legalization view: BEL grid and cell buckets (optional)
    cell_type: Optional[np.ndarray] = None
    bel_x: Optional[np.ndarray] = None
    bel_y: Optional[np.ndarray] = None
    bel_type: Optional[np.ndarray] = None
    init_bel: Optional[np.ndarray] = None
    name: str = ""
    # optional per-net weights of the WA gradient (common.rs:204-261)
    net_weight: Optional[np.ndarray] = None

    def __post_init__(self):
        self.x0 = np.ascontiguousarray(self.x0, dtype=np.float64)
        self.y0 = np.ascontiguousarray(self.y0, dtype=np.float64)
        self.fixed_x = np.ascontiguousarray(self.fixed_x, dtype=np.float64)
        self.fixed_y = np.ascontiguousarray(self.fixed_y, dtype=np.float64)
        self.net_off = np.ascontiguousarray(self.net_off, dtype=np.uint32)
        self.pin_cell = np.ascontiguousarray(self.pin_cell, dtype=np.uint32)
        if self.area is not None:
            self.area = np.ascontiguousarray(self.area, dtype=np.float32)
        if self.pinw is not None:
            self.pinw = np.ascontiguousarray(self.pinw, dtype=np.float64)
        if self.net_weight is not None:
            self.net_weight = np.ascontiguousarray(self.net_weight, dtype=np.float64)
        if not self.Bx:
            self.Bx = self.W
        if not self.By:
            self.By = self.H

    @property
    def n_mov(self) -> int:
        return int(self.x0.shape[0])

    @property
    def n_fixed(self) -> int:
        return int(self.fixed_x.shape[0])

    @property
    def n_nets(self) -> int:
        return int(self.net_off.shape[0]) - 1

    @property
    def n_pins(self) -> int:
        return int(self.pin_cell.shape[0])

    def to_c(self) -> L.EpDesign:
        d = L.EpDesign()
        d.n_mov, d.n_fixed, d.n_nets = self.n_mov, self.n_fixed, self.n_nets
        d.W, d.H, d.Bx, d.By = self.W, self.H, self.Bx, self.By
        d.n_pins = self.n_pins
        d.x0, d.y0 = L.ptr(self.x0), L.ptr(self.y0)
        d.area = L.ptr(self.area)
        d.fixed_x, d.fixed_y = L.ptr(self.fixed_x), L.ptr(self.fixed_y)
        d.net_off, d.pin_cell = L.ptr(self.net_off), L.ptr(self.pin_cell)
        d.pinw = L.ptr(self.pinw)
        d.net_weight = L.ptr(self.net_weight)
        return d

    def nbytes(self, precision: str = "fast") -> int:
        return int(L.lib().ep_design_bytes(C.byref(self.to_c()), _prec(precision)))


@dataclass
class LegalProblem:
    """Legalization view of a design (include/eplace.h ep_legal_problem):
    BELs with their buckets (bucket order = BEL index order,
    context/buckets.rs:25-43), the bucket of each movable cell, BELs bound by
    cells the placer never unbinds, and the cluster children placed with
    their roots (common.rs:319-374).  A child with pins is the fixed cell
    `child_fixed` of the design (-1 without pins)."""

    bel_x: np.ndarray
    bel_y: np.ndarray
    bel_bucket: np.ndarray
    cell_bucket: np.ndarray
    bel_taken: Optional[np.ndarray] = None
    child_root: Optional[np.ndarray] = None
    child_dx: Optional[np.ndarray] = None
    child_dy: Optional[np.ndarray] = None
    child_bucket: Optional[np.ndarray] = None
    child_fixed: Optional[np.ndarray] = None
    net_has_driver: Optional[np.ndarray] = None

    def __post_init__(self):
        i32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.int32)
        u8 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.uint8)
        self.bel_x, self.bel_y = i32(self.bel_x), i32(self.bel_y)
        self.bel_bucket, self.cell_bucket = i32(self.bel_bucket), i32(self.cell_bucket)
        self.bel_taken, self.net_has_driver = u8(self.bel_taken), u8(self.net_has_driver)
        n = 0 if self.child_root is None else len(self.child_root)
        z = np.zeros(n, np.int32)
        self.child_root = i32(self.child_root) if n else z
        self.child_dx = i32(self.child_dx) if n else z
        self.child_dy = i32(self.child_dy) if n else z
        self.child_bucket = i32(self.child_bucket) if n else z
        self.child_fixed = i32(self.child_fixed) if n else z

    @property
    def n_children(self) -> int:
        return int(self.child_root.shape[0])

    def to_c(self) -> L.EpLegalProblem:
        c = L.EpLegalProblem()
        c.n_bels, c.n_children = int(self.bel_x.shape[0]), self.n_children
        c.bel_x, c.bel_y, c.bel_bucket = L.ptr(self.bel_x), L.ptr(self.bel_y), L.ptr(self.bel_bucket)
        c.bel_taken, c.cell_bucket = L.ptr(self.bel_taken), L.ptr(self.cell_bucket)
        c.child_root, c.child_dx, c.child_dy = L.ptr(self.child_root), L.ptr(self.child_dx), L.ptr(self.child_dy)
        c.child_bucket, c.child_fixed = L.ptr(self.child_bucket), L.ptr(self.child_fixed)
        c.net_has_driver = L.ptr(self.net_has_driver)
        return c

    @staticmethod
    def of(d: "FlatDesign") -> "LegalProblem":
        """The generator's BEL grid of a design (no clusters)."""
        if d.bel_x is None:
            raise PlacementFailed("design has no BEL grid to legalize onto")
        return LegalProblem(d.bel_x, d.bel_y, d.bel_type, d.cell_type)


def legalize_problem(lp: LegalProblem, x, y):
    """legalize_electro with cluster children (ep_legalize_problem).
    Returns (cell_bel, child_bel, displacement)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    cb = np.full(x.shape[0], -1, np.int32)
    kb = np.full(lp.n_children, -1, np.int32)
    disp = C.c_double()
    _check(L.lib().ep_legalize_problem(C.byref(lp.to_c()), x.shape[0], L.ptr(x), L.ptr(y), L.ptr(cb),
                                       L.ptr(kb), C.byref(disp)), None, "ep_legalize_problem")
    return cb, kb, disp.value


def legal_hpwl(d: "FlatDesign", lp: LegalProblem, cell_bel, child_bel=None) -> float:
    """total_hpwl (metrics/wirelength.rs:10-51) of a legalized binding (ep_legal_hpwl)."""
    cb = np.ascontiguousarray(cell_bel, np.int32)
    kb = np.ascontiguousarray(child_bel if child_bel is not None else np.zeros(0), np.int32)
    h = C.c_double()
    _check(L.lib().ep_legal_hpwl(C.byref(d.to_c()), C.byref(lp.to_c()), L.ptr(cb), L.ptr(kb), C.byref(h)),
           None, "ep_legal_hpwl")
    return h.value


def _prec(p: str) -> int:
    return {"fp64": L.EP_PREC_FP64, "fast": L.EP_PREC_FAST}[p]


def generate(n_cells: int, n_nets: int, W: int, H: Optional[int] = None, *, n_fixed: int = 0,
             Bx: int = 0, By: int = 0, n_clusters: int = 32, area_mix: bool = False,
             n_types: int = 1, type1_frac: float = 0.0, n_high_fanout: int = 0,
             hf_min: int = 500, hf_max: int = 5000, locality: float = 0.8, seed: int = 0,
             name: str = "") -> FlatDesign:
    """Synthetic generic-FPGA design (the library's seeded generator)."""
    H = W if H is None else H
    p = L.EpGenParams(seed, n_cells, n_fixed, n_nets, W, H, Bx, By, n_clusters, int(area_mix),
                      n_types, n_high_fanout, hf_min, hf_max, locality, type1_frac)
    g = C.POINTER(L.EpGenDesign)()
    _check(L.lib().ep_generate(C.byref(p), C.byref(g)), None, "ep_generate")
    try:
        G = g.contents
        d = G.d

        def arr(ptr, n, ct, dt):
            if n == 0 or not ptr:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).astype(dt).copy()

        nm = d.n_mov
        fd = FlatDesign(
            W=d.W, H=d.H, Bx=d.Bx, By=d.By,
            x0=arr(d.x0, nm, C.c_double, np.float64), y0=arr(d.y0, nm, C.c_double, np.float64),
            fixed_x=arr(d.fixed_x, d.n_fixed, C.c_double, np.float64),
            fixed_y=arr(d.fixed_y, d.n_fixed, C.c_double, np.float64),
            net_off=arr(d.net_off, d.n_nets + 1, C.c_uint32, np.uint32),
            pin_cell=arr(d.pin_cell, d.n_pins, C.c_uint32, np.uint32),
            area=arr(d.area, nm, C.c_float, np.float32) if d.area else None,
            cell_type=arr(G.cell_type, nm, C.c_int32, np.int32),
            bel_x=arr(G.bel_x, G.n_bels, C.c_int32, np.int32),
            bel_y=arr(G.bel_y, G.n_bels, C.c_int32, np.int32),
            bel_type=arr(G.bel_type, G.n_bels, C.c_int32, np.int32),
            init_bel=arr(G.init_bel, nm, C.c_int32, np.int32),
            name=name,
        )
    finally:
        L.lib().ep_gen_free(g)
    return fd


# -------------------------------------------------------------- workloads --

def config1(seed: int = 0) -> FlatDesign:
    """1,500 movable cells (1,200 LUT4-like, 300 DFF-like), 64 IO on the
    perimeter, 1,700 nets, 64x64 tiles = 64x64 bins, 32 clusters."""
    return generate(1500, 1700, 64, n_fixed=64, n_clusters=32, n_types=2, type1_frac=0.2,
                    seed=seed, name="config1")


def config2(seed: int = 0) -> FlatDesign:
    """200,000 cells (areas 1/2/4), 220,000 nets, 1,024 clusters, 800 IO,
    600x600 tiles, 512x512 bins."""
    return generate(200_000, 220_000, 600, n_fixed=800, Bx=512, By=512, n_clusters=1024,
                    area_mix=True, seed=seed, name="config2")


def config3(seed: int = 0) -> FlatDesign:
    """1,000,000 cells, 1,100,000 nets plus 0.1 % (1,100) nets of 500-5,000
    pins, 1,600 IO, 1,200x1,200 tiles, 1,024x1,024 bins."""
    return generate(1_000_000, 1_100_000, 1200, n_fixed=1600, Bx=1024, By=1024,
                    n_clusters=1024, area_mix=True, n_high_fanout=1100, hf_min=500,
                    hf_max=5000, seed=seed, name="config3")


def bins_for_side(side: int) -> int:
    """Nearest power of two (log2 rounding) to the grid side, in [64, 512]."""
    b = 2 ** int(round(math.log2(max(side, 1))))
    return int(min(512, max(64, b)))


def with_spacers(d: FlatDesign, target_util: float = 0.7, spacer_area: float = 1.0,
                 seed: int = 0):
    """Spacer (filler) cells up to `target_util` of the grid, the reference
    design spec's section 1f (docs/superpowers/specs/2026-03-13-placer-
    optimization-design.md:91-97), which `ElectroPlaceCfg.target_util`
    (config.rs) names but the shipped placer never inserts:

        spacer_count = (grid area * target_util - movable area) / spacer_area

    Spacers are movable cells of `spacer_area` without pins (they take part in
    the density, not in the wirelength), appended after the design's movable
    cells at seeded uniform positions; the fixed cells' pin indices move up by
    their number.  The first `d.n_mov` positions of a placement are the
    design's own cells.  Returns (design with spacers, spacer count)."""
    area = d.area.astype(np.float64) if d.area is not None else np.ones(d.n_mov)
    n_sp = int(math.floor((d.W * d.H * float(target_util) - float(area.sum())) / float(spacer_area)))
    if n_sp <= 0:
        return d, 0
    rng = np.random.default_rng(seed)
    pins = d.pin_cell.astype(np.int64)
    pins = np.where(pins >= d.n_mov, pins + n_sp, pins).astype(np.uint32)
    # a spacer's footprint is its area's square; keep it inside the grid
    half = 0.5 * math.sqrt(spacer_area)
    sx = rng.uniform(half, max(half, d.W - half), n_sp)
    sy = rng.uniform(half, max(half, d.H - half), n_sp)
    out = FlatDesign(
        W=d.W, H=d.H, Bx=d.Bx, By=d.By,
        x0=np.concatenate([d.x0, sx]), y0=np.concatenate([d.y0, sy]),
        fixed_x=d.fixed_x, fixed_y=d.fixed_y, net_off=d.net_off, pin_cell=pins,
        area=np.concatenate([area, np.full(n_sp, spacer_area)]).astype(np.float32),
        pinw=None if d.pinw is None else np.concatenate([d.pinw, np.zeros(n_sp)]),
        net_weight=d.net_weight, name=d.name + "+spacers")
    return out, n_sp


def timing_net_weights(crit, timing_weight: float = 5.0) -> np.ndarray:
    """Net weights of timing-driven placement, 1 + timing_weight * crit^2 per
    net (the reference design spec's section 1i,
    docs/superpowers/specs/2026-03-13-placer-optimization-design.md:109-113;
    `ElectroPlaceCfg.timing_driven` / `timing_weight` in config.rs name it,
    the shipped placer never applies it).  `crit` holds each net's
    criticality in [0, 1] from the caller's timing analysis; the result is
    `FlatDesign.net_weight`, which the WA gradient multiplies into every pin
    of the net (common.rs:246-261)."""
    c = np.asarray(crit, dtype=np.float64)
    if c.size and (not np.all(np.isfinite(c)) or c.min() < 0.0 or c.max() > 1.0):
        raise ValueError("criticalities must lie in [0, 1]")
    return 1.0 + float(timing_weight) * c * c


def sized_design(n_cells: int, seed: int, name: str = "") -> FlatDesign:
    """Batch/serving rule: side = round(sqrt(N / 0.7)), nets = 1.1 N,
    IO = perimeter / 3, clusters of ~200 cells, areas 1/2/4."""
    side = int(round(math.sqrt(n_cells / 0.7)))
    B = bins_for_side(side)
    return generate(n_cells, int(round(1.1 * n_cells)), side, n_fixed=max(16, (4 * side) // 3),
                    Bx=B, By=B, n_clusters=max(1, n_cells // 200), area_mix=True, seed=seed,
                    name=name or f"n{n_cells}_s{seed}")


def batch_sizes(n: int = 64, lo: int = 2_000, hi: int = 100_000, seed: int = 0) -> list:
    """Log-uniform design sizes of the 64-design batch."""
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(v))) for v in rng.uniform(math.log(lo), math.log(hi), n)]


def batch_designs(n: int = 64, lo: int = 2_000, hi: int = 100_000, seed: int = 0) -> list:
    return [sized_design(s, seed + i) for i, s in enumerate(batch_sizes(n, lo, hi, seed))]


# ------------------------------------------------------------------ device --

@dataclass
class PlaceResult:
    status: int
    iters: int
    converged: bool
    overflow: float
    hpwl: float
    lam: float
    x: Optional[np.ndarray] = None
    y: Optional[np.ndarray] = None
    log: list = field(default_factory=list)
    t_submit_ms: float = 0.0
    t_admit_ms: float = 0.0
    t_done_ms: float = 0.0


def _result(r: L.EpResult) -> PlaceResult:
    return PlaceResult(r.status, r.iters, bool(r.converged), r.overflow, r.hpwl, r.lambda_,
                       t_submit_ms=r.t_submit_ms, t_admit_ms=r.t_admit_ms, t_done_ms=r.t_done_ms)


class Device:
    """One placer context on one HIP device (ep_ctx)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._h = C.c_void_p()
        _check(L.lib().ep_create(C.byref(self._h), device, hbm_budget_bytes), None, "ep_create")
        self.designs: dict = {}

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            L.lib().ep_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, d: FlatDesign, precision: str = "fast") -> int:
        i = C.c_uint32()
        _check(L.lib().ep_design_upload(self._h, C.byref(d.to_c()), _prec(precision), C.byref(i)),
               self._h, "ep_design_upload")
        self.designs[i.value] = d
        return i.value

    def free(self, design_id: int):
        _check(L.lib().ep_design_free(self._h, design_id), self._h, "ep_design_free")
        self.designs.pop(design_id, None)

    def place(self, design_id: int, cfg: ElectroPlaceCfg, positions: bool = True,
              log: bool = False) -> PlaceResult:
        return self.place_batch([design_id], cfg, positions, log)[0]

    def place_batch(self, ids: Sequence[int], cfg: ElectroPlaceCfg, positions: bool = True,
                    log: bool = False) -> list:
        n = len(ids)
        arr = (C.c_uint32 * n)(*ids)
        res = (L.EpResult * n)()
        rc = L.lib().ep_place_batch(self._h, arr, n, C.byref(cfg.to_c()), res)
        if rc not in (L.EP_OK, L.EP_ERR_NONFINITE):
            _check(rc, self._h, "ep_place_batch")
        out = []
        for k in range(n):
            r = _result(res[k])
            if positions:
                r.x, r.y = self.positions(ids[k])
            if log:
                r.log = self.iter_log(ids[k])
            out.append(r)
        if rc == L.EP_ERR_NONFINITE:
            _check(rc, self._h, "ep_place_batch")
        return out

    def place_noalloc(self, design_id: int, cfg: ElectroPlaceCfg, positions: bool = True,
                      log: bool = False) -> PlaceResult:
        """ep_place_noalloc: the same placement as place(), with no device
        allocation (launch descriptors in the library's preallocated buffer,
        eager launches instead of a graph)."""
        r = L.EpResult()
        rc = L.lib().ep_place_noalloc(self._h, design_id, C.byref(cfg.to_c()), C.byref(r))
        if rc not in (L.EP_OK, L.EP_ERR_NONFINITE):
            _check(rc, self._h, "ep_place_noalloc")
        out = _result(r)
        if positions:
            out.x, out.y = self.positions(design_id)
        if log:
            out.log = self.iter_log(design_id)
        if rc == L.EP_ERR_NONFINITE:
            _check(rc, self._h, "ep_place_noalloc")
        return out

    def place_legalized(self, design_id: int, cfg: ElectroPlaceCfg, lp: LegalProblem,
                        log: bool = False):
        """place_electro end to end (ep_place_legalized): the continuous loop on
        the device, a legalized and scored snapshot at every check iteration,
        and the final binding legalize(best).  Returns (result, cell_bel,
        child_bel, info) with info = {best_iter, legalizations, best_hpwl,
        final_displacement}."""
        n = self.designs[design_id].n_mov
        cb = np.full(n, -1, np.int32)
        kb = np.full(lp.n_children, -1, np.int32)
        r = L.EpResult()
        lr = L.EpLegalResult()
        _check(L.lib().ep_place_legalized(self._h, design_id, C.byref(cfg.to_c()), C.byref(lp.to_c()),
                                          C.byref(r), C.byref(lr), L.ptr(cb), L.ptr(kb)),
               self._h, "ep_place_legalized")
        out = _result(r)
        out.x, out.y = self.positions(design_id)
        if log:
            out.log = self.iter_log(design_id)
        info = {f[0]: getattr(lr, f[0]) for f in L.EpLegalResult._fields_}
        return out, cb, kb, info

    def positions(self, design_id: int):
        n = self.designs[design_id].n_mov
        x = np.zeros(n)
        y = np.zeros(n)
        _check(L.lib().ep_positions(self._h, design_id, L.ptr(x), L.ptr(y)), self._h, "ep_positions")
        return x, y

    def iter_log(self, design_id: int) -> list:
        cap = 4096
        recs = (L.EpIterRec * cap)()
        n = C.c_uint32()
        _check(L.lib().ep_iter_log(self._h, design_id, recs, cap, C.byref(n)), self._h, "ep_iter_log")
        keys = [f[0] for f in L.EpIterRec._fields_]
        return [{k.rstrip("_"): getattr(recs[i], k) for k in keys} for i in range(min(n.value, cap))]

    def profile(self, ids: Sequence[int], cfg: ElectroPlaceCfg, warm: int = 8, reps: int = 5) -> dict:
        """Average HIP-event time (ms) of each phase of one batched iteration."""
        n = len(ids)
        arr = (C.c_uint32 * n)(*ids)
        out = (L.EpPhaseTime * 32)()
        k = C.c_uint32()
        _check(L.lib().ep_profile_iteration(self._h, arr, n, C.byref(cfg.to_c()), warm, reps, out, 32,
                                            C.byref(k)), self._h, "ep_profile_iteration")
        return {out[i].name.decode(): out[i].avg_ms for i in range(k.value)}

    ENGINES = {0: "mfma-k32", 1: "fma", 2: "mfma-splitk", 3: "mfma-k64", 4: "mfma-kslices", 5: "fft"}

    def autotune(self, ids: Sequence[int], reps: int = 5, cache_path: Optional[str] = None) -> dict:
        """Pick the fastest transform engine per bin shape (ep_autotune);
        returns the tuning table {"BxxBy": engine name}."""
        n = len(ids)
        arr = (C.c_uint32 * n)(*ids)
        path = cache_path.encode() if cache_path else None
        _check(L.lib().ep_autotune(self._h, arr, n, reps, path), self._h, "ep_autotune")
        return self.tune_table()

    def tune_table(self) -> dict:
        out = (C.c_int32 * 3 * 64)()
        k = C.c_uint32()
        _check(L.lib().ep_tune_table(self._h, C.cast(out, C.POINTER(C.c_int32)), 64, C.byref(k)),
               self._h, "ep_tune_table")
        flat = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_int32)), shape=(64 * 3,))
        return {f"{flat[3 * i]}x{flat[3 * i + 1]}": self.ENGINES.get(int(flat[3 * i + 2]), "?")
                for i in range(min(k.value, 64))}

    def tune_set(self, bx: int, by: int, engine: int):
        """Pin the transform engine of one bin shape (ep_tune_set; < 0 clears)."""
        _check(L.lib().ep_tune_set(self._h, bx, by, engine), self._h, "ep_tune_set")

    # -- single-kernel evaluation --
    def eval_config(self, cfg: ElectroPlaceCfg):
        _check(L.lib().ep_eval_config(self._h, C.byref(cfg.to_c())), self._h, "ep_eval_config")

    def wa_grad(self, design_id, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        gx, gy = np.zeros_like(x), np.zeros_like(y)
        _check(L.lib().ep_eval_wa_grad(self._h, design_id, L.ptr(x), L.ptr(y), L.ptr(gx), L.ptr(gy)),
               self._h, "ep_eval_wa_grad")
        return gx, gy

    def density(self, design_id, x, y):
        d = self.designs[design_id]
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        dens = np.zeros(d.Bx * d.By)
        ov = C.c_double()
        _check(L.lib().ep_eval_density(self._h, design_id, L.ptr(x), L.ptr(y), L.ptr(dens),
                                       C.byref(ov)), self._h, "ep_eval_density")
        return dens.reshape(d.By, d.Bx), ov.value

    def bin_ranges(self, design_id, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        r = np.zeros((x.shape[0], 4), np.int32)
        _check(L.lib().ep_eval_bin_ranges(self._h, design_id, L.ptr(x), L.ptr(y), L.ptr(r)),
               self._h, "ep_eval_bin_ranges")
        return r

    def field(self, design_id, dens):
        d = self.designs[design_id]
        dens = np.ascontiguousarray(dens, np.float64).reshape(-1)
        ex = np.zeros(d.Bx * d.By)
        ey = np.zeros(d.Bx * d.By)
        _check(L.lib().ep_eval_field(self._h, design_id, L.ptr(dens), L.ptr(ex), L.ptr(ey)),
               self._h, "ep_eval_field")
        return ex.reshape(d.By, d.Bx), ey.reshape(d.By, d.Bx)

    def density_grad(self, design_id, x, y, ex, ey):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        ex = np.ascontiguousarray(ex, np.float64).reshape(-1)
        ey = np.ascontiguousarray(ey, np.float64).reshape(-1)
        gx, gy = np.zeros_like(x), np.zeros_like(y)
        _check(L.lib().ep_eval_density_grad(self._h, design_id, L.ptr(x), L.ptr(y), L.ptr(ex),
                                            L.ptr(ey), L.ptr(gx), L.ptr(gy)),
               self._h, "ep_eval_density_grad")
        return gx, gy

    def hpwl(self, design_id, x, y) -> float:
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        h = C.c_double()
        _check(L.lib().ep_eval_hpwl(self._h, design_id, L.ptr(x), L.ptr(y), C.byref(h)),
               self._h, "ep_eval_hpwl")
        return h.value


# ---------------------------------------------------------------- serving --

class Server:
    """Continuous-batching placement service on one Device (ep_serve_*)."""

    def __init__(self, dev: Device, max_slots: int = 128, iters_per_sync: int = 5,
                 hbm_budget_bytes: int = 0):
        self.dev = dev
        s = L.EpServeConfig(max_slots, iters_per_sync, hbm_budget_bytes)
        _check(L.lib().ep_serve_start(dev.handle, C.byref(s)), dev.handle, "ep_serve_start")
        self.sizes: dict = {}

    def submit(self, d: FlatDesign, cfg: ElectroPlaceCfg) -> int:
        t = C.c_uint64()
        _check(L.lib().ep_serve_submit(self.dev.handle, C.byref(d.to_c()), C.byref(cfg.to_c()),
                                       C.byref(t)), self.dev.handle, "ep_serve_submit")
        self.sizes[t.value] = d.n_mov
        return t.value

    def poll(self, ticket: int):
        r = L.EpResult()
        rc = L.lib().ep_serve_poll(self.dev.handle, ticket, C.byref(r))
        if rc == L.EP_PENDING:
            return None
        res = _result(r)
        res.status = rc
        return res

    def positions(self, ticket: int):
        n = self.sizes[ticket]
        x, y = np.zeros(n), np.zeros(n)
        _check(L.lib().ep_serve_positions(self.dev.handle, ticket, L.ptr(x), L.ptr(y)),
               self.dev.handle, "ep_serve_positions")
        return x, y

    def release(self, ticket: int):
        _check(L.lib().ep_serve_release(self.dev.handle, ticket), self.dev.handle, "ep_serve_release")
        self.sizes.pop(ticket, None)

    def cancel(self, ticket: int):
        _check(L.lib().ep_serve_cancel(self.dev.handle, ticket), self.dev.handle, "ep_serve_cancel")

    def stats(self) -> dict:
        """Serving counters (ep_serve_stats_get): admissions, finishes, batch
        rebuilds, graph captures, chunks, and the status of a failure that
        stopped the serving thread (0 if none)."""
        st = L.EpServeStats()
        _check(L.lib().ep_serve_stats_get(self.dev.handle, C.byref(st)), self.dev.handle,
               "ep_serve_stats_get")
        return {f[0]: getattr(st, f[0]) for f in L.EpServeStats._fields_ if f[0] != "reserved"}

    def stop(self):
        L.lib().ep_serve_stop(self.dev.handle)


# ------------------------------------------------------------ legalization --

def legalize(x, y, cell_bucket, bel_x, bel_y, bel_bucket, bel_taken=None):
    """Greedy nearest-free-BEL legalization with the reference's rule
    (electro_place/mod.rs:206-256).  Returns (cell_bel, displacement)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    cb = np.ascontiguousarray(cell_bucket, np.int32)
    bx = np.ascontiguousarray(bel_x, np.int32)
    by = np.ascontiguousarray(bel_y, np.int32)
    bb = np.ascontiguousarray(bel_bucket, np.int32)
    taken = (np.zeros(bx.shape[0], np.uint8) if bel_taken is None
             else np.ascontiguousarray(bel_taken, np.uint8).copy())
    cell_bel = np.full(x.shape[0], -1, np.int32)
    disp = C.c_double()
    _check(L.lib().ep_legalize(x.shape[0], L.ptr(x), L.ptr(y), L.ptr(cb), bx.shape[0], L.ptr(bx),
                               L.ptr(by), L.ptr(bb), L.ptr(taken), L.ptr(cell_bel), C.byref(disp)),
           None, "ep_legalize")
    return cell_bel, disp.value


def total_hpwl_int(d: FlatDesign, ix, iy) -> float:
    """metrics/wirelength.rs:10-51 on integer locations (driver = first pin)."""
    off = d.net_off.astype(np.int64)
    deg = np.diff(off)
    px = np.concatenate([np.asarray(ix, np.int64), d.fixed_x.astype(np.int64)])[d.pin_cell]
    py = np.concatenate([np.asarray(iy, np.int64), d.fixed_y.astype(np.int64)])[d.pin_cell]
    keep = deg >= 2
    starts = off[:-1][keep]
    if starts.size == 0:
        return 0.0
    hx = np.maximum.reduceat(px, starts) - np.minimum.reduceat(px, starts)
    hy = np.maximum.reduceat(py, starts) - np.minimum.reduceat(py, starts)
    return float((hx + hy).sum())


# --------------------------------------------------------------- placers ----

class PlacerElectro:
    """`impl Placer for PlacerElectro` (electro_place/mod.rs:47-63)."""

    def __init__(self, device: Optional[Device] = None, precision: str = "fast"):
        self.device = device
        self.precision = precision

    def place(self, ctx: "Context", cfg: ElectroPlaceCfg) -> None:
        place_electro(ctx, cfg, device=self.device, precision=self.precision)

    def place_cells(self, ctx: "Context", cfg: ElectroPlaceCfg, cells) -> None:
        """Incremental placement (mod.rs:54-62): every cell outside `cells`
        is held fixed at its current location for the duration of the call."""
        keep = np.zeros(ctx.design.n_mov, bool)
        keep[np.asarray(cells, np.int64)] = True
        sub, back = ctx.design_subset(keep)
        legal = None
        if ctx.legal is not None:
            # the others keep their BELs (locked for the call, common.rs:502-530)
            taken = (np.zeros(len(ctx.legal.bel_x), np.uint8) if ctx.legal.bel_taken is None
                     else ctx.legal.bel_taken.copy())
            if ctx.cell_bel is not None:
                taken[ctx.cell_bel[~keep]] = 1
            legal = LegalProblem(ctx.legal.bel_x, ctx.legal.bel_y, ctx.legal.bel_bucket,
                                 ctx.legal.cell_bucket[keep], bel_taken=taken)
        sub_ctx = Context(sub, legal=legal)
        place_electro(sub_ctx, cfg, device=self.device, precision=self.precision)
        ctx.x[back] = sub_ctx.x
        ctx.y[back] = sub_ctx.y
        if legal is not None:
            if ctx.cell_bel is None:
                ctx.cell_bel = np.full(ctx.design.n_mov, -1, np.int32)
            ctx.cell_bel[back] = sub_ctx.cell_bel


class Context:
    """Flattened stand-in for the reference Context at the placer boundary:
    continuous positions, the BEL bindings of the movable cells and cluster
    children, and `place()` with the reference's string selection
    (crates/npnr-python/src/lib.rs:245-306).  `legal` is the legalization
    view (BELs, buckets, cluster children); by default the design's own BEL
    grid."""

    def __init__(self, design: FlatDesign, device: Optional[Device] = None,
                 legal: Optional[LegalProblem] = None):
        self.design = design
        self.device = device
        self.legal = legal if legal is not None else (
            LegalProblem.of(design) if design.bel_x is not None else None)
        self.x = design.x0.copy()
        self.y = design.y0.copy()
        self.cell_bel = design.init_bel.copy() if design.init_bel is not None else None
        self.child_bel = None
        self.best: Optional[dict] = None  # best snapshot of the last placement
        self.last_result: Optional[PlaceResult] = None

    def place(self, *, placer: str = "heap", seed: int = 1, max_iters: Optional[int] = None,
              precision: str = "fast", field_mode: str = "reference", **_ignored) -> None:
        if placer != "electro":
            raise ValueError(f"Unknown placer: {placer}. Available here: electro")
        cfg = ElectroPlaceCfg(seed=seed, field_mode=field_mode)
        if max_iters is not None:
            cfg.max_iters = max_iters
        try:
            PlacerElectro(self.device, precision).place(self, cfg)
        except PlacerError as e:
            raise RuntimeError(f"ElectroPlace error: {e}") from e

    def legalize(self) -> float:
        if self.legal is None:
            raise PlacementFailed("design has no BEL grid to legalize onto")
        self.cell_bel, self.child_bel, disp = legalize_problem(self.legal, self.x, self.y)
        return disp

    def total_hpwl(self) -> float:
        """metrics/wirelength.rs:46-51 on the current binding."""
        if self.cell_bel is None or self.legal is None:
            return 0.0
        kb = self.child_bel if self.child_bel is not None else np.zeros(self.legal.n_children, np.int32)
        return legal_hpwl(self.design, self.legal, self.cell_bel, kb)

    def design_subset(self, keep: np.ndarray):
        """Design whose movable cells are `keep`; the others become fixed at
        their current positions (common.rs:502-530 with_locked_others)."""
        d = self.design
        idx = np.nonzero(keep)[0]
        other = np.nonzero(~keep)[0]
        remap = np.empty(d.n_mov + d.n_fixed, np.int64)
        remap[idx] = np.arange(idx.size)
        nf = d.n_fixed
        remap[d.n_mov:] = idx.size + np.arange(nf)
        remap[other] = idx.size + nf + np.arange(other.size)
        sub = FlatDesign(
            W=d.W, H=d.H, Bx=d.Bx, By=d.By, x0=self.x[idx], y0=self.y[idx],
            fixed_x=np.concatenate([d.fixed_x, self.x[other]]),
            fixed_y=np.concatenate([d.fixed_y, self.y[other]]),
            net_off=d.net_off, pin_cell=remap[d.pin_cell].astype(np.uint32),
            area=None if d.area is None else d.area[idx],
            cell_type=None if d.cell_type is None else d.cell_type[idx],
            bel_x=d.bel_x, bel_y=d.bel_y, bel_type=d.bel_type, name=d.name + "/subset")
        return sub, idx


def place_electro(ctx: Context, cfg: ElectroPlaceCfg, device: Optional[Device] = None,
                  precision: str = "fast") -> PlaceResult:
    """mod.rs:65-204.  With a legalization view (the usual case) the whole
    of it runs through ep_place_legalized: the continuous loop on the GPU,
    at every check iteration (`legalize_interval`, the last one, and a stop)
    the positions legalized on the host and scored by total_hpwl, the best
    kept (common.rs:485-495), and the final binding legalize(best)
    (mod.rs:199).  Without one, only the continuous loop runs."""
    d = ctx.design
    if d.n_mov == 0:
        return PlaceResult(0, 0, False, 0.0, 0.0, 0.0)  # mod.rs:80-82
    own = device is None
    dev = device or Device()
    try:
        i = dev.upload(d, precision)
        try:
            if ctx.legal is not None:
                r, cb, kb, info = dev.place_legalized(i, cfg, ctx.legal, log=True)
                ctx.cell_bel, ctx.child_bel, ctx.best = cb, kb, info
            else:
                r = dev.place(i, cfg, positions=True, log=True)
        finally:
            dev.free(i)
    finally:
        if own:
            dev.close()
    ctx.x, ctx.y = r.x.copy(), r.y.copy()
    ctx.last_result = r
    if r.status == L.EP_ERR_NONFINITE:
        raise NonFiniteError("non-finite placement state")
    return r
