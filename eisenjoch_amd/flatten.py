"""Flattening of a placement context into the library's problem (the glue of
the drop-in: the Rust side's `flatten_for_gpu`).

`ContextView` holds what `place_electro` reads from the reference's
`Context` (crates/nextpnr/src/context/): the cells in slot order with their
type (bucket), BEL, binding strength and cluster, the nets in slot order
with driver and users, the cluster table, and the chip's BELs.
`flatten_for_gpu` turns it into an `ep_design` (`FlatDesign`) and an
`ep_legal_problem` (`LegalProblem`) following the reference rules:

* solver index = alive cells in slot order that are not locked and are not
  a cluster's non-root cells (`collect_movable_cells`, common.rs:155-174);
* initial positions = BEL location, the grid centre (W/2, H/2) when
  unplaced (`init_positions_from_bels`, common.rs:179-199);
* nets = alive nets in slot order (every one kept, so net k of the problem
  is the k-th alive net); pins = driver first, then users
  (`add_wa_wirelength_gradient`, common.rs:228-240); a pin on a movable cell
  names its solver index, a pin on a placed non-movable cell names a fixed
  cell at that BEL's location, a pin on an unplaced non-movable cell (or an
  invalid user) is dropped (`collect_pin_position_xy`, common.rs:268-291);
* fixed cells = the placed non-movable cells that carry pins, in order of
  first appearance in that pin walk;
* pin weights = sum over alive nets of degree >= 2 of 1 / degree, degree =
  driver + every user, dropped pins included (`compute_pin_weights`,
  common.rs:381-407), so they are passed explicitly;
* cluster children = for each movable root in solver order, its alive,
  unlocked constraint children in their order (`unbind_movable_cells`,
  `place_cluster_children`, common.rs:294-374); a child that carries pins is
  the fixed cell of its pins;
* BELs bound on entry = those of every bound cell the legalization does not
  unbind (locked cells, and cells that are neither movable nor children).

A cluster child's pins are read at its BEL, and legalization gives an
unplaced child one, after which the reference counts its pins: a child that
carries pins must therefore be placed when the problem is flattened, as
`initial_placement` (common.rs:27-117) leaves every alive cell before the
loop; `flatten_for_gpu` refuses an unplaced one (PlacementFailed).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .placer import FlatDesign, LegalProblem, PlacementFailed


@dataclass
class ChipView:
    """The chip's BELs in BEL index order (= bucket order within a bucket)."""
    W: int
    H: int
    bel_x: np.ndarray
    bel_y: np.ndarray
    bel_bucket: np.ndarray


@dataclass
class CellView:
    cell_type: int              # bucket
    bel: int = -1               # bound BEL, -1 if unplaced
    locked: bool = False        # bel_strength.is_locked(): Fixed, Locked or User
    cluster: int = -1           # slot of the cluster root, -1 if in no cluster
    constr_x: int = 0           # offset from the root (cluster children)
    constr_y: int = 0
    alive: bool = True


@dataclass
class NetView:
    driver: Optional[int]       # cell slot, None when the net has no driver
    users: List[Optional[int]]  # cell slots; None for an invalid user entry
    alive: bool = True


@dataclass
class ContextView:
    chip: ChipView
    cells: List[CellView]
    nets: List[NetView]
    clusters: Dict[int, List[int]] = field(default_factory=dict)  # root slot -> children slots


@dataclass
class Flattened:
    design: FlatDesign
    legal: LegalProblem
    idx_to_cell: np.ndarray     # cell slot of each solver index
    fixed_cell: np.ndarray      # cell slot of each fixed cell
    net_slot: np.ndarray        # net slot of each problem net
    child_cell: np.ndarray      # cell slot of each cluster child (legal.child_* order)

    def write_back(self, ctx: ContextView, cell_bel, child_bel) -> None:
        """Binds the result (the `bind_bel` calls of legalize_electro): each
        movable cell and cluster child to its BEL."""
        for i, b in zip(self.idx_to_cell, cell_bel):
            ctx.cells[int(i)].bel = int(b)
        for k, b in zip(self.child_cell, child_bel):
            ctx.cells[int(k)].bel = int(b)


def flatten_for_gpu(ctx: ContextView, net_weight: Optional[Dict[int, float]] = None,
                    name: str = "") -> Flattened:
    chip = ctx.chip
    W, H = int(chip.W), int(chip.H)
    bel_x = np.asarray(chip.bel_x, np.int64)
    bel_y = np.asarray(chip.bel_y, np.int64)
    # collect_movable_cells
    idx_to_cell: List[int] = []
    cell_to_idx: Dict[int, int] = {}
    for slot, c in enumerate(ctx.cells):
        if not c.alive or c.locked:
            continue
        if c.cluster >= 0 and c.cluster != slot:
            continue
        cell_to_idx[slot] = len(idx_to_cell)
        idx_to_cell.append(slot)
    n = len(idx_to_cell)
    # init_positions_from_bels
    x0 = np.empty(n)
    y0 = np.empty(n)
    for i, slot in enumerate(idx_to_cell):
        b = ctx.cells[slot].bel
        if b >= 0:
            x0[i], y0[i] = float(bel_x[b]), float(bel_y[b])
        else:
            x0[i], y0[i] = W / 2.0, H / 2.0
    # cluster children of the movable roots, in solver order
    child_cell, child_root, cdx, cdy, cbk = [], [], [], [], []
    for i, root in enumerate(idx_to_cell):
        for ch in ctx.clusters.get(root, []):
            c = ctx.cells[ch]
            if not c.alive or c.locked:
                continue  # a locked child keeps its BEL and is a plain fixed cell
            child_cell.append(ch)
            child_root.append(i)
            cdx.append(c.constr_x)
            cdy.append(c.constr_y)
            cbk.append(c.cell_type)
    # nets: driver first, then users; fixed cells in first-appearance order
    children = set(child_cell)
    fixed_of: Dict[int, int] = {}
    fixed_cell: List[int] = []
    net_off, pin_cell, net_slot, has_drv, nw = [0], [], [], [], []
    pinw = np.zeros(n)
    for e, net in enumerate(ctx.nets):
        if not net.alive:
            continue
        pins = ([net.driver] if net.driver is not None else []) + list(net.users)
        for cs in pins:
            if cs is None:
                continue  # an invalid user names no cell
            if cs in cell_to_idx:
                pin_cell.append(cell_to_idx[cs])
            elif cs in children and ctx.cells[cs].bel < 0:
                raise PlacementFailed(f"cluster child in cell slot {cs} carries pins but is unplaced: "
                                      "run the initial placement first")
            elif ctx.cells[cs].bel >= 0:
                if cs not in fixed_of:
                    fixed_of[cs] = len(fixed_cell)
                    fixed_cell.append(cs)
                pin_cell.append(n + fixed_of[cs])
        net_off.append(len(pin_cell))
        net_slot.append(e)
        has_drv.append(net.driver is not None)
        nw.append(1.0 if net_weight is None else float(net_weight.get(e, 1.0)))
        # compute_pin_weights: every user counts towards the degree
        degree = (net.driver is not None) + len(net.users)
        if degree >= 2:
            w = 1.0 / degree
            for cs in pins:
                if cs is not None and cs in cell_to_idx:
                    pinw[cell_to_idx[cs]] += w
    fixed_bel = [ctx.cells[s].bel for s in fixed_cell]
    fx = bel_x[fixed_bel].astype(np.float64) if fixed_cell else np.zeros(0)
    fy = bel_y[fixed_bel].astype(np.float64) if fixed_cell else np.zeros(0)
    # BELs the legalization never unbinds
    unbound = set(idx_to_cell) | set(child_cell)
    taken = np.zeros(len(bel_x), np.uint8)
    for slot, c in enumerate(ctx.cells):
        if c.alive and c.bel >= 0 and slot not in unbound:
            taken[c.bel] = 1
    design = FlatDesign(
        W=W, H=H, x0=x0, y0=y0, fixed_x=fx, fixed_y=fy,
        net_off=np.array(net_off, np.uint32), pin_cell=np.array(pin_cell, np.uint32),
        pinw=pinw, net_weight=None if net_weight is None else np.array(nw), name=name,
        cell_type=np.array([ctx.cells[s].cell_type for s in idx_to_cell], np.int32),
        bel_x=np.asarray(chip.bel_x, np.int32), bel_y=np.asarray(chip.bel_y, np.int32),
        bel_type=np.asarray(chip.bel_bucket, np.int32),
        init_bel=np.array([ctx.cells[s].bel for s in idx_to_cell], np.int32))
    legal = LegalProblem(
        chip.bel_x, chip.bel_y, chip.bel_bucket, design.cell_type, bel_taken=taken,
        child_root=np.array(child_root, np.int32), child_dx=np.array(cdx, np.int32),
        child_dy=np.array(cdy, np.int32), child_bucket=np.array(cbk, np.int32),
        child_fixed=np.array([fixed_of.get(ch, -1) for ch in child_cell], np.int32),
        net_has_driver=np.array(has_drv, np.uint8))
    return Flattened(design, legal, np.array(idx_to_cell, np.int64), np.array(fixed_cell, np.int64),
                     np.array(net_slot, np.int64), np.array(child_cell, np.int64))
