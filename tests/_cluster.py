"""Designs with cluster children for the legalization tests.

clustered_config1(): the config-1 design plus `n_pairs` LUT+FF clusters: a
LUT-type movable cell (bucket 0) is the root, a new FF-type cell (bucket 1)
its child at (dx, dy) = (0, 0).  A child is not a solver cell: it is
appended as a fixed cell at its root's initial location and gets two new
2-pin nets (child-root, child-random cell), so the WA gradient reads its
location, which follows it to every new BEL."""
from __future__ import annotations

import numpy as np

import eisenjoch_amd as ep


def clustered_config1(n_pairs: int = 60, seed: int = 3):
    d = ep.config1()
    rng = np.random.default_rng(seed)
    luts = np.nonzero(d.cell_type == 0)[0]
    roots = np.sort(rng.choice(luts, n_pairs, replace=False)).astype(np.int32)
    F0 = d.n_fixed
    fixed_x = np.concatenate([d.fixed_x, d.x0[roots]])
    fixed_y = np.concatenate([d.fixed_y, d.y0[roots]])
    # pin ids of fixed cells start at n_mov: the children are n_mov + F0 + k
    off = list(d.net_off.astype(np.int64))
    pins = list(d.pin_cell.astype(np.int64))
    for k, r in enumerate(roots):
        child = d.n_mov + F0 + k
        other = int(rng.integers(0, d.n_mov))
        for a, b in ((child, int(r)), (child, other)):
            pins += [a, b]
            off.append(len(pins))
    cd = ep.FlatDesign(W=d.W, H=d.H, x0=d.x0, y0=d.y0, fixed_x=fixed_x, fixed_y=fixed_y,
                       net_off=np.array(off, np.uint32), pin_cell=np.array(pins, np.uint32),
                       cell_type=d.cell_type, bel_x=d.bel_x, bel_y=d.bel_y, bel_type=d.bel_type,
                       init_bel=d.init_bel, name="config1+clusters")
    lp = ep.LegalProblem(d.bel_x, d.bel_y, d.bel_type, d.cell_type, child_root=roots,
                         child_dx=np.zeros(n_pairs), child_dy=np.zeros(n_pairs),
                         child_bucket=np.ones(n_pairs), child_fixed=F0 + np.arange(n_pairs))
    return cd, lp


def clustered_context(n_pairs: int = 40, seed: int = 5):
    """The config-1 design as a placement context (ContextView): its cells
    on their initial BELs, its IO cells locked on BELs of their own bucket
    at the perimeter locations, and `n_pairs` LUT+FF clusters whose FF
    children sit on free FF BELs and carry two nets each."""
    from eisenjoch_amd.flatten import CellView, ChipView, ContextView, NetView
    d = ep.config1()
    rng = np.random.default_rng(seed)
    nb = len(d.bel_x)
    bel_x = np.concatenate([d.bel_x, d.fixed_x.astype(np.int32)])
    bel_y = np.concatenate([d.bel_y, d.fixed_y.astype(np.int32)])
    bel_b = np.concatenate([d.bel_type, np.full(d.n_fixed, 2, np.int32)])
    chip = ChipView(d.W, d.H, bel_x, bel_y, bel_b)
    cells = [CellView(int(t), bel=int(b)) for t, b in zip(d.cell_type, d.init_bel)]
    cells += [CellView(2, bel=nb + j, locked=True) for j in range(d.n_fixed)]
    used = set(int(b) for b in d.init_bel)
    free_ff = [b for b in range(nb) if d.bel_type[b] == 1 and b not in used]
    luts = np.nonzero(d.cell_type == 0)[0]
    roots = np.sort(rng.choice(luts, n_pairs, replace=False))
    clusters = {}
    nets = []
    for e in range(d.n_nets):
        p = [int(c) for c in d.pin_cell[d.net_off[e]:d.net_off[e + 1]]]
        nets.append(NetView(p[0], p[1:]))
    for k, r in enumerate(roots):
        ch = len(cells)
        cells.append(CellView(1, bel=int(free_ff[k]), cluster=int(r)))
        cells[int(r)].cluster = int(r)
        clusters[int(r)] = [ch]
        nets.append(NetView(ch, [int(r)]))
        nets.append(NetView(int(rng.integers(0, d.n_mov)), [ch]))
    return ContextView(chip, cells, nets, clusters)
