"""The drop-in glue (flatten_for_gpu) against a brute-force restatement of
the reference's rules, written here independently, on contexts that hold
every case the rules distinguish: dead cells and nets, locked cells placed
and unplaced, movable cells placed and unplaced, clusters with pinned,
pinless, locked, dead and unplaced children, driverless nets and invalid
users (common.rs:155-199 collect_movable_cells / init_positions_from_bels,
common.rs:205-291 add_wa_wirelength_gradient / collect_pin_position_xy,
common.rs:381-407 compute_pin_weights, mod.rs:206-256 with common.rs:294-374
legalization, wirelength.rs:10-51 total_hpwl)."""
import numpy as np
import pytest

import _oracle as O
import eisenjoch_amd as ep
from eisenjoch_amd.flatten import CellView, ChipView, ContextView, NetView, flatten_for_gpu


def random_context(seed):
    rng = np.random.default_rng(seed)
    W = H = 14
    bx, by, bb = [], [], []
    for y in range(H):
        for x in range(W):
            for t in (0, 1):
                bx.append(x); by.append(y); bb.append(t)
    chip = ChipView(W, H, np.array(bx), np.array(by), np.array(bb))
    free = {0: list(rng.permutation([b for b in range(len(bb)) if bb[b] == 0])),
            1: list(rng.permutation([b for b in range(len(bb)) if bb[b] == 1]))}
    cells = []
    for s in range(160):
        t = int(rng.integers(0, 2))
        c = CellView(cell_type=t)
        u = rng.random()
        if u < 0.05:
            c.alive = False
        elif u < 0.10:
            c.locked = True
            c.bel = int(free[t].pop())
        elif u < 0.12:
            c.locked = True  # locked but unplaced: its pins are dropped
        elif u < 0.17:
            pass  # movable, unplaced: starts at the grid centre
        else:
            c.bel = int(free[t].pop())
        cells.append(c)
    clusters = {}
    roots = [s for s in range(0, 160, 9) if cells[s].alive and not cells[s].locked]
    k = 0
    for r in roots:
        kids = []
        for _ in range(int(rng.integers(1, 3))):
            ch = len(cells)
            c = CellView(cell_type=1, cluster=r, constr_x=int(rng.integers(0, 2)), constr_y=0)
            kind = k % 6
            k += 1
            if kind == 4:
                c.alive = False
            elif kind == 5:
                c.locked = True
                c.bel = int(free[1].pop())
            elif kind != 3:  # kind 3: unplaced child
                c.bel = int(free[1].pop())
            cells.append(c)
            kids.append(ch)
        cells[r].cluster = r
        clusters[r] = kids
    # an unplaced child carries no pins (flatten_for_gpu refuses one that does)
    alive = [s for s, c in enumerate(cells) if c.alive and not (c.cluster >= 0 and c.cluster != s
                                                                  and c.bel < 0 and not c.locked)]
    nets = []
    for e in range(230):
        deg = int(rng.integers(1, 7))
        pins = [int(v) for v in rng.choice(alive, deg, replace=False)]
        drv = None if rng.random() < 0.1 else pins[0]
        users = pins[1:] if drv is not None else pins
        users = [None if rng.random() < 0.04 else u for u in users]
        nets.append(NetView(drv, users, alive=rng.random() > 0.05))
    return ContextView(chip, cells, nets, clusters)


# ---- the reference's rules, restated directly on the context ----

def ref_movable(ctx):
    out = []
    for s, c in enumerate(ctx.cells):
        if not c.alive or c.locked:
            continue
        if c.cluster >= 0 and c.cluster != s:
            continue
        out.append(s)
    return out


def ref_loc(ctx, s):
    b = ctx.cells[s].bel
    return (float(ctx.chip.bel_x[b]), float(ctx.chip.bel_y[b])) if b >= 0 else None


def ref_wa_gradient(ctx, xs, ys, a=0.5):
    mov = ref_movable(ctx)
    idx = {s: i for i, s in enumerate(mov)}
    gx, gy = np.zeros(len(mov)), np.zeros(len(mov))
    for net in ctx.nets:
        if not net.alive:
            continue
        px, py, pi = [], [], []
        for cs in ([net.driver] if net.driver is not None else []) + net.users:
            if cs is None:
                continue
            if cs in idx:
                px.append(xs[idx[cs]]); py.append(ys[idx[cs]]); pi.append(idx[cs])
            elif ref_loc(ctx, cs) is not None:
                lx, ly = ref_loc(ctx, cs)
                px.append(lx); py.append(ly); pi.append(-1)
        if len(px) < 2:
            continue
        ax, ay = np.array(px), np.array(py)
        ngx, ngy = np.zeros(len(px)), np.zeros(len(px))
        O.lib().orc_wa_axis_grad(O.P(ax), len(px), a, O.P(ngx))
        O.lib().orc_wa_axis_grad(O.P(ay), len(py), a, O.P(ngy))
        for k, i in enumerate(pi):
            if i >= 0:
                gx[i] += ngx[k]
                gy[i] += ngy[k]
    return gx, gy


def ref_pin_weights(ctx):
    mov = ref_movable(ctx)
    idx = {s: i for i, s in enumerate(mov)}
    w = np.zeros(len(mov))
    for net in ctx.nets:
        if not net.alive:
            continue
        degree = (net.driver is not None) + len(net.users)
        if degree < 2:
            continue
        if net.driver is not None and net.driver in idx:
            w[idx[net.driver]] += 1.0 / degree
        for u in net.users:
            if u is not None and u in idx:
                w[idx[u]] += 1.0 / degree
    return w


def ref_legalize(ctx, xs, ys):
    """legalize_electro: unbind movable cells and their unlocked children,
    then bind in solver order, children right after their root."""
    mov = ref_movable(ctx)
    bound = {}
    unbind = set(mov)
    for r in mov:
        for ch in ctx.clusters.get(r, []):
            if ctx.cells[ch].alive and not ctx.cells[ch].locked:
                unbind.add(ch)
    for s, c in enumerate(ctx.cells):
        if c.alive and c.bel >= 0 and s not in unbind:
            bound[c.bel] = s
    chip = ctx.chip
    result = {}
    for i, s in enumerate(mov):
        best, bc = None, np.inf
        for b in range(len(chip.bel_x)):
            if chip.bel_bucket[b] != ctx.cells[s].cell_type or b in bound:
                continue
            c = (chip.bel_x[b] - xs[i]) ** 2 + (chip.bel_y[b] - ys[i]) ** 2
            if c < bc:
                best, bc = b, c
        bound[best] = s
        result[s] = best
        for ch in ctx.clusters.get(s, []):
            cc = ctx.cells[ch]
            if not cc.alive or cc.locked:
                continue
            tx, ty = chip.bel_x[best] + cc.constr_x, chip.bel_y[best] + cc.constr_y
            cand = [b for b in range(len(chip.bel_x)) if chip.bel_bucket[b] == cc.cell_type and
                    b not in bound and chip.bel_x[b] == tx and chip.bel_y[b] == ty]
            if not cand:
                cand = [b for b in range(len(chip.bel_x)) if chip.bel_bucket[b] == cc.cell_type and
                        b not in bound]
            bound[cand[0]] = ch
            result[ch] = cand[0]
    return result


def ref_total_hpwl(ctx, bels):
    """wirelength.rs:10-51 with every cell at bels[slot] (or its own BEL)."""
    total = 0.0
    for net in ctx.nets:
        if not net.alive or not net.users or net.driver is None:
            continue
        pts = []
        for cs in [net.driver] + [u for u in net.users if u is not None]:
            b = bels.get(cs, ctx.cells[cs].bel)
            if b >= 0:
                pts.append((ctx.chip.bel_x[b], ctx.chip.bel_y[b]))
        if not pts:
            continue
        xs, ys = [p[0] for p in pts], [p[1] for p in pts]
        total += (max(xs) - min(xs)) + (max(ys) - min(ys))
    return total


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_flatten_matches_the_reference_rules(seed):
    ctx = random_context(seed)
    f = flatten_for_gpu(ctx)
    d, lp = f.design, f.legal
    mov = ref_movable(ctx)
    assert list(f.idx_to_cell) == mov
    # initial positions: BEL location, grid centre when unplaced
    for i, s in enumerate(mov):
        loc = ref_loc(ctx, s) or (ctx.chip.W / 2.0, ctx.chip.H / 2.0)
        assert (d.x0[i], d.y0[i]) == loc
    # pin weights count every user, dropped pins included
    np.testing.assert_array_equal(d.pinw, ref_pin_weights(ctx))
    # nets: every alive net, in slot order
    assert list(f.net_slot) == [e for e, n in enumerate(ctx.nets) if n.alive]
    assert list(lp.net_has_driver) == [int(ctx.nets[e].driver is not None) for e in f.net_slot]
    # the WA gradient of the flattened problem is the reference's on the context
    rng = np.random.default_rng(seed + 10)
    xs = rng.uniform(0, ctx.chip.W - 1, len(mov))
    ys = rng.uniform(0, ctx.chip.H - 1, len(mov))
    gx, gy = O.wa_gradient(d, xs, ys)
    rx, ry = ref_wa_gradient(ctx, xs, ys)
    np.testing.assert_allclose(gx, rx, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(gy, ry, rtol=1e-13, atol=1e-13)
    # legalization through the problem equals legalize_electro on the context
    cb, kb, _ = ep.legalize_problem(lp, xs, ys)
    ref = ref_legalize(ctx, xs, ys)
    assert {int(f.idx_to_cell[i]): int(b) for i, b in enumerate(cb)} == \
        {s: b for s, b in ref.items() if s in set(mov)}
    assert {int(f.child_cell[k]): int(b) for k, b in enumerate(kb)} == \
        {s: b for s, b in ref.items() if s not in set(mov)}
    # and its legalized HPWL is total_hpwl of the context so bound
    assert ep.legal_hpwl(d, lp, cb, kb) == ref_total_hpwl(ctx, ref)
    # write-back binds every movable cell and child
    f.write_back(ctx, cb, kb)
    assert all(ctx.cells[s].bel == b for s, b in ref.items())


def test_flatten_edge_cases_are_present():
    # the random contexts do exercise every rule
    ctx = random_context(0)
    f = flatten_for_gpu(ctx)
    cells = ctx.cells
    assert any(not c.alive for c in cells) and any(c.locked and c.bel >= 0 for c in cells)
    assert any(c.locked and c.bel < 0 for c in cells)
    assert (f.legal.child_fixed >= 0).any() and (f.legal.child_fixed < 0).any()
    assert f.legal.bel_taken.sum() > 0
    assert any(n.driver is None for n in ctx.nets) and any(None in n.users for n in ctx.nets)
    # some pin weights differ from the ones the CSR alone would give
    csr = np.zeros(f.design.n_mov)
    off, pc = f.design.net_off.astype(int), f.design.pin_cell
    for e in range(len(off) - 1):
        p = pc[off[e]:off[e + 1]]
        if len(p) >= 2:
            for c in p[p < f.design.n_mov]:
                csr[c] += 1.0 / len(p)
    assert not np.allclose(csr, f.design.pinw)


def test_unplaced_child_with_pins_is_refused():
    ctx = random_context(1)
    child = next(s for s, c in enumerate(ctx.cells)
                 if c.alive and c.cluster >= 0 and c.cluster != s and c.bel < 0 and not c.locked)
    ctx.nets.append(NetView(child, [0, 1]))
    with pytest.raises(ep.PlacementFailed, match="unplaced"):
        flatten_for_gpu(ctx)
