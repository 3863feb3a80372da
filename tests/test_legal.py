"""Legalization with cluster children, the legalized HPWL and the
best-snapshot loop: the host library (ep_legalize_problem, ep_legal_hpwl)
against the oracle, and the oracle pinned by hand-worked cases and by a
brute-force restatement of the best-snapshot rule from separately pinned
parts (mod.rs:176-199, common.rs:319-374 and 485-495, wirelength.rs:10-51)."""
import numpy as np
import pytest

import _oracle as O
import eisenjoch_amd as ep
from _cluster import clustered_config1


def tiny():
    # 2 movable cells (bucket 0), one cluster child (bucket 1) of cell 0 at
    # (+1, 0) that is fixed cell 1; fixed cell 0 sits at (0, 3).  BELs: bucket
    # 0 at (0,0) (1,0) (2,0); bucket 1 at (1,0) (3,0) (2,0) in that index order.
    bel_x = np.array([0, 1, 2, 1, 3, 2])
    bel_y = np.zeros(6)
    bel_b = np.array([0, 0, 0, 1, 1, 1])
    d = ep.FlatDesign(W=4, H=4, x0=np.array([0.4, 0.6]), y0=np.zeros(2),
                      fixed_x=np.array([0.0, 9.0]), fixed_y=np.array([3.0, 9.0]),
                      # net 0: cell 0 -> cell 1; net 1: fixed 0 -> child; net 2: cell 1, fixed 0
                      net_off=np.array([0, 2, 4, 6]), pin_cell=np.array([0, 1, 2, 3, 1, 2]))
    lp = ep.LegalProblem(bel_x, bel_y, bel_b, np.array([0, 0]), child_root=np.array([0]),
                         child_dx=np.array([1]), child_dy=np.array([0]), child_bucket=np.array([1]),
                         child_fixed=np.array([1]))
    return d, lp


def test_tiny_legalization_by_hand():
    d, lp = tiny()
    # cell 0 at 0.4: BEL 0 (cost 0.16); its child wants (1, 0): BEL 3.
    # cell 1 at 0.6: BEL 0 taken, BEL 1 (cost 0.16) beats BEL 2 (1.96).
    for cb, kb, disp in (ep.legalize_problem(lp, d.x0, d.y0), O.legalize_full(d, lp, d.x0, d.y0)[1:]):
        assert list(cb) == [0, 1] and list(kb) == [3]
        assert disp == pytest.approx(0.32, abs=1e-15)
    # cell 0 at 1.2 now takes BEL 1, its child wants (2, 0): BEL 5 (exact
    # location beats the earlier free BEL 4); cell 1 at 1.4 then takes BEL 2
    x = np.array([1.2, 1.4])
    for cb, kb, _ in (ep.legalize_problem(lp, x, d.y0), O.legalize_full(d, lp, x, d.y0)[1:]):
        assert list(cb) == [1, 2] and list(kb) == [5]
    # with BEL 5 bound on entry, the child falls back to the first free BEL
    # of its bucket in BEL order (BEL 3, not the nearer-looking 4)
    lp2 = ep.LegalProblem(lp.bel_x, lp.bel_y, lp.bel_bucket, lp.cell_bucket,
                          bel_taken=np.array([0, 0, 0, 0, 0, 1]), child_root=lp.child_root,
                          child_dx=lp.child_dx, child_dy=lp.child_dy, child_bucket=lp.child_bucket,
                          child_fixed=lp.child_fixed)
    for cb, kb, _ in (ep.legalize_problem(lp2, x, d.y0), O.legalize_full(d, lp2, x, d.y0)[1:]):
        assert list(kb) == [3]


def test_tiny_legal_hpwl_by_hand():
    d, lp = tiny()
    cb, kb = np.array([0, 1]), np.array([4])  # cells at x 0, 1; child at (3, 0)
    # net 0: (0,0)-(1,0) -> 1; net 1: (0,3)-(3,0) -> 6; net 2: (1,0)-(0,3) -> 4
    assert ep.legal_hpwl(d, lp, cb, kb) == 11.0
    assert O.legal_hpwl(d, lp, cb, kb) == 11.0
    # a driverless net counts 0 (wirelength.rs:16-18)
    lp.net_has_driver = np.array([1, 0, 1], np.uint8)
    assert ep.legal_hpwl(d, lp, cb, kb) == 5.0
    assert O.legal_hpwl(d, lp, cb, kb) == 5.0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ring_search_with_children_equals_full_scan(seed):
    d, lp = clustered_config1(seed=seed)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, d.W - 1, d.n_mov)
    y = rng.uniform(0, d.H - 1, d.n_mov)
    x[::7] = np.round(x[::7]) + 0.5  # exact ties between two BELs
    cb, kb, disp = ep.legalize_problem(lp, x, y)
    rc, ocb, okb, odisp = O.legalize_full(d, lp, x, y)
    assert rc == 0
    np.testing.assert_array_equal(cb, ocb)
    np.testing.assert_array_equal(kb, okb)
    assert disp == odisp
    assert len(set(cb) | set(kb)) == d.n_mov + lp.n_children  # every BEL bound once
    assert ep.legal_hpwl(d, lp, cb, kb) == O.legal_hpwl(d, lp, cb, kb)


def test_best_snapshot_equals_brute_force():
    # the oracle's best-snapshot loop against a restatement from pinned parts:
    # without clusters the continuous trajectory ignores legalization, so the
    # snapshot of check iteration k is the final position of a run of k + 1
    # iterations; each is legalized and scored, the first strict minimum wins
    d = ep.config1()
    lp = ep.LegalProblem.of(d)
    n = 60
    o = O.place_legalized(d, lp, max_iters=n)
    assert o["rc"] == 0
    checks = [k for k in range(n) if k % 5 == 0 or k == n - 1]
    assert o["legalizations"] == len(checks)
    best, best_k, best_xy = np.inf, -1, (d.x0, d.y0)
    for k in checks:
        r = O.place(d, max_iters=k + 1)
        rc, cb, kb, _ = O.legalize_full(d, lp, r["x"], r["y"])
        h = O.legal_hpwl(d, lp, cb, kb)
        if h < best:
            best, best_k, best_xy = h, k, (r["x"], r["y"])
    assert o["best_iter"] == best_k and o["best_hpwl"] == best
    rc, cb, kb, _ = O.legalize_full(d, lp, *best_xy)
    np.testing.assert_array_equal(o["cell_bel"], cb)


def test_legalization_reports_no_bels():
    d, lp = tiny()
    full = ep.LegalProblem(lp.bel_x, lp.bel_y, lp.bel_bucket, lp.cell_bucket,
                           bel_taken=np.array([1, 1, 0, 0, 0, 0]))
    with pytest.raises(ep.NoBelsAvailable):
        ep.legalize_problem(full, d.x0, d.y0)
    assert O.legalize_full(d, full, d.x0, d.y0)[0] == -2
