"""place_electro end to end on the GPU (ep_place_legalized): legalized and
scored snapshots at every check iteration, the best kept, the final binding
legalize(best) (mod.rs:176-199, common.rs:485-495).  The FP64 path is
compared with the oracle's orc_place_legalized bit for bit (BEL binding,
best iteration, best legalized HPWL, number of legalizations); the FAST
path by the validity of its binding and its HPWL (within 3 % of the
oracle's best: the trajectories differ by FP32 rounding)."""
import numpy as np
import pytest

import _oracle as O
import eisenjoch_amd as ep
from _cluster import clustered_config1
from conftest import log_case

pytestmark = pytest.mark.gpu


def _valid(d, lp, cb, kb):
    bels = np.concatenate([cb, kb])
    assert (bels >= 0).all() and len(np.unique(bels)) == len(bels)
    assert (lp.bel_bucket[cb] == lp.cell_bucket).all()
    if lp.n_children:
        assert (lp.bel_bucket[kb] == lp.child_bucket).all()


@pytest.mark.parametrize("field_mode,iters", [("reference", 300), ("idxst", 300)])
def test_fp64_best_snapshot_equals_oracle(device, field_mode, iters):
    # config 1 in the reference field mode runs all 300 iterations (it never
    # converges), so the best snapshot, not the last iterate, is the result
    d = ep.config1()
    lp = ep.LegalProblem.of(d)
    cfg = ep.ElectroPlaceCfg(max_iters=iters, field_mode=field_mode)
    i = device.upload(d, "fp64")
    try:
        log_case(f"ep_place_legalized fp64 config1 {field_mode} {iters}")
        r, cb, kb, info = device.place_legalized(i, cfg, lp)
    finally:
        device.free(i)
    o = O.place_legalized(d, lp, max_iters=iters, field_mode=1 if field_mode == "idxst" else 0)
    assert o["rc"] == 0
    assert r.iters == o["iters"] and r.converged == o["converged"]
    assert info["legalizations"] == o["legalizations"]
    assert info["best_iter"] == o["best_iter"]
    assert info["best_hpwl"] == o["best_hpwl"]
    np.testing.assert_array_equal(cb, o["cell_bel"])
    _valid(d, lp, cb, kb)
    if field_mode == "reference":
        assert r.iters == iters and info["legalizations"] == 61
        assert info["best_iter"] < iters - 1  # the last iterate is not the best


def test_fp64_clusters_equal_oracle(device):
    # cluster children move with their roots at every legalization, and the
    # WA gradient of the following iterations sees them there
    d, lp = clustered_config1()
    cfg = ep.ElectroPlaceCfg(max_iters=120, field_mode="idxst")
    i = device.upload(d, "fp64")
    try:
        log_case("ep_place_legalized fp64 config1 with clusters")
        r, cb, kb, info = device.place_legalized(i, cfg, lp)
        # the design's own fixed locations are back on the device afterwards
        r2 = device.place(i, cfg, positions=False)
    finally:
        device.free(i)
    o = O.place_legalized(d, lp, max_iters=120, field_mode=1)
    assert o["rc"] == 0
    assert r.iters == o["iters"] and info["best_iter"] == o["best_iter"]
    assert info["best_hpwl"] == o["best_hpwl"]
    np.testing.assert_array_equal(cb, o["cell_bel"])
    np.testing.assert_array_equal(kb, o["child_bel"])
    _valid(d, lp, cb, kb)
    # children stay with their roots (dx = dy = 0) wherever a FF BEL was free there
    same = (lp.bel_x[kb] == lp.bel_x[cb[lp.child_root]]) & (lp.bel_y[kb] == lp.bel_y[cb[lp.child_root]])
    assert same.mean() > 0.9
    o2 = O.place(d, max_iters=120, field_mode=1)
    assert r2.iters == o2["iters"]


def test_fast_best_snapshot_and_context(device):
    d = ep.config1()
    lp = ep.LegalProblem.of(d)
    o = O.place_legalized(d, lp, max_iters=300, field_mode=1)
    ctx = ep.Context(d, device=device)
    log_case("Context.place electro fast idxst")
    ctx.place(placer="electro", max_iters=300, field_mode="idxst")
    _valid(d, lp, ctx.cell_bel, np.zeros(0, np.int64))
    assert ctx.best["best_hpwl"] == ctx.total_hpwl()
    assert abs(ctx.total_hpwl() - o["best_hpwl"]) <= 0.03 * o["best_hpwl"]


def test_flattened_context_places_like_the_oracle(device):
    # the drop-in path end to end: a context flattened in reference order,
    # placed with the best snapshot on the GPU (FP64), equals the oracle run
    # on the same problem, and its binding writes back onto the context
    from _cluster import clustered_context
    from eisenjoch_amd.flatten import flatten_for_gpu
    ctx = clustered_context()
    f = flatten_for_gpu(ctx)
    cfg = ep.ElectroPlaceCfg(max_iters=100, field_mode="idxst")
    i = device.upload(f.design, "fp64")
    try:
        log_case("flattened context, ep_place_legalized fp64")
        r, cb, kb, info = device.place_legalized(i, cfg, f.legal)
    finally:
        device.free(i)
    o = O.place_legalized(f.design, f.legal, max_iters=100, field_mode=1)
    assert o["rc"] == 0 and info["best_iter"] == o["best_iter"]
    np.testing.assert_array_equal(cb, o["cell_bel"])
    np.testing.assert_array_equal(kb, o["child_bel"])
    f.write_back(ctx, cb, kb)
    bels = [c.bel for c in ctx.cells if c.alive]
    assert min(bels) >= 0 and len(set(bels)) == len(bels)
