"""GPU parity of every kernel and of the placement loop against the FP64
oracle, through the C ABI (libeplace.so).

Tolerances (stated per test):
  FP64 path: single kernels agree with the oracle to ~1e-12 relative (the
    only differences are summation order, the max-shifted WA exponentials and
    dense-matrix vs FFT transforms).
  FAST path: FP32 bulk arithmetic and BF16 hi+lo transforms; single kernels
    agree to ~1e-5 of the largest magnitude; integer bin indices are
    bit-exact for the same positions.
"""
import math
import os

import numpy as np
import pytest

import _oracle as O
import eisenjoch_amd as ep
from conftest import log_case

pytestmark = pytest.mark.gpu


def f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def rand_pos(d, seed, edges=True):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, d.W - 1, d.n_mov)
    y = rng.uniform(0, d.H - 1, d.n_mov)
    if edges and d.n_mov >= 8:
        x[:4] = [0.0, d.W - 1, 0.5, d.W / 2 + 0.5]
        y[:4] = [d.H - 1, 0.0, 0.5, 1.0]
        x[4:8] = np.floor(x[4:8])
    return f32(x), f32(y)


@pytest.fixture(scope="module")
def designs():
    small = ep.generate(3000, 3300, 66, n_fixed=64, Bx=64, By=64, n_clusters=16, area_mix=True,
                        n_high_fanout=3, hf_min=70, hf_max=300, seed=11, name="small_hf")
    return {"config1": ep.config1(), "small_hf": small}


@pytest.mark.parametrize("prec", ["fp64", "fast"])
@pytest.mark.parametrize("name", ["config1", "small_hf"])
def test_wa_gradient_parity(device, designs, prec, name):
    d = designs[name]
    log_case(f"wa_gradient {name} {prec}")
    i = device.upload(d, prec)
    try:
        x, y = rand_pos(d, 1)
        gx, gy = device.wa_grad(i, x, y)
        ox, oy = O.wa_gradient(d, x, y, 0.5)
        scale = max(np.abs(ox).max(), np.abs(oy).max())
        tol = 1e-12 if prec == "fp64" else 2e-5  # relative to the largest component
        assert np.abs(gx - ox).max() <= tol * scale
        assert np.abs(gy - oy).max() <= tol * scale
    finally:
        device.free(i)


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_wa_gradient_net_weights_parity(device, designs, prec):
    # optional per-net weights (common.rs:246-261), high-fanout nets included
    d = designs["small_hf"]
    w = np.random.default_rng(5).uniform(0.0, 6.0, d.n_nets)
    dw = ep.FlatDesign(W=d.W, H=d.H, x0=d.x0, y0=d.y0, fixed_x=d.fixed_x, fixed_y=d.fixed_y,
                       net_off=d.net_off, pin_cell=d.pin_cell, area=d.area, Bx=d.Bx, By=d.By,
                       net_weight=w)
    log_case(f"wa_gradient weighted {prec}")
    i = device.upload(dw, prec)
    try:
        x, y = rand_pos(d, 3)
        gx, gy = device.wa_grad(i, x, y)
        ox, oy = O.wa_gradient(dw, x, y, 0.5)
        scale = max(np.abs(ox).max(), np.abs(oy).max())
        tol = 1e-12 if prec == "fp64" else 2e-5  # relative to the largest component
        assert np.abs(gx - ox).max() <= tol * scale
        assert np.abs(gy - oy).max() <= tol * scale
    finally:
        device.free(i)


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_bin_ranges_bit_exact(device, prec):
    # bins != tiles and areas 1/2/4: integer ranges equal the oracle's exactly
    d = ep.generate(5000, 5500, 150, n_fixed=32, Bx=128, By=64, area_mix=True, seed=4)
    log_case(f"bin_ranges {prec}")
    i = device.upload(d, prec)
    try:
        x, y = rand_pos(d, 2)
        # positions on bin and tile boundaries
        x[8:16] = f32(np.arange(8) * d.W / d.Bx)
        y[8:16] = f32(np.arange(8) * d.H / d.By + 0.5)
        got = device.bin_ranges(i, x, y)
        np.testing.assert_array_equal(got, O.bin_ranges(d, x, y))
    finally:
        device.free(i)
    d = designs_unit = ep.config1()
    i = device.upload(designs_unit, prec)
    try:
        x, y = rand_pos(d, 3)
        np.testing.assert_array_equal(device.bin_ranges(i, x, y), O.bin_ranges(d, x, y))
    finally:
        device.free(i)


@pytest.mark.parametrize("prec", ["fp64", "fast"])
@pytest.mark.parametrize("name", ["config1", "small_hf"])
def test_density_and_overflow_parity(device, designs, prec, name):
    d = designs[name]
    log_case(f"density {name} {prec}")
    i = device.upload(d, prec)
    try:
        for seed in (5, 6):
            x, y = rand_pos(d, seed)
            dens, ov = device.density(i, x, y)
            od = O.density(d, x, y)
            oov = O.overlap(od)
            if prec == "fp64":
                np.testing.assert_allclose(dens, od, rtol=1e-13, atol=1e-13)
                assert ov == pytest.approx(oov, rel=1e-12)
            else:  # fixed-point 2^-31 sums, stored as float
                np.testing.assert_allclose(dens, od, rtol=2e-7, atol=1e-8)
                assert ov == pytest.approx(oov, rel=1e-8, abs=1e-10)
            # invariant: total = cell area inside the grid
            assert dens.sum() == pytest.approx(od.sum(), rel=1e-6)
    finally:
        device.free(i)


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_density_large_cells_atomic_path(device, prec):
    # macro-sized cells (footprints wider than 9 bins) take the atomic scatter
    d = ep.generate(4000, 4400, 90, n_fixed=16, Bx=128, By=128, area_mix=True, seed=12)
    d.area = d.area.copy()
    d.area[:40] = 30.0
    log_case(f"density atomic path {prec}")
    i = device.upload(d, prec)
    try:
        x, y = rand_pos(d, 13)
        dens, ov = device.density(i, x, y)
        od = O.density(d, x, y)
        tol = 1e-13 if prec == "fp64" else 2e-7
        np.testing.assert_allclose(dens, od, rtol=tol, atol=tol)
        assert ov == pytest.approx(O.overlap(od), rel=1e-8)
    finally:
        device.free(i)


def closed_form(W, H, p, q, mode):
    xs = np.arange(W); ys = np.arange(H)
    cx = np.cos(np.pi * p * (2 * xs + 1) / (2 * W))
    cy = np.cos(np.pi * q * (2 * ys + 1) / (2 * H))
    lp = 2 * (1 - np.cos(np.pi * p / W)); lq = 2 * (1 - np.cos(np.pi * q / H))
    phi = 1.0 / (lp + lq)
    sh, sg = (1, 1.0) if mode == 0 else (0, -1.0)
    sx = np.sin(np.pi * (p + sh) * (2 * xs + 1) / (2 * W))
    sy = np.sin(np.pi * (q + sh) * (2 * ys + 1) / (2 * H))
    ex = sg * phi * (np.pi * p / W) * np.outer(cy, sx) * (0.5 if q == 0 else 1.0)
    ey = sg * phi * (np.pi * q / H) * np.outer(sy, cx) * (0.5 if p == 0 else 1.0)
    return np.outer(cy, cx) + 0.7, ex, ey


@pytest.mark.parametrize("prec", ["fp64", "fast"])
@pytest.mark.parametrize("mode", ["reference", "idxst"])
@pytest.mark.parametrize("B,p,q", [(64, 3, 5), (512, 1, 0), (512, 37, 200), (1024, 5, 900)])
def test_field_single_mode_full_size(device, prec, mode, B, p, q):
    # closed form of a single DCT mode, size-independent (see test_oracle.py)
    d = ep.generate(64, 64, B, n_fixed=0, Bx=B, By=B, seed=1)
    log_case(f"field closed form B={B} p={p} q={q} {prec} {mode}")
    device.eval_config(ep.ElectroPlaceCfg(field_mode=mode))
    i = device.upload(d, prec)
    try:
        rho, eex, eey = closed_form(B, B, p, q, 0 if mode == "reference" else 1)
        ex, ey = device.field(i, rho)
        scale = max(np.abs(eex).max(), np.abs(eey).max())
        tol = 1e-10 if prec == "fp64" else 1e-4  # of the largest field value
        assert np.abs(ex - eex).max() <= tol * scale
        assert np.abs(ey - eey).max() <= tol * scale
    finally:
        device.free(i)
        device.eval_config(ep.ElectroPlaceCfg())


@pytest.mark.parametrize("prec", ["fp64", "fast"])
@pytest.mark.parametrize("mode", ["reference", "idxst"])
@pytest.mark.parametrize("Bx,By", [(64, 64), (128, 64), (40, 24)])
def test_field_parity_random(device, prec, mode, Bx, By):
    d = ep.generate(64, 64, Bx, By, n_fixed=0, Bx=Bx, By=By, seed=1)
    log_case(f"field random {Bx}x{By} {prec} {mode}")
    device.eval_config(ep.ElectroPlaceCfg(field_mode=mode))
    i = device.upload(d, prec)
    try:
        rng = np.random.default_rng(Bx * 7 + By)
        dens = rng.uniform(0, 3, (By, Bx))
        ex, ey = device.field(i, dens)
        oex, oey = O.field(dens, Bx, By, 0 if mode == "reference" else 1)
        scale = max(np.abs(oex).max(), np.abs(oey).max())
        tol = 1e-10 if prec == "fp64" else 1e-4
        assert np.abs(ex - oex).max() <= tol * scale
        assert np.abs(ey - oey).max() <= tol * scale
    finally:
        device.free(i)
        device.eval_config(ep.ElectroPlaceCfg())


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_density_gradient_parity(device, designs, prec):
    d = designs["small_hf"]
    log_case(f"density_grad {prec}")
    i = device.upload(d, prec)
    try:
        x, y = rand_pos(d, 8)
        rng = np.random.default_rng(9)
        ex = f32(rng.standard_normal((d.By, d.Bx)))
        ey = f32(rng.standard_normal((d.By, d.Bx)))
        gx, gy = device.density_grad(i, x, y, ex, ey)
        ox, oy = O.density_gradient(d, x, y, ex, ey)
        tol = 1e-13 if prec == "fp64" else 1e-6
        np.testing.assert_allclose(gx, ox, atol=tol * np.abs(ox).max())
        np.testing.assert_allclose(gy, oy, atol=tol * np.abs(oy).max())
        # constant field -> that constant
        gx, gy = device.density_grad(i, x, y, np.full((d.By, d.Bx), 0.75), np.full((d.By, d.Bx), -2.0))
        np.testing.assert_allclose(gx, 0.75, rtol=1e-6)
        np.testing.assert_allclose(gy, -2.0, rtol=1e-6)
    finally:
        device.free(i)


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_density_gradient_linear_field(device, designs, prec):
    # a field linear in the bin indices: for interior unit cells on bins =
    # tiles the overlap weights average the bin index to centre - 1/2, so the
    # gradient is the field at (x - 1/2, y - 1/2); distinct coefficients per
    # axis and component catch an index, axis or component mix-up
    d = designs["config1"]
    log_case(f"density_grad linear field {prec}")
    i = device.upload(d, prec)
    try:
        x, y = rand_pos(d, 12, edges=False)
        x = np.clip(x, 1.0, d.W - 2.0)
        y = np.clip(y, 1.0, d.H - 2.0)
        by_, bx_ = np.mgrid[0:d.By, 0:d.Bx].astype(np.float64)
        ex = 0.25 * bx_ - 0.375 * by_ + 1.0  # dyadic: exact in FP32
        ey = -0.125 * bx_ + 0.5 * by_ + 0.5
        gx, gy = device.density_grad(i, x, y, ex, ey)
        tol = 1e-13 if prec == "fp64" else 1e-6
        np.testing.assert_allclose(gx, 0.25 * (x - 0.5) - 0.375 * (y - 0.5) + 1.0, rtol=tol, atol=tol)
        np.testing.assert_allclose(gy, -0.125 * (x - 0.5) + 0.5 * (y - 0.5) + 0.5, rtol=tol, atol=tol)
    finally:
        device.free(i)


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_hpwl_parity(device, designs, prec):
    d = designs["small_hf"]
    log_case(f"hpwl {prec}")
    i = device.upload(d, prec)
    try:
        x, y = rand_pos(d, 10)
        assert device.hpwl(i, x, y) == pytest.approx(O.hpwl(d, x, y), rel=1e-12)
    finally:
        device.free(i)


def _compare_logs(glog, olog, n, rtol, what):
    assert len(glog) >= n and len(olog) >= n, (len(glog), len(olog))
    for k in range(n):
        g, o = glog[k], olog[k]
        assert g["iter"] == o["iter"] == k
        assert g["checked"] == o["checked"]
        for key in ("overflow", "lambda", "step_size_x", "step_size_y"):
            assert g[key] == pytest.approx(o[key], rel=rtol, abs=rtol * 1e-3), (what, k, key, g[key], o[key])


@pytest.mark.parametrize("mode", ["reference", "idxst"])
def test_trajectory_fp64_matches_oracle(device, mode):
    # per-iteration overflow, penalty and BB step sizes of the FP64 GPU placer
    # against the oracle on the config-1 workload
    d = ep.config1()
    log_case(f"trajectory fp64 {mode}")
    iters = 40
    cfg = ep.ElectroPlaceCfg(max_iters=iters, field_mode=mode)
    i = device.upload(d, "fp64")
    try:
        r = device.place(i, cfg, log=True)
    finally:
        device.free(i)
    o = O.place(d, max_iters=iters, field_mode=0 if mode == "reference" else 1)
    assert r.iters == o["iters"] and r.converged == o["converged"]
    _compare_logs(r.log, o["log"], r.iters, 1e-6, "fp64")
    assert r.hpwl == pytest.approx(o["hpwl"], rel=1e-6)
    np.testing.assert_allclose(r.x, o["x"], atol=1e-5 * d.W)


@pytest.mark.parametrize("n,seed", [(5000, 3), (20000, 21)])
def test_trajectory_fp64_whole_run_idxst(device, n, seed):
    # the whole IDXST run of a batch-rule design (70-90 iterations): same
    # iteration count, every iteration's overflow, penalty and step sizes
    # within 1e-3 relative (measured: 1e-4 at most, growing from 1e-14 in
    # the first 20 iterations), final HPWL within 1e-5
    d = ep.sized_design(n, seed)
    log_case(f"whole trajectory fp64 idxst n={n}")
    i = device.upload(d, "fp64")
    try:
        r = device.place(i, ep.ElectroPlaceCfg(max_iters=300, field_mode="idxst"), log=True)
    finally:
        device.free(i)
    o = O.place(d, max_iters=300, field_mode=1)
    assert r.iters == o["iters"] and r.converged and o["converged"]
    _compare_logs(r.log, o["log"], 20, 1e-9, "fp64 first 20")
    _compare_logs(r.log, o["log"], r.iters, 1e-3, "fp64 whole run")
    assert r.hpwl == pytest.approx(o["hpwl"], rel=1e-5)


def test_trajectory_fast_reference_field(device):
    # the reference field is chaotic: rounding differences grow about
    # tenfold every two iterations (FP64: 3e-13 by iteration 20, 3e-8 by 40,
    # O(1) by 80; FAST: 7e-4 by 20).  Its first 20 FAST iterations follow
    # the oracle within 2e-3
    d = ep.config1()
    log_case("trajectory fast reference field")
    i = device.upload(d, "fast")
    try:
        r = device.place(i, ep.ElectroPlaceCfg(max_iters=60), log=True)
    finally:
        device.free(i)
    o = O.place(d, max_iters=60)
    _compare_logs(r.log, o["log"], 20, 2e-3, "fast reference")


def test_trajectory_fast_tracks_oracle(device):
    # the FAST path follows the FP64 oracle's early trajectory and reaches the
    # overflow target within one check interval of it, with HPWL within 0.5 %
    d = ep.config1()
    log_case("trajectory fast idxst")
    cfg = ep.ElectroPlaceCfg(max_iters=300, field_mode="idxst")
    i = device.upload(d, "fast")
    try:
        r = device.place(i, cfg, log=True)
    finally:
        device.free(i)
    o = O.place(d, max_iters=300, field_mode=1)
    # the whole run (21 iterations) within 2e-4 (measured 5e-5)
    _compare_logs(r.log, o["log"], min(r.iters, o["iters"]), 2e-4, "fast")
    assert r.converged and o["converged"]
    assert abs(r.iters - o["iters"]) <= 10
    assert r.hpwl == pytest.approx(o["hpwl"], rel=5e-3)
    # a longer run (a 20k-cell design, about 90 iterations): the first 20
    # iterations within 1e-4 (measured 9e-6), then FP32 rounding makes the
    # runs drift apart by a few percent; they stop within two iterations of
    # each other and their HPWL agrees within 0.5 %
    d = ep.sized_design(20000, 21)
    i = device.upload(d, "fast")
    try:
        r = device.place(i, cfg, log=True)
    finally:
        device.free(i)
    o = O.place(d, max_iters=300, field_mode=1)
    _compare_logs(r.log, o["log"], 20, 1e-4, "fast 20k")
    assert abs(r.iters - o["iters"]) <= 2
    assert r.hpwl == pytest.approx(o["hpwl"], rel=5e-3)


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_reference_mode_does_not_converge(device, prec):
    # exact reference field: no convergence within 150 iterations, as the oracle
    d = ep.config1()
    log_case(f"reference mode 150 iters {prec}")
    i = device.upload(d, prec)
    try:
        r = device.place(i, ep.ElectroPlaceCfg(max_iters=150), log=False)
    finally:
        device.free(i)
    o = O.place(d, max_iters=150)
    assert not r.converged and not o["converged"]
    assert r.iters == o["iters"] == 150
    assert r.overflow > 0.5


def test_batch_equals_alone_bitwise(device):
    # FAST path: fixed-point sums and per-design fixed-order reductions make a
    # design's trajectory independent of the batch it runs in
    ds = [ep.sized_design(n, 100 + k) for k, n in enumerate([2500, 4000, 9000])]
    ds.append(ep.config1())
    cfg = ep.ElectroPlaceCfg(max_iters=200, field_mode="idxst")
    log_case("batch vs alone")
    ids = [device.upload(d, "fast") for d in ds]
    try:
        alone = [device.place(i, cfg) for i in ids]
        batch = device.place_batch(ids, cfg)
        for a, b in zip(alone, batch):
            assert a.iters == b.iters and a.converged == b.converged
            np.testing.assert_array_equal(a.x, b.x)
            np.testing.assert_array_equal(a.y, b.y)
            assert a.hpwl == b.hpwl
        assert all(b.converged for b in batch)
    finally:
        for i in ids:
            device.free(i)


def test_batch_done_times_follow_iterations(device):
    # a batch advances its designs in lockstep, so the device-clock time from
    # the reset to a design's stop grows with the iteration it stopped at
    ds = [ep.sized_design(n, 200 + k) for k, n in enumerate([2000, 5000, 12000, 30000])]
    cfg = ep.ElectroPlaceCfg(max_iters=300, field_mode="idxst", log_hpwl=False, stop_every_iter=True)
    log_case("batch done times")
    ids = [device.upload(d, "fast") for d in ds]
    try:
        rs = device.place_batch(ids, cfg, positions=False)
        t = np.array([r.t_done_ms for r in rs])
        it = np.array([r.iters for r in rs])
        assert np.all(np.isfinite(t)) and np.all(t > 0.0)
        order = np.argsort(it, kind="stable")
        assert np.all(np.diff(t[order]) >= -1e-3 * t.max() - 1e-3)  # non-decreasing in iterations
        assert np.all(np.array([r.t_submit_ms for r in rs]) == 0.0)
    finally:
        for i in ids:
            device.free(i)


def test_graph_replay_equals_eager(device):
    d = ep.sized_design(6000, 7)
    cfg = ep.ElectroPlaceCfg(max_iters=120, field_mode="idxst")
    log_case("graph vs eager")
    i = device.upload(d, "fast")
    try:
        g = device.place(i, cfg)
        os.environ["EP_NO_GRAPH"] = "1"
        try:
            e = device.place(i, cfg)
        finally:
            del os.environ["EP_NO_GRAPH"]
        g2 = device.place(i, cfg)
        np.testing.assert_array_equal(g.x, e.x)
        np.testing.assert_array_equal(g.x, g2.x)
        assert g.iters == e.iters == g2.iters
    finally:
        device.free(i)


def test_fast_within_half_percent_of_fp64(device):
    # the north-star quality bar on a batch-rule design
    d = ep.sized_design(20000, 21)
    cfg = ep.ElectroPlaceCfg(max_iters=500, field_mode="idxst")
    log_case("fast vs fp64 hpwl")
    a = device.upload(d, "fp64")
    b = device.upload(d, "fast")
    try:
        r64 = device.place(a, cfg)
        r32 = device.place(b, cfg)
    finally:
        device.free(a)
        device.free(b)
    assert r64.converged and r32.converged
    assert r32.overflow <= 0.10
    assert abs(r32.hpwl - r64.hpwl) / r64.hpwl < 5e-3


def test_nonfinite_and_invalid_inputs_are_rejected(device):
    d = ep.config1()
    bad = ep.FlatDesign(W=d.W, H=d.H, x0=d.x0.copy(), y0=d.y0.copy(), fixed_x=d.fixed_x,
                        fixed_y=d.fixed_y, net_off=d.net_off, pin_cell=d.pin_cell)
    bad.x0[3] = np.nan
    with pytest.raises(ep.NonFiniteError):
        device.upload(bad, "fast")
    bad = ep.FlatDesign(W=d.W, H=d.H, x0=d.x0, y0=d.y0, fixed_x=d.fixed_x, fixed_y=d.fixed_y,
                        net_off=d.net_off, pin_cell=d.pin_cell.copy())
    bad.pin_cell[0] = 10 ** 6
    with pytest.raises(ValueError):
        device.upload(bad, "fast")
    i = device.upload(d, "fast")
    try:
        with pytest.raises(ValueError):
            device.place(i, ep.ElectroPlaceCfg(max_iters=10 ** 7))
    finally:
        device.free(i)


def test_empty_and_tiny_designs(device):
    # electro_place_tests.rs:7-95: 0, 1, 2 and 4 cells
    for n in (1, 2, 4):
        d = ep.FlatDesign(W=2, H=2, x0=np.arange(n) % 2, y0=np.arange(n) // 2,
                          fixed_x=np.zeros(0), fixed_y=np.zeros(0),
                          net_off=np.array([0, n], np.uint32), pin_cell=np.arange(n, dtype=np.uint32))
        i = device.upload(d, "fast")
        try:
            r = device.place(i, ep.ElectroPlaceCfg(max_iters=10, legalize_interval=2))
            o = O.place(d, max_iters=10, interval=2)
            assert r.iters == o["iters"]
            assert np.all(np.isfinite(r.x))
        finally:
            device.free(i)
    d = ep.FlatDesign(W=4, H=4, x0=np.zeros(0), y0=np.zeros(0), fixed_x=np.zeros(0),
                      fixed_y=np.zeros(0), net_off=np.array([0], np.uint32),
                      pin_cell=np.zeros(0, np.uint32))
    i = device.upload(d, "fast")
    try:
        r = device.place(i, ep.ElectroPlaceCfg())
        assert r.iters == 0
    finally:
        device.free(i)
