"""The five workloads at their stated sizes on the GPU, against the oracle.

Tolerances, per kernel and precision (relative to the largest magnitude of
the oracle's output unless stated):
  WA gradient: FP64 1e-12, FAST 2e-5 (FP32 exponentials and sums);
  density: FP64 1e-13 absolute, FAST 2e-7 relative + 1e-8 absolute
    (2^-31 fixed point); overflow FP64 1e-12, FAST 1e-7 relative;
  field (from the oracle's density): FP64 1e-10, FAST 1e-4 (BF16 hi+lo MFMA);
  density gradient (from the oracle's field): FP64 1e-13, FAST 1e-6;
  placement: FAST and plain FP64 both reach overflow <= 0.10 and their
    continuous HPWL differs by at most 0.5 %.
Positions are rounded to FP32 before both sides see them, so the FAST path's
inputs are exact."""
import math
import time

import numpy as np
import pytest

import _oracle as O
import eisenjoch_amd as ep
from conftest import log_case

pytestmark = pytest.mark.gpu

PLACE = ep.ElectroPlaceCfg(max_iters=500, field_mode="idxst", log_hpwl=False, stop_every_iter=True)
TOL_WA = {"fp64": 1e-12, "fast": 2e-5}


def f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def with_nets(d, sizes, seed):
    """d plus one net of each given size: distinct random movable cells, the
    first pin on a fixed cell when the design has one."""
    rng = np.random.default_rng(seed)
    off = list(d.net_off.astype(np.int64))
    pins = list(d.pin_cell.astype(np.int64))
    for s in sizes:
        p = rng.choice(d.n_mov, s, replace=False)
        if d.n_fixed:
            p[0] = d.n_mov + int(rng.integers(0, d.n_fixed))
        pins += [int(v) for v in p]
        off.append(len(pins))
    return ep.FlatDesign(W=d.W, H=d.H, Bx=d.Bx, By=d.By, x0=d.x0, y0=d.y0, fixed_x=d.fixed_x,
                         fixed_y=d.fixed_y, net_off=np.array(off, np.uint32),
                         pin_cell=np.array(pins, np.uint32), area=d.area, name=d.name + "+hf")


def shifted_wa_gradient(d, x, y, a=0.5):
    """wa.rs:51-101 with the exponentials shifted by each net's extreme
    instead of its mean (equal in exact arithmetic, no overflow), in numpy
    FP64: the check for cells whose reference gradient overflows."""
    off = d.net_off.astype(np.int64)
    deg = np.diff(off)
    keep = np.repeat(deg >= 2, deg)
    net = np.repeat(np.arange(len(deg)), deg)[keep]
    cell = d.pin_cell.astype(np.int64)[keep]
    starts = np.concatenate([[0], np.cumsum(deg[deg >= 2])[:-1]])
    out = []
    for pos, fixed in ((x, d.fixed_x), (y, d.fixed_y)):
        v = np.concatenate([pos, fixed])[cell]
        hi = np.maximum.reduceat(v, starts)[np.searchsorted(np.unique(net), net)]
        lo = np.minimum.reduceat(v, starts)[np.searchsorted(np.unique(net), net)]
        u, w = v - hi, v - lo
        e, f = np.exp(a * u), np.exp(-a * w)
        seg = lambda t: np.add.reduceat(t, starts)[np.searchsorted(np.unique(net), net)]
        S, U, Q, V = seg(e), seg(u * e), seg(f), seg(w * f)
        g = e / S * (1 + a * (u - U / S)) - f / Q * (1 - a * (w - V / Q))
        gc = np.zeros(d.n_mov)
        mov = cell < d.n_mov
        np.add.at(gc, cell[mov], g[mov])
        out.append(gc)
    return out


def check_wa(device, i, d, x, y, prec):
    gx, gy = device.wa_grad(i, x, y)
    ox, oy = O.wa_gradient(d, x, y, 0.5)
    ok = np.isfinite(ox) & np.isfinite(oy)
    scale = max(np.abs(ox[ok]).max(), np.abs(oy[ok]).max())
    assert np.abs(gx - ox)[ok].max() <= TOL_WA[prec] * scale
    assert np.abs(gy - oy)[ok].max() <= TOL_WA[prec] * scale
    if not ok.all():
        # the reference's mean-centred exponentials overflow FP64 on nets
        # spanning more than ~700 tiles (a = 0.5); there the shifted form,
        # itself checked against the oracle on every finite cell, is the bar
        sx, sy = shifted_wa_gradient(d, x, y)
        np.testing.assert_allclose(sx[ok], ox[ok], rtol=0, atol=1e-11 * scale)
        np.testing.assert_allclose(sy[ok], oy[ok], rtol=0, atol=1e-11 * scale)
        assert np.isfinite(gx).all() and np.isfinite(gy).all()
        assert np.abs(gx - sx)[~ok].max() <= TOL_WA[prec] * scale
        assert np.abs(gy - sy)[~ok].max() <= TOL_WA[prec] * scale
    return int((~ok).sum())


def check_density_chain(device, i, d, x, y, prec):
    dens, ov = device.density(i, x, y)
    od = O.density(d, x, y)
    oov = O.overlap(od)
    if prec == "fp64":
        np.testing.assert_allclose(dens, od, rtol=0, atol=1e-13)
        assert ov == pytest.approx(oov, rel=1e-12)
    else:
        np.testing.assert_allclose(dens, od, rtol=2e-7, atol=1e-8)
        assert ov == pytest.approx(oov, rel=1e-7)
    device.eval_config(ep.ElectroPlaceCfg(field_mode="idxst"))
    ex, ey = device.field(i, od)
    oex, oey = O.field(od, d.Bx, d.By, mode=1)
    tol = 1e-10 if prec == "fp64" else 1e-4
    scale = max(np.abs(oex).max(), np.abs(oey).max())
    assert np.abs(ex - oex).max() <= tol * scale
    assert np.abs(ey - oey).max() <= tol * scale
    gx, gy = device.density_grad(i, x, y, oex, oey)
    ogx, ogy = O.density_gradient(d, x, y, oex, oey)
    tol = 1e-13 if prec == "fp64" else 1e-6
    np.testing.assert_allclose(gx, ogx, atol=tol * np.abs(ogx).max())
    np.testing.assert_allclose(gy, ogy, atol=tol * np.abs(ogy).max())


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_wa_high_fanout_parity(device, prec):
    # k_wa_hf: a net of up to 5,120 pins keeps each thread's pins (up to ten)
    # in registers; a larger one reads them again.  Nets of 600 and 5,000
    # pins (config 3's range), both sides of 5,120 and 7,000.
    base = ep.sized_design(12000, 41)
    d = with_nets(base, [600, 5000, 5120, 5121, 7000], 42)
    log_case(f"wa high fanout {prec}")
    i = device.upload(d, prec)
    try:
        rng = np.random.default_rng(43)
        x = f32(rng.uniform(0, d.W - 1, d.n_mov))
        y = f32(rng.uniform(0, d.H - 1, d.n_mov))
        check_wa(device, i, d, x, y, prec)
        # gradients of the cells on the big nets alone: differences vs the
        # design without them are the big nets' contributions
        gx, _ = device.wa_grad(i, x, y)
        ob, _ = O.wa_gradient(base, x, y, 0.5)
        od, _ = O.wa_gradient(d, x, y, 0.5)
        hf = np.abs(od - ob) > 0
        assert hf.sum() > 10000  # the big nets touch most cells
        np.testing.assert_allclose((gx - ob)[hf], (od - ob)[hf],
                                   atol=TOL_WA[prec] * 50 * np.abs(od).max())
    finally:
        device.free(i)


@pytest.fixture(scope="module")
def config2():
    return ep.config2()


@pytest.fixture(scope="module")
def config3():
    return ep.config3()


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_config2_kernels_parity(device, config2, prec):
    d = config2
    log_case(f"config2 kernels {prec} (initial positions)")
    i = device.upload(d, prec)
    try:
        x, y = f32(d.x0), f32(d.y0)
        check_wa(device, i, d, x, y, prec)
        check_density_chain(device, i, d, x, y, prec)
        # positions from inside a placement (40 FAST iterations)
        log_case(f"config2 kernels {prec} (iterated positions)")
        r = device.place(i, ep.ElectroPlaceCfg(max_iters=40, field_mode="idxst"))
        x, y = f32(r.x), f32(r.y)
        check_wa(device, i, d, x, y, prec)
        check_density_chain(device, i, d, x, y, prec)
    finally:
        device.free(i)


def _place_both(device, d, what):
    out = {}
    for prec in ("fp64", "fast"):
        log_case(f"{what} place {prec}")
        i = device.upload(d, prec)
        try:
            out[prec] = device.place(i, PLACE, positions=False)
        finally:
            device.free(i)
    return out["fp64"], out["fast"]


def test_config2_place_fast_vs_fp64(device, config2):
    r64, r32 = _place_both(device, config2, "config2")
    assert r64.converged and r32.converged
    assert r32.overflow <= 0.10 and r64.overflow <= 0.10
    assert abs(r32.hpwl - r64.hpwl) / r64.hpwl <= 5e-3


@pytest.mark.parametrize("prec", ["fp64", "fast"])
def test_config3_kernels_parity(device, config3, prec):
    # 1M cells, 1,100 nets of 500-5,000 pins, 1,024^2 bins
    d = config3
    off = d.net_off.astype(np.int64)
    deg = np.diff(off)
    assert (deg > 512).sum() >= 1000 and deg.max() <= 5000
    log_case(f"config3 kernels {prec}")
    i = device.upload(d, prec)
    try:
        x, y = f32(d.x0), f32(d.y0)
        assert check_wa(device, i, d, x, y, prec) > 0  # the reference overflows on some nets
        check_density_chain(device, i, d, x, y, prec)
    finally:
        device.free(i)


def test_config3_place_fast_vs_fp64(device, config3):
    r64, r32 = _place_both(device, config3, "config3")
    assert r64.converged and r32.converged
    assert r32.overflow <= 0.10
    assert abs(r32.hpwl - r64.hpwl) / r64.hpwl <= 5e-3


def test_config4_batch_of_64(device):
    # 64 designs of 2k-100k cells in one batch: all converge, each within
    # 0.5 % HPWL of the plain FP64 placer, and sampled designs are bitwise
    # the same placed alone
    ds = ep.batch_designs(64)
    log_case("config4 batch upload")
    ids = [device.upload(d, "fast") for d in ds]
    try:
        log_case("config4 batch place")
        rs = device.place_batch(ids, PLACE, positions=True)
        assert all(r.converged and r.overflow <= 0.10 for r in rs)
        for k in (0, 17, 40, 63):
            log_case(f"config4 alone {k}")
            a = device.place(ids[k], PLACE, positions=True)
            assert a.iters == rs[k].iters and a.hpwl == rs[k].hpwl
            np.testing.assert_array_equal(a.x, rs[k].x)
            np.testing.assert_array_equal(a.y, rs[k].y)
    finally:
        for i in ids:
            device.free(i)
    gaps = []
    for k, d in enumerate(ds):
        j = device.upload(d, "fp64")
        try:
            p = device.place(j, PLACE, positions=False)
        finally:
            device.free(j)
        assert p.converged
        gaps.append(abs(rs[k].hpwl - p.hpwl) / p.hpwl)
    assert max(gaps) <= 5e-3, max(gaps)


def test_config5_poisson_serving_window(device):
    # Poisson arrivals (seed 1) at 50 requests/s for 2 s, sizes log-uniform
    # 1k-500k cells by the serving rule, 128 slots, 200 GB budget: every
    # request converges; a sample is bitwise the same placed alone; every
    # served HPWL is within 0.5 % of the plain FP64 placer's
    rng = np.random.default_rng(1)
    sizes = [int(round(math.exp(v))) for v in rng.uniform(math.log(1000), math.log(500_000), 8)]
    pool = [ep.sized_design(n, 1000 + k) for k, n in enumerate(sizes)]
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / 50.0)
        if t > 2.0:
            break
        arrivals.append((t, int(rng.integers(0, len(pool)))))
    assert len(arrivals) > 60
    with ep.Device(0, hbm_budget_bytes=200 << 30) as dev:
        log_case("config5: alone runs of the pool")
        alone = []
        for d in pool:
            i = dev.upload(d, "fast")
            alone.append(dev.place(i, PLACE, positions=True))
            dev.free(i)
        plain = []
        for d in pool:
            i = dev.upload(d, "fp64")
            plain.append(dev.place(i, PLACE, positions=False).hpwl)
            dev.free(i)
        log_case(f"config5: serving {len(arrivals)} requests")
        srv = ep.Server(dev, max_slots=128, iters_per_sync=5)
        try:
            tickets, t0, k = [], time.perf_counter(), 0
            while k < len(arrivals):
                now = time.perf_counter() - t0
                while k < len(arrivals) and arrivals[k][0] <= now:
                    tickets.append((srv.submit(pool[arrivals[k][1]], PLACE), arrivals[k][1]))
                    k += 1
                time.sleep(0.001)
            done = {}
            deadline = time.time() + 120
            while len(done) < len(tickets) and time.time() < deadline:
                for tk, p in tickets:
                    if tk not in done:
                        r = srv.poll(tk)
                        if r is not None:
                            done[tk] = r
                time.sleep(0.002)
            assert len(done) == len(tickets)
            st = srv.stats()
            assert st["failed"] == 0 and st["finished"] == len(tickets)
            for n, (tk, p) in enumerate(tickets):
                r = done[tk]
                assert r.status == 0 and r.converged
                assert r.iters == alone[p].iters and r.hpwl == alone[p].hpwl
                assert abs(r.hpwl - plain[p]) / plain[p] <= 5e-3
                if n % 7 == 0:
                    x, y = srv.positions(tk)
                    np.testing.assert_array_equal(x, alone[p].x)
                    np.testing.assert_array_equal(y, alone[p].y)
                srv.release(tk)
        finally:
            srv.stop()
