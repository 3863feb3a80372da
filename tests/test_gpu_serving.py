"""Serving: continuous batching on one GPU.  Requests that arrive while
others run are admitted as slots and memory free up; each result is the same
alone and in the mix (bitwise on the FAST path, whose sums are fixed-point
or fixed-order per design)."""
import time

import numpy as np
import pytest

import eisenjoch_amd as ep
from conftest import log_case

pytestmark = pytest.mark.gpu

CFG = ep.ElectroPlaceCfg(max_iters=300, field_mode="idxst", log_hpwl=False)


def wait(srv, tickets, timeout=120.0):
    out = {}
    t0 = time.time()
    while len(out) < len(tickets):
        for t in tickets:
            if t in out:
                continue
            r = srv.poll(t)
            if r is not None:
                out[t] = r
        if time.time() - t0 > timeout:
            raise TimeoutError(f"{len(out)}/{len(tickets)} requests finished")
        time.sleep(0.002)
    return out


def test_serving_results_equal_alone():
    sizes = [1500, 3000, 6000, 2500, 12000, 4000]
    ds = [ep.sized_design(n, 300 + k) for k, n in enumerate(sizes)]
    with ep.Device(0) as dev:
        log_case("serving: alone runs")
        alone = []
        for d in ds:
            i = dev.upload(d, "fast")
            alone.append(dev.place(i, CFG))
            dev.free(i)
        log_case("serving: mixed arrivals, 3 slots")
        srv = ep.Server(dev, max_slots=3, iters_per_sync=5)
        try:
            tickets = []
            for k, d in enumerate(ds):
                tickets.append(srv.submit(d, CFG))
                time.sleep(0.003 * (k % 3))  # staggered arrivals while others run
            res = wait(srv, tickets)
            for k, t in enumerate(tickets):
                r = res[t]
                assert r.status == 0
                x, y = srv.positions(t)
                assert r.iters == alone[k].iters and r.converged == alone[k].converged
                np.testing.assert_array_equal(x, alone[k].x)
                np.testing.assert_array_equal(y, alone[k].y)
                assert r.hpwl == alone[k].hpwl
                assert r.t_submit_ms <= r.t_admit_ms <= r.t_done_ms
                srv.release(t)
        finally:
            srv.stop()


def test_serving_slots_and_cancel():
    ds = [ep.sized_design(n, 400 + k) for k, n in enumerate([3000, 3000, 3000, 3000])]
    with ep.Device(0) as dev:
        log_case("serving: slot limit 1 and cancel")
        srv = ep.Server(dev, max_slots=1, iters_per_sync=5)
        try:
            t = [srv.submit(d, CFG) for d in ds]
            srv.cancel(t[3])  # still queued behind a single slot
            res = wait(srv, t)
            assert res[t[3]].status == ep._lib.EP_ERR_CANCELLED
            done = [res[k] for k in t[:3]]
            assert all(r.status == 0 and r.converged for r in done)
            # one slot: each admission waits for the previous finish
            assert done[1].t_admit_ms >= done[0].t_done_ms - 1.0
            assert done[2].t_admit_ms >= done[1].t_done_ms - 1.0
        finally:
            srv.stop()


def test_serving_admission_refused_over_budget():
    d = ep.sized_design(20000, 5)
    with ep.Device(0, hbm_budget_bytes=16 << 20) as dev:
        log_case("serving: budget refusal")
        srv = ep.Server(dev, max_slots=4)
        try:
            big = ep.sized_design(200_000, 6)
            with pytest.raises(ep.AdmissionRefused):
                srv.submit(big, CFG)
            t = srv.submit(d, CFG)
            r = wait(srv, [t])[t]
            assert r.status == 0
        finally:
            srv.stop()


def test_serving_rejects_invalid_designs_at_submit():
    # submit validates like ep_design_upload: a pin naming a cell that does
    # not exist never reaches the device
    d = ep.sized_design(2000, 9)
    bad_pins = d.pin_cell.copy()
    bad_pins[5] = d.n_mov + d.n_fixed + 10
    bad = ep.FlatDesign(W=d.W, H=d.H, x0=d.x0, y0=d.y0, fixed_x=d.fixed_x, fixed_y=d.fixed_y,
                        net_off=d.net_off, pin_cell=bad_pins, area=d.area, Bx=d.Bx, By=d.By)
    with ep.Device(0) as dev:
        log_case("serving: invalid design at submit")
        srv = ep.Server(dev, max_slots=2)
        try:
            with pytest.raises(ValueError, match="pin_cell out of range"):
                srv.submit(bad, CFG)
            t = srv.submit(d, CFG)  # the server keeps serving
            assert wait(srv, [t])[t].status == 0
        finally:
            srv.stop()


def test_serving_failure_ends_every_ticket(monkeypatch):
    # A failure inside the serving thread (forced from the host before the
    # fourth chunk launches, so no kernel is involved) must end every queued
    # and running request with the error, refuse later submissions and leave
    # nothing pending.  A context needs a device, hence the gpu mark.
    ds = [ep.sized_design(n, 500 + k) for k, n in enumerate([3000, 4000, 2500, 5000, 3500])]
    monkeypatch.setenv("EP_SERVE_FAIL_AFTER", "3")
    with ep.Device(0) as dev:
        log_case("serving: forced failure after 3 chunks")
        srv = ep.Server(dev, max_slots=2, iters_per_sync=5)
        try:
            # the serving thread runs while the designs are submitted, so the
            # failure may come before the last submissions: those are refused
            # with it, and every one accepted before it ends with it
            tickets, refused = [], 0
            for d in ds:
                try:
                    tickets.append(srv.submit(d, CFG))
                except ep.DeviceError as e:
                    assert "forced serving failure" in str(e)
                    refused += 1
                else:
                    assert refused == 0  # no submission is accepted after a refusal
            assert tickets
            res = wait(srv, tickets, timeout=60.0)  # nothing hangs
            # 15 iterations ran before the failure: no design can have stopped (min 20)
            assert all(res[t].status == ep._lib.EP_ERR_DEVICE for t in tickets)
            with pytest.raises(ep.DeviceError, match="forced serving failure"):
                srv.submit(ds[0], CFG)
            st = srv.stats()
            assert st["failed"] == ep._lib.EP_ERR_DEVICE and st["chunks"] == 3
            assert st["finished"] == 0 and 1 <= st["admitted"] <= min(2, len(tickets))
            with pytest.raises(ep.PlacerError):
                srv.positions(tickets[0])
            for t in tickets:
                srv.release(t)
        finally:
            srv.stop()
    # a new context serves normally again
    monkeypatch.delenv("EP_SERVE_FAIL_AFTER")
    with ep.Device(0) as dev:
        srv = ep.Server(dev, max_slots=2)
        try:
            t = srv.submit(ds[2], CFG)
            assert wait(srv, [t])[t].status == 0
            assert srv.stats()["failed"] == 0
        finally:
            srv.stop()
