"""Transform engines and the autotuner.

Every engine computes the same spectral field (split along K or not): on a
random density each one matches the FP64 oracle's field to 1e-4 of its largest value (the BF16
hi+lo tolerance of the FAST path; the FMA engine is FP32).  The autotuner
fills one table entry per bin shape with an engine it timed, reuses a cache
file, and a batch placed with the tuned table still meets the quality bar
(overflow <= 0.10, HPWL within 0.5 % of the plain FP64 placer).
"""
import numpy as np
import pytest

import _oracle as O
import eisenjoch_amd as ep
from conftest import log_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("B", [64, 128, 256, 512])
def test_every_engine_gives_the_field(device, engine, B):
    # at 512 bins engine 4 splits each product into four K slices whose
    # partial planes the next pass (and the final plane sum) add
    d = ep.generate(64, 64, B, n_fixed=0, Bx=B, By=B, seed=2)
    log_case(f"engine {engine} field B={B}")
    device.eval_config(ep.ElectroPlaceCfg(field_mode="idxst"))
    device.tune_set(B, B, engine)
    i = device.upload(d, "fast")
    try:
        dens = np.random.default_rng(B + engine).uniform(0, 3, (B, B))
        ex, ey = device.field(i, dens)
        oex, oey = O.field(dens, B, B, 1)
        scale = max(np.abs(oex).max(), np.abs(oey).max())
        assert np.abs(ex - oex).max() <= 1e-4 * scale
        assert np.abs(ey - oey).max() <= 1e-4 * scale
    finally:
        device.free(i)
        device.tune_set(B, B, -1)
        device.eval_config(ep.ElectroPlaceCfg())


def test_tune_set_rejects_bad_engines(device):
    # invalid arguments raise ValueError, as every EP_ERR_INVALID_ARG does
    with pytest.raises(ValueError, match="unknown transform engine"):
        device.tune_set(64, 64, 7)
    with pytest.raises(ValueError, match="multiples of 64"):
        device.tune_set(40, 24, 0)
    device.tune_set(40, 24, 1)
    device.tune_set(40, 24, -1)


def test_autotune_table_cache_and_quality(device, tmp_path):
    ds = [ep.sized_design(n, 50 + k) for k, n in enumerate([3000, 12000, 40000])]
    assert sorted({d.Bx for d in ds}) == [64, 128, 256]
    cfg = ep.ElectroPlaceCfg(max_iters=500, field_mode="idxst", log_hpwl=False)
    ids = [device.upload(d, "fast") for d in ds]
    plain = [device.upload(d, "fp64") for d in ds]
    cache = tmp_path / "tune.txt"
    try:
        log_case("autotune 64/128/256")
        table = device.autotune(ids, reps=3, cache_path=str(cache))
        shapes = {f"{d.Bx}x{d.By}" for d in ds}
        assert shapes <= set(table)
        assert all(v in ep.Device.ENGINES.values() for v in table.values())
        lines = cache.read_text().split()
        assert len(lines) == 3 * len(shapes)
        # a second run takes the shapes from the cache file
        log_case("autotune from cache")
        again = device.autotune(ids, reps=3, cache_path=str(cache))
        assert again == table
        assert cache.read_text().split() == lines
        log_case("batch with tuned engines vs plain fp64")
        res = device.place_batch(ids, cfg, positions=False)
        ref = [device.place(p, cfg, positions=False) for p in plain]
        for r, q in zip(res, ref):
            assert r.converged and r.overflow <= 0.10
            assert abs(r.hpwl - q.hpwl) / q.hpwl < 5e-3
    finally:
        for i in ids + plain:
            device.free(i)
        for d in ds:
            device.tune_set(d.Bx, d.By, -1)


def test_kslice_engine_placement_matches_unsplit(device):
    # a 512-bin design placed with its products split into K slices (engine
    # 4) and unsplit (engine 0): both reach overflow <= 0.10 with HPWL within
    # 0.1 % of each other (the engines differ only in FP32 summation order)
    d = ep.sized_design(100000, 31)
    assert d.Bx == 512
    cfg = ep.ElectroPlaceCfg(max_iters=500, field_mode="idxst", log_hpwl=False)
    i = device.upload(d, "fast")
    try:
        out = {}
        for e in (0, 4):
            log_case(f"placement with engine {e} at 512 bins")
            device.tune_set(512, 512, e)
            out[e] = device.place(i, cfg, positions=False)
        for r in out.values():
            assert r.converged and r.overflow <= 0.10
        assert abs(out[4].hpwl - out[0].hpwl) / out[0].hpwl < 1e-3
    finally:
        device.free(i)
        device.tune_set(512, 512, -1)


def test_reused_arena_gives_the_fresh_result(device):
    # a freed arena goes to the pool and the next design of about its size
    # reuses it: the result must equal the same design placed in a fresh
    # context, bit for bit (the arena is zeroed and refilled on reuse)
    a = ep.sized_design(30000, 41)
    b = ep.sized_design(29000, 42)
    cfg = ep.ElectroPlaceCfg(max_iters=300, field_mode="idxst", log_hpwl=False)
    log_case("arena reuse")
    ia = device.upload(a, "fast")
    device.place(ia, cfg, positions=False)
    device.free(ia)
    ib = device.upload(b, "fast")
    try:
        reused = device.place(ib, cfg)
    finally:
        device.free(ib)
    with ep.Device(0) as fresh:
        j = fresh.upload(b, "fast")
        ref = fresh.place(j, cfg)
        fresh.free(j)
    assert reused.iters == ref.iters
    np.testing.assert_array_equal(reused.x, ref.x)
    np.testing.assert_array_equal(reused.y, ref.y)
