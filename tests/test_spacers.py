"""Spacer (filler) insertion to the target utilisation and timing-driven net
weights, from the reference design spec
(docs/superpowers/specs/2026-03-13-placer-optimization-design.md): :91-97,
spacer_count = (grid area x 0.7 - movable area) / spacer area, spacers take
part in the density, not the wirelength; :109-113, net weight 1 + 5 crit^2."""
import math

import numpy as np
import pytest

import eisenjoch_amd as ep
from conftest import log_case


def test_spacer_count_area_and_pin_remap():
    d = ep.config1()
    s, n = ep.with_spacers(d, target_util=0.7)
    assert n == math.floor(64 * 64 * 0.7 - d.n_mov)
    assert s.n_mov == d.n_mov + n and s.n_fixed == d.n_fixed
    assert abs(float(s.area.sum()) - 64 * 64 * 0.7) < 1.0
    # the design's own cells keep their indices and positions; pins of the
    # fixed cells move up by the spacer count; no pin lands on a spacer
    assert np.array_equal(s.x0[:d.n_mov], d.x0) and np.array_equal(s.y0[:d.n_mov], d.y0)
    old, new = d.pin_cell.astype(np.int64), s.pin_cell.astype(np.int64)
    assert np.array_equal(new, np.where(old >= d.n_mov, old + n, old))
    assert not np.any((new >= d.n_mov) & (new < d.n_mov + n))
    assert np.array_equal(s.net_off, d.net_off)
    sx, sy = s.x0[d.n_mov:], s.y0[d.n_mov:]
    assert sx.min() >= 0.5 and sx.max() <= d.W - 0.5 and sy.min() >= 0.5 and sy.max() <= d.H - 0.5


def test_timing_net_weights():
    w = ep.timing_net_weights([0.0, 0.5, 1.0])
    assert np.allclose(w, [1.0, 2.25, 6.0])
    assert np.allclose(ep.timing_net_weights([1.0], timing_weight=2.0), [3.0])
    with pytest.raises(ValueError):
        ep.timing_net_weights([1.5])
    # the weights are the design's per-net weights
    d = ep.config1()
    d.net_weight = ep.timing_net_weights(np.linspace(0.0, 1.0, d.n_nets))
    assert d.to_c().net_weight


def test_no_spacers_when_full():
    d = ep.config1()
    s, n = ep.with_spacers(d, target_util=0.2)  # 0.2 x 4096 < 1500 cells
    assert n == 0 and s is d


@pytest.mark.gpu
def test_placement_with_spacers_converges(device):
    d = ep.config1()
    s, n = ep.with_spacers(d, target_util=0.7, seed=2)
    cfg = ep.ElectroPlaceCfg(max_iters=400, field_mode="idxst", log_hpwl=False, stop_every_iter=True)
    out = {}
    for name, des in (("plain", d), ("spacers", s)):
        log_case(f"spacers: config1 {name}, {des.n_mov} movable cells")
        i = device.upload(des, "fast")
        try:
            out[name] = device.place(i, cfg, positions=True)
        finally:
            device.free(i)
    r = out["spacers"]
    assert r.converged and r.overflow <= 0.10
    assert np.all(np.isfinite(r.x)) and np.all(np.isfinite(r.y))
    # the spacers fill the grid's free area, so the design's own cells end
    # up in a smaller region than without them
    bbox = lambda x, y: (x.max() - x.min()) * (y.max() - y.min())
    own = bbox(r.x[:d.n_mov], r.y[:d.n_mov])
    alone = bbox(out["plain"].x, out["plain"].y)
    assert own < alone
