"""ep_place_noalloc against ep_place: the allocation-free entry point (launch
descriptors in the library's preallocated device buffer, eager launches)
must give the ep_place result bit for bit: iterations, convergence,
overflow, HPWL, the iteration log and every position.  Two calls in a row,
the second on a smaller design, so the buffer's earlier contents cannot
leak into a later call."""
import numpy as np
import pytest

import eisenjoch_amd as ep
from conftest import log_case

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.status == b.status
    assert a.iters == b.iters
    assert a.converged == b.converged
    assert a.overflow == b.overflow
    assert a.hpwl == b.hpwl
    assert a.lam == b.lam
    assert np.array_equal(a.x, b.x)
    assert np.array_equal(a.y, b.y)
    assert len(a.log) == len(b.log)
    for ra, rb in zip(a.log, b.log):
        assert ra == rb


@pytest.mark.parametrize("prec", ["fast", "fp64"])
def test_noalloc_equals_place_two_calls(device, prec):
    big = ep.sized_design(20_000, 7)
    small = ep.config1()
    cfg = ep.ElectroPlaceCfg(max_iters=200, field_mode="idxst")
    ids = [device.upload(big, prec), device.upload(small, prec)]
    try:
        ref = []
        for i in ids:
            log_case(f"ep_place reference {prec} design {i}")
            ref.append(device.place(i, cfg, positions=True, log=True))
        for k, i in enumerate(ids):  # big first, then the smaller one
            log_case(f"ep_place_noalloc {prec} design {i}")
            r = device.place_noalloc(i, cfg, positions=True, log=True)
            _same(r, ref[k])
        assert ref[0].converged and ref[1].converged
    finally:
        for i in ids:
            device.free(i)


def test_noalloc_reference_field_mode(device):
    # the reference field (does not converge): runs to max_iters on both paths
    d = ep.config1()
    cfg = ep.ElectroPlaceCfg(max_iters=60)
    i = device.upload(d, "fp64")
    try:
        log_case("ep_place vs ep_place_noalloc, reference field")
        a = device.place(i, cfg, positions=True, log=True)
        b = device.place_noalloc(i, cfg, positions=True, log=True)
        _same(b, a)
        assert a.iters == 60
    finally:
        device.free(i)
