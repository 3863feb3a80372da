"""The FFT transform engine (engine 5: FAST path, bin counts that are powers
of two).

Every pass of the spectral solve runs as batched row FFTs: the DCT-II by
Makhoul's reordering, the inverse cosine transform and both sine transforms
(the reference field's DST-III and the IDXST one) as a DCT-III of a
reordered row.  On random densities its field matches the FP64 oracle's
field to 2e-6 of the field's largest value (FP32 arithmetic; the MFMA
engines' BF16 hi+lo products are held to 1e-4), for square and non-square
shapes from 16 to 512 bins and both field modes.  A placement on it
converges like one on the MFMA engine.
"""
import numpy as np
import pytest

import _oracle as O
import eisenjoch_amd as ep
from conftest import log_case

pytestmark = pytest.mark.gpu

FFT = 5


@pytest.mark.parametrize("mode", ["idxst", "reference"])
@pytest.mark.parametrize("Bx,By", [(16, 16), (16, 32), (64, 64), (128, 64), (64, 128), (32, 256),
                                   (256, 256), (512, 512), (512, 16)])
def test_fft_engine_gives_the_field(device, mode, Bx, By):
    W = max(Bx, By)
    d = ep.generate(64, 64, W, n_fixed=0, Bx=Bx, By=By, seed=3)
    log_case(f"fft field {Bx}x{By} {mode}")
    device.eval_config(ep.ElectroPlaceCfg(field_mode=mode))
    device.tune_set(Bx, By, FFT)
    i = device.upload(d, "fast")
    try:
        dens = np.random.default_rng(Bx * 7 + By).uniform(0, 3, (By, Bx))
        ex, ey = device.field(i, dens)
        oex, oey = O.field(dens, Bx, By, 1 if mode == "idxst" else 0)
        scale = max(np.abs(oex).max(), np.abs(oey).max())
        err = max(np.abs(ex - oex).max(), np.abs(ey - oey).max()) / scale
        print(f"  relative error {err:.2e}", flush=True)
        assert err <= 2e-6
    finally:
        device.free(i)
        device.tune_set(Bx, By, -1)
        device.eval_config(ep.ElectroPlaceCfg())


def test_fft_engine_single_modes(device):
    # a density of one cosine mode (u, v) gives Ex, Ey proportional to one
    # sine-cosine mode: checked against the closed form through the oracle
    # on a mode near the top of the spectrum (the last twiddles of every
    # stage) and on the constant mode (zero field)
    B = 256
    d = ep.generate(64, 64, B, n_fixed=0, Bx=B, By=B, seed=4)
    device.eval_config(ep.ElectroPlaceCfg(field_mode="idxst"))
    device.tune_set(B, B, FFT)
    i = device.upload(d, "fast")
    try:
        x = (np.arange(B) + 0.5) / B
        for (u, v) in [(B - 1, B - 2), (1, 0), (0, 0)]:
            log_case(f"fft single mode ({u}, {v})")
            dens = 1.0 + np.outer(np.cos(np.pi * v * x), np.cos(np.pi * u * x))
            ex, ey = device.field(i, dens)
            oex, oey = O.field(dens, B, B, 1)
            scale = max(np.abs(oex).max(), np.abs(oey).max(), 1e-30)
            assert np.abs(ex - oex).max() <= 2e-6 * max(scale, 1.0)
            assert np.abs(ey - oey).max() <= 2e-6 * max(scale, 1.0)
    finally:
        device.free(i)
        device.tune_set(B, B, -1)
        device.eval_config(ep.ElectroPlaceCfg())


def test_fft_engine_placement_matches_mfma(device):
    # one 256-bin design placed on the FFT engine and on engine 0: both reach
    # overflow <= 0.10 in as many iterations as the plain FP64 placer, with
    # HPWL within 0.25 % of its (half the 0.5 % bar of the FAST path; the two
    # FAST runs end 0.01-0.1 % from it, differences in rounding moving the
    # end point of the trajectory by that much) and of each other
    d = ep.sized_design(40000, 33)
    assert d.Bx == 256
    cfg = ep.ElectroPlaceCfg(max_iters=500, field_mode="idxst", log_hpwl=False)
    i64 = device.upload(d, "fp64")
    i = device.upload(d, "fast")
    try:
        log_case("plain FP64 placement at 256 bins")
        r64 = device.place(i64, cfg, positions=False)
        out = {}
        for e in (0, FFT):
            log_case(f"placement with engine {e} at 256 bins")
            device.tune_set(256, 256, e)
            out[e] = device.place(i, cfg, positions=False)
        for e, r in out.items():
            assert r.converged and r.overflow <= 0.10
            assert abs(r.iters - r64.iters) <= 2
            dev = abs(r.hpwl - r64.hpwl) / r64.hpwl
            print(f"  engine {e}: HPWL {dev:.2e} from FP64", flush=True)
            assert dev < 2.5e-3
        assert abs(out[FFT].hpwl - out[0].hpwl) / out[0].hpwl < 2.5e-3
    finally:
        device.free(i)
        device.free(i64)
        device.tune_set(256, 256, -1)


def test_fft_engine_rejects_other_sizes(device):
    with pytest.raises(ValueError, match="powers of two"):
        device.tune_set(96, 64, FFT)
    with pytest.raises(ValueError, match="powers of two"):
        device.tune_set(1024, 64, FFT)
    with pytest.raises(ValueError, match="powers of two"):
        device.tune_set(8, 64, FFT)
    device.tune_set(16, 512, FFT)
    device.tune_set(16, 512, -1)
