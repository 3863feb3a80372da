"""Writes field_golden.npz: density -> (Ex, Ey) vectors of the reference's
spectral solve (crates/nextpnr/src/placer/electro_place/density.rs:96-166)
computed with scipy.fft, an implementation independent of the oracle's.

scipy's unnormalised dct(type=2), dct(type=3) and dst(type=3) are exactly
twice rustdct 0.7.1's DCT2, DCT3 and DST3, so each transform is divided by 2.
Run from the repository root: python tests/golden/make_golden.py
"""
from pathlib import Path

import numpy as np
import scipy
import scipy.fft as sf


def field_scipy(d, mode):
    By, Bx = d.shape
    rho = d - 1.0 * max(d.mean(), 1e-30)
    A = sf.dct(sf.dct(rho, type=2, axis=1) / 2, type=2, axis=0) / 2
    kx = np.arange(Bx)[None, :]
    ky = np.arange(By)[:, None]
    s = 4.0 / (Bx * By) * np.where(kx == 0, 0.5, 1.0) * np.where(ky == 0, 0.5, 1.0)
    a = A * s
    wx = np.pi * kx / Bx
    wy = np.pi * ky / By
    den = 2 * (1 - np.cos(wx)) + 2 * (1 - np.cos(wy))
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = np.where(den < 1e-30, 0.0, a / den)
    exh = phi * wx
    eyh = phi * wy
    if mode == 1:  # index-shifted, negated spectrum (textbook IDXST)
        exh = np.concatenate([-exh[:, 1:], np.zeros((By, 1))], axis=1)
        eyh = np.concatenate([-eyh[1:, :], np.zeros((1, Bx))], axis=0)
    ex = sf.dct(sf.dst(exh, type=3, axis=1) / 2, type=3, axis=0) / 2
    ey = sf.dst(sf.dct(eyh, type=3, axis=1) / 2, type=3, axis=0) / 2
    return ex, ey


def main():
    rng = np.random.default_rng(20261203)
    out = {"scipy_version": np.array(scipy.__version__)}
    shapes = [(2, 2), (4, 4), (8, 8), (64, 64), (4, 8), (6, 10)]
    for (By, Bx) in shapes:
        d = rng.uniform(0.0, 2.0, size=(By, Bx))
        out[f"dens_{By}x{Bx}"] = d
        for mode in (0, 1):
            ex, ey = field_scipy(d, mode)
            out[f"ex_{By}x{Bx}_m{mode}"] = ex
            out[f"ey_{By}x{Bx}_m{mode}"] = ey
    for n in (1, 2, 3, 5, 8, 12, 64, 100):
        v = rng.standard_normal(n)
        out[f"vec_{n}"] = v
        out[f"dct2_{n}"] = sf.dct(v, type=2) / 2
        out[f"dct3_{n}"] = sf.dct(v, type=3) / 2
        out[f"dst3_{n}"] = sf.dst(v, type=3) / 2
    path = Path(__file__).resolve().parent / "field_golden.npz"
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
