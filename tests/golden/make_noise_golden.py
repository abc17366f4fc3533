"""Writes tests/golden/fft_filter_24x20.json: a zero-padded Gaussian FFT filter
(noise.jl:244-270 semantics) of a seeded 24x20 field, computed with numpy's
pocketfft.  Run from the repo root: python tests/golden/make_noise_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(2024)
    Nz, Nx, pad, h = 24, 20, 8, 0.1
    sigma = 0.15
    psi = rng.standard_normal((Nz, Nx))
    hp = pad // 2
    a = np.zeros((Nz + pad, Nx), dtype=complex)
    a[hp:hp + Nz] = psi
    kz = 2 * np.pi * np.fft.fftfreq(Nz + pad, d=h)
    kx = 2 * np.pi * np.fft.fftfreq(Nx, d=h)
    G = np.exp(-sigma ** 2 * (kz[:, None] ** 2 + kx[None, :] ** 2))
    out = np.real(np.fft.ifft2(np.fft.fft2(a) * G))[hp:hp + Nz]
    out = (out - out.mean()) * psi.std(ddof=1) / out.std(ddof=1) + psi.mean()
    with open(os.path.join(HERE, "fft_filter_24x20.json"), "w") as fh:
        json.dump({"pad": pad, "h": h, "sigma": sigma, "psi": psi.tolist(), "out": out.tolist()}, fh)


if __name__ == "__main__":
    main()
