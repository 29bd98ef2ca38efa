"""Writes tests/golden/fft_vectors.npz: numpy (pocketfft) DFTs of seeded data.

numpy's FFT is an independent implementation of the DFT definition; the
fixtures pin the oracle's transforms and its perm packing
(src/mod_fft.cc:79-86) without reference to the oracle's own code.
Run: python tests/golden/make_fft_vectors.py
"""
import os

import numpy as np

SIZES = [4, 6, 8, 10, 12, 20, 24, 30, 48, 64, 96, 192, 384]


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for n in SIZES:
        x = rng.standard_normal(n)
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x_{n}"] = x
        out[f"rfft_{n}"] = np.fft.rfft(x)
        out[f"z_{n}"] = z
        out[f"fft_{n}"] = np.fft.fft(z)
        out[f"ifft_unscaled_{n}"] = np.fft.ifft(z) * n
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
