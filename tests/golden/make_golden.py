"""Writes tests/golden/fft_golden.npz: seeded real fields and their numpy.fft.rfftn transforms.

numpy's pocketfft is used only here, as an independent library routine, to produce committed
FFT vectors that pin the oracle's own Stockham FFT (the reference's cuFFT cannot run in this
image and ships no vectors).  Run from the repository root: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(16, 16, 16), (24, 24, 24), (96, 8, 12), (8, 48, 6)]


def main():
    out = {}
    rng = np.random.default_rng(20261202)
    for i, n in enumerate(SHAPES):
        x = rng.standard_normal(n)
        out[f"x{i}"] = x
        out[f"X{i}"] = np.fft.rfftn(x)  # [kx][ky][kz], unnormalised
        out[f"shape{i}"] = np.array(n)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
