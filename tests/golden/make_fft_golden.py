"""Write small FFT golden vectors with numpy.fft (a textbook library routine,
not the reference): seeded real inputs and their numpy.fft.rfftn spectra.

Run from the repo root:  python tests/golden/make_fft_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(4, 6, 8), (6, 4, 4), (8, 8, 8), (12, 6, 12), (4, 12, 24)]


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for s in SHAPES:
        x = rng.standard_normal(s)
        key = "x_%dx%dx%d" % s
        out[key] = x
        out["X_%dx%dx%d" % s] = np.fft.rfftn(x)
    np.savez_compressed(os.path.join(HERE, "fft_numpy.npz"), **out)


if __name__ == "__main__":
    main()
