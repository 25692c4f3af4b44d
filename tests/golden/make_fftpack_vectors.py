"""Writes tests/golden/fftpack_filter.npz: filter_mod.F90's row filter
(rfft1f -> 0/1 mask -> rfft1b, filter_mod.F90:119-124) computed by the
reference's own FFTPACK 5.1 (lib/fft, compiled into oracle/_ref by
oracle/Makefile) on seeded rows, for row lengths with different radix
factorisations (rffti1.f tries 4, 2, 3, 5).  Also stores the raw forward
coefficients so the FFTPACK5 layout/normalisation is pinned.

Run in the container that holds the reference tree:
    make -C oracle ref && python tests/golden/make_fftpack_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import RefFFTPACK  # noqa: E402


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for n in (12, 16, 24, 90, 180, 360, 720):
        fft = RefFFTPACK(n)
        x = rng.standard_normal(n)
        out[f"x_{n}"] = x
        out[f"coef_{n}"] = fft.forward(x)
        cuts = [c for c in (0, 1, 2, 3, n // 8, n // 4, n // 2 - 1) if 2 * c + 2 <= n]
        ys = []
        for c in cuts:
            mask = np.zeros(n)
            mask[: 2 * c + 2] = 1.0
            ys.append(fft.filter(x, mask))
        out[f"cuts_{n}"] = np.array(cuts, dtype=np.int32)
        out[f"y_{n}"] = np.array(ys)
    np.savez_compressed(os.path.join(HERE, "fftpack_filter.npz"), **out)
    print("wrote", os.path.join(HERE, "fftpack_filter.npz"))


if __name__ == "__main__":
    main()
