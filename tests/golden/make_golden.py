"""Writes tests/golden/golden_fft.npz: numpy.fft (pocketfft, bundled with
numpy 2.2) forward and inverse transforms of seeded SplitMix64 inputs.

These fixtures are an independent FFT implementation's outputs; they pin the
CPU oracle (tests/test_oracle.py), which in turn is the checker for the GPU
path.  Inputs are regenerated from the seed; each entry stores the input's
checksum, the forward output and the inverse output (numpy's ifft, which
divides by N like the reference's INVERSE).

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "tope-fft_amd", "python"))
import workloads  # noqa: E402

CASES_1D = [1, 2, 3, 4, 5, 7, 8, 12, 16, 17, 32, 64, 100, 128, 256, 512, 1024, 2048, 4096]
CASES_ND = [(4, 4), (8, 4), (4, 16), (16, 16), (64, 32), (2, 2, 2), (4, 4, 4), (8, 16, 4), (16, 8, 32)]


def case_seed(dims):
    h = 0x60D
    for d in dims:
        h = (h * 1000003 + d) & 0xFFFFFFFF
    return h


def main():
    data = {}
    for dims in [(n,) for n in CASES_1D] + CASES_ND:
        n = int(np.prod(dims))
        x = workloads.complex_uniform(case_seed(dims), n)
        shape = tuple(reversed(dims))  # x fastest
        xs = x.reshape(shape)
        fwd = np.fft.fftn(xs).reshape(-1)
        inv = np.fft.ifftn(xs).reshape(-1)
        key = "x".join(str(d) for d in dims)
        data[f"in_sum_{key}"] = np.array([x.sum()])
        data[f"fwd_{key}"] = fwd
        data[f"inv_{key}"] = inv
    out = os.path.join(HERE, "golden_fft.npz")
    np.savez_compressed(out, **data)
    print("wrote", out, len(data) // 3, "cases")


if __name__ == "__main__":
    main()
