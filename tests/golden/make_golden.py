"""Writes tests/golden/fft_golden.npz: numpy.fft (pocketfft) FP64 vectors that pin
the oracle's FFT, xtok and ktox independently of the oracle's own code.

Run from the repository root:  python tests/golden/make_golden.py
Contents (all float64 / complex128):
  c{n}_in, c{n}_fwd, c{n}_inv   1-D complex, n in 8,16,32,64,128,256,12
                                (fwd = np.fft.fft, inv = n * np.fft.ifft)
  r{nx}x{ny}_in, r{nx}x{ny}_rfft2   2-D real field and np.fft.rfft2 of it
  k{nx}x{ny}_in, k{nx}x{ny}_irfft2  padded half spectrum and nx*ny*np.fft.irfft2
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20241203)
    out = {}
    for n in (8, 16, 32, 64, 128, 256, 12):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"c{n}_in"] = x
        out[f"c{n}_fwd"] = np.fft.fft(x)
        out[f"c{n}_inv"] = np.fft.ifft(x) * n
    for nx, ny in ((32, 48), (64, 64)):
        f = rng.standard_normal((nx, ny))
        out[f"r{nx}x{ny}_in"] = f
        out[f"r{nx}x{ny}_rfft2"] = np.fft.rfft2(f)
        P = rng.standard_normal((nx, ny // 2 + 1)) + 1j * rng.standard_normal((nx, ny // 2 + 1))
        out[f"k{nx}x{ny}_in"] = P
        out[f"k{nx}x{ny}_irfft2"] = np.fft.irfft2(P, s=(nx, ny)) * nx * ny
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
