"""Write the small golden vectors used to pin the oracle.

The reference ships no stored vectors (its sample drivers check analytic
identities at run time) and its 1D transforms come from FFTW, which is not
available here.  These fixtures come from an independent FP64 FFT (numpy /
scipy pocketfft) with the FFTW conventions the reference relies on:
unnormalised transforms, forward sign e^{-i}, r2c along the first Fortran
index (x), c2r dropping the imaginary parts of kx=0 and kx=nx/2, DCT-I
(REDFT00) and DST-I (RODFT00) applied to the real and imaginary parts.

Arrays are stored in numpy order (nz, ny, nx) = Fortran (nx, ny, nz).
Run:  python tests/golden/gen_golden.py
"""
import os

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    # (nx, ny, nz): powers of two, radix 3/5/7, odd and prime lengths
    shapes = [(8, 6, 10), (16, 16, 16), (12, 9, 5), (7, 11, 13), (10, 14, 6), (6, 4, 17)]
    for (nx, ny, nz) in shapes:
        tag = "%dx%dx%d" % (nx, ny, nz)
        a = rng.random((nz, ny, nx))
        out["r2c_in_" + tag] = a
        out["r2c_out_" + tag] = np.fft.rfftn(a)                 # 'fft': x r2c, y, z c2c
        spec = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
        out["c2r_in_" + tag] = spec
        out["c2r_out_" + tag] = np.fft.irfftn(spec, s=(nz, ny, nx), axes=(0, 1, 2)) * (nx * ny * nz)   # 'tff'
        # Z = DCT-I / DST-I on re and im ('ffc' / 'ffs'), Z none ('ffn')
        xy = np.fft.fft(np.fft.rfft(a, axis=2), axis=1)
        if nz >= 2:
            out["ffc_out_" + tag] = sf.dct(xy.real, type=1, axis=0) + 1j * sf.dct(xy.imag, type=1, axis=0)
        out["ffs_out_" + tag] = sf.dst(xy.real, type=1, axis=0) + 1j * sf.dst(xy.imag, type=1, axis=0)
        out["ffn_out_" + tag] = xy
        # backward with Z = DCT-I ('cff') / DST-I ('sff'), then Y ifft, X c2r
        if nz >= 2:
            zc = sf.dct(spec.real, type=1, axis=0) + 1j * sf.dct(spec.imag, type=1, axis=0)
            out["cff_out_" + tag] = np.fft.irfft(np.fft.ifft(zc, axis=1) * ny, n=nx, axis=2) * nx
        zs = sf.dst(spec.real, type=1, axis=0) + 1j * sf.dst(spec.imag, type=1, axis=0)
        out["sff_out_" + tag] = np.fft.irfft(np.fft.ifft(zs, axis=1) * ny, n=nx, axis=2) * nx
    # 1D vectors
    for n in (1, 2, 3, 5, 7, 8, 12, 17, 30, 64, 97, 128, 257):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out["fft1_in_%d" % n] = x
        out["fft1_out_%d" % n] = np.fft.fft(x)
        out["ifft1_out_%d" % n] = np.fft.ifft(x) * n
        r = rng.standard_normal(n)
        out["real1_in_%d" % n] = r
        if n >= 2:
            out["dct1_out_%d" % n] = sf.dct(r, type=1)
        out["dst1_out_%d" % n] = sf.dst(r, type=1)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
