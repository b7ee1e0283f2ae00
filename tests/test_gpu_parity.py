"""GPU parity of the HIP path against the FP64 CPU oracle, through the C-ABI.

Tolerances (relative L2 unless stated):
  FP64 library vs FP64 oracle:  1e-12
  FP32 library (compensated hi+lo twiddles) vs FP64 oracle:  2e-6
  round trips additionally meet the reference drivers' absolute bound
  prec * N / 4 with prec = 1e-14 (FP64) or 1e-5 (FP32)
  (sample/FORTRAN/driver_sine.F90:304-314).
"""
import contextlib

import numpy as np
import pytest

import p3dfft_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {8: 1e-12, 4: 2e-6}
PREC = {8: 1e-14, 4: 1e-5}


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert p3dfft_amd.device_count() > 0, "no GPU visible to the HIP runtime"


@contextlib.contextmanager
def lib_setup(precision, nx, ny, nz, nxc=None, nyc=None, nzc=None, ow=True, stride1=False):
    P = p3dfft_amd.P3DFFT(precision)
    P.setup((1, 1), nx, ny, nz, nxc, nyc, nzc, overwrite=ow, stride1=stride1)
    try:
        yield P
    finally:
        P.clean()


def oplan(nx, ny, nz, nxc=None, nyc=None, nzc=None, ow=True, stride1=False):
    return O.Plan(nx, ny, nz, nxc or 0, nyc or 0, nzc or 0, ow=ow, stride1=stride1)


# ---------------------------------------------------------------- C1: 64^3 FP64
def test_c1_64cube_fp64_against_oracle():
    nx = ny = nz = 64
    a = np.random.default_rng(1).random((nz, ny, nx))
    ref = oplan(nx, ny, nz).ftran_r2c(a)
    with lib_setup(8, nx, ny, nz) as P:
        assert P.memsize == (64, 64, 66)
        assert P.get_dims(2) == ((1, 1, 1), (33, 64, 64), (33, 64, 64))
        s = P.ftran_r2c(a, "fft")
        assert rel(s, ref) < 1e-12
        back = P.btran_c2r(s / (nx * ny * nz), "tff")
    assert np.abs(back - a).max() < PREC[8] * nx * ny * nz * 0.25
    assert rel(back, a) < 1e-13


# ---------------------------------------------------------------- every Z option
SHAPES = [(8, 6, 10), (16, 16, 16), (12, 9, 5), (7, 11, 13), (10, 14, 6), (6, 4, 17), (32, 24, 20)]


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape", SHAPES)
def test_forward_ops(precision, shape):
    nx, ny, nz = shape
    a = np.random.default_rng(sum(shape)).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz)
    with lib_setup(precision, nx, ny, nz) as P:
        for op in ("fft", "fff", "ffc", "ffs", "ffn", "ff0"):
            if op == "ffc" and nz < 2:
                continue
            assert rel(P.ftran_r2c(a, op), o.ftran_r2c(a, op)) < TOL[precision], op


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("ow", [True, False])
def test_backward_ops(precision, shape, ow):
    nx, ny, nz = shape
    rng = np.random.default_rng(sum(shape) + 1)
    s = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    o = oplan(nx, ny, nz, ow=ow)
    with lib_setup(precision, nx, ny, nz, ow=ow) as P:
        for op in ("tff", "fff", "cff", "sff", "nff", "0ff"):
            assert rel(P.btran_c2r(s, op), o.btran_c2r(s, op)[0]) < TOL[precision], op


# ---------------------------------------------------------------- transform lengths
@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape", [(14, 26, 38), (22, 39, 17), (96, 160, 48), (30, 20, 44), (19, 23, 29),
                                   (128, 40, 56), (34, 2, 3), (2, 3, 2)])
def test_radix_coverage_round_trip(precision, shape):
    """radix 2/3/4/5/7/8/11/13, generic primes (17, 19, 23, 29) and the
    reference's uneven grid 14 26 38 (extra/makejob.py:142-157)."""
    nx, ny, nz = shape
    N = nx * ny * nz
    a = np.random.default_rng(N).random((nz, ny, nx))
    o = oplan(nx, ny, nz)
    ref = o.ftran_r2c(a)
    with lib_setup(precision, nx, ny, nz) as P:
        s = P.ftran_r2c(a)
        assert rel(s, ref) < TOL[precision]
        back = P.btran_c2r(s, "tff") / N
    assert np.abs(back - a).max() < PREC[precision] * N * 0.25


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape", [(64, 48, 32), (96, 128, 160), (192, 40, 65), (384, 20, 33)])
def test_generic_and_specialised_kernels_agree(precision, shape):
    """Lengths with compile-time specialised kernels, run through both the
    specialised and the run-time-generic engine, for every Z option."""
    nx, ny, nz = shape
    a = np.random.default_rng(7).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz)
    lib = p3dfft_amd.load(precision)
    try:
        for spec_on in (1, 0):
            lib.p3dfft_amd_set_specialised(spec_on)
            with lib_setup(precision, nx, ny, nz) as P:
                for op in ("fft", "ffc", "ffs"):
                    s = P.ftran_r2c(a, op)
                    assert rel(s, o.ftran_r2c(a, op)) < TOL[precision], (op, spec_on)
                    bop = op[2] + "ff"
                    assert rel(P.btran_c2r(s, bop), o.btran_c2r(s.astype(np.complex128), bop)[0]) \
                        < TOL[precision], (bop, spec_on)
    finally:
        lib.p3dfft_amd_set_specialised(1)


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("stride1", [False, True])
def test_chunked_xy_passes(precision, stride1):
    """X/Y passes run z-plane chunk by chunk (chunk of 3 planes here, with a
    ragged last chunk) give the oracle's result, also with pruning, in place
    on the device and for every Z option."""
    nx, ny, nz = 24, 20, 17
    rng = np.random.default_rng(12)
    a = rng.standard_normal((nz, ny, nx))
    lib = p3dfft_amd.load(precision)
    plane = (nx // 2 + 1) * ny * (16 if precision == 8 else 8)
    lib.p3dfft_amd_set_chunk_bytes(3 * plane)
    try:
        for cut in ((None, None, None), (16, 14, 12)):
            o = oplan(nx, ny, nz, *cut, stride1=stride1)
            with lib_setup(precision, nx, ny, nz, *cut, stride1=stride1) as P:
                for op in ("fft", "ffc", "ffs", "ffn"):
                    s = P.ftran_r2c(a, op)
                    assert rel(s, o.ftran_r2c(a, op)) < TOL[precision], (cut, op)
                    bop = op[2] + "ff"
                    assert rel(P.btran_c2r(s, bop), o.btran_c2r(s.astype(np.complex128), bop)[0]) \
                        < TOL[precision], (cut, bop)
                if cut[0] is None:
                    buf = P.alloc(max(P.real_bytes(), P.spec_bytes()))
                    buf.upload(a.astype(P.rdtype))
                    P.ftran_r2c_dev(buf.ptr, buf.ptr, "fft")
                    s = buf.download(P.spec_shape(), P.cdtype)
                    assert rel(s, o.ftran_r2c(a)) < TOL[precision]
                    P.btran_c2r_dev(buf.ptr, buf.ptr, "tff")
                    back = buf.download(P.real_shape(), P.rdtype) / (nx * ny * nz)
                    assert rel(back, a) < TOL[precision] * 10
                    buf.free()
    finally:
        lib.p3dfft_amd_set_chunk_bytes(0)  # the library default: no chunking


# ---------------------------------------------------------------- pruning, layout, OW
CUTS = [(12, 12, 10), (8, 12, 10), (6, 8, 6), (10, 9, 7), (12, 7, 8), (7, 12, 9)]


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("cut", CUTS)
@pytest.mark.parametrize("stride1", [False, True])
def test_pruned_forward(precision, cut, stride1):
    nx, ny, nz = 12, 12, 10
    a = np.random.default_rng(3).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz, *cut, stride1=stride1)
    with lib_setup(precision, nx, ny, nz, *cut, stride1=stride1) as P:
        assert P.get_dims(2) == o.get_dims(2)
        assert P.get_dims(3) == o.get_dims(3)
        for op in ("fft", "ffc", "ffs", "ffn"):
            assert rel(P.ftran_r2c(a, op), o.ftran_r2c(a, op)) < TOL[precision], op


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("cut", CUTS)
@pytest.mark.parametrize("stride1", [False, True])
@pytest.mark.parametrize("ow", [True, False])
def test_pruned_backward(precision, cut, stride1, ow):
    nx, ny, nz = 12, 12, 10
    o = oplan(nx, ny, nz, *cut, ow=ow, stride1=stride1)
    rng = np.random.default_rng(4)
    shp = o.spec_shape()
    s = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    with lib_setup(precision, nx, ny, nz, *cut, ow=ow, stride1=stride1) as P:
        for op in ("tff", "cff", "sff", "nff"):
            assert rel(P.btran_c2r(s, op), o.btran_c2r(s, op)[0]) < TOL[precision], op


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("ow", [True, False])
def test_device_api_overwrite_and_in_place(precision, ow):
    """btran with overwrite and nz==nzc transforms the caller's (device) input
    in place along Z (build/btran.F90:525-550), as the oracle restates;
    without overwrite the input is preserved.  IN==OUT works both ways."""
    nx, ny, nz = 16, 12, 10
    rng = np.random.default_rng(6)
    a = rng.standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz, ow=ow)
    ref_s = o.ftran_r2c(a)
    ref_b, ref_in_after = o.btran_c2r(ref_s.copy(), "tff", keep_input=False)
    with lib_setup(precision, nx, ny, nz, ow=ow) as P:
        nb = max(P.real_bytes(), P.spec_bytes())
        buf = P.alloc(nb)
        buf.upload(a.astype(P.rdtype))
        P.ftran_r2c_dev(buf.ptr, buf.ptr, "fft")          # in place
        s = buf.download(P.spec_shape(), P.cdtype)
        assert rel(s, ref_s) < TOL[precision]
        out = P.alloc(nb)
        P.btran_c2r_dev(buf.ptr, out.ptr, "tff")
        P.synchronize()
        assert rel(out.download(P.real_shape(), P.rdtype), ref_b) < TOL[precision]
        after = buf.download(P.spec_shape(), P.cdtype)
        if ow:
            # the destroyed input holds the Z result with the unnormalised
            # inverse Y transform applied in place (the reference leaves the
            # Z result there; both are documented as overwritten)
            assert rel(after, ny * np.fft.ifft(ref_in_after, axis=1)) < TOL[precision]
        else:
            assert rel(after, ref_in_after) < TOL[precision]
        if not ow:
            assert np.array_equal(after, s)
        buf.free(); out.free()


# ---------------------------------------------------------------- Chebyshev
def cheby_tol(precision, nzc, Lz):
    """The derivative recurrence d_k = sum_m (4/Lz) m a_m (m > k, m - k odd)
    weighs each coefficient error by up to (4/Lz) nzc, so the FP32 bound is the
    FFT bound times that factor; FP64 stays at 1e-12."""
    return TOL[precision] if precision == 8 else TOL[4] * (4.0 / Lz) * nzc


def cheby_points(nz, Lz):
    return np.cos(np.pi * np.arange(nz) / (nz - 1)) * 2.0 / Lz


@pytest.mark.parametrize("precision", [8, 4])
def test_driver_cheby(precision):
    """sample/FORTRAN/driver_cheby.F90:218-285 restated: cheby of sin(z) on
    Gauss-Lobatto points, end coefficients doubled, halved, btran 'cff' = cos(z)."""
    nx, ny, nz, Lz = 32, 32, 33, 2.0
    z = cheby_points(nz, Lz)
    a = np.broadcast_to(np.sin(z)[:, None, None], (nz, ny, nx)).copy()
    o = oplan(nx, ny, nz)
    with lib_setup(precision, nx, ny, nz) as P:
        c = P.cheby(a, Lz)
        assert rel(c, o.cheby(a, Lz)) < cheby_tol(precision, nz, Lz)
        c = c.astype(np.complex128)
        c[0] *= 2.0
        c[nz - 1] *= 2.0
        c *= 0.5
        d = P.btran_c2r(c, "cff")
    assert np.abs(d - np.cos(z)[:, None, None]).max() < PREC[precision] * nx * ny * nz * 0.25


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("stride1", [False, True])
def test_cheby_random_polynomial(precision, stride1):
    """The Chebyshev configuration at reduced size: trig modes in x,y times a
    random degree-<=64 Chebyshev polynomial in z (coefficients N(0,1)/(1+n)^2)."""
    nx, ny, nz, Lz = 64, 64, 129, 2.0
    rng = np.random.default_rng(44)
    c = rng.standard_normal(65) / (1 + np.arange(65)) ** 2
    zz = cheby_points(nz, Lz)
    g = np.polynomial.chebyshev.chebval(zz, c)
    xs = np.arange(nx) / nx
    ys = np.arange(ny) / ny
    xy = sum(np.cos(2 * np.pi * (rng.integers(1, 8) * xs[None, :] + rng.integers(1, 8) * ys[:, None])
                    + rng.random() * 2 * np.pi) for _ in range(4))
    a = g[:, None, None] * xy[None]
    o = oplan(nx, ny, nz, stride1=stride1)
    with lib_setup(precision, nx, ny, nz, stride1=stride1) as P:
        assert rel(P.cheby(a, Lz), o.cheby(a, Lz)) < cheby_tol(precision, nz, Lz)


def test_c4_cheby_full_size_fp64():
    """The Chebyshev configuration at its full size, 256 x 256 x 257 FP64: a
    degree-64 Chebyshev polynomial in z (interpolated exactly on the 257
    Gauss-Lobatto points) times trig modes in x, y.  The GPU result against the
    oracle, and the driver_cheby round trip (end coefficients doubled, halved,
    btran 'cff') against numpy's chebder of the polynomial."""
    nx, ny, nz, Lz = 256, 256, 257, 2.0
    rng = np.random.default_rng(45)
    c = rng.standard_normal(65) / (1 + np.arange(65)) ** 2
    zz = cheby_points(nz, Lz)
    g = np.polynomial.chebyshev.chebval(zz, c)
    dg = np.polynomial.chebyshev.chebval(zz, np.polynomial.chebyshev.chebder(c))
    xs = np.arange(nx) / nx
    ys = np.arange(ny) / ny
    xy = np.cos(2 * np.pi * (3 * xs[None, :] + 5 * ys[:, None])) + 0.25
    a = g[:, None, None] * xy[None]
    o = oplan(nx, ny, nz)
    with lib_setup(8, nx, ny, nz) as P:
        got = P.cheby(a, Lz)
        assert rel(got, o.cheby(a, Lz)) < TOL[8]
        got[0] *= 2.0
        got[nz - 1] *= 2.0
        got *= 0.5
        d = P.btran_c2r(got, "cff")
    assert rel(d, dg[:, None, None] * xy[None]) < 1e-10


# ---------------------------------------------------------------- errors, timers
def test_error_paths():
    P = p3dfft_amd.P3DFFT(8)
    P.setup((1, 1), 8, 8, 8)
    try:
        with pytest.raises(p3dfft_amd.P3DFFTError, match="already initialized"):
            P.setup((1, 1), 8, 8, 8)
        with pytest.raises(p3dfft_amd.P3DFFTError, match="Unknown transform type"):
            P.ftran_r2c(np.zeros((8, 8, 8)), "ffx")
        with pytest.raises(p3dfft_amd.P3DFFTError, match="Unknown transform type"):
            P.btran_c2r(np.zeros((8, 8, 5), np.complex128), "xff")
    finally:
        P.clean()
    # a new setup is allowed after clean (module.F90:424 resets mpi_set)
    P.setup((1, 1), 4, 4, 4)
    P.clean()


@pytest.mark.parametrize("plane", [0, 1])
def test_timers_accumulate(plane):
    """Stage timers in the reference's slots (5 X, 7 Y, 8 Z forward; 9 Z,
    10 Y, 12 X backward); with the plane kernels (X and Y of a z-plane in
    one launch) the fused pass is timed in the X slot and the Y slots stay 0."""
    nx = ny = nz = 32
    a = np.random.default_rng(0).random((nz, ny, nx))
    lib = p3dfft_amd.load(8)
    lib.p3dfft_amd_set_plane_kernels(plane)
    try:
        with lib_setup(8, nx, ny, nz) as P:
            P.set_timers()
            assert P.get_timers() == [0.0] * 12
            s = P.ftran_r2c(a)
            P.btran_c2r(s)
            t = P.get_timers()
            busy = (5, 8, 9, 12) if plane else (5, 7, 8, 9, 10, 12)
            for slot in busy:
                assert t[slot - 1] > 0.0, slot
            for slot in range(1, 13):
                if slot not in busy:
                    assert t[slot - 1] == 0.0, slot
            P.set_timers()
            assert P.get_timers() == [0.0] * 12
    finally:
        lib.p3dfft_amd_set_plane_kernels(2)


# ---------------------------------------------------------------- C2 / C3 at full size
def modes_field(n, seed, dtype):
    """sum of 8 random-phase Fourier modes (integer wavenumbers in [1,32]) plus
    N(0, 1e-3^2) noise, evaluated in FP64."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 33, size=(8, 3))
    ph = rng.random(8) * 2 * np.pi
    x = np.arange(n) / n
    f = np.zeros((n, n, n))
    for m in range(8):
        f += np.cos(2 * np.pi * (k[m, 0] * x[None, None, :] + k[m, 1] * x[None, :, None]
                                 + k[m, 2] * x[:, None, None]) + ph[m])
    f += rng.normal(0.0, 1e-3, size=f.shape)
    return f.astype(dtype), k, ph


@pytest.mark.parametrize("precision", [8, 4])
def test_c2_512cube_modes_and_round_trip(precision):
    n = 512
    N = n ** 3
    f, k, ph = modes_field(n, 2026, np.float64 if precision == 8 else np.float32)
    with lib_setup(precision, n, n, n) as P:
        buf = P.alloc(P.spec_bytes())
        buf.upload(f)
        spec_d = P.alloc(P.spec_bytes())
        P.ftran_r2c_dev(buf.ptr, spec_d.ptr, "fft")
        s = spec_d.download(P.spec_shape(), P.cdtype)
        # each mode cos(2 pi k.x + ph) contributes N/2 e^{i ph} at +k (kx>0)
        for m in range(8):
            v = s[k[m, 2], k[m, 1], k[m, 0]]
            assert abs(v - N / 2 * np.exp(1j * ph[m])) < 1e-3 * N, m
        P.btran_c2r_dev(spec_d.ptr, buf.ptr, "tff")
        back = buf.download(P.real_shape(), P.rdtype).astype(np.float64) / N
        buf.free(); spec_d.free()
    err = back - f.astype(np.float64)
    assert np.abs(err).max() < PREC[precision] * N * 0.25
    assert np.linalg.norm(err) / np.linalg.norm(f.astype(np.float64)) < (1e-14 if precision == 8 else 2e-6)


def test_c2_512cube_fp64_spectrum_against_oracle():
    n = 512
    f, _, _ = modes_field(n, 7, np.float64)
    ref = oplan(n, n, n).ftran_r2c(f)
    with lib_setup(8, n, n, n) as P:
        s = P.ftran_r2c(f)
    assert rel(s, ref) < 1e-12


def test_c3_384cube_against_oracle():
    """The radix-3 grid 384^3 against the FP64 oracle, so that the length-384
    column and X kernels are checked against an independent FFT and not only
    against the GPU's own FP64 path: the FP64 spectrum within 1e-12, the FP32
    spectrum within 2e-6, and the FP64 round trip of the oracle's spectrum."""
    n = 384
    a = np.random.default_rng(385).standard_normal((n, n, n))
    o = oplan(n, n, n)
    ref = o.ftran_r2c(a)
    with lib_setup(8, n, n, n) as P:
        assert rel(P.ftran_r2c(a), ref) < 1e-12
        b = P.btran_c2r(ref.copy())
    assert rel(b, o.btran_c2r(ref)[0]) < 1e-12
    with lib_setup(4, n, n, n) as P:
        assert rel(P.ftran_r2c(a.astype(np.float32)), ref) < 2e-6


@pytest.mark.parametrize("n", [384, 512])
def test_c3_precision_fp32_against_fp64(n):
    """FP32 bulk with compensated twiddles vs the FP64 path on i.i.d. N(0,1)."""
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n))
    with lib_setup(8, n, n, n) as P:
        s64 = P.ftran_r2c(a)
        b64 = P.btran_c2r(s64) / n ** 3
    with lib_setup(4, n, n, n) as P:
        s32 = P.ftran_r2c(a.astype(np.float32))
        b32 = P.btran_c2r(s32.copy()) / n ** 3
    assert rel(s32, s64) < 2e-6
    assert rel(b32, a) < 2e-6
    assert rel(b64, a) < 1e-14


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape", [(12, 9, 5), (64, 48, 40), (35, 16, 8)])
def test_ftran_r2c_1d(precision, shape):
    """p3dfft_ftran_r2c_1d (build/ftran.F90:787-814): X r2c of every row, all
    nx/2+1 modes even under a cut setup, against numpy's rfft."""
    nx, ny, nz = shape
    P = p3dfft_amd.P3DFFT(precision)
    P.setup((1, 1), nx, ny, nz, nxc=max(1, nx - 3))
    try:
        rng = np.random.default_rng(5)
        f = rng.standard_normal((nz, ny, nx))
        got = P.ftran_r2c_1d(f)
        ref = np.fft.rfft(f.astype(P.rdtype).astype(np.float64), axis=2)
        tol = 1e-12 if precision == 8 else 2e-6
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol
    finally:
        P.clean()


@pytest.mark.parametrize("precision", [8, 4])
def test_tile_order_bit_identical(precision):
    """Walking the tiles of alternate passes in reverse changes only the
    schedule: fwd+bwd results are bit-identical either way."""
    shape = (64, 48, 40)
    nx, ny, nz = shape
    f = np.random.default_rng(9).standard_normal((nz, ny, nx))
    res = []
    for alt in (1, 0):
        P = p3dfft_amd.P3DFFT(precision)
        P.lib.p3dfft_amd_set_tile_order(alt)
        P.setup((1, 1), nx, ny, nz)
        try:
            s = P.ftran_r2c(f, "fft")
            b = P.btran_c2r(s, "tff")
            res.append((s, b))
        finally:
            P.clean()
            P.lib.p3dfft_amd_set_tile_order(1)
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,ops", [((64, 48, 40), ("fft", "tff")), ((254, 512, 16), ("fft", "tff")),
                                       ((254, 512, 17), ("ffc", "cff"))])
def test_flat_order_bit_identical(precision, shape, ops):
    """The Z pass's tiles in XCD-aware order (0), dispatch order (1) or the
    order chosen by shape (2) change only the schedule: identical bits.  At
    254 x 512 the Z view's rows are 128*512 complex apart, 512 KiB (FP32) or
    1 MiB (FP64), where the shape rule picks dispatch order."""
    nx, ny, nz = shape
    f = np.random.default_rng(31).standard_normal((nz, ny, nx))
    res = []
    for mode in (0, 1, 2):
        P = p3dfft_amd.P3DFFT(precision)
        P.lib.p3dfft_amd_set_flat_order(mode)
        P.setup((1, 1), nx, ny, nz)
        try:
            s = P.ftran_r2c(f, ops[0])
            res.append((s, P.btran_c2r(s, ops[1])))
        finally:
            P.clean()
            P.lib.p3dfft_amd_set_flat_order(2)
    for s, b in res[1:]:
        assert np.array_equal(res[0][0], s) and np.array_equal(res[0][1], b)
    scale = float(nx * ny * (2 * (nz - 1) if ops[0][2] == "c" else nz))
    tol = 1e-4 if precision == 4 else 1e-11
    assert np.abs(res[0][1] / scale - f).max() < tol * np.abs(f).max()


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,ops", [((64, 48, 40), ("fft", "tff")), ((36, 20, 30), ("fft", "tff"))])
def test_wide_offsets_bit_identical(precision, shape, ops):
    """The 64-bit-offset column kernels (the path of views over 4 GiB) give
    the same bits as the 32-bit-offset ones."""
    nx, ny, nz = shape
    f = np.random.default_rng(23).standard_normal((nz, ny, nx))
    res = []
    for wide in (0, 1):
        P = p3dfft_amd.P3DFFT(precision)
        P.lib.p3dfft_amd_set_wide_offsets(wide)
        P.setup((1, 1), nx, ny, nz)
        try:
            s = P.ftran_r2c(f, ops[0])
            res.append((s, P.btran_c2r(s, ops[1])))
        finally:
            P.clean()
            P.lib.p3dfft_amd_set_wide_offsets(0)
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,op", [((64, 48, 40), "fft"), ((30, 18, 22), "ffc"), ((36, 20, 30), "ffn")])
def test_x_into_out_bit_identical(precision, shape, op):
    """X r2c straight into the output with Y in place there computes every
    line as the work-buffer schedule does: identical spectra."""
    nx, ny, nz = shape
    f = np.random.default_rng(29).standard_normal((nz, ny, nx))
    res = []
    for on in (1, 0):
        P = p3dfft_amd.P3DFFT(precision)
        P.lib.p3dfft_amd_set_x_into_out(on)
        P.setup((1, 1), nx, ny, nz)
        try:
            res.append(P.ftran_r2c(f, op))
        finally:
            P.clean()
            P.lib.p3dfft_amd_set_x_into_out(1)
    assert np.array_equal(res[0], res[1])


def test_in_place_backward_matches_out_of_place():
    """btran with the output over its input (in-place use of one HBM array)
    runs the passes in an order that is safe for the overlap; results match
    the out-of-place call."""
    nx = ny = 256
    nz = 16
    f = np.random.default_rng(43).standard_normal((nz, ny, nx))
    _in_place_backward(nx, ny, nz, f)


def _in_place_backward(nx, ny, nz, f):
    with lib_setup(4, nx, ny, nz) as P:
        s = P.ftran_r2c(f, "fft")
        out_of_place = P.btran_c2r(s.copy(), "tff")
        L = P.lib
        nbytes = (nx // 2 + 1) * ny * nz * 8
        buf = L.p3dfft_amd_malloc(nbytes)
        try:
            sc = np.ascontiguousarray(s.astype(np.complex64))
            L.p3dfft_amd_memcpy_h2d(buf, sc.ctypes.data, nbytes)
            assert L.p3dfft_btran_c2r_dev(buf, buf, b"tff", None) == 0
            assert L.p3dfft_amd_synchronize() == 0
            back = np.empty((nz, ny, nx), np.float32)
            L.p3dfft_amd_memcpy_d2h(back.ctypes.data, buf, back.nbytes)
        finally:
            L.p3dfft_amd_free(buf)
    assert rel(back, out_of_place) < TOL[4] / 10


@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("shape", [(64, 48, 40), (60, 36, 50)])
def test_measured_plans_match_oracle(precision, shape):
    """Plan-time measurement (the counterpart of the reference's FFTW_MEASURE)
    chooses per length among the specialised kernels, generic kernels with
    three tile sizes and three column segments, and other radix orders, then
    the schedule switches; whatever it keeps computes the same transforms."""
    import ctypes
    nx, ny, nz = shape
    f = np.random.default_rng(47).standard_normal((nz, ny, nx))
    L = p3dfft_amd.load(precision)
    L.p3dfft_amd_set_plan_mode(1)
    try:
        with lib_setup(precision, nx, ny, nz) as P:
            s = P.ftran_r2c(f, "fft")
            b = P.btran_c2r(s, "tff")
        buf = ctypes.create_string_buffer(2048)
        assert L.p3dfft_amd_plan_report(buf, 2048) > 0
        rep = buf.value.decode()
        for n in set(shape):
            assert "n=%d:" % n in rep, rep
    finally:
        L.p3dfft_amd_set_plan_mode(0)
        assert L.p3dfft_amd_reset_plans() == 0
    ref = oplan(nx, ny, nz).ftran_r2c(f)
    assert rel(s, ref) < TOL[precision]
    assert np.abs(b / (nx * ny * nz) - f).max() < PREC[precision] * nx * ny * nz * 0.25


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,stride1", [((64, 64, 64), False), ((48, 40, 30), False), ((36, 20, 26), True),
                                           ((128, 96, 80), False), ((31, 17, 23), False)])
def test_power_spectrum_matches_oracle(precision, shape, stride1):
    """The drivers' power spectrum (sample/C/driver_spec.c:298-383) of the
    library's own spectrum, against the oracle's restatement on the same
    stored values: shells summed in FP64 in both, so they differ only by
    summation order; repeated calls are bit-identical; a kmax below the
    drivers' bound is refused (driver_spec.F90:272-275)."""
    nx, ny, nz = shape
    f = np.random.default_rng(53).standard_normal((nz, ny, nx))
    kmax = O.spectrum_kmax(nx, ny, nz)
    with lib_setup(precision, nx, ny, nz, stride1=stride1) as P:
        s = P.ftran_r2c(f, "fft")
        E = P.power_spectrum(s, kmax)
        E2 = P.power_spectrum(s, kmax)
        with pytest.raises(p3dfft_amd.P3DFFTError, match="kmax is less than it should be"):
            P.power_spectrum(s, 1)
    ref = O.power_spectrum(s.astype(np.complex128), (nx, ny, nz), kmax, stride1=stride1)
    assert np.array_equal(E, E2)
    assert np.abs(E - ref).max() <= 1e-13 * ref.sum()
    assert E.sum() > 0


def power_law_field(n, seed):
    """C3's power-law field: half-spectrum amplitude |k|^(-11/6) (shell energy
    ~ k^(-5/3) in 3D), uniform random phases, k = 0 set to 0, synthesised in
    FP64 by an inverse real FFT and scaled to unit RMS."""
    rng = np.random.default_rng(seed)
    kx = np.arange(n // 2 + 1)[None, None, :]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :, None]
    kz = np.fft.fftfreq(n, 1.0 / n)[:, None, None]
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    amp = np.where(k > 0, np.power(np.where(k > 0, k, 1.0), -11.0 / 6.0), 0.0)
    spec = amp * np.exp(2j * np.pi * rng.random(amp.shape))
    f = np.fft.irfftn(spec, s=(n, n, n), axes=(0, 1, 2))
    return f / np.sqrt(np.mean(f * f))


@pytest.mark.parametrize("n", [384, 512])
def test_c3_power_law_fp32_against_fp64(n):
    """FP32 bulk with compensated twiddles vs the FP64 path on the power-law
    field (energy concentrated in the lowest shells): forward spectra and
    round trips within the stated 2e-6 relative L2."""
    a = power_law_field(n, 3 * n)
    with lib_setup(8, n, n, n) as P:
        s64 = P.ftran_r2c(a)
        b64 = P.btran_c2r(s64.copy()) / n ** 3
    with lib_setup(4, n, n, n) as P:
        s32 = P.ftran_r2c(a.astype(np.float32))
        b32 = P.btran_c2r(s32.copy()) / n ** 3
    e_spec, e_rt = rel(s32, s64), rel(b32, a)
    print("power-law n=%d: FP32 vs FP64 spectrum %.3g, FP32 round trip %.3g, FP64 round trip %.3g"
          % (n, e_spec, e_rt, rel(b64, a)))
    assert e_spec < 2e-6
    assert e_rt < 2e-6
    assert rel(b64, a) < 1e-14


@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("shape", [(64, 48, 40), (60, 36, 50), (64, 64, 64)])
def test_captured_transforms_replay_bit_exact(precision, shape):
    """ftran_r2c_dev + btran_c2r_dev captured from a stream into a HIP graph
    (kernel launches only: no allocation, attribute call or memset) replay to
    the eager results bit for bit, three times, with eager transforms and
    uploads between the replays (64^3: the plane kernels)."""
    import ctypes
    nx, ny, nz = shape
    rng = np.random.default_rng(59)
    f = rng.standard_normal((nz, ny, nx))
    g2 = rng.standard_normal((nz, ny, nx))
    with lib_setup(precision, nx, ny, nz) as P:
        L = P.lib
        st = L.p3dfft_amd_stream_create()
        fin, spec, out = P.alloc(P.real_bytes()), P.alloc(P.spec_bytes()), P.alloc(P.real_bytes())
        other, ospec = P.alloc(P.real_bytes()), P.alloc(P.spec_bytes())
        graph = ctypes.c_void_p()
        try:
            fin.upload(f.astype(P.rdtype))
            P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft", st)
            L.p3dfft_amd_stream_synchronize(st)
            s0 = spec.download(P.spec_shape(), P.cdtype)
            P.btran_c2r_dev(spec.ptr, out.ptr, "tff", st)
            L.p3dfft_amd_stream_synchronize(st)
            b0 = out.download(P.real_shape(), P.rdtype)
            assert L.p3dfft_amd_capture_begin(st) == 0
            P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft", st)
            P.btran_c2r_dev(spec.ptr, out.ptr, "tff", st)
            assert L.p3dfft_amd_capture_end(st, ctypes.byref(graph)) == 0
            for rep in range(3):
                other.upload(g2.astype(P.rdtype))
                P.ftran_r2c_dev(other.ptr, ospec.ptr, "fft", st)  # eager work between replays
                out.upload(np.zeros((nz, ny, nx), P.rdtype))
                assert L.p3dfft_amd_graph_launch(graph, st) == 0
                L.p3dfft_amd_stream_synchronize(st)
                assert np.array_equal(out.download(P.real_shape(), P.rdtype), b0), rep
            # the spectrum the graph's ftran produced, before its btran overwrote it in place
            assert L.p3dfft_amd_capture_begin(st) == 0
            P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft", st)
            g1 = ctypes.c_void_p()
            assert L.p3dfft_amd_capture_end(st, ctypes.byref(g1)) == 0
            assert L.p3dfft_amd_graph_launch(g1, st) == 0
            L.p3dfft_amd_stream_synchronize(st)
            assert np.array_equal(spec.download(P.spec_shape(), P.cdtype), s0)
            L.p3dfft_amd_graph_destroy(g1)
        finally:
            if graph.value:
                L.p3dfft_amd_graph_destroy(graph)
            for b in (fin, spec, out, other, ospec):
                b.free()
            L.p3dfft_amd_stream_destroy(st)


@pytest.mark.parametrize("precision", [4, 8])
def test_batched_transforms_of_mixed_shapes_bit_identical(precision):
    """p3dfft_amd_batch_*: forward and backward transforms of several grids of
    different shapes as one launch per pass and transform length give, grid
    by grid, bit for bit the spectra and round trips of p3dfft_setup +
    ftran_r2c / btran_c2r of that grid; a length without a specialised kernel
    is refused."""
    import ctypes
    shapes = [(64, 48, 32), (32, 64, 96), (48, 48, 48), (128, 32, 64), (96, 96, 32)]
    rng = np.random.default_rng(61)
    fields = [rng.standard_normal((nz, ny, nx)) for (nx, ny, nz) in shapes]
    want_s, want_b = [], []
    for (nx, ny, nz), f in zip(shapes, fields):
        with lib_setup(precision, nx, ny, nz) as P:
            s = P.ftran_r2c(f, "fft")
            want_s.append(s)
            want_b.append(P.btran_c2r(s.copy(), "tff"))
    L = p3dfft_amd.load(precision)
    P = p3dfft_amd.P3DFFT(precision)
    rdt, cdt = P.rdtype, P.cdtype
    n = len(shapes)
    reals = [p3dfft_amd.DeviceBuffer(L, nx * ny * nz * np.dtype(rdt).itemsize) for (nx, ny, nz) in shapes]
    specs = [p3dfft_amd.DeviceBuffer(L, (nx // 2 + 1) * ny * nz * np.dtype(cdt).itemsize) for (nx, ny, nz) in shapes]
    sh = (ctypes.c_int * (3 * n))(*[v for s in shapes for v in s])
    rp = (ctypes.c_void_p * n)(*[b.ptr for b in reals])
    sp = (ctypes.c_void_p * n)(*[b.ptr for b in specs])
    fwd, bwd = ctypes.c_void_p(), ctypes.c_void_p()
    try:
        assert L.p3dfft_amd_batch_create(1, n, sh, rp, sp, ctypes.byref(fwd)) == 0
        assert L.p3dfft_amd_batch_create(-1, n, sh, rp, sp, ctypes.byref(bwd)) == 0
        for b, f in zip(reals, fields):
            b.upload(f.astype(rdt))
        assert L.p3dfft_amd_batch_launch(fwd, None) == 0
        assert L.p3dfft_amd_synchronize() == 0
        for i, (nx, ny, nz) in enumerate(shapes):
            assert np.array_equal(specs[i].download((nz, ny, nx // 2 + 1), cdt), want_s[i]), shapes[i]
        assert L.p3dfft_amd_batch_launch(bwd, None) == 0
        assert L.p3dfft_amd_synchronize() == 0
        for i, (nx, ny, nz) in enumerate(shapes):
            assert np.array_equal(reals[i].download((nz, ny, nx), rdt), want_b[i]), shapes[i]
        bad = (ctypes.c_int * 3)(36, 32, 32)
        h = ctypes.c_void_p()
        assert L.p3dfft_amd_batch_create(1, 1, bad, rp, sp, ctypes.byref(h)) == -1
        assert b"specialised" in L.p3dfft_amd_last_error()
        L.p3dfft_amd_clear_error()
    finally:
        for h in (fwd, bwd):
            if h.value:
                L.p3dfft_amd_batch_destroy(h)
        for b in reals + specs:
            b.free()


def test_clean_and_work_growth_refused_while_a_captured_graph_lives():
    """A captured transform graph replays launches on the setup's work
    buffers: while one exists, p3dfft_clean and a _many call that would grow
    those buffers are refused (with a message) and leave the setup intact;
    after the graph is destroyed both work."""
    import ctypes
    nx, ny, nz = 16, 12, 10
    f = np.random.default_rng(81).standard_normal((nz, ny, nx))
    P = p3dfft_amd.P3DFFT(4)
    P.setup((1, 1), nx, ny, nz)
    L = P.lib
    st = L.p3dfft_amd_stream_create()
    fin, spec, out = P.alloc(P.real_bytes()), P.alloc(P.spec_bytes()), P.alloc(P.real_bytes())
    graph = ctypes.c_void_p()
    try:
        want = P.ftran_r2c(f, "fft")
        fin.upload(f.astype(P.rdtype))
        P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft", st)
        L.p3dfft_amd_stream_synchronize(st)
        assert L.p3dfft_amd_capture_begin(st) == 0
        P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft", st)
        assert L.p3dfft_amd_capture_end(st, ctypes.byref(graph)) == 0
        spec.upload(np.zeros(P.spec_shape(), P.cdtype))
        L.p3dfft_amd_clear_error()
        L.p3dfft_clean()
        assert "destroy the captured graphs" in P.last_error()
        L.p3dfft_amd_clear_error()
        assert rel(P.ftran_r2c(f, "fft"), want) == 0  # the setup is still there
        n_r, n_c = nx * ny * nz, (nx // 2 + 1) * ny * nz
        big_in, big_out = P.alloc(2 * n_r * 4), P.alloc(2 * n_c * 8)
        with pytest.raises(p3dfft_amd.P3DFFTError, match="destroy the captured graphs"):
            P.ftran_r2c_many_dev(big_in.ptr, n_r, big_out.ptr, n_c, 2, "fft")
        assert L.p3dfft_amd_graph_launch(graph, st) == 0
        L.p3dfft_amd_stream_synchronize(st)
        assert np.array_equal(spec.download(P.spec_shape(), P.cdtype), want.astype(P.cdtype))
        L.p3dfft_amd_graph_destroy(graph)
        graph = ctypes.c_void_p()
        P.ftran_r2c_many_dev(big_in.ptr, n_r, big_out.ptr, n_c, 2, "fft")
        P.synchronize()
        big_in.free(); big_out.free()
    finally:
        if graph.value:
            L.p3dfft_amd_graph_destroy(graph)
        for b in (fin, spec, out):
            b.free()
        L.p3dfft_amd_stream_destroy(st)
        P.clean()
    assert not P.last_error()


@pytest.mark.parametrize("precision", [8, 4])
def test_nz_8192_c2c_thin_grid(precision):
    """nz = 8192, the engine's largest one-workgroup length, for the c2c Z
    ops: fwd against numpy's rfftn, and the round trip.  (A DCT-I along it
    needs length 16382, which runs in several passes: test_gpu_long.py.)"""
    nx, ny, nz = 4, 3, 8192
    f = np.random.default_rng(82).standard_normal((nz, ny, nx))
    with lib_setup(precision, nx, ny, nz) as P:
        s = P.ftran_r2c(f, "fft")
        assert rel(s, np.fft.rfftn(f)) < TOL[precision]
        b = P.btran_c2r(s, "tff") / (nx * ny * nz)
        assert np.abs(b - f).max() < PREC[precision] * 100



@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,ops", [((64, 64, 64), ("fft", "tff")), ((32, 32, 20), ("ffc", "cff")),
                                       ((128, 64, 8), ("fft", "tff")), ((64, 128, 6), ("ffs", "sff")),
                                       ((32, 64, 12), ("fft", "tff")), ((128, 128, 16), ("fft", "tff"))])
def test_plane_kernels_against_oracle(precision, shape, ops):
    """X and Y of each z-plane in one workgroup (plane kernels forced on)
    against the oracle and against the separate passes (forced off); the
    forward's Y timer stays zero when the plane kernel ran (its time is in
    the X slot).  FP64 128 x 128 planes do not fit the LDS and take the
    separate passes either way."""
    nx, ny, nz = shape
    a = np.random.default_rng(nx + ny + nz).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz)
    ref_f = o.ftran_r2c(a, ops[0])
    lib = p3dfft_amd.load(precision)
    got = []
    try:
        for mode in (1, 0):
            lib.p3dfft_amd_set_plane_kernels(mode)
            with lib_setup(precision, nx, ny, nz) as P:
                P.set_timers()
                s = P.ftran_r2c(a, ops[0])
                t = P.get_timers()
                got.append((s, P.btran_c2r(s, ops[1]), t))
    finally:
        lib.p3dfft_amd_set_plane_kernels(2)
    ref_b = o.btran_c2r(got[0][0].astype(np.complex128), ops[1])[0]
    for s, b, _ in got:
        assert rel(s, ref_f) < TOL[precision]
        assert rel(b, ref_b) < TOL[precision]
    assert rel(got[0][0], got[1][0]) < (1e-14 if precision == 8 else 1e-6)
    fits = not (precision == 8 and nx == 128 and ny == 128)
    assert (got[0][2][6] == 0.0) == fits and got[0][2][4] > 0.0
    assert got[1][2][6] > 0.0


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape", [(126, 40, 2048), (126, 2048, 40)])
def test_length_2048_columns_against_oracle(precision, shape):
    """Length-2048 column passes over more tiles than the GPU has CUs (the
    share-mode Y and Z lengths): the Z pass at (126, 40, 2048) and the Y pass
    at (126, 2048, 40) against the oracle, and the round trip."""
    nx, ny, nz = shape
    a = np.random.default_rng(2048 + ny).standard_normal((nz, ny, nx))
    ref = oplan(nx, ny, nz).ftran_r2c(a)
    with lib_setup(precision, nx, ny, nz) as P:
        s = P.ftran_r2c(a, "fft")
        b = P.btran_c2r(s, "tff")
    assert rel(s, ref) < TOL[precision]
    assert np.abs(b / (nx * ny * nz) - a).max() < PREC[precision] * 100 * np.abs(a).max()


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,nzc,stride1", [((16, 12, 33), None, False), ((24, 20, 129), None, True),
                                               ((20, 16, 65), 40, False), ((8, 6, 257), None, False)])
def test_cheby_recurrence_in_the_dct_pass(precision, shape, nzc, stride1):
    """p3dfft_cheby with the derivative recurrence applied inside the DCT-I
    Z pass (specialised kernels, DCT lengths 32 ... 512) against the oracle
    and against the separate recurrence kernel (run-time-generic kernels),
    also in the STRIDE1 layout and with a Z cut."""
    nx, ny, nz = shape
    Lz = 2.0
    rng = np.random.default_rng(nz + nx)
    a = rng.standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz, nzc=nzc, stride1=stride1)
    ref = o.cheby(a, Lz)
    lib = p3dfft_amd.load(precision)
    got = []
    try:
        for spec_on in (1, 0):
            lib.p3dfft_amd_set_specialised(spec_on)
            with lib_setup(precision, nx, ny, nz, nzc=nzc, stride1=stride1) as P:
                got.append(P.cheby(a, Lz))
    finally:
        lib.p3dfft_amd_set_specialised(1)
    tol = cheby_tol(precision, nzc or nz, Lz)
    assert rel(got[0], ref) < tol
    assert rel(got[1], ref) < tol
    assert rel(got[0], got[1]) < tol


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,nxc", [((512, 6, 4), None), ((512, 3, 3), None), ((512, 5, 2), 300),
                                       ((512, 64, 8), None), ((512, 160, 129), None)])
def test_x_wave_kernels_against_oracle(precision, shape, nxc):
    """X rows of 512 points through the wave kernels (one wavefront per row
    pair, cross-lane exchanges; switched on) and through the LDS-tile
    kernels (the default) against the oracle: an odd row count (a last pair
    with one row), an X cut (fewer stored modes forward, zero modes
    backward), and more row pairs than the wave kernels' grid (each
    workgroup walks several)."""
    nx, ny, nz = shape
    a = np.random.default_rng(nx + ny + nz).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz, nxc=nxc)
    ref_f = o.ftran_r2c(a, "fft")
    ref_b = o.btran_c2r(ref_f, "tff")[0]
    lib = p3dfft_amd.load(precision)
    try:
        for wave in (1, 0):
            lib.p3dfft_amd_set_x_wave(wave)
            with lib_setup(precision, nx, ny, nz, nxc=nxc) as P:
                assert rel(P.ftran_r2c(a, "fft"), ref_f) < TOL[precision], wave
                assert rel(P.btran_c2r(ref_f, "tff"), ref_b) < TOL[precision], wave
    finally:
        lib.p3dfft_amd_set_x_wave(0)


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,nzc", [((16, 12, 33), None), ((32, 24, 65), 40), ((20, 10, 129), None),
                                       ((12, 8, 257), 100), ((10, 6, 17), None)])
def test_dct1_direct_and_lds_passes_against_oracle(precision, shape, nzc):
    """DCT-I Z passes ('ffc' forward, 'cff' backward) through the direct-load
    kernel reading the mirrored rows (the default) and through the LDS
    extension pass, against the oracle; with and without a Z cut (length 17:
    extension 32, specialised; the others 64 ... 512)."""
    nx, ny, nz = shape
    a = np.random.default_rng(nz + nx).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz, nzc=nzc)
    ref_f = o.ftran_r2c(a, "ffc")
    ref_b = o.btran_c2r(ref_f, "cff")[0]
    lib = p3dfft_amd.load(precision)
    try:
        for direct in (1, 0):
            lib.p3dfft_amd_set_dct_direct(direct)
            with lib_setup(precision, nx, ny, nz, nzc=nzc) as P:
                assert rel(P.ftran_r2c(a, "ffc"), ref_f) < TOL[precision], direct
                assert rel(P.btran_c2r(ref_f, "cff"), ref_b) < TOL[precision], direct
    finally:
        lib.p3dfft_amd_set_dct_direct(1)


@pytest.mark.parametrize("nvec", [1, 255, 257, 1000003])
def test_stream_copy_exact(nvec):
    """The streaming copy kernel (the copy-bandwidth reference of the bench
    and the traffic-counter calibration) copies every byte, also when the
    last workgroup is partial, and leaves the bytes after the copy alone;
    sizes that are not a multiple of 16 bytes are refused."""
    lib = p3dfft_amd.load(4)
    nbytes = 16 * nvec
    src = np.random.default_rng(nvec).integers(0, 256, nbytes + 64, dtype=np.uint8)
    a, b = p3dfft_amd.DeviceBuffer(lib, nbytes + 64), p3dfft_amd.DeviceBuffer(lib, nbytes + 64)
    try:
        a.upload(src)
        b.upload(np.zeros(nbytes + 64, dtype=np.uint8))
        assert lib.p3dfft_amd_stream_copy(a.ptr, b.ptr, nbytes, None) == 0
        assert lib.p3dfft_amd_synchronize() == 0
        got = b.download(nbytes + 64, np.uint8)
        assert np.array_equal(got[:nbytes], src[:nbytes])
        assert not got[nbytes:].any()
        assert lib.p3dfft_amd_stream_copy(a.ptr, b.ptr, nbytes + 8, None) == -1
    finally:
        a.free()
        b.free()


@pytest.mark.parametrize("precision", [8, 4])
def test_col_pipe_z2048_against_oracle(precision):
    """Z passes of length 2048 (126 x 36 x 2048: 288 FP32 / 576 FP64 tiles,
    more than one per CU) through the persistent column kernel
    (p3dfft_amd_set_col_pipe(1)), in the XCD-aware and the dispatch tile
    order, against the oracle and equal to rounding to one workgroup per tile
    (the same stage code, contracted to FMAs differently inside the loop)."""
    nx, ny, nz = 126, 36, 2048
    a = np.random.default_rng(2048).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz)
    ref_f = o.ftran_r2c(a, "fft")
    ref_b = o.btran_c2r(ref_f, "tff")[0]
    lib = p3dfft_amd.load(precision)
    res = {}
    try:
        for pipe, order in ((0, 2), (1, 0), (1, 1)):
            lib.p3dfft_amd_set_col_pipe(pipe)
            lib.p3dfft_amd_set_flat_order(order)
            with lib_setup(precision, nx, ny, nz) as P:
                f = P.ftran_r2c(a, "fft")
                b = P.btran_c2r(ref_f, "tff")
            assert rel(f, ref_f) < TOL[precision], (pipe, order)
            assert rel(b, ref_b) < TOL[precision], (pipe, order)
            res[(pipe, order)] = (f, b)
        for key in ((1, 0), (1, 1)):
            assert rel(res[key][0], res[(0, 2)][0]) < TOL[precision] / 10, key
            assert rel(res[key][1], res[(0, 2)][1]) < TOL[precision] / 10, key
    finally:
        lib.p3dfft_amd_set_col_pipe(1)
        lib.p3dfft_amd_set_flat_order(2)
