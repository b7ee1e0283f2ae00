"""Pins the FP64 CPU oracle (oracle/p3dfft_oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails one of them:
  * golden vectors from an independent FFT (numpy/scipy pocketfft,
    tests/golden/gen_golden.py) for every Z option of ftran and btran;
  * the reference's own known-answer drivers restated: driver_sine's closed
    form spectrum (sample/FORTRAN/driver_sine.F90), driver_inverse's four
    +-N/4 spikes (sample/FORTRAN/driver_inverse.F90:277-343), driver_cheby's
    derivative check (sample/FORTRAN/driver_cheby.F90:218-285);
  * mathematical identities: Parseval, shift theorem, DCT-I / DST-I
    self-inverse, Chebyshev differentiation against numpy.polynomial;
  * the decomposition arithmetic of build/setup.F90 on hand-worked cases;
  * the pack/unpack maps of the transposes by emulating every rank of a
    P1 x P2 grid and comparing with the single-rank transform.
"""
import numpy as np
import pytest

from oracle import oracle as O

RTOL64 = 1e-12


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


SHAPES = [(8, 6, 10), (16, 16, 16), (12, 9, 5), (7, 11, 13), (10, 14, 6), (6, 4, 17)]


# --------------------------------------------------------------------- 1D
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 17, 30, 64, 97, 128, 257])
def test_1d_against_golden(golden, n):
    x = golden["fft1_in_%d" % n]
    assert rel(O.fft(x, -1), golden["fft1_out_%d" % n]) < 1e-14
    assert rel(O.fft(x, +1), golden["ifft1_out_%d" % n]) < 1e-14
    assert rel(O.dft_naive(x, -1), golden["fft1_out_%d" % n]) < 1e-14
    r = golden["real1_in_%d" % n]
    assert rel(O.r2c(r), np.fft.rfft(r)) < 1e-14
    if n >= 2:
        assert rel(O.dct1_naive(r), golden["dct1_out_%d" % n]) < 1e-14
        assert rel(O.dct1(r).real, golden["dct1_out_%d" % n]) < 1e-14
    assert rel(O.dst1_naive(r), golden["dst1_out_%d" % n]) < 1e-14
    assert rel(O.dst1(r).real, golden["dst1_out_%d" % n]) < 1e-14


def test_c2r_drops_imag_of_dc_and_nyquist():
    rng = np.random.default_rng(3)
    for n in (8, 9):
        X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
        Y = X.copy()
        Y[0] = Y[0].real
        if n % 2 == 0:
            Y[-1] = Y[-1].real
        a = O.c2r(X, n)
        b = O.c2r(Y, n)
        assert np.abs(a - b).max() == 0.0
        assert rel(a, np.fft.irfft(Y, n) * n) < 1e-14


@pytest.mark.parametrize("n", [16384, 12000, 9000, 8209])
def test_oracle_long_lengths_against_numpy(n):
    """The oracle at lengths above 8192 (those the GPU engine runs in several
    passes: four-step 128x128, 100x120, 90x100, and the prime 8209) against
    numpy's pocketfft, an independent implementation: complex forward and
    backward, r2c and c2r."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(O.fft(x, -1), np.fft.fft(x)) < 1e-13
    assert rel(O.fft(x, +1), np.fft.ifft(x) * n) < 1e-13
    r = x.real.copy()
    assert rel(O.r2c(r), np.fft.rfft(r)) < 1e-13
    X = np.fft.rfft(r)
    assert rel(O.c2r(X, n), r * n) < 1e-13


@pytest.mark.parametrize("n", [4200, 4500, 8192])
def test_oracle_long_dct1_dst1_against_scipy(n):
    """DCT-I / DST-I whose extensions (2(n-1), 2(n+1)) exceed 8192 points,
    against scipy.fft's type-1 transforms (FFTW's REDFT00 / RODFT00)."""
    import scipy.fft as sf
    r = np.random.default_rng(n).standard_normal(n)
    assert rel(O.dct1(r).real, sf.dct(r, type=1)) < 1e-13
    assert rel(O.dst1(r).real, sf.dst(r, type=1)) < 1e-13


@pytest.mark.parametrize("n", [2, 3, 9, 33, 65])
def test_dct1_dst1_self_inverse(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    assert rel(O.dct1_naive(O.dct1_naive(x)), 2 * (n - 1) * x) < 1e-13
    assert rel(O.dst1_naive(O.dst1_naive(x)), 2 * (n + 1) * x) < 1e-13


# --------------------------------------------------------------------- 3D vs golden
@pytest.mark.parametrize("shape", SHAPES)
def test_3d_forward_golden(golden, shape):
    nx, ny, nz = shape
    tag = "%dx%dx%d" % shape
    p = O.Plan(nx, ny, nz)
    a = golden["r2c_in_" + tag]
    assert rel(p.ftran_r2c(a, "fft"), golden["r2c_out_" + tag]) < RTOL64
    assert rel(p.ftran_r2c(a, "fff"), golden["r2c_out_" + tag]) < RTOL64
    assert rel(p.ftran_r2c(a, "ffs"), golden["ffs_out_" + tag]) < RTOL64
    assert rel(p.ftran_r2c(a, "ffn"), golden["ffn_out_" + tag]) < RTOL64
    assert rel(p.ftran_r2c(a, "ff0"), golden["ffn_out_" + tag]) < RTOL64
    if nz >= 2:
        assert rel(p.ftran_r2c(a, "ffc"), golden["ffc_out_" + tag]) < RTOL64


@pytest.mark.parametrize("shape", SHAPES)
def test_3d_backward_golden(golden, shape):
    nx, ny, nz = shape
    tag = "%dx%dx%d" % shape
    for ow in (True, False):
        p = O.Plan(nx, ny, nz, ow=ow)
        s = golden["c2r_in_" + tag]
        assert rel(p.btran_c2r(s, "tff")[0], golden["c2r_out_" + tag]) < RTOL64
        assert rel(p.btran_c2r(s, "sff")[0], golden["sff_out_" + tag]) < RTOL64
        if nz >= 2:
            assert rel(p.btran_c2r(s, "cff")[0], golden["cff_out_" + tag]) < RTOL64


def test_btran_overwrite_semantics():
    """With ow and nz==nzc the reference transforms the caller's spectrum in
    place along Z (build/btran.F90:525-550); otherwise the input survives."""
    rng = np.random.default_rng(5)
    nx, ny, nz = 8, 6, 10
    s = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    p = O.Plan(nx, ny, nz, ow=True)
    _, after = p.btran_c2r(s.copy(), "tff", keep_input=False)
    assert rel(after, np.fft.ifft(s, axis=0) * nz) < 1e-14
    p = O.Plan(nx, ny, nz, ow=False)
    _, after = p.btran_c2r(s.copy(), "tff", keep_input=False)
    assert np.array_equal(after, s)


def test_unknown_op_rejected():
    p = O.Plan(8, 8, 8)
    with pytest.raises(ValueError):
        p.ftran_r2c(np.zeros((8, 8, 8)), "ffx")
    with pytest.raises(ValueError):
        p.btran_c2r(np.zeros((8, 8, 5), complex), "xff")


# --------------------------------------------------------------------- reference drivers restated
def sine_field(nx, ny, nz):
    x = np.sin(2 * np.pi * np.arange(nx) / nx)
    y = np.sin(2 * np.pi * np.arange(ny) / ny)
    z = np.sin(2 * np.pi * np.arange(nz) / nz)
    return z[:, None, None] * y[None, :, None] * x[None, None, :]


@pytest.mark.parametrize("shape", [(16, 16, 16), (12, 10, 8), (64, 64, 64)])
def test_driver_sine_closed_form(shape):
    """sample/FORTRAN/driver_sine.F90: the forward spectrum of
    sin(2 pi x/nx) sin(2 pi y/ny) sin(2 pi z/nz) is nonzero only at
    kx=1, ky in {1, ny-1}, kz in {1, nz-1}, with value prod(+-n/(2i))."""
    nx, ny, nz = shape
    p = O.Plan(nx, ny, nz)
    s = p.ftran_r2c(sine_field(nx, ny, nz), "fft")
    expect = np.zeros_like(s)
    one = lambda n: {1: n / 2j, n - 1: -n / 2j}
    for kz, vz in one(nz).items():
        for ky, vy in one(ny).items():
            expect[kz, ky, 1] = (nx / 2j) * vy * vz
    assert np.abs(s - expect).max() < 1e-12 * nx * ny * nz
    # round trip as the driver checks it (tolerance 1e-14 * N / 4)
    back, _ = p.btran_c2r(s / (nx * ny * nz), "tff")
    assert np.abs(back - sine_field(nx, ny, nz)).max() < 1e-14 * nx * ny * nz * 0.25


@pytest.mark.parametrize("shape", [(16, 12, 10), (32, 32, 32)])
def test_driver_inverse_spikes(shape):
    """sample/FORTRAN/driver_inverse.F90:277-406: btran of
    e^{2 pi i (kx)/nx} sin(2*2 pi ky/ny) sin(3*2 pi kz/nz) gives -N/4 at
    (nx,3,4), +N/4 at (nx,3,nz-2) and (nx,ny-1,4), -N/4 at (nx,ny-1,nz-2)
    (1-based), zero elsewhere."""
    nx, ny, nz = shape
    N = nx * ny * nz
    kx = np.arange(nx // 2 + 1)
    ky = np.arange(ny)
    kz = np.arange(nz)
    spec = (np.exp(2j * np.pi * kx / nx)[None, None, :] * np.sin(2 * 2 * np.pi * ky / ny)[None, :, None]
            * np.sin(3 * 2 * np.pi * kz / nz)[:, None, None])
    p = O.Plan(nx, ny, nz)
    f, _ = p.btran_c2r(spec, "tff")
    expect = np.zeros((nz, ny, nx))
    expect[3, 2, nx - 1] = -N / 4
    expect[nz - 3, 2, nx - 1] = N / 4
    expect[3, ny - 2, nx - 1] = N / 4
    expect[nz - 3, ny - 2, nx - 1] = -N / 4
    assert np.abs(f - expect).max() < 1e-14 * N * 0.25


def cheby_points(nz, Lz):
    return np.cos(np.pi * np.arange(nz) / (nz - 1)) * 2.0 / Lz


def cheby_check(p, spec, nz):
    """driver_cheby.F90:246-256: double the end coefficients, halve all, btran 'cff'."""
    s = spec.copy()
    s[0] *= 2.0
    s[nz - 1] *= 2.0
    s *= 0.5
    return p.btran_c2r(s, "cff")[0]


def test_driver_cheby_sin_to_cos():
    nx, ny, nz, Lz = 32, 32, 33, 2.0
    z = cheby_points(nz, Lz)
    a = np.broadcast_to(np.sin(z)[:, None, None], (nz, ny, nx)).copy()
    p = O.Plan(nx, ny, nz)
    d = cheby_check(p, p.cheby(a, Lz), nz)
    err = np.abs(d - np.cos(z)[:, None, None]).max()
    assert err < 1e-14 * nx * ny * nz * 0.25
    assert err < 1e-10


def test_cheby_polynomial_derivative_exact():
    """A degree-12 Chebyshev polynomial in z times trig modes in x, y:
    the derivative after the round trip matches numpy.polynomial.chebyshev."""
    rng = np.random.default_rng(11)
    nx, ny, nz, Lz = 16, 12, 17, 2.0
    c = rng.standard_normal(13) / (1 + np.arange(13)) ** 2
    z = cheby_points(nz, Lz)
    g = np.polynomial.chebyshev.chebval(z, c)
    dg = np.polynomial.chebyshev.chebval(z, np.polynomial.chebyshev.chebder(c))
    xs = np.arange(nx) / nx
    ys = np.arange(ny) / ny
    xy = np.cos(2 * np.pi * (2 * xs[None, :] + 3 * ys[:, None]) + 0.3) + 0.5
    p = O.Plan(nx, ny, nz)
    d = cheby_check(p, p.cheby(g[:, None, None] * xy[None], Lz), nz)
    assert rel(d, dg[:, None, None] * xy[None]) < 1e-12


def test_parseval_and_shift():
    rng = np.random.default_rng(7)
    nx, ny, nz = 12, 10, 9
    a = rng.standard_normal((nz, ny, nx))
    p = O.Plan(nx, ny, nz)
    s = p.ftran_r2c(a)
    w = np.full(nx // 2 + 1, 2.0)
    w[0] = 1.0
    if nx % 2 == 0:
        w[-1] = 1.0
    assert abs((w * np.abs(s) ** 2).sum() / (nx * ny * nz) - (a ** 2).sum()) < 1e-10 * (a ** 2).sum()
    s2 = p.ftran_r2c(np.roll(a, 1, axis=2))
    ph = np.exp(-2j * np.pi * np.arange(nx // 2 + 1) / nx)
    assert rel(s2, s * ph[None, None, :]) < 1e-13


# --------------------------------------------------------------------- pruning
def keep_idx(n, nc):
    """FFT pruning semantics: the low ceil(nc/2) and the top floor(nc/2) modes."""
    lo = (nc + 1) // 2
    return np.r_[np.arange(lo), np.arange(n - (nc - lo), n)]


@pytest.mark.parametrize("cut", [(8, 12, 10), (6, 8, 6), (10, 9, 7), (12, 7, 8)])
def test_pruned_forward_is_mode_subset(cut):
    rng = np.random.default_rng(1)
    nx, ny, nz = 12, 12, 10
    nxc, nyc, nzc = cut
    a = rng.standard_normal((nz, ny, nx))
    full = O.Plan(nx, ny, nz).ftran_r2c(a)
    p = O.Plan(nx, ny, nz, nxc, nyc, nzc)
    s = p.ftran_r2c(a)
    ref = full[keep_idx(nz, nzc)][:, keep_idx(ny, nyc)][:, :, : nxc // 2 + 1]
    assert rel(s, ref) < RTOL64
    ps = O.Plan(nx, ny, nz, nxc, nyc, nzc, stride1=True)
    assert rel(ps.ftran_r2c(a), ref.transpose(2, 1, 0)) < RTOL64


@pytest.mark.parametrize("cut", [(8, 12, 10), (6, 8, 6), (12, 8, 8)])
def test_pruned_backward_is_zero_padded_inverse(cut):
    """Even cut sizes: backward of a pruned spectrum equals the unpruned
    backward of the zero-padded spectrum (both layouts, both OW settings)."""
    rng = np.random.default_rng(2)
    nx, ny, nz = 12, 12, 10
    nxc, nyc, nzc = cut
    nxhpc = nxc // 2 + 1
    s = rng.standard_normal((nzc, nyc, nxhpc)) + 1j * rng.standard_normal((nzc, nyc, nxhpc))
    pad = np.zeros((nz, ny, nx // 2 + 1), complex)
    kz, ky = keep_idx(nz, nzc), keep_idx(ny, nyc)
    pad[np.ix_(kz, ky, np.arange(nxhpc))] = s
    ref, _ = O.Plan(nx, ny, nz).btran_c2r(pad)
    for ow in (True, False):
        got, _ = O.Plan(nx, ny, nz, nxc, nyc, nzc, ow=ow).btran_c2r(s)
        assert rel(got, ref) < RTOL64
        got, _ = O.Plan(nx, ny, nz, nxc, nyc, nzc, ow=ow, stride1=True).btran_c2r(s.transpose(2, 1, 0).copy())
        assert rel(got, ref) < RTOL64


def test_stride1_layout_is_transpose():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((10, 6, 8))
    s = O.Plan(8, 6, 10).ftran_r2c(a)
    s1 = O.Plan(8, 6, 10, stride1=True).ftran_r2c(a)
    assert np.array_equal(s1, s.transpose(2, 1, 0))


# --------------------------------------------------------------------- decomposition
def test_map_data_to_proc_remainder_to_last():
    assert O.map_data_to_proc(10, 3) == ([1, 4, 7], [3, 6, 10], [3, 3, 4])
    st, en, sz = O.map_data_to_proc(1025, 16)
    assert sz == [64] * 15 + [65] and en[-1] == 1025 and st[-1] == 961
    assert O.map_data_to_proc(7, 7)[2] == [1] * 7
    assert O.map_data_to_proc(5, 8)[2] == [0, 0, 0, 1, 1, 1, 1, 1]


def test_setup_sizes():
    p = O.Plan(64, 64, 64)
    assert p.memsize == (64, 64, 66)
    assert p.get_dims(1) == ((1, 1, 1), (64, 64, 64), (64, 64, 64))
    assert p.get_dims(2) == ((1, 1, 1), (33, 64, 64), (33, 64, 64))
    assert p.get_dims(3) == ((0, 0, 0), (64, 64, 66), (64, 64, 66))
    p = O.Plan(512, 512, 512)
    assert p.nm == 257 * 512 * 512 and p.memsize == (512, 512, 514)
    # one rank of 2048^3 on a 16x16 grid (the last row rank holds 65 kx)
    p = O.Plan(2048, 2048, 2048, dims=(16, 16), ipid=15, jpid=3)
    assert (p.iisize, p.jisize, p.jjsize, p.kjsize) == (65, 128, 128, 128)
    assert p.get_dims(1) == ((1, 1 + 15 * 128, 1 + 3 * 128), (2048, 16 * 128, 4 * 128), (2048, 128, 128))
    assert p.get_dims(2) == ((961, 385, 1), (1025, 512, 2048), (65, 128, 2048))
    s = O.Plan(8, 6, 10, stride1=True)
    assert s.get_dims(2) == ((1, 1, 1), (10, 6, 5), (10, 6, 5))


# --------------------------------------------------------------------- share-mode maps
def _slices(p):
    return (slice(p.jistart - 1, p.jiend), slice(p.kjstart - 1, p.kjend))


def emulate_grid(nx, ny, nz, p1, p2, a, nyc=None):
    """Run the reference's multi-rank forward+backward path for every rank of
    a p1 x p2 grid with the oracle's pack/unpack maps and numpy 1D FFTs, and
    exchange blocks as MPI_Alltoallv would (counts and displacements of
    build/setup.F90:480-518)."""
    nyc = nyc or ny
    P = {(i, j): O.Plan(nx, ny, nz, nyc=nyc, dims=(p1, p2), ipid=i, jpid=j) for i in range(p1) for j in range(p2)}
    C = {k: v.share_counts() for k, v in P.items()}
    # X r2c per rank: buf2 (nxhp, jisize, kjsize) stored (kjsize, jisize, nxhp)
    send1 = {}
    for (i, j), p in P.items():
        ys, zs = _slices(p)
        buf2 = np.fft.rfft(a[zs, ys, :], axis=2)
        send1[(i, j)] = p.fcomm1_pack(buf2)
    out = {}
    for (r, j), p in P.items():
        recv = np.zeros(sum(C[(r, j)]["frcv_ct"]), complex)
        for i in range(p1):
            cs, cc = C[(i, j)]["fsnd_st"][r], C[(i, j)]["fsnd_ct"][r]
            rs = C[(r, j)]["frcv_st"][i]
            recv[rs:rs + cc] = send1[(i, j)][cs:cs + cc]
        ypen = p.fcomm1_unpack(recv).reshape(p.kjsize, ny, p.iisize)
        ypen = np.fft.fft(ypen, axis=1)
        out[(r, j)] = p.fcomm2_pack(ypen)
    spec = {}
    for (i, r), p in P.items():
        o = np.zeros(p.iisize * p.jjsize * nz, complex)
        for j in range(p2):
            cs, cc = C[(i, j)]["ksnd_st"][r], C[(i, j)]["ksnd_ct"][r]
            rs = C[(i, r)]["krcv_st"][j]
            o[rs:rs + cc] = out[(i, j)][cs:cs + cc]
        spec[(i, r)] = np.fft.fft(o.reshape(nz, p.jjsize, p.iisize), axis=0)
    # backward
    send_b = {}
    for (i, j), p in P.items():
        zb = np.fft.ifft(spec[(i, j)], axis=0) * nz
        send_b[(i, j)] = zb.reshape(-1)
    back = {}
    for (i, r), p in P.items():
        recv = np.zeros(p.iisize * ny * p.kjsize, complex)
        for j in range(p2):
            # bcomm1 sends z-block kjst(r)..kjen(r) from (iisize, jjsize, nz) of rank (i, j)
            cs, cc = C[(i, j)]["krcv_st"][r], C[(i, j)]["krcv_ct"][r]
            rs = C[(i, r)]["ksnd_st"][j]
            recv[rs:rs + cc] = send_b[(i, j)][cs:cs + cc]
        yp = p.bcomm1_unpack(recv).reshape(p.kjsize, ny, p.iisize)
        yp = np.fft.ifft(yp, axis=1) * ny
        back[(i, r)] = p.bcomm2_pack(yp)
    field = np.zeros((nz, ny, nx))
    for (r, j), p in P.items():
        recv = np.zeros(p.nxhp * p.jisize * p.kjsize, complex)
        for i in range(p1):
            cs, cc = C[(i, j)]["frcv_st"][r], C[(i, j)]["frcv_ct"][r]
            rs = C[(r, j)]["fsnd_st"][i]
            recv[rs:rs + cc] = back[(i, j)][cs:cs + cc]
        xp = p.bcomm2_unpack(recv).reshape(p.kjsize, p.jisize, p.nxhp)
        ys, zs = _slices(p)
        field[zs, ys, :] = np.fft.irfft(xp, n=nx, axis=2) * nx
    return P, spec, field


@pytest.mark.parametrize("grid", [(2, 3), (3, 2), (4, 4), (1, 3), (3, 1)])
def test_share_maps_reproduce_global_transform(grid):
    rng = np.random.default_rng(9)
    nx, ny, nz = 14, 26, 38   # the reference's uneven test grid (extra/makejob.py:142-157)
    a = rng.standard_normal((nz, ny, nx))
    P, spec, field = emulate_grid(nx, ny, nz, grid[0], grid[1], a)
    full = O.Plan(nx, ny, nz).ftran_r2c(a)
    for (i, j), p in P.items():
        ref = full[:, p.jjstart - 1:p.jjend, p.iistart - 1:p.iiend]
        assert rel(spec[(i, j)], ref) < RTOL64
    assert rel(field, a * nx * ny * nz) < RTOL64


def test_share_maps_with_y_pruning():
    rng = np.random.default_rng(10)
    nx, ny, nz, nyc = 12, 16, 10, 8
    a = rng.standard_normal((nz, ny, nx))
    P, spec, field = emulate_grid(nx, ny, nz, 2, 3, a, nyc=nyc)
    full = O.Plan(nx, ny, nz, nyc=nyc).ftran_r2c(a)
    for (i, j), p in P.items():
        ref = full[:, p.jjstart - 1:p.jjend, p.iistart - 1:p.iiend]
        assert rel(spec[(i, j)], ref) < RTOL64
    ref_back, _ = O.Plan(nx, ny, nz, nyc=nyc).btran_c2r(full)
    assert rel(field, ref_back) < RTOL64


# ---------------------------------------------------------------- power spectrum
def test_power_spectrum_single_mode_closed_form():
    """One stored mode B(x, y, z) = A: E is k2*|A|^2 at shell int(|k| + 1/2)
    and zero elsewhere; y and z above n/2 fold to n - j
    (sample/C/driver_spec.c:336-344)."""
    nx, ny, nz = 16, 12, 10
    kmax = O.spectrum_kmax(nx, ny, nz)
    for (x, y, z) in [(3, 2, 1), (5, 9, 7), (0, 0, 0), (8, 6, 5)]:
        B = np.zeros((nz, ny, nx // 2 + 1), complex)
        B[z, y, x] = 2.0 - 1.5j
        ky = y if y <= ny // 2 else ny - y
        kz = z if z <= nz // 2 else nz - z
        k2 = x * x + ky * ky + kz * kz
        E = O.power_spectrum(B, (nx, ny, nz), kmax)
        want = np.zeros(kmax + 1)
        want[int(np.sqrt(k2) + 0.5)] = k2 * 6.25
        assert np.array_equal(E, want), (x, y, z)


def test_power_spectrum_matches_loop_restatement_and_total():
    """The vectorised restatement equals the drivers' loop nest element by
    element on a small random spectrum, and its shells add up to the total
    sum of k2*|B|^2; the STRIDE1 form of the same array gives the same E."""
    rng = np.random.default_rng(5)
    nx, ny, nz = 10, 7, 9
    kmax = O.spectrum_kmax(nx, ny, nz)
    B = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    E = O.power_spectrum(B, (nx, ny, nz), kmax)
    L = O.power_spectrum_loops(B, (nx, ny, nz), kmax)
    assert np.allclose(E, L, rtol=1e-13, atol=0)
    x = np.arange(nx // 2 + 1)[None, None, :]
    y = np.minimum(np.arange(ny), ny - np.arange(ny))[None, :, None]
    z = np.minimum(np.arange(nz), nz - np.arange(nz))[:, None, None]
    total = ((x * x + y * y + z * z) * np.abs(B) ** 2).sum()
    assert abs(E.sum() - total) < 1e-12 * total
    Es = O.power_spectrum(np.ascontiguousarray(B.transpose(2, 1, 0)), (nx, ny, nz), kmax, stride1=True)
    assert np.allclose(Es, E, rtol=1e-13, atol=0)


# Send buffers of rank (ipid, jpid) = (1, 1) of an 8 x 6 x 5 grid on a 2 x 2
# processor grid, written out by hand from the reference's pack loops
# (tests/golden/hand_pack_2x2.json; not produced by any code).  nxhp = 5
# splits over the row as iisz = (2, 3) (remainder to the last rank:
# iist = (1, 3), iien = (2, 5)); ny = 6 gives jisz = (3, 3) (jist = (1, 4));
# nz = 5 gives kjsize = 3 for jpid = 1.
# fcomm1.F90:239-253: for each row peer i, for z = 1..kjsize, y = 1..jisize,
# x = iist(i)..iien(i): buf1(pos++) = source(x, y, z), peer blocks back to back
# (IfSndStrt = 0, 18); source (nxhp, jisize, kjsize) = (5, 3, 3).
# bcomm2.F90:233-273 (non-STRIDE1 branch): for each row peer i, for
# z = 1..kjsize, position = IfRcvStrt(i) + (z-1) iisize jisz(i), then
# y = jist(i)..jien(i), x = 1..iisize: buf1(pos++) = source(x, y, z)
# (IfRcvStrt = 0, 27); source (iisize, ny, kjsize) = (3, 6, 3).
def hand_pack_fixture():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "hand_pack_2x2.json")) as f:
        return json.load(f)


def test_share_pack_order_against_hand_worked_buffers():
    """The oracle's fcomm1 and bcomm2 packs reproduce the send buffers worked
    out by hand from the reference loops, element by element: this pins the
    order inside each peer block, which the all-ranks round trip cannot see
    (a pack and an unpack reordered the same way would cancel)."""
    h = hand_pack_fixture()
    o = O.Plan(*h["grid"], dims=tuple(h["dims"]), ipid=h["rank"][0], jpid=h["rank"][1])
    assert (o.iisize, o.jisize, o.kjsize) == (3, 3, 3)
    src = np.arange(5 * 3 * 3, dtype=np.float64) + 0j
    assert np.array_equal(o.fcomm1_pack(src).real, np.array(h["fcomm1_send"], dtype=np.float64))
    ypen = np.arange(3 * 6 * 3, dtype=np.float64) + 0j
    assert np.array_equal(o.bcomm2_pack(ypen).real, np.array(h["bcomm2_send"], dtype=np.float64))


@pytest.mark.parametrize("grid", [(2, 3), (3, 2), (1, 4), (4, 1), (2, 2)])
def test_ranks_path_reproduces_single_rank(grid):
    """The distributed CPU path (ranks as threads, exchanges as copies of the
    pack/unpack maps' buffers) on the reference's uneven 14 x 26 x 38 grid:
    the assembled spectrum equals the single-rank oracle's and numpy's
    rfftn, the round trip returns N times the field, for every grid shape
    (each rank's X, Y and Z pencils and all four exchanges)."""
    nx, ny, nz = 14, 26, 38
    a = np.random.default_rng(sum(grid)).standard_normal((nz, ny, nx))
    spec, back, tf, tb = O.ranks_fwd_bwd(a, *grid, reps=2)
    ref = O.Plan(nx, ny, nz).ftran_r2c(a)
    assert rel(spec, ref) < 1e-13
    assert rel(spec, np.fft.rfftn(a)) < 1e-13
    assert rel(back, a * (nx * ny * nz)) < 1e-13
    assert np.all(tf > 0) and np.all(tb > 0)


def test_ranks_path_timer_slots():
    """The 12 per-rank timer slots of the distributed path: every slot the
    reference fills on a 2 x 2 grid is timed (slots 1-12; the exchanges 1-4
    include the barrier wait) and non-negative, and the slots of one
    repetition add up to no more than that repetition's wall time."""
    a = np.random.default_rng(3).standard_normal((40, 36, 32))
    reps = 3
    sl = np.full((4, reps, 12), -1.0)
    _, _, tf, tb = O.ranks_fwd_bwd(a, 2, 2, reps=reps, spectrum=False, back=False, timers=sl)
    assert np.all(sl >= 0)
    # compute slots on every rank and every repetition
    assert np.all(sl[:, :, [4, 6, 7, 8, 9, 11]] > 0)
    assert np.all(sl.sum(axis=2)[0] <= (tf + tb) * 1.05 + 1e-4)
    with pytest.raises(ValueError):
        O.ranks_fwd_bwd(a, 2, 2, reps=reps, timers=np.zeros((4, reps, 11)))


def test_ranks_path_rejects_grids_without_a_pencil_per_rank():
    """More ranks along an axis than grid lines is refused (no rank may hold
    an empty pencil), not run."""
    a = np.zeros((4, 4, 4))
    with pytest.raises(ValueError):
        O.ranks_fwd_bwd(a, 8, 1)
    with pytest.raises(ValueError):
        O.ranks_fwd_bwd(a, 1, 8)


def test_dims_create_like_the_drivers():
    assert O.dims_create(16) == (4, 4)
    assert O.dims_create(12) == (3, 4)
    assert O.dims_create(8) == (2, 4)
    assert O.dims_create(7) == (1, 7)
