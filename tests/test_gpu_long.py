"""Transform lengths above the one-workgroup limit (8192 points) and DCT-I /
DST-I extensions above it, through the C-ABI against the FP64 oracle.

Such lengths run in several passes over a scratch buffer: four-step when the
length splits into two factors <= 8192 whose primes are at most 127, else
Bluestein's chirp-z convolution through a 5-smooth length M >= 2m-1.  The
reference accepts any length (FFTW plans it, build/init_plan.F90:100-128).

Tolerances: FP64 1e-12 relative L2 against the oracle; FP32 2e-6 (the FP32
Bluestein runs three FP32 FFTs of length M, its filter's transform included).
The oracle itself is pinned at these lengths against numpy / scipy in
tests/test_oracle.py (test_oracle_long_lengths_against_numpy).
"""
import numpy as np
import pytest

from test_gpu_parity import PREC, TOL, lib_setup, oplan, rel

pytestmark = pytest.mark.gpu

# (nx, ny, nz), what the long axis runs as
LONG = [
    ((4, 3, 16384), "z four-step 128 x 128 (specialised sub-lengths)"),
    ((4, 3, 12000), "z four-step 100 x 120 (generic sub-lengths)"),
    ((3, 2, 8209), "z prime: Bluestein, M = 16875 = 125 x 135"),
    ((16384, 3, 2), "x four-step, two-for-one rows"),
    ((8209, 2, 3), "x odd prime: Bluestein"),
    ((4, 9000, 2), "y four-step 90 x 100"),
]


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("shape,what", LONG)
def test_long_lengths_fwd_bwd(precision, shape, what):
    nx, ny, nz = shape
    f = np.random.default_rng(nx + 7 * ny + 13 * nz).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz)
    ref = o.ftran_r2c(f, "fft")
    with lib_setup(precision, nx, ny, nz) as P:
        s = P.ftran_r2c(f, "fft")
        assert rel(s, ref) < TOL[precision], what
        b = P.btran_c2r(ref, "tff")
    refb = o.btran_c2r(ref, "tff")[0]
    assert rel(b, refb) < TOL[precision], what
    n = nx * ny * nz
    assert np.abs(b / n - f).max() < PREC[precision] * n * 0.25, what


@pytest.mark.parametrize("precision", [8, 4])
def test_long_dct1_dst1_extensions(precision):
    """DCT-I of nz = 4200 (extension 8398 = 38 x 221, four-step) and DST-I of
    the same nz (extension 8402 = 2 x 4201: Bluestein), forward and backward,
    against the oracle; nz = 8192 'ffc' (extension 16382 = 2 x 8191:
    Bluestein, M = 32768), which setup once refused."""
    for (nx, ny, nz) in ((3, 2, 4200), (4, 3, 8192)):
        f = np.random.default_rng(nz).standard_normal((nz, ny, nx))
        o = oplan(nx, ny, nz)
        with lib_setup(precision, nx, ny, nz) as P:
            for fop, bop in (("ffc", "cff"), ("ffs", "sff")):
                ref = o.ftran_r2c(f, fop)
                assert rel(P.ftran_r2c(f, fop), ref) < TOL[precision], (nz, fop)
                assert rel(P.btran_c2r(ref, bop), o.btran_c2r(ref, bop)[0]) < TOL[precision], (nz, bop)


@pytest.mark.parametrize("precision", [8, 4])
def test_long_cheby(precision):
    """p3dfft_cheby along nz = 4500 (DCT-I extension 8998 = 2 x 11 x 409:
    Bluestein), the recurrence in its own pass, against the oracle.  The
    field is i.i.d. N(0,1): the derivative recurrence multiplies coefficient
    m by up to (4/Lz) m, so a field whose coefficients decay (a smooth
    polynomial) leaves the round-off of thousands of near-zero coefficients,
    amplified ~nzc^2, as the whole difference between two correct results
    (1.5e-10 in FP64 for a degree-40 polynomial at this nzc); with
    coefficients of every degree the signal is amplified alike and the
    relative error is the transform's."""
    nx, ny, nz, Lz = 4, 2, 4500, 2.0
    a = np.random.default_rng(46).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz)
    with lib_setup(precision, nx, ny, nz) as P:
        assert rel(P.cheby(a, Lz), o.cheby(a, Lz)) < TOL[precision]


@pytest.mark.parametrize("precision", [8, 4])
def test_long_pruned_and_stride1(precision):
    """Long Z and Y lengths with cuts (the load/store row maps of the
    multi-pass form), and the STRIDE1 layout (length-fastest views)."""
    cases = [dict(shape=(4, 3, 16384), nzc=10001), dict(shape=(4, 9000, 2), nyc=5001),
             dict(shape=(4, 3, 12000), stride1=True), dict(shape=(4, 3, 16384), nzc=9000, stride1=True)]
    for cs in cases:
        nx, ny, nz = cs["shape"]
        kw = dict(nyc=cs.get("nyc"), nzc=cs.get("nzc"), stride1=cs.get("stride1", False))
        f = np.random.default_rng(nz + ny).standard_normal((nz, ny, nx))
        o = oplan(nx, ny, nz, **kw)
        ref = o.ftran_r2c(f, "fft")
        with lib_setup(precision, nx, ny, nz, **kw) as P:
            assert rel(P.ftran_r2c(f, "fft"), ref) < TOL[precision], cs
            assert rel(P.btran_c2r(ref, "tff"), o.btran_c2r(ref, "tff")[0]) < TOL[precision], cs


@pytest.mark.parametrize("precision", [8, 4])
def test_long_many_and_1d(precision):
    """The multivariable entry points with a long Z length (nv = 2, padded
    dim_in / dim_out) against the oracle variable by variable, and the X
    r2c stage alone (ftran_r2c_1d) of a long X length against numpy."""
    import p3dfft_amd
    nx, ny, nz = 4, 3, 16384
    o = oplan(nx, ny, nz)
    nr, nc = nx * ny * nz, int(np.prod(o.spec_shape()))
    nv, pad = 2, 5
    a = np.random.default_rng(7).standard_normal(nv * (nr + pad))
    with lib_setup(precision, nx, ny, nz) as P:
        s = P.ftran_r2c_many(a, nr + pad, nc + pad, nv, "fft")
        for j in range(nv):
            ref = o.ftran_r2c(a[j * (nr + pad):j * (nr + pad) + nr], "fft").reshape(-1)
            assert rel(s[j * (nc + pad):j * (nc + pad) + nc], ref) < TOL[precision], j
    nx, ny, nz = 9000, 2, 3
    f = np.random.default_rng(8).standard_normal((nz, ny, nx))
    P = p3dfft_amd.P3DFFT(precision)
    P.setup((1, 1), nx, ny, nz)
    try:
        got = P.ftran_r2c_1d(f).reshape(nz, ny, nx // 2 + 1)
    finally:
        P.clean()
    assert rel(got, np.fft.rfft(f, axis=2)) < TOL[precision]


@pytest.mark.parametrize("precision", [8, 4])
def test_long_capture_after_eager_run(precision):
    """A transform with a long Z length captured into a graph after it ran
    once eagerly (plans and scratch exist) replays bit for bit, three times
    with eager transforms between the replays."""
    import ctypes
    nx, ny, nz = 4, 3, 16384
    rng = np.random.default_rng(60)
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
            P.btran_c2r_dev(spec.ptr, out.ptr, "tff", st)
            L.p3dfft_amd_stream_synchronize(st)
            b0 = out.download(P.real_shape(), P.rdtype)
            assert L.p3dfft_amd_capture_begin(st) == 0
            P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft", st)
            P.btran_c2r_dev(spec.ptr, out.ptr, "tff", st)
            assert L.p3dfft_amd_capture_end(st, ctypes.byref(graph)) == 0
            for rep in range(3):
                other.upload(g2.astype(P.rdtype))
                P.ftran_r2c_dev(other.ptr, ospec.ptr, "fft", st)
                out.upload(np.zeros((nz, ny, nx), P.rdtype))
                assert L.p3dfft_amd_graph_launch(graph, st) == 0
                L.p3dfft_amd_stream_synchronize(st)
                assert np.array_equal(out.download(P.real_shape(), P.rdtype), b0), rep
        finally:
            if graph.value:
                L.p3dfft_amd_graph_destroy(graph)
            for b in (fin, spec, out, other, ospec):
                b.free()
            L.p3dfft_amd_stream_destroy(st)
