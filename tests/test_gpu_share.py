"""Share mode on the GPU: one rank of a P1 x P2 grid.

* pack/unpack kernels against the oracle's restatement of fcomm1/fcomm2/
  bcomm1/bcomm2 on integer-coded buffers: bit-exact, including positions the
  reference leaves unwritten (sentinel-filled on both sides);
* every rank of small grids run stage by stage on the GPU with the exchanges
  done on the host as MPI_Alltoallv would: the spectra equal the single-rank
  transform and the round trip returns the field;
* one rank of 2048^3 on a 16x16 grid at its real pencil shape (2048,128,128).
"""
import numpy as np
import pytest

import p3dfft_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = {8: 1e-12, 4: 2e-6}
SENT = -7.0 - 3.0j


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert p3dfft_amd.device_count() > 0, "no GPU visible to the HIP runtime"


def coded(n, seed):
    v = np.arange(1, n + 1, dtype=np.float64) + 0.25 * seed
    return v - 1j * v


def dev(S, arr):
    b = S.alloc(max(arr.size, 1) * np.dtype(S.cdtype).itemsize)
    b.upload(np.ascontiguousarray(arr, dtype=S.cdtype))
    return b


def get(S, b, n):
    return b.download((n,), S.cdtype)


GRIDS = [((2, 3), (14, 26, 38), (None, None, None)),
         ((3, 2), (14, 26, 38), (None, None, None)),
         ((4, 4), (16, 12, 10), (None, None, None)),
         ((1, 3), (8, 9, 10), (None, None, None)),
         ((2, 3), (12, 16, 10), (12, 9, 10)),
         ((3, 3), (12, 15, 9), (10, 7, 9)),
         ((16, 2), (24, 20, 6), (None, None, None))]


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("grid", GRIDS)
def test_pack_unpack_bit_exact(precision, grid):
    dims, (nx, ny, nz), cut = grid
    S = p3dfft_amd.Share(precision)
    for ip in range(dims[0]):
        for jp in range(dims[1]):
            S.setup_share(dims, ip, jp, nx, ny, nz, *cut)
            try:
                o = O.Plan(nx, ny, nz, *(c or 0 for c in cut), dims=dims, ipid=ip, jpid=jp)
                cases = [
                    (p3dfft_amd.PACK_FCOMM1, S.n_buf2(), S.n_send1(), o.fcomm1_pack),
                    (p3dfft_amd.UNPACK_FCOMM1, S.iisize * ny * S.kjsize, S.n_ypen(), o.fcomm1_unpack),
                    (p3dfft_amd.PACK_FCOMM2, S.n_ypen(), S.n_send2(), o.fcomm2_pack),
                    (p3dfft_amd.UNPACK_BCOMM1, S.n_send2(), S.n_ypen(), o.bcomm1_unpack),
                    (p3dfft_amd.PACK_BCOMM2, S.n_ypen(), S.n_ypen(), o.bcomm2_pack),
                    (p3dfft_amd.UNPACK_BCOMM2, S.n_send1(), S.n_buf2(), o.bcomm2_unpack),
                ]
                for which, nin, nout, ref_fn in cases:
                    src = coded(nin, which)
                    a = dev(S, src)
                    b = dev(S, np.full(nout, SENT))
                    S.pack(which, a, b)
                    got = get(S, b, nout)
                    ref = ref_fn(src, fill=SENT)
                    assert ref.size == nout
                    assert np.array_equal(got.astype(np.complex128), ref), (which, ip, jp)
                    a.free(); b.free()
            finally:
                S.clean()


def run_grid_on_gpu(precision, dims, nx, ny, nz, a, op="fft"):
    """All ranks of a dims grid, stage by stage on the GPU; exchanges on the host."""
    p1, p2 = dims
    S = p3dfft_amd.Share(precision)
    ranks = [(i, j) for i in range(p1) for j in range(p2)]
    info, C = {}, {}

    def stage(r, fn):
        S.setup_share(dims, r[0], r[1], nx, ny, nz)
        try:
            info[r] = dict(n_send1=S.n_send1(), n_send2=S.n_send2(), n_ypen=S.n_ypen(), n_zpen=S.n_zpen(),
                           n_real=S.n_real(), jist=S.get_dims(1)[0][1], kjst=S.get_dims(1)[0][2],
                           jisize=S.jisize, kjsize=S.kjsize)
            C[r] = S.counts()
            return fn()
        finally:
            S.clean()

    def run_x(r):
        def f():
            d = info[r]
            loc = a[d["kjst"] - 1:d["kjst"] - 1 + d["kjsize"], d["jist"] - 1:d["jist"] - 1 + d["jisize"], :]
            inp = S.alloc(max(loc.size, 1) * np.dtype(S.rdtype).itemsize)
            inp.upload(np.ascontiguousarray(loc, dtype=S.rdtype))
            out = S.alloc(max(d["n_send1"], 1) * np.dtype(S.cdtype).itemsize)
            S.ftran_x(inp, out)
            return get(S, out, d["n_send1"]).astype(np.complex128)
        return f

    send1 = {r: stage(r, run_x(r)) for r in ranks}
    send2 = {}
    for (r, j) in ranks:
        nrecv = sum(C[(r, j)]["frcv_ct"])
        recv = np.zeros(nrecv, complex)
        for i in range(p1):
            cs, cc = C[(i, j)]["fsnd_st"][r], C[(i, j)]["fsnd_ct"][r]
            rs = C[(r, j)]["frcv_st"][i]
            recv[rs:rs + cc] = send1[(i, j)][cs:cs + cc]

        def fy(recv=recv, r=r, j=j):
            b = dev(S, recv)
            out = S.alloc(max(info[(r, j)]["n_send2"], 1) * np.dtype(S.cdtype).itemsize)
            S.ftran_y(b, out)
            return get(S, out, info[(r, j)]["n_send2"]).astype(np.complex128)
        send2[(r, j)] = stage((r, j), fy)
    spec = {}
    for (i, r) in ranks:
        n = info[(i, r)]["n_zpen"]
        recv = np.zeros(n, complex)
        for j in range(p2):
            cs, cc = C[(i, j)]["ksnd_st"][r], C[(i, j)]["ksnd_ct"][r]
            rs = C[(i, r)]["krcv_st"][j]
            recv[rs:rs + cc] = send2[(i, j)][cs:cs + cc]

        def fz(recv=recv, n=n):
            b = dev(S, recv)
            S.ftran_z(b, b, op)
            return get(S, b, n).astype(np.complex128)
        spec[(i, r)] = stage((i, r), fz)
    # backward
    sendb = {}
    for r in ranks:
        def fbz(r=r):
            b = dev(S, spec[r])
            S.btran_z(b, b, op[2] + "ff")
            return get(S, b, info[r]["n_zpen"]).astype(np.complex128)
        sendb[r] = stage(r, fbz)
    sendy = {}
    for (i, r) in ranks:
        recv = np.zeros(info[(i, r)]["n_send2"], complex)
        for j in range(p2):
            cs, cc = C[(i, j)]["krcv_st"][r], C[(i, j)]["krcv_ct"][r]
            rs = C[(i, r)]["ksnd_st"][j]
            recv[rs:rs + cc] = sendb[(i, j)][cs:cs + cc]

        def fby(recv=recv, i=i, r=r):
            b = dev(S, recv)
            out = S.alloc(max(info[(i, r)]["n_ypen"], 1) * np.dtype(S.cdtype).itemsize)
            S.btran_y(b, out)
            return get(S, out, info[(i, r)]["n_ypen"]).astype(np.complex128)
        sendy[(i, r)] = stage((i, r), fby)
    field = np.zeros((nz, ny, nx))
    for (r, j) in ranks:
        recv = np.zeros(info[(r, j)]["n_send1"], complex)
        for i in range(p1):
            cs, cc = C[(i, j)]["frcv_st"][r], C[(i, j)]["frcv_ct"][r]
            rs = C[(r, j)]["fsnd_st"][i]
            recv[rs:rs + cc] = sendy[(i, j)][cs:cs + cc]

        def fbx(recv=recv, r=r, j=j):
            d = info[(r, j)]
            b = dev(S, recv)
            out = S.alloc(max(d["n_real"], 1) * np.dtype(S.rdtype).itemsize)
            S.btran_x(b, out)
            return out.download((d["kjsize"], d["jisize"], nx), S.rdtype)
        d = info[(r, j)]
        field[d["kjst"] - 1:d["kjst"] - 1 + d["kjsize"], d["jist"] - 1:d["jist"] - 1 + d["jisize"], :] = \
            stage((r, j), fbx)
    return spec, field, info


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("dims", [(2, 3), (3, 2), (1, 4), (4, 1)])
def test_all_ranks_stage_by_stage(precision, dims):
    nx, ny, nz = 14, 26, 38
    a = np.random.default_rng(5).standard_normal((nz, ny, nx))
    spec, field, info = run_grid_on_gpu(precision, dims, nx, ny, nz, a)
    full = O.Plan(nx, ny, nz).ftran_r2c(a)
    for (i, j), s in spec.items():
        o = O.Plan(nx, ny, nz, dims=dims, ipid=i, jpid=j)
        ref = full[:, o.jjstart - 1:o.jjend, o.iistart - 1:o.iiend].reshape(-1)
        assert rel(s, ref) < TOL[precision], (i, j)
    assert rel(field, a * nx * ny * nz) < TOL[precision] * 10


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("dims", [(2, 3), (4, 2)])
def test_all_ranks_fused_stages(precision, dims):
    """Lengths with specialised kernels (64, 48, 32): the X and Y stages read
    and write the exchange buffers through the maps directly.  Results match
    the single-rank transform and the separate pack/unpack passes (to the
    last bits: the fused kernels are other instantiations of the butterflies,
    which the compiler may contract into FMAs differently)."""
    nx, ny, nz = 64, 48, 32
    a = np.random.default_rng(6).standard_normal((nz, ny, nx))
    runs = []
    for fused in (1, 0):
        p3dfft_amd.load(precision).p3dfft_share_set_fused(fused)
        try:
            runs.append(run_grid_on_gpu(precision, dims, nx, ny, nz, a))
        finally:
            p3dfft_amd.load(precision).p3dfft_share_set_fused(1)
    (spec, field, _), (spec0, field0, _) = runs
    full = O.Plan(nx, ny, nz).ftran_r2c(a)
    for (i, j), s in spec.items():
        o = O.Plan(nx, ny, nz, dims=dims, ipid=i, jpid=j)
        ref = full[:, o.jjstart - 1:o.jjend, o.iistart - 1:o.iiend].reshape(-1)
        assert rel(s, ref) < TOL[precision], (i, j)
        assert rel(s, spec0[(i, j)]) < TOL[precision] / 10, (i, j)
    assert rel(field, a * nx * ny * nz) < TOL[precision] * 10
    assert rel(field, field0) < TOL[precision] / 10


@pytest.mark.parametrize("precision", [8, 4])
def test_all_ranks_y2048_persistent_table_kernel(precision):
    """A 2048-point Y axis on a 2 x 2 grid: every rank's fused Y stages run
    the table-addressed persistent column kernel (fewer tiles than XCDs, the
    last tile partial), in both precisions.  All ranks against the
    single-rank oracle, and to rounding against one workgroup per tile."""
    nx, ny, nz, dims = 10, 2048, 12, (2, 2)
    a = np.random.default_rng(7).standard_normal((nz, ny, nx))
    lib = p3dfft_amd.load(precision)
    runs = []
    for pipe in (1, 0):
        lib.p3dfft_amd_set_col_pipe(pipe)
        try:
            runs.append(run_grid_on_gpu(precision, dims, nx, ny, nz, a))
        finally:
            lib.p3dfft_amd_set_col_pipe(1)
    (spec, field, _), (spec0, field0, _) = runs
    full = O.Plan(nx, ny, nz).ftran_r2c(a)
    for (i, j), s in spec.items():
        o = O.Plan(nx, ny, nz, dims=dims, ipid=i, jpid=j)
        ref = full[:, o.jjstart - 1:o.jjend, o.iistart - 1:o.iiend].reshape(-1)
        assert rel(s, ref) < TOL[precision], (i, j)
        assert rel(s, spec0[(i, j)]) < TOL[precision] / 10, (i, j)
    assert rel(field, a * nx * ny * nz) < TOL[precision] * 10
    assert rel(field, field0) < TOL[precision] / 10


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("grid", [((2, 3), (32, 48, 16), (20, 30, 12)), ((3, 2), (64, 32, 48), (40, 20, 30))])
def test_fused_stages_with_cuts_match_separate(precision, grid):
    """Pruned grids (fewer stored x modes than nx/2+1, Y and Z cuts): on
    every rank the six stages through the fused table-addressed kernels equal
    to rounding the separate pack/unpack passes on the same random buffers;
    those passes' maps are bit-exact against the oracle's
    (test_pack_unpack_bit_exact), their FFTs the plain column and X kernels."""
    dims, (nx, ny, nz), cut = grid
    S = p3dfft_amd.Share(precision)
    rng = np.random.default_rng(sum(grid[1]) + precision)

    def crandn(m):
        return (rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(S.cdtype)

    def run(fn, inp, nout, real_in=False, real_out=False):
        outs = []
        for fused in (1, 0):
            S.lib.p3dfft_share_set_fused(fused)
            a = S.alloc(max(inp.size, 1) * inp.itemsize)
            a.upload(inp)
            isz = np.dtype(S.rdtype if real_out else S.cdtype).itemsize
            b = S.alloc(max(nout, 1) * isz)
            b.upload(np.zeros(nout, dtype=S.rdtype if real_out else S.cdtype))
            fn(a, b)
            outs.append(b.download(nout, S.rdtype if real_out else S.cdtype).astype(np.complex128))
            a.free(); b.free()
        return outs

    try:
        for ip in range(dims[0]):
            for jp in range(dims[1]):
                S.setup_share(dims, ip, jp, nx, ny, nz, *cut)
                try:
                    stages = [
                        ("ftran_x", lambda a, b: S.ftran_x(a, b),
                         rng.standard_normal(S.n_real()).astype(S.rdtype), S.n_send1(), False),
                        ("ftran_y", lambda a, b: S.ftran_y(a, b), crandn(S.n_ypen()), S.n_send2(), False),
                        ("btran_y", lambda a, b: S.btran_y(a, b), crandn(S.n_send2()), S.n_ypen(), False),
                        ("btran_x", lambda a, b: S.btran_x(a, b), crandn(S.n_send1()), S.n_real(), True),
                    ]
                    for name, fn, inp, nout, real_out in stages:
                        u, v = run(fn, inp, nout, real_out=real_out)
                        assert np.linalg.norm(u) > 0, (name, ip, jp)
                        assert rel(u, v) < TOL[precision] / 10, (name, ip, jp)
                finally:
                    S.clean()
    finally:
        S.lib.p3dfft_share_set_fused(1)


def test_one_rank_of_2048cube_on_16x16_fp32():
    """Rank (15, 3) of 2048^3 on 16x16 (local X pencil (2048,128,128), the last
    row rank holds 65 kx): each stage's send buffer against numpy FFTs and the
    oracle's pack maps, on synthetic receive buffers."""
    n, dims, ip, jp = 2048, (16, 16), 15, 3
    S = p3dfft_amd.Share(4)
    S.setup_share(dims, ip, jp, n, n, n)
    o = O.Plan(n, n, n, dims=dims, ipid=ip, jpid=jp)
    rng = np.random.default_rng(2048)
    try:
        assert (S.iisize, S.jisize, S.jjsize, S.kjsize) == (65, 128, 128, 128)
        loc = rng.standard_normal((128, 128, n)).astype(np.float32)
        inp = S.alloc(loc.nbytes)
        inp.upload(loc)
        send1 = S.alloc(S.n_send1() * 8)
        S.ftran_x(inp, send1)
        got = get(S, send1, S.n_send1())
        ref = o.fcomm1_pack(np.fft.rfft(loc.astype(np.float64), axis=2))
        assert rel(got, ref) < TOL[4]
        inp.free(); send1.free()
        recv1 = (rng.standard_normal(S.n_ypen()) + 1j * rng.standard_normal(S.n_ypen())).astype(np.complex64)
        r1 = dev(S, recv1)
        send2 = S.alloc(S.n_send2() * 8)
        S.ftran_y(r1, send2)
        got = get(S, send2, S.n_send2())
        ypen = o.fcomm1_unpack(recv1.astype(np.complex128)).reshape(128, n, 65)
        ref = o.fcomm2_pack(np.fft.fft(ypen, axis=1))
        assert rel(got, ref) < TOL[4]
        r1.free(); send2.free()
        zp = (rng.standard_normal(S.n_zpen()) + 1j * rng.standard_normal(S.n_zpen())).astype(np.complex64)
        z = dev(S, zp)
        S.ftran_z(z, z, "fft")
        got = get(S, z, S.n_zpen()).reshape(n, 128, 65)
        assert rel(got, np.fft.fft(zp.astype(np.complex128).reshape(n, 128, 65), axis=0)) < TOL[4]
        z.free()
    finally:
        S.clean()


def test_one_rank_of_2048cube_on_16x16_fp32_backward():
    """Rank (15, 3) of 2048^3 on 16x16, backward stages at the real pencil
    shape: btran_z against numpy's unnormalised inverse FFT along z, btran_y
    against the oracle's bcomm1 unpack (bcomm1.F90:309-380) + inverse FFT along
    y + bcomm2 pack (bcomm2.F90:233-273), btran_x against the oracle's bcomm2
    unpack (bcomm2.F90:290-313) + numpy irfft, on synthetic receive buffers."""
    n, dims, ip, jp = 2048, (16, 16), 15, 3
    S = p3dfft_amd.Share(4)
    S.setup_share(dims, ip, jp, n, n, n)
    o = O.Plan(n, n, n, dims=dims, ipid=ip, jpid=jp)
    rng = np.random.default_rng(2049)

    def crandn(m):
        return (rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(np.complex64)
    try:
        assert (S.iisize, S.jisize, S.jjsize, S.kjsize) == (65, 128, 128, 128)
        spec = crandn(S.n_zpen())
        z = dev(S, spec)
        send = S.alloc(S.n_zpen() * 8)
        S.btran_z(z, send, "tff")
        got = get(S, send, S.n_zpen()).reshape(n, 128, 65)
        ref = np.fft.ifft(spec.astype(np.complex128).reshape(n, 128, 65), axis=0) * n
        assert rel(got, ref) < TOL[4]
        z.free(); send.free()
        del got, ref, spec
        recv = crandn(S.n_send2())
        r = dev(S, recv)
        sendy = S.alloc(S.n_ypen() * 8)
        S.btran_y(r, sendy)
        got = get(S, sendy, S.n_ypen())
        ypen = o.bcomm1_unpack(recv.astype(np.complex128)).reshape(128, n, 65)
        ref = o.bcomm2_pack(np.fft.ifft(ypen, axis=1) * n)
        assert ref.size == S.n_ypen()
        assert rel(got, ref) < TOL[4]
        r.free(); sendy.free()
        del got, ref, ypen, recv
        recv = crandn(S.n_send1())
        r = dev(S, recv)
        out = S.alloc(S.n_real() * 4)
        S.btran_x(r, out)
        got = out.download((128, 128, n), S.rdtype)
        xpen = o.bcomm2_unpack(recv.astype(np.complex128)).reshape(128, 128, n // 2 + 1)
        ref = np.fft.irfft(xpen, n=n, axis=2) * n
        assert rel(got, ref) < TOL[4]
        r.free(); out.free()
    finally:
        S.clean()


def test_col_pipe_2048cube_rank_fp32():
    """The persistent column kernel (p3dfft_amd_set_col_pipe(1): one workgroup
    per CU looping over tiles, the next tile's loads in flight while the
    current one computes) against one workgroup per tile at rank (15, 3) of
    2048^3 on 16x16: ftran_y / btran_y (table-addressed rows) and ftran_z /
    btran_z (1040 tiles of length 2048) equal to rounding (the same stage
    code, contracted to FMAs differently inside the loop), ftran_z against
    numpy."""
    n, dims, ip, jp = 2048, (16, 16), 15, 3
    S = p3dfft_amd.Share(4)
    S.setup_share(dims, ip, jp, n, n, n)
    rng = np.random.default_rng(2050)

    def crandn(m):
        return (rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(np.complex64)
    recv1, recv2, zp = crandn(S.n_ypen()), crandn(S.n_send2()), crandn(S.n_zpen())
    outs = {}
    try:
        for pipe in (0, 1):
            S.lib.p3dfft_amd_set_col_pipe(pipe)
            r, s2 = dev(S, recv1), S.alloc(S.n_send2() * 8)
            S.ftran_y(r, s2)
            fy = get(S, s2, S.n_send2())
            r.free(); s2.free()
            r, sy = dev(S, recv2), S.alloc(S.n_ypen() * 8)
            S.btran_y(r, sy)
            by = get(S, sy, S.n_ypen())
            r.free(); sy.free()
            z = dev(S, zp)
            S.ftran_z(z, z, "fft")
            fz = get(S, z, S.n_zpen())
            z.free()
            z, snd = dev(S, zp), S.alloc(S.n_zpen() * 8)
            S.btran_z(z, snd, "tff")
            bz = get(S, snd, S.n_zpen())
            z.free(); snd.free()
            outs[pipe] = (fy, by, fz, bz)
        for name, u, v in zip(("ftran_y", "btran_y", "ftran_z", "btran_z"), outs[0], outs[1]):
            assert rel(u, v) < TOL[4] / 10, name
        ref = np.fft.fft(zp.astype(np.complex128).reshape(n, 128, 65), axis=0)
        assert rel(outs[1][2].reshape(n, 128, 65), ref) < TOL[4]
    finally:
        S.lib.p3dfft_amd_set_col_pipe(1)
        S.clean()


def coded_unique(n, seed):
    """n distinct complex values exact in FP32: the index split into two
    22-bit halves (a permutation error cannot hide between equal values)."""
    i = np.arange(n, dtype=np.int64)
    return (i % (1 << 22) + 1).astype(np.float64) - 1j * ((i >> 22) + 1 + 0.25 * seed)


def test_pack_unpack_bit_exact_2048cube_rank_fp32():
    """The six separate pack/unpack kernels at the real shape of rank (15, 3)
    of 2048^3 on 16x16 (65-wide last row rank, buffers of 2^24 elements and
    more) against the oracle's maps, bit-exact on index-coded buffers."""
    n, dims, ip, jp = 2048, (16, 16), 15, 3
    S = p3dfft_amd.Share(4)
    S.setup_share(dims, ip, jp, n, n, n)
    o = O.Plan(n, n, n, dims=dims, ipid=ip, jpid=jp)
    try:
        cases = [
            (p3dfft_amd.PACK_FCOMM1, S.n_buf2(), S.n_send1(), o.fcomm1_pack),
            (p3dfft_amd.UNPACK_FCOMM1, S.iisize * n * S.kjsize, S.n_ypen(), o.fcomm1_unpack),
            (p3dfft_amd.PACK_FCOMM2, S.n_ypen(), S.n_send2(), o.fcomm2_pack),
            (p3dfft_amd.UNPACK_BCOMM1, S.n_send2(), S.n_ypen(), o.bcomm1_unpack),
            (p3dfft_amd.PACK_BCOMM2, S.n_ypen(), S.n_ypen(), o.bcomm2_pack),
            (p3dfft_amd.UNPACK_BCOMM2, S.n_send1(), S.n_buf2(), o.bcomm2_unpack),
        ]
        for which, nin, nout, ref_fn in cases:
            src = coded_unique(nin, which)
            a = dev(S, src)
            b = dev(S, np.full(nout, SENT))
            S.pack(which, a, b)
            got = get(S, b, nout)
            ref = ref_fn(src, fill=SENT)
            assert ref.size == nout
            assert np.array_equal(got.astype(np.complex128), ref), which
            a.free(); b.free()
            del src, got, ref
    finally:
        S.clean()


@pytest.mark.parametrize("precision", [8, 4])
def test_pack_order_against_hand_worked_buffers(precision):
    """The GPU fcomm1 and bcomm2 pack kernels reproduce the send buffers of
    rank (1, 1) of 8 x 6 x 5 on 2 x 2 that were worked out by hand from
    fcomm1.F90:239-253 and bcomm2.F90:233-273 (tests/golden/hand_pack_2x2.json):
    the order inside each peer block, independent of the oracle."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "hand_pack_2x2.json")) as f:
        h = json.load(f)
    S = p3dfft_amd.Share(precision)
    S.setup_share(tuple(h["dims"]), h["rank"][0], h["rank"][1], *h["grid"])
    try:
        for which, nin, ref in ((p3dfft_amd.PACK_FCOMM1, S.n_buf2(), h["fcomm1_send"]),
                                (p3dfft_amd.PACK_BCOMM2, S.n_ypen(), h["bcomm2_send"])):
            a = dev(S, np.arange(nin, dtype=np.float64) + 0j)
            b = dev(S, np.full(len(ref), SENT))
            S.pack(which, a, b)
            got = get(S, b, len(ref))
            assert np.array_equal(got.real.astype(np.float64), np.array(ref, dtype=np.float64)), which
            assert np.all(got.imag == 0), which
            a.free(); b.free()
    finally:
        S.clean()


def test_proc_dims_and_get_proc_parts():
    """The processor-grid queries (build/setup.F90:224-229,551-574 and
    get_proc_parts, build/module.F90:888-1054) on share-mode setups: the
    proc_dims table seen from one rank equals what every rank's own setup
    reports through p3dfft_get_dims; rank ids follow the reference's
    Cartesian order; get_proc_parts matches a hand-worked split and, on
    random boxes, returns pieces in i-then-j order that each lie inside their
    rank's extents and tile the box exactly."""
    S = p3dfft_amd.Share(8)
    # hand-worked: 8 x 6 x 5 on 2 x 2; y splits (1-3 | 4-6) over ipid, z
    # splits (1-2 | 3-5) over jpid; rank id = jpid * 2 + ipid
    S.setup_share((2, 2), 1, 1, 8, 6, 5)
    try:
        ierr, parts, _ = S.get_proc_parts((1, 2, 2), (8, 4, 3), 1)
        assert ierr == 0
        assert parts == [(0, (1, 2, 2), (8, 2, 1)), (2, (1, 2, 3), (8, 2, 2)),
                         (1, (1, 4, 2), (8, 2, 1)), (3, (1, 4, 3), (8, 2, 2))]
        assert S.get_proc_parts((1, 6, 1), (8, 2, 1), 1)[0] == -1
        assert S.get_proc_parts((1, 1, 1), (1, 1, 1), 3)[0] == 1
    finally:
        S.clean()
    rng = np.random.default_rng(90)
    for dims, (nx, ny, nz) in (((3, 4), (14, 26, 38)), ((4, 2), (24, 20, 9)), ((1, 3), (8, 9, 10))):
        own = {}
        for ip in range(dims[0]):
            for jp in range(dims[1]):
                S.setup_share(dims, ip, jp, nx, ny, nz)
                try:
                    own[(ip, jp)] = (S.get_dims(1), S.get_dims(2))
                    pd, coords = S.proc_dims()
                finally:
                    S.clean()
        S.setup_share(dims, dims[0] - 1, 0, nx, ny, nz)
        try:
            pd, coords = S.proc_dims()
            P = dims[0] * dims[1]
            assert pd.shape == (P, 2, 9)
            ext = {}
            for p in range(P):
                ip, jp = coords[p]
                assert p == jp * dims[0] + ip
                for conf in (1, 2):
                    st, en, sz = own[(ip, jp)][conf - 1]
                    assert tuple(pd[p, conf - 1]) == st + en + sz, (dims, p, conf)
                    ext[(p, conf)] = (np.array(st), np.array(en))
            for conf in (1, 2):
                full = np.array([nx, ny, nz]) if conf == 1 else np.array([nx // 2 + 1, ny, nz])
                for _ in range(20):
                    lo = np.array([rng.integers(1, full[a] + 1) for a in range(3)])
                    hi = np.array([rng.integers(lo[a], full[a] + 1) for a in range(3)])
                    ierr, parts, _ = S.get_proc_parts(lo, hi - lo + 1, conf)
                    assert ierr == 0
                    cover = np.zeros(tuple(hi - lo + 1), dtype=np.int32)
                    ia = 1 if conf == 1 else 0
                    ja = 2 if conf == 1 else 1
                    prev = None
                    for rank, b, s in parts:
                        b, s = np.array(b), np.array(s)
                        st, en = ext[(rank, conf)]
                        assert np.all(b >= st) and np.all(b + s - 1 <= en), (conf, rank, b, s)
                        key = (b[ia], b[ja])
                        assert prev is None or key > prev  # i outer, then j
                        prev = key
                        cover[tuple(slice(b[a] - lo[a], b[a] - lo[a] + s[a]) for a in range(3))] += 1
                    assert np.all(cover == 1), (dims, conf, lo, hi)
        finally:
            S.clean()
