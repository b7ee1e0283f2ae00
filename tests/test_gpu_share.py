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
