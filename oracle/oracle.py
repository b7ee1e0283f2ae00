"""ctypes wrapper of the FP64 CPU oracle (oracle/p3dfft_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker that the HIP library is compared
against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module; the product (p3dfft_amd) never does.

Arrays follow the reference's Fortran layout: a real field (nx, ny, nz) with
x fastest is a numpy array of shape (nz, ny, nx); a complex spectrum
(nxhpc, nyc, nzc) is a complex128 array of shape (nzc, nyc, nxhpc); under
STRIDE1 the spectrum (nzc, nyc, nxhpc) is shaped (nxhpc, nyc, nzc).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libp3dfft_oracle.so")
_lib = None

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_ll_p = ctypes.POINTER(ctypes.c_longlong)


class OrcPlan(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "nx", "ny", "nz", "nxc", "nyc", "nzc", "ow", "stride1", "p1", "p2",
        "ipid", "jpid", "nxh", "nxhp", "nxhc", "nxhpc", "nyh", "nzh", "nyhc",
        "nzhc", "nycph", "nzcph", "iistart", "iiend", "iisize", "jistart",
        "jiend", "jisize", "jjstart", "jjend", "jjsize", "kjstart", "kjend",
        "kjsize", "padi")] + [("nm", ctypes.c_longlong), ("memsize", ctypes.c_int * 3)]


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_setup.argtypes = [ctypes.POINTER(OrcPlan), _c_int_p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_setup.restype = ctypes.c_int
        L.orc_get_dims.argtypes = [ctypes.POINTER(OrcPlan), ctypes.c_int, _c_int_p, _c_int_p, _c_int_p]
        L.orc_map_data_to_proc.argtypes = [ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p]
        for name in ("orc_ftran_r2c", "orc_btran_c2r"):
            f = getattr(L, name)
            f.argtypes = [ctypes.POINTER(OrcPlan), _c_dbl_p, _c_dbl_p, ctypes.c_char_p]
            f.restype = ctypes.c_int
        L.orc_cheby.argtypes = [ctypes.POINTER(OrcPlan), _c_dbl_p, _c_dbl_p, ctypes.c_double]
        L.orc_cheby.restype = ctypes.c_int
        L.orc_fft.argtypes = [_c_dbl_p, ctypes.c_int, ctypes.c_int]
        L.orc_dft_naive.argtypes = [_c_dbl_p, _c_dbl_p, ctypes.c_int, ctypes.c_int]
        L.orc_r2c.argtypes = [_c_dbl_p, _c_dbl_p, ctypes.c_int]
        L.orc_c2r.argtypes = [_c_dbl_p, _c_dbl_p, ctypes.c_int]
        L.orc_dct1.argtypes = [_c_dbl_p, ctypes.c_int, ctypes.c_int]
        L.orc_dst1.argtypes = [_c_dbl_p, ctypes.c_int, ctypes.c_int]
        L.orc_dct1_naive.argtypes = [_c_dbl_p, _c_dbl_p, ctypes.c_int]
        L.orc_dst1_naive.argtypes = [_c_dbl_p, _c_dbl_p, ctypes.c_int]
        for name in ("orc_fcomm1_pack", "orc_fcomm1_unpack", "orc_fcomm2_pack",
                     "orc_bcomm1_unpack", "orc_bcomm2_pack", "orc_bcomm2_unpack"):
            f = getattr(L, name)
            f.argtypes = [ctypes.POINTER(OrcPlan), _c_dbl_p, _c_dbl_p]
        L.orc_share_counts.argtypes = [ctypes.POINTER(OrcPlan)] + [_c_ll_p] * 8
        L.orc_ranks_fwd_bwd.argtypes = [ctypes.c_int] * 5 + [_c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.c_int,
                                                             _c_dbl_p, _c_dbl_p, _c_dbl_p]
        L.orc_ranks_fwd_bwd.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def dims_create(nprocs):
    """The sample drivers' processor grid: MPI_Dims_create(nprocs, 2) (the
    most balanced factorisation, larger factor first), then dims(1) <=
    dims(2) (sample/FORTRAN/driver_rand.F90:115-137)."""
    a = int(nprocs ** 0.5)
    while nprocs % a:
        a -= 1
    return (a, nprocs // a)


def ranks_fwd_bwd(a, p1, p2, reps=1, spectrum=True, back=True, timers=None):
    """The distributed CPU path ('fft' then 'tff', no cuts) on a p1 x p2
    grid with the ranks as threads.  a: real field (nz, ny, nx).  Returns
    (spectrum (nz, ny, nx/2+1) or None, back (nz, ny, nx) or None, forward
    times, backward times) in seconds.  timers: optional float64 array of
    shape (p1*p2, reps, 12), filled with each rank's 12 timer slots per
    repetition (the reference's timers(1..12) meaning, seconds)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    nz, ny, nx = a.shape
    spec = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128) if spectrum else None
    out = np.zeros_like(a) if back else None
    tf = np.zeros(reps)
    tb = np.zeros(reps)
    if timers is not None:
        if timers.shape != (p1 * p2, reps, 12) or timers.dtype != np.float64:
            raise ValueError("timers must be float64 of shape (p1*p2, reps, 12)")
    rc = lib().orc_ranks_fwd_bwd(nx, ny, nz, p1, p2, _dp(a),
                                 _dp(spec.view(np.float64)) if spectrum else None,
                                 _dp(out) if back else None, reps, _dp(tf), _dp(tb),
                                 _dp(timers) if timers is not None else None)
    if rc:
        raise ValueError("oracle ranks_fwd_bwd failed (%d)" % rc)
    return spec, out, tf, tb


def _dp(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_c_dbl_p)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def map_data_to_proc(data, proc):
    st = (ctypes.c_int * proc)()
    en = (ctypes.c_int * proc)()
    sz = (ctypes.c_int * proc)()
    lib().orc_map_data_to_proc(data, proc, st, en, sz)
    return list(st), list(en), list(sz)


class Plan:
    """One oracle setup (the reference's module state, restated per object)."""

    def __init__(self, nx, ny, nz, nxc=0, nyc=0, nzc=0, ow=True, stride1=False,
                 dims=(1, 1), ipid=0, jpid=0):
        self.p = OrcPlan()
        d = (ctypes.c_int * 2)(*dims)
        rc = lib().orc_setup(ctypes.byref(self.p), d, ipid, jpid, nx, ny, nz, nxc, nyc, nzc,
                             1 if ow else 0, 1 if stride1 else 0)
        if rc != 0:
            raise ValueError("oracle setup rejected the arguments")

    def __getattr__(self, k):
        return getattr(self.p, k)

    @property
    def memsize(self):
        return tuple(self.p.memsize)

    def get_dims(self, conf):
        s = (ctypes.c_int * 3)(); e = (ctypes.c_int * 3)(); z = (ctypes.c_int * 3)()
        lib().orc_get_dims(ctypes.byref(self.p), conf, s, e, z)
        return tuple(s), tuple(e), tuple(z)

    def real_shape(self):
        return (self.p.nz, self.p.ny, self.p.nx)

    def spec_shape(self):
        p = self.p
        if p.stride1:
            return (p.nxhpc, p.nyc, p.nzc)
        return (p.nzc, p.nyc, p.nxhpc)

    def ftran_r2c(self, a, op="fft"):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.real_shape())
        out = np.zeros(self.spec_shape(), dtype=np.complex128)
        rc = lib().orc_ftran_r2c(ctypes.byref(self.p), _dp(a), _dp(out.view(np.float64)),
                                 op.encode()[:3].ljust(3))
        if rc:
            raise ValueError("oracle ftran_r2c failed (%d) for op %r" % (rc, op))
        return out

    def btran_c2r(self, spec, op="tff", keep_input=True):
        """Returns the real field; with keep_input the caller's array is not
        modified (the oracle works on a copy)."""
        s = np.array(spec, dtype=np.complex128, order="C", copy=keep_input).reshape(self.spec_shape())
        out = np.zeros(self.real_shape(), dtype=np.float64)
        rc = lib().orc_btran_c2r(ctypes.byref(self.p), _dp(s.view(np.float64)), _dp(out),
                                 op.encode()[:3].ljust(3))
        if rc:
            raise ValueError("oracle btran_c2r failed (%d) for op %r" % (rc, op))
        return (out, s)

    def cheby(self, a, Lz):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.real_shape())
        out = np.zeros(self.spec_shape(), dtype=np.complex128)
        rc = lib().orc_cheby(ctypes.byref(self.p), _dp(a), _dp(out.view(np.float64)), float(Lz))
        if rc:
            raise ValueError("oracle cheby failed (%d)" % rc)
        return out

    # ---- share mode maps (complex element buffers) ----
    def share_counts(self):
        p1, p2 = self.p.p1, self.p.p2
        arrs = [(ctypes.c_longlong * n)() for n in (p1, p1, p1, p1, p2, p2, p2, p2)]
        lib().orc_share_counts(ctypes.byref(self.p), *arrs)
        keys = ("fsnd_st", "fsnd_ct", "frcv_st", "frcv_ct", "ksnd_st", "ksnd_ct", "krcv_st", "krcv_ct")
        return {k: list(a) for k, a in zip(keys, arrs)}

    def _map(self, name, src, nout, fill=0.0):
        src = np.ascontiguousarray(src, dtype=np.complex128)
        out = np.full(nout, fill, dtype=np.complex128)
        getattr(lib(), name)(ctypes.byref(self.p), _dp(src.view(np.float64)), _dp(out.view(np.float64)))
        return out

    def fcomm1_pack(self, buf2, fill=0.0):
        c = self.share_counts()
        return self._map("orc_fcomm1_pack", buf2, sum(c["fsnd_ct"]), fill)

    def fcomm1_unpack(self, recv, fill=0.0):
        p = self.p
        return self._map("orc_fcomm1_unpack", recv, p.iisize * p.ny * p.kjsize, fill)

    def fcomm2_pack(self, src, fill=0.0):
        c = self.share_counts()
        return self._map("orc_fcomm2_pack", src, sum(c["ksnd_ct"]), fill)

    def bcomm1_unpack(self, recv, fill=0.0):
        p = self.p
        return self._map("orc_bcomm1_unpack", recv, p.iisize * p.ny * p.kjsize, fill)

    def bcomm2_pack(self, src, fill=0.0):
        c = self.share_counts()
        return self._map("orc_bcomm2_pack", src, sum(c["frcv_ct"]), fill)

    def bcomm2_unpack(self, recv, fill=0.0):
        p = self.p
        return self._map("orc_bcomm2_unpack", recv, p.nxhp * p.jisize * p.kjsize, fill)


# ---- 1D helpers ----
def fft(x, sign=-1):
    a = np.array(x, dtype=np.complex128, copy=True)
    lib().orc_fft(_dp(a.view(np.float64)), a.size, sign)
    return a


def dft_naive(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(a)
    lib().orc_dft_naive(_dp(a.view(np.float64)), _dp(out.view(np.float64)), a.size, sign)
    return out


def r2c(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(a.size // 2 + 1, dtype=np.complex128)
    lib().orc_r2c(_dp(a), _dp(out.view(np.float64)), a.size)
    return out


def c2r(X, n):
    a = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.zeros(n, dtype=np.float64)
    lib().orc_c2r(_dp(a.view(np.float64)), _dp(out), n)
    return out


def dct1(x):
    a = np.array(x, dtype=np.complex128, copy=True)
    lib().orc_dct1(_dp(a.view(np.float64)), a.size, 1)
    return a


def dst1(x):
    a = np.array(x, dtype=np.complex128, copy=True)
    lib().orc_dst1(_dp(a.view(np.float64)), a.size, 1)
    return a


def dct1_naive(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(a)
    lib().orc_dct1_naive(_dp(a), _dp(out), a.size)
    return out


def dst1_naive(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(a)
    lib().orc_dst1_naive(_dp(a), _dp(out), a.size)
    return out


# ---- power spectrum (sample/C/driver_spec.c:298-383; the FORTRAN driver's
# compute_spectrum, sample/FORTRAN/driver_spec.F90:263-340, is the same sum
# except that for odd ng its 1-based test `y > ng/2` folds one index earlier)
def _fold(j, n):
    """|wavenumber| of storage index j on an axis of n points (driver_spec.c:336-344)."""
    j = np.asarray(j)
    return np.where(j > n // 2, n - j, j)


def power_spectrum(spec, ng, kmax, stride1=False):
    """E[ik] = sum over the stored half spectrum of k2*|B|^2, ik = int(sqrt(k2)+0.5),
    k2 = kx^2 + ky^2 + kz^2 with kx the stored x index and ky, kz folded (FP64).
    spec is shaped as this module's spectra: (nzc, nyc, nxhpc), or
    (nxhpc, nyc, nzc) under STRIDE1; ng = (nx, ny, nz).  Terms with ik > kmax
    are dropped (the reference would index past its array)."""
    B = np.asarray(spec, dtype=np.complex128)
    nx, ny, nz = ng
    if stride1:
        s2, s1, s0 = B.shape  # x slowest, z fastest
        kx = np.arange(s2)[:, None, None]
        ky = _fold(np.arange(s1), ny)[None, :, None]
        kz = _fold(np.arange(s0), nz)[None, None, :]
    else:
        s2, s1, s0 = B.shape  # z slowest, x fastest
        kz = _fold(np.arange(s2), nz)[:, None, None]
        ky = _fold(np.arange(s1), ny)[None, :, None]
        kx = np.arange(s0)[None, None, :]
    k2 = (kx * kx + ky * ky + kz * kz).astype(np.int64)
    ik = (np.sqrt(k2.astype(np.float64)) + 0.5).astype(np.int64)
    w = k2 * (B.real ** 2 + B.imag ** 2)
    keep = ik <= kmax
    return np.bincount(ik[keep].ravel(), weights=w[keep].ravel(), minlength=kmax + 1)[: kmax + 1]


def power_spectrum_loops(spec, ng, kmax):
    """The driver's loop nest element by element (standard layout), for tiny
    arrays: the brute-force check of power_spectrum."""
    B = np.asarray(spec, dtype=np.complex128)
    nzc, nyc, nxh = B.shape
    E = [0.0] * (kmax + 1)
    for z in range(nzc):
        kz = z if z <= ng[2] // 2 else ng[2] - z
        for y in range(nyc):
            ky = y if y <= ng[1] // 2 else ng[1] - y
            for x in range(nxh):
                k2 = x * x + ky * ky + kz * kz
                ik = int(k2 ** 0.5 + 0.5)
                if ik <= kmax:
                    v = B[z, y, x]
                    E[ik] += k2 * (v.real * v.real + v.imag * v.imag)
    return np.array(E)


def spectrum_kmax(nx, ny, nz):
    """The drivers' shell count (driver_spec.c:226): sqrt(nx^2+ny^2+nz^2)/2 + 1/2."""
    return int((nx * nx + ny * ny + nz * nz) ** 0.5 * 0.5 + 0.5)
