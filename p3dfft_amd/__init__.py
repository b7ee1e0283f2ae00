"""p3dfft_amd -- MI355X (gfx950) implementation of P3DFFT's forward/backward
3D FFT path behind P3DFFT's own C-ABI.

The product is the shared library built from p3dfft_amd/csrc:
  lib/libp3dfft_amd.so     double precision (reference default build)
  lib/libp3dfft_amd_sp.so  single precision (reference SINGLE_PREC build)
Both export the reference's BIND(C) entry points (p3dfft_setup,
p3dfft_get_dims, p3dfft_ftran_r2c, p3dfft_btran_c2r, p3dfft_cheby,
p3dfft_clean, get_timers, set_timers) plus device-pointer extensions.

This module is a thin ctypes mirror of that interface for Python callers
(tests, benchmark, serving loop).  It never falls back to a CPU path: if the
library is missing it raises.

Layout conventions (Fortran order of the reference, first index fastest): a
real field (nx, ny, nz) is a numpy array of shape (nz, ny, nx); the complex
spectrum (nxhpc, nyc, nzc) has shape (nzc, nyc, nxhpc); under STRIDE1 the
spectrum (nzc, nyc, nxhpc) has shape (nxhpc, nyc, nzc).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")
_LIBNAMES = {8: "libp3dfft_amd.so", 4: "libp3dfft_amd_sp.so"}
_libs = {}

_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class P3DFFTError(RuntimeError):
    pass


def lib_path(precision=8):
    return os.path.join(LIBDIR, _LIBNAMES[precision])


def load(precision=8):
    """Load (once) the library of the given precision (8 = double, 4 = single)."""
    if precision in _libs:
        return _libs[precision]
    path = lib_path(precision)
    if not os.path.exists(path):
        raise P3DFFTError("p3dfft_amd library not built: %s (run __graft_entry__.build())" % path)
    L = ctypes.CDLL(path)
    rp = ctypes.POINTER(ctypes.c_double if precision == 8 else ctypes.c_float)
    rt = ctypes.c_double if precision == 8 else ctypes.c_float
    L.p3dfft_setup.argtypes = [_ip] * 10
    L.p3dfft_setup.restype = None
    L.p3dfft_get_dims.argtypes = [_ip] * 4
    for n in ("p3dfft_ftran_r2c", "p3dfft_btran_c2r"):
        getattr(L, n).argtypes = [_vp, _vp, ctypes.c_char_p]
        getattr(L, n).restype = None
    L.p3dfft_cheby.argtypes = [_vp, _vp, rp]
    L.p3dfft_clean.argtypes = []
    L.get_timers.argtypes = [ctypes.POINTER(ctypes.c_double)]
    L.set_timers.argtypes = []
    for n in ("p3dfft_ftran_r2c_dev", "p3dfft_btran_c2r_dev"):
        getattr(L, n).argtypes = [_vp, _vp, ctypes.c_char_p, _vp]
        getattr(L, n).restype = ctypes.c_int
    L.p3dfft_cheby_dev.argtypes = [_vp, _vp, rt, _vp]
    L.p3dfft_cheby_dev.restype = ctypes.c_int
    L.p3dfft_ftran_r2c_1d.argtypes = [_vp, _vp]
    L.p3dfft_ftran_r2c_1d.restype = None
    L.p3dfft_ftran_r2c_1d_dev.argtypes = [_vp, _vp, _vp]
    L.p3dfft_ftran_r2c_1d_dev.restype = ctypes.c_int
    for n in ("p3dfft_ftran_r2c_many", "p3dfft_btran_c2r_many"):
        getattr(L, n).argtypes = [_vp, _ip, _vp, _ip, _ip, ctypes.c_char_p]
        getattr(L, n).restype = None
    L.p3dfft_cheby_many.argtypes = [_vp, _ip, _vp, _ip, _ip, rp]
    L.p3dfft_cheby_many.restype = None
    _ll = ctypes.c_longlong
    for n in ("p3dfft_ftran_r2c_many_dev", "p3dfft_btran_c2r_many_dev"):
        getattr(L, n).argtypes = [_vp, _ll, _vp, _ll, ctypes.c_int, ctypes.c_char_p, _vp]
        getattr(L, n).restype = ctypes.c_int
    L.p3dfft_cheby_many_dev.argtypes = [_vp, _ll, _vp, _ll, ctypes.c_int, rt, _vp]
    L.p3dfft_cheby_many_dev.restype = ctypes.c_int
    L.p3dfft_amd_set_stride1.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_error_mode.argtypes = [ctypes.c_int]
    L.p3dfft_amd_last_error.restype = ctypes.c_char_p
    L.p3dfft_amd_set_timing.argtypes = [ctypes.c_int]
    L.p3dfft_amd_precision.restype = ctypes.c_int
    L.p3dfft_amd_set_specialised.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_tile_order.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_wide_offsets.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_flat_order.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_plane_kernels.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_x_wave.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_dct_direct.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_col_pipe.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_x_into_out.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_plan_mode.argtypes = [ctypes.c_int]
    L.p3dfft_amd_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.POINTER(_vp)]
    L.p3dfft_amd_batch_create.restype = ctypes.c_int
    L.p3dfft_amd_batch_launch.argtypes = [_vp, _vp]
    L.p3dfft_amd_batch_launch.restype = ctypes.c_int
    L.p3dfft_amd_batch_destroy.argtypes = [_vp]
    L.p3dfft_amd_batch_destroy.restype = ctypes.c_int
    L.p3dfft_amd_capture_begin.argtypes = [_vp]
    L.p3dfft_amd_capture_begin.restype = ctypes.c_int
    L.p3dfft_amd_capture_end.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.p3dfft_amd_capture_end.restype = ctypes.c_int
    L.p3dfft_amd_graph_launch.argtypes = [_vp, _vp]
    L.p3dfft_amd_graph_launch.restype = ctypes.c_int
    L.p3dfft_amd_graph_destroy.argtypes = [_vp]
    L.p3dfft_amd_graph_destroy.restype = ctypes.c_int
    L.p3dfft_amd_stream_create.restype = _vp
    L.p3dfft_amd_stream_destroy.argtypes = [_vp]
    L.p3dfft_amd_stream_synchronize.argtypes = [_vp]
    L.p3dfft_power_spectrum.argtypes = [_vp, ctypes.c_int, _vp]
    L.p3dfft_power_spectrum.restype = ctypes.c_int
    L.p3dfft_power_spectrum_dev.argtypes = [_vp, ctypes.c_int, _vp, _vp]
    L.p3dfft_power_spectrum_dev.restype = ctypes.c_int
    L.p3dfft_power_spectrum_dev_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp]
    L.p3dfft_power_spectrum_dev_noalloc.restype = ctypes.c_int
    L.p3dfft_amd_plan_report.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.p3dfft_amd_plan_report.restype = ctypes.c_int
    L.p3dfft_amd_reset_plans.restype = ctypes.c_int
    L.p3dfft_amd_proc_dims.argtypes = [_vp, _vp]
    L.p3dfft_amd_proc_dims.restype = ctypes.c_int
    L.p3dfft_amd_get_proc_parts.argtypes = [ctypes.c_int] * 7 + [_vp, _vp]
    L.p3dfft_amd_get_proc_parts.restype = ctypes.c_int
    L.p3dfft_share_set_fused.argtypes = [ctypes.c_int]
    L.p3dfft_amd_set_chunk_bytes.argtypes = [ctypes.c_size_t]
    L.p3dfft_amd_malloc.argtypes = [ctypes.c_size_t]
    L.p3dfft_amd_malloc.restype = _vp
    L.p3dfft_amd_fill_normal.argtypes = [_vp, ctypes.c_longlong, ctypes.c_ulonglong, _vp]
    L.p3dfft_amd_fill_normal.restype = ctypes.c_int
    L.p3dfft_amd_free.argtypes = [_vp]
    for n in ("p3dfft_amd_memcpy_h2d", "p3dfft_amd_memcpy_d2h", "p3dfft_amd_memcpy_d2d"):
        getattr(L, n).argtypes = [_vp, _vp, ctypes.c_size_t]
        getattr(L, n).restype = ctypes.c_int
    L.p3dfft_amd_memset.argtypes = [_vp, ctypes.c_int, ctypes.c_size_t]
    L.p3dfft_amd_synchronize.restype = ctypes.c_int
    L.p3dfft_amd_stream_copy.argtypes = [_vp, _vp, ctypes.c_size_t, _vp]
    L.p3dfft_amd_stream_copy.restype = ctypes.c_int
    L.p3dfft_amd_device_count.restype = ctypes.c_int
    L.p3dfft_amd_event_create.restype = _vp
    L.p3dfft_amd_event_record.argtypes = [_vp, _vp]
    L.p3dfft_amd_event_elapsed_ms.argtypes = [_vp, _vp]
    L.p3dfft_amd_event_elapsed_ms.restype = ctypes.c_float
    L.p3dfft_amd_event_destroy.argtypes = [_vp]
    _ip8 = ctypes.POINTER(ctypes.c_longlong)
    L.p3dfft_setup_share.argtypes = [_ip] * 11
    L.p3dfft_setup_share.restype = ctypes.c_int
    L.p3dfft_share_counts.argtypes = [_ip8] * 8
    L.p3dfft_share_counts.restype = ctypes.c_int
    L.p3dfft_share_pack.argtypes = [ctypes.c_int, _vp, _vp, _vp]
    L.p3dfft_share_pack.restype = ctypes.c_int
    for n in ("p3dfft_share_ftran_x", "p3dfft_share_ftran_y", "p3dfft_share_btran_y", "p3dfft_share_btran_x"):
        getattr(L, n).argtypes = [_vp, _vp, _vp]
        getattr(L, n).restype = ctypes.c_int
    for n in ("p3dfft_share_ftran_z", "p3dfft_share_btran_z"):
        getattr(L, n).argtypes = [_vp, _vp, ctypes.c_char_p, _vp]
        getattr(L, n).restype = ctypes.c_int
    if L.p3dfft_amd_precision() != precision:
        raise P3DFFTError("library %s reports precision %d" % (path, L.p3dfft_amd_precision()))
    _libs[precision] = L
    return L


class DeviceBuffer:
    """HBM allocation owned by the library's HIP runtime."""

    def __init__(self, lib, nbytes):
        self.lib = lib
        self.nbytes = int(nbytes)
        self.ptr = lib.p3dfft_amd_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise P3DFFTError("device allocation of %d bytes failed" % self.nbytes)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise P3DFFTError("upload larger than buffer")
        if self.lib.p3dfft_amd_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes):
            raise P3DFFTError("H2D copy failed")

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise P3DFFTError("download larger than buffer")
        if self.lib.p3dfft_amd_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes):
            raise P3DFFTError("D2H copy failed")
        return out

    def free(self):
        if self.ptr:
            self.lib.p3dfft_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class P3DFFT:
    """Mirror of the reference's API for one setup (the library keeps one
    active setup, as the reference does)."""

    def __init__(self, precision=8):
        self.precision = precision
        self.lib = load(precision)
        self.rdtype = np.float64 if precision == 8 else np.float32
        self.cdtype = np.complex128 if precision == 8 else np.complex64
        self.active = False

    # -- reference API ----------------------------------------------------
    def setup(self, dims, nx, ny, nz, nxc=None, nyc=None, nzc=None, overwrite=True, stride1=False,
              strict=True):
        """p3dfft_setup (build/setup.F90:83-104).  Returns memsize."""
        L = self.lib
        L.p3dfft_amd_set_error_mode(1 if strict else 0)
        L.p3dfft_amd_clear_error()
        L.p3dfft_amd_set_stride1(1 if stride1 else 0)
        d = (ctypes.c_int * 2)(*dims)
        ms = (ctypes.c_int * 3)()
        args = [nx, ny, nz, 0, nxc or nx, nyc or ny, nzc or nz, 1 if overwrite else 0]
        cargs = [ctypes.c_int(v) for v in args]
        L.p3dfft_setup(d, *[ctypes.byref(c) for c in cargs], ms)
        self._raise_if_error()
        self.active = True
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nxc, self.nyc, self.nzc = nxc or nx, nyc or ny, nzc or nz
        self.stride1 = stride1
        self.nxhpc = self.nxc // 2 + 1
        self.memsize = tuple(ms)
        return self.memsize

    def get_dims(self, conf):
        s = (ctypes.c_int * 3)(); e = (ctypes.c_int * 3)(); z = (ctypes.c_int * 3)()
        c = ctypes.c_int(conf)
        self.lib.p3dfft_get_dims(s, e, z, ctypes.byref(c))
        return tuple(s), tuple(e), tuple(z)

    def proc_dims(self):
        """(proc_dims as an array (P, 2, 9): rank, conf, field; id2coords (P, 2))
        of the current setup (the reference's module arrays)."""
        P = self.lib.p3dfft_amd_proc_dims(None, None)
        if P < 0:
            self._raise_if_error(force=True)
        d = np.zeros(18 * P, dtype=np.int32)
        c = np.zeros(2 * P, dtype=np.int32)
        self.lib.p3dfft_amd_proc_dims(d.ctypes.data, c.ctypes.data)
        return d.reshape(P, 9, 2).transpose(0, 2, 1).copy(), c.reshape(P, 2)

    def get_proc_parts(self, base, size, conf):
        """get_proc_parts (build/module.F90:888-1054): (ierr, parts) with parts a
        list of (rank, (base x, y, z), (size x, y, z))."""
        P = self.lib.p3dfft_amd_proc_dims(None, None)
        if P < 0:
            self._raise_if_error(force=True)
        a = np.zeros(7 * P, dtype=np.int32)
        ierr = ctypes.c_int(0)
        n = self.lib.p3dfft_amd_get_proc_parts(*[int(v) for v in base], *[int(v) for v in size], int(conf),
                                               a.ctypes.data, ctypes.byref(ierr))
        t = a.reshape(7, P)
        parts = [(int(t[0, q]), tuple(int(v) for v in t[1:4, q]), tuple(int(v) for v in t[4:7, q]))
                 for q in range(max(n, 0))]
        return ierr.value, parts, t

    def real_shape(self):
        return (self.nz, self.ny, self.nx)

    def spec_shape(self):
        if self.stride1:
            return (self.nxhpc, self.nyc, self.nzc)
        return (self.nzc, self.nyc, self.nxhpc)

    def ftran_r2c(self, a, op="fft"):
        """Host arrays in and out (copied to/from HBM inside the call)."""
        a = np.ascontiguousarray(a, dtype=self.rdtype).reshape(self.real_shape())
        out = np.empty(self.spec_shape(), dtype=self.cdtype)
        self.lib.p3dfft_ftran_r2c(a.ctypes.data, out.ctypes.data, _op(op))
        self._raise_if_error()
        return out

    def power_spectrum(self, spec, kmax):
        """The drivers' k2-weighted shell sums (sample/C/driver_spec.c:298-383)
        of a host spectrum in this setup's layout: kmax+1 FP64 values."""
        b = np.ascontiguousarray(spec, dtype=self.cdtype).reshape(self.spec_shape())
        E = np.zeros(int(kmax) + 1, dtype=np.float64)
        if self.lib.p3dfft_power_spectrum(b.ctypes.data, int(kmax), E.ctypes.data):
            self._raise_if_error(force=True)
        return E

    def power_spectrum_dev(self, spec_ptr, kmax, e_ptr, stream=None):
        if self.lib.p3dfft_power_spectrum_dev(spec_ptr, int(kmax), e_ptr, stream):
            self._raise_if_error(force=True)

    def ftran_r2c_1d(self, a):
        """p3dfft_ftran_r2c_1d (build/ftran.F90:787-814): X r2c only, all
        nx/2+1 modes, (nz, ny, nx) real -> (nz, ny, nx/2+1) complex."""
        a = np.ascontiguousarray(a, dtype=self.rdtype).reshape(self.real_shape())
        nz, ny, nx = self.real_shape()
        out = np.empty((nz, ny, nx // 2 + 1), dtype=self.cdtype)
        self.lib.p3dfft_ftran_r2c_1d(a.ctypes.data, out.ctypes.data)
        self._raise_if_error()
        return out

    def ftran_r2c_1d_dev(self, in_ptr, out_ptr, stream=None):
        if self.lib.p3dfft_ftran_r2c_1d_dev(in_ptr, out_ptr, stream):
            self._raise_if_error()
            raise P3DFFTError("p3dfft_ftran_r2c_1d_dev failed")

    def btran_c2r(self, spec, op="tff"):
        s = np.ascontiguousarray(spec, dtype=self.cdtype).reshape(self.spec_shape())
        out = np.empty(self.real_shape(), dtype=self.rdtype)
        self.lib.p3dfft_btran_c2r(s.ctypes.data, out.ctypes.data, _op(op))
        self._raise_if_error()
        return out

    def cheby(self, a, Lz):
        a = np.ascontiguousarray(a, dtype=self.rdtype).reshape(self.real_shape())
        out = np.empty(self.spec_shape(), dtype=self.cdtype)
        lz = (ctypes.c_double if self.precision == 8 else ctypes.c_float)(Lz)
        self.lib.p3dfft_cheby(a.ctypes.data, out.ctypes.data, ctypes.byref(lz))
        self._raise_if_error()
        return out

    # -- multivariable (build/ftran.F90:104, build/btran.F90:104) ----------
    def ftran_r2c_many(self, a, dim_in, dim_out, nv, op="fft"):
        """a: flat real array of nv*dim_in; returns a flat complex array of nv*dim_out."""
        a = np.ascontiguousarray(a, dtype=self.rdtype).reshape(-1)
        assert a.size >= nv * dim_in
        out = np.zeros(nv * dim_out, dtype=self.cdtype)
        ci, co, cn = ctypes.c_int(dim_in), ctypes.c_int(dim_out), ctypes.c_int(nv)
        self.lib.p3dfft_ftran_r2c_many(a.ctypes.data, ctypes.byref(ci), out.ctypes.data, ctypes.byref(co),
                                       ctypes.byref(cn), _op(op))
        self._raise_if_error()
        return out

    def btran_c2r_many(self, s, dim_in, dim_out, nv, op="tff"):
        s = np.ascontiguousarray(s, dtype=self.cdtype).reshape(-1)
        assert s.size >= nv * dim_in
        out = np.zeros(nv * dim_out, dtype=self.rdtype)
        ci, co, cn = ctypes.c_int(dim_in), ctypes.c_int(dim_out), ctypes.c_int(nv)
        self.lib.p3dfft_btran_c2r_many(s.ctypes.data, ctypes.byref(ci), out.ctypes.data, ctypes.byref(co),
                                       ctypes.byref(cn), _op(op))
        self._raise_if_error()
        return out

    def cheby_many(self, a, dim_in, dim_out, nv, Lz):
        a = np.ascontiguousarray(a, dtype=self.rdtype).reshape(-1)
        out = np.zeros(nv * dim_out, dtype=self.cdtype)
        ci, co, cn = ctypes.c_int(dim_in), ctypes.c_int(dim_out), ctypes.c_int(nv)
        lz = (ctypes.c_double if self.precision == 8 else ctypes.c_float)(Lz)
        self.lib.p3dfft_cheby_many(a.ctypes.data, ctypes.byref(ci), out.ctypes.data, ctypes.byref(co),
                                   ctypes.byref(cn), ctypes.byref(lz))
        self._raise_if_error()
        return out

    def ftran_r2c_many_dev(self, in_ptr, dim_in, out_ptr, dim_out, nv, op="fft", stream=None):
        if self.lib.p3dfft_ftran_r2c_many_dev(in_ptr, dim_in, out_ptr, dim_out, nv, _op(op), stream):
            self._raise_if_error(force=True)

    def btran_c2r_many_dev(self, in_ptr, dim_in, out_ptr, dim_out, nv, op="tff", stream=None):
        if self.lib.p3dfft_btran_c2r_many_dev(in_ptr, dim_in, out_ptr, dim_out, nv, _op(op), stream):
            self._raise_if_error(force=True)

    def clean(self):
        self.lib.p3dfft_clean()
        self.active = False

    def get_timers(self):
        t = (ctypes.c_double * 12)()
        self.lib.get_timers(t)
        return list(t)

    def set_timers(self):
        self.lib.set_timers()

    # -- device-pointer extensions -----------------------------------------
    def ftran_r2c_dev(self, in_ptr, out_ptr, op="fft", stream=None):
        rc = self.lib.p3dfft_ftran_r2c_dev(in_ptr, out_ptr, _op(op), stream)
        if rc:
            self._raise_if_error(force=True)

    def btran_c2r_dev(self, in_ptr, out_ptr, op="tff", stream=None):
        rc = self.lib.p3dfft_btran_c2r_dev(in_ptr, out_ptr, _op(op), stream)
        if rc:
            self._raise_if_error(force=True)

    def cheby_dev(self, in_ptr, out_ptr, Lz, stream=None):
        rc = self.lib.p3dfft_cheby_dev(in_ptr, out_ptr, Lz, stream)
        if rc:
            self._raise_if_error(force=True)

    def alloc(self, nbytes):
        return DeviceBuffer(self.lib, nbytes)

    def real_bytes(self):
        return self.nx * self.ny * self.nz * np.dtype(self.rdtype).itemsize

    def spec_bytes(self):
        return self.nxhpc * self.nyc * self.nzc * np.dtype(self.cdtype).itemsize

    def synchronize(self):
        if self.lib.p3dfft_amd_synchronize():
            raise P3DFFTError("device synchronize failed")

    def last_error(self):
        v = self.lib.p3dfft_amd_last_error()
        return v.decode() if v else ""

    def _raise_if_error(self, force=False):
        msg = self.last_error()
        if msg or force:
            self.lib.p3dfft_amd_clear_error()
            raise P3DFFTError(msg or "p3dfft_amd call failed")


def _op(op):
    b = op.encode() if isinstance(op, str) else bytes(op)
    return (b + b"\0\0\0")[:4]


def device_count():
    return load(8).p3dfft_amd_device_count()


PACK_FCOMM1, UNPACK_FCOMM1, PACK_FCOMM2, UNPACK_BCOMM1, PACK_BCOMM2, UNPACK_BCOMM2 = range(6)


class Share(P3DFFT):
    """One rank (ipid, jpid) of a p1 x p2 processor grid ("share mode"): the
    reference's per-rank stages with the exchange buffers exposed.  Buffers
    are DeviceBuffers holding complex elements in the reference's layouts."""

    def setup_share(self, dims, ipid, jpid, nx, ny, nz, nxc=None, nyc=None, nzc=None, overwrite=True):
        L = self.lib
        L.p3dfft_amd_set_error_mode(1)
        L.p3dfft_amd_clear_error()
        L.p3dfft_amd_set_stride1(0)
        d = (ctypes.c_int * 2)(*dims)
        ms = (ctypes.c_int * 3)()
        vals = [ipid, jpid, nx, ny, nz, nxc or nx, nyc or ny, nzc or nz, 1 if overwrite else 0]
        cv = [ctypes.c_int(v) for v in vals]
        rc = L.p3dfft_setup_share(d, *[ctypes.byref(c) for c in cv], ms)
        self._raise_if_error(force=rc != 0)
        self.active = True
        self.dims, self.ipid, self.jpid = tuple(dims), ipid, jpid
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nxc, self.nyc, self.nzc = nxc or nx, nyc or ny, nzc or nz
        self.stride1 = False
        self.nxhpc = self.nxc // 2 + 1
        self.nxhp = nx // 2 + 1
        self.memsize = tuple(ms)
        s1, _, z1 = self.get_dims(1)
        s2, _, z2 = self.get_dims(2)
        self.jisize, self.kjsize = z1[1], z1[2]
        self.iisize, self.jjsize = z2[0], z2[1]
        return self.memsize

    def counts(self):
        p1, p2 = self.dims
        arrs = [(ctypes.c_longlong * n)() for n in (p1, p1, p1, p1, p2, p2, p2, p2)]
        if self.lib.p3dfft_share_counts(*arrs):
            self._raise_if_error(force=True)
        keys = ("fsnd_st", "fsnd_ct", "frcv_st", "frcv_ct", "ksnd_st", "ksnd_ct", "krcv_st", "krcv_ct")
        return {k: list(a) for k, a in zip(keys, arrs)}

    # element counts of the stage arrays
    def n_real(self):
        return self.nx * self.jisize * self.kjsize

    def n_buf2(self):
        return self.nxhp * self.jisize * self.kjsize

    def n_ypen(self):
        return self.iisize * self.ny * self.kjsize

    def n_send1(self):
        return self.nxhpc * self.jisize * self.kjsize

    def n_send2(self):
        return self.iisize * self.nyc * self.kjsize

    def n_zpen(self):
        return self.iisize * self.jjsize * self.nz

    def n_spec(self):
        return self.iisize * self.jjsize * self.nzc

    def pack(self, which, src, dst, stream=None):
        if self.lib.p3dfft_share_pack(which, src.ptr, dst.ptr, stream):
            self._raise_if_error(force=True)

    def _call(self, name, *args):
        if getattr(self.lib, name)(*args):
            self._raise_if_error(force=True)

    def ftran_x(self, inp, send1, stream=None):
        self._call("p3dfft_share_ftran_x", inp.ptr, send1.ptr, stream)

    def ftran_y(self, recv1, send2, stream=None):
        self._call("p3dfft_share_ftran_y", recv1.ptr, send2.ptr, stream)

    def ftran_z(self, recv2, out, op="fft", stream=None):
        self._call("p3dfft_share_ftran_z", recv2.ptr, out.ptr, _op(op), stream)

    def btran_z(self, inp, send, op="tff", stream=None):
        self._call("p3dfft_share_btran_z", inp.ptr, send.ptr, _op(op), stream)

    def btran_y(self, recv, send, stream=None):
        self._call("p3dfft_share_btran_y", recv.ptr, send.ptr, stream)

    def btran_x(self, recv, out, stream=None):
        self._call("p3dfft_share_btran_x", recv.ptr, out.ptr, stream)
