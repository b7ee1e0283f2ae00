"""Continuous-batching service of spectral Poisson solves on one MI355X
(the loop itself is native: p3dfft_serve_run in csrc/serve.cpp).

Requests (mixed grid shapes) arrive over time with their right-hand sides
resident in HBM.  A request is admitted as soon as it has arrived, a stream
slot is free and its work area fits the HBM budget (first-fit arena); its
solve (r2c, -1/|k|^2 multiply, c2r: p3dfft_poisson_solve) is queued on the
slot's stream, so admitted solves of different shapes overlap on the device.
When the oldest waiting request does not fit, later arrived ones that do may
go ahead of it a bounded number of times (no head-of-line blocking, no
starvation).
Completion is detected with HIP events, whose GPU timestamps give each
request's latency (arrival to completion).  The same loop with one slot and a
synchronise after every launch is the one-at-a-time baseline.
"""
import ctypes
import time

import numpy as np

from . import P3DFFTError, load

GRID_SIZES = (32, 48, 64, 96, 128, 160, 192, 256, 320, 384)


def _bind(L):
    if getattr(L, "_serve_bound", False):
        return L
    vp = ctypes.c_void_p
    L.p3dfft_poisson_work_bytes.argtypes = [ctypes.c_int] * 3
    L.p3dfft_poisson_work_bytes.restype = ctypes.c_size_t
    L.p3dfft_poisson_solve.argtypes = [ctypes.c_int] * 3 + [vp, vp, ctypes.c_size_t, vp]
    L.p3dfft_poisson_solve.restype = ctypes.c_int
    L.p3dfft_amd_fill_normal.argtypes = [vp, ctypes.c_longlong, ctypes.c_ulonglong, vp]
    L.p3dfft_amd_fill_normal.restype = ctypes.c_int
    L.p3dfft_amd_stream_create.restype = vp
    L.p3dfft_amd_stream_destroy.argtypes = [vp]
    L.p3dfft_amd_stream_synchronize.argtypes = [vp]
    L.p3dfft_amd_event_query.argtypes = [vp]
    L.p3dfft_amd_event_query.restype = ctypes.c_int
    L.p3dfft_poisson_set_fused.argtypes = [ctypes.c_int]
    L.p3dfft_poisson_multiply.argtypes = [vp] + [ctypes.c_int] * 3 + [vp]
    L.p3dfft_poisson_multiply.restype = ctypes.c_int
    L.p3dfft_amd_prepare_length.argtypes = [ctypes.c_int]
    L.p3dfft_amd_prepare_length.restype = ctypes.c_int
    L.p3dfft_poisson_graph_create.argtypes = [ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.p3dfft_poisson_graph_create.restype = ctypes.c_int
    L.p3dfft_poisson_graph_launch.argtypes = [vp, vp]
    L.p3dfft_poisson_graph_launch.restype = ctypes.c_int
    L.p3dfft_poisson_graph_destroy.argtypes = [vp]
    L.p3dfft_poisson_graph_destroy.restype = ctypes.c_int
    L.p3dfft_poisson_batch_create.argtypes = [ctypes.c_int, vp, vp, vp, vp, ctypes.POINTER(vp)]
    L.p3dfft_poisson_batch_create.restype = ctypes.c_int
    L.p3dfft_poisson_batch_launch.argtypes = [vp, vp]
    L.p3dfft_poisson_batch_launch.restype = ctypes.c_int
    L.p3dfft_poisson_batch_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.p3dfft_poisson_batch_info.restype = ctypes.c_int
    L.p3dfft_poisson_batch_destroy.argtypes = [vp]
    L.p3dfft_poisson_batch_destroy.restype = ctypes.c_int
    L.p3dfft_serve_run.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_int, vp, ctypes.c_size_t,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_size_t)]
    L.p3dfft_serve_run.restype = ctypes.c_int
    L._serve_bound = True
    return L


class Arena:
    """First-fit allocator over one HBM allocation (offsets in bytes, 256-aligned)."""

    def __init__(self, lib, nbytes):
        self.lib = lib
        self.size = int(nbytes)
        self.base = lib.p3dfft_amd_malloc(self.size) if self.size else None
        if self.size and not self.base:
            raise P3DFFTError("arena allocation of %d bytes failed" % self.size)
        self.free_list = [(0, self.size)]
        self.used = 0
        self.peak = 0

    def alloc(self, n):
        n = (int(n) + 255) & ~255
        for i, (off, sz) in enumerate(self.free_list):
            if sz >= n:
                if sz == n:
                    self.free_list.pop(i)
                else:
                    self.free_list[i] = (off + n, sz - n)
                self.used += n
                self.peak = max(self.peak, self.used)
                return (off, n)
        return None

    def release(self, blk):
        off, n = blk
        self.used -= n
        fl = self.free_list
        fl.append((off, n))
        fl.sort()
        merged = [fl[0]]
        for o, s in fl[1:]:
            po, ps = merged[-1]
            if po + ps == o:
                merged[-1] = (po, ps + s)
            else:
                merged.append((o, s))
        self.free_list = merged

    def ptr(self, blk):
        return self.base + blk[0]

    def close(self):
        if self.base:
            self.lib.p3dfft_amd_free(self.base)
            self.base = None


def make_requests(n, seed=2000, sizes=GRID_SIZES):
    """Grid shapes drawn uniformly per axis from `sizes` (mixed x, y, z)."""
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(sizes, 3)) for _ in range(n)]


def exp_arrivals(n, mean_gap, seed=2001):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(mean_gap, size=n)) - 0.0


class PoissonServer:
    """Fields of all requests resident in one HBM allocation; solves run on
    `slots` streams with work areas from a first-fit arena of `work_budget`
    bytes (the HBM left for in-flight solves)."""

    def __init__(self, shapes, precision=4, slots=4, work_budget=None, seed=7):
        self.lib = _bind(load(precision))
        self.precision = precision
        self.rsize = 4 if precision == 4 else 8
        self.shapes = list(shapes)
        self.nelem = [a * b * c for (a, b, c) in self.shapes]
        self.field_off = np.concatenate([[0], np.cumsum([((e * self.rsize + 255) & ~255) for e in self.nelem])])
        self.fields = self.lib.p3dfft_amd_malloc(int(self.field_off[-1]))
        if not self.fields:
            raise P3DFFTError("field allocation failed")
        self.work = [int(self.lib.p3dfft_poisson_work_bytes(*s)) for s in self.shapes]
        if work_budget is None:
            work_budget = slots * max(self.work)
        if work_budget < max(self.work):
            raise P3DFFTError("work budget below the largest request's work area")
        self.arena = Arena(self.lib, work_budget)
        self.slots = slots
        self.streams = [self.lib.p3dfft_amd_stream_create() for _ in range(slots)]
        self.events = [self.lib.p3dfft_amd_event_create() for _ in range(slots)]
        self.t0ev = self.lib.p3dfft_amd_event_create()
        self.seed = seed
        # plans and twiddles of every length are built before serving
        for n in sorted({v for sh in self.shapes for v in sh}):
            if self.lib.p3dfft_amd_prepare_length(n):
                raise P3DFFTError("cannot prepare length %d" % n)

    def field_ptr(self, i):
        return self.fields + int(self.field_off[i])

    def fill(self, idx=None):
        """Right-hand sides: i.i.d. N(0,1) per request (counter-based, seeded by
        request id).  The solve zeroes the k=0 mode, which removes the mean."""
        for i in (range(len(self.shapes)) if idx is None else idx):
            self.lib.p3dfft_amd_fill_normal(self.field_ptr(i), self.nelem[i], self.seed * 1000003 + i, None)
        self.lib.p3dfft_amd_synchronize()

    def download(self, i):
        nx, ny, nz = self.shapes[i]
        dt = np.float32 if self.precision == 4 else np.float64
        out = np.empty((nz, ny, nx), dtype=dt)
        self.lib.p3dfft_amd_memcpy_d2h(out.ctypes.data, self.field_ptr(i), out.nbytes)
        return out

    def solve_alone(self, i):
        blk = self.arena.alloc(self.work[i])
        try:
            rc = self.lib.p3dfft_poisson_solve(*self.shapes[i], self.field_ptr(i), self.arena.ptr(blk),
                                               self.work[i], self.streams[0])
            if rc:
                raise P3DFFTError(self.lib.p3dfft_amd_last_error().decode())
            self.lib.p3dfft_amd_stream_synchronize(self.streams[0])
        finally:
            self.arena.release(blk)

    def run(self, order, arrivals, sequential=False, lookahead=16, max_bypass=8):
        """Serve requests `order` (indices) arriving at `arrivals` seconds,
        by the native loop p3dfft_serve_run over this server's fields, work
        arena and `slots` streams.  Admission is first-come first-served,
        except that when the oldest waiting request does not fit the free HBM,
        up to `lookahead` later arrived requests that do fit may be admitted
        ahead of it, at most `max_bypass` times per waiting request, so a large
        request neither blocks the small ones behind it nor starves.
        sequential: one request at a time, synchronised after each launch.
        Returns per-request latency (s), the makespan (s) and the peak work
        bytes in use."""
        L = self.lib
        n = len(order)
        if self.arena.used:
            raise P3DFFTError("the work arena must be free when serving")
        nf = len(self.shapes)
        shapes = (ctypes.c_int * (3 * nf))(*[v for sh in self.shapes for v in sh])
        fields = (ctypes.c_void_p * nf)(*[self.field_ptr(i) for i in range(nf)])
        wb = (ctypes.c_size_t * nf)(*self.work)
        od = (ctypes.c_int * n)(*[int(i) for i in order])
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        lat = np.zeros(n)
        mk = ctypes.c_double()
        pk = ctypes.c_size_t()
        rc = L.p3dfft_serve_run(n, shapes, fields, wb, od, arr.ctypes.data, self.slots, self.arena.base,
                                self.arena.size, 1 if sequential else 0, int(lookahead), int(max_bypass),
                                lat.ctypes.data, ctypes.byref(mk), ctypes.byref(pk))
        if rc:
            raise P3DFFTError(L.p3dfft_amd_last_error().decode())
        self.arena.peak = max(self.arena.peak, int(pk.value))
        return lat, mk.value, int(pk.value)

    def run_api(self, order):
        """One request at a time through the reference's entry points, as a
        caller of its single global plan does: p3dfft_setup for the request's
        grid, ftran_r2c 'fft' into a spectrum buffer, the multiply,
        btran_c2r 'tff' back into the field, p3dfft_clean, synchronise.
        Returns the makespan (s)."""
        from . import P3DFFT
        L = self.lib
        P = P3DFFT(precision=self.precision)
        spec_bytes = max((a // 2 + 1) * b * c for (a, b, c) in self.shapes) * 2 * self.rsize
        spec = L.p3dfft_amd_malloc(spec_bytes)
        if not spec:
            raise P3DFFTError("spectrum allocation failed")
        try:
            L.p3dfft_amd_synchronize()
            t0 = time.perf_counter()
            for i in order:
                nx, ny, nz = self.shapes[i]
                P.setup((1, 1), nx, ny, nz)
                P.ftran_r2c_dev(self.field_ptr(i), spec, "fft")
                if L.p3dfft_poisson_multiply(spec, nx, ny, nz, None):
                    raise P3DFFTError(L.p3dfft_amd_last_error().decode())
                P.btran_c2r_dev(spec, self.field_ptr(i), "tff")
                P.clean()
                L.p3dfft_amd_synchronize()
            return time.perf_counter() - t0
        finally:
            L.p3dfft_amd_free(spec)

    def run_plain(self, order):
        """One request at a time through the plain transform path, plans
        prepared before the timed loop: the forward r2c 'fft' (X, Y and Z
        passes) into a spectrum buffer, the per-mode multiply, the backward
        c2r 'tff' (Z, Y and X passes) back into the field, synchronise; seven
        HBM passes per solve and no setup or clean per request (the grids'
        plans are p3dfft_amd_batch_* objects of one grid each).  Returns the
        makespan (s)."""
        L = self.lib
        spec_bytes = max((a // 2 + 1) * b * c for (a, b, c) in self.shapes) * 2 * self.rsize
        spec = L.p3dfft_amd_malloc(spec_bytes)
        if not spec:
            raise P3DFFTError("spectrum allocation failed")
        plans = {}
        try:
            for i in sorted(set(int(j) for j in order)):
                sh = (ctypes.c_int * 3)(*self.shapes[i])
                rp = (ctypes.c_void_p * 1)(self.field_ptr(i))
                sp = (ctypes.c_void_p * 1)(spec)
                f, b = ctypes.c_void_p(), ctypes.c_void_p()
                if L.p3dfft_amd_batch_create(1, 1, sh, rp, sp, ctypes.byref(f)):
                    raise P3DFFTError(L.p3dfft_amd_last_error().decode())
                plans[i] = (f, None)
                if L.p3dfft_amd_batch_create(-1, 1, sh, rp, sp, ctypes.byref(b)):
                    raise P3DFFTError(L.p3dfft_amd_last_error().decode())
                plans[i] = (f, b)
            L.p3dfft_amd_synchronize()
            t0 = time.perf_counter()
            for i in order:
                nx, ny, nz = self.shapes[i]
                f, b = plans[i]
                if L.p3dfft_amd_batch_launch(f, None) or L.p3dfft_poisson_multiply(spec, nx, ny, nz, None) \
                        or L.p3dfft_amd_batch_launch(b, None):
                    raise P3DFFTError(L.p3dfft_amd_last_error().decode())
                L.p3dfft_amd_synchronize()
            return time.perf_counter() - t0
        finally:
            L.p3dfft_amd_synchronize()
            for f, b in plans.values():
                for h in (f, b):
                    if h is not None and h.value:
                        L.p3dfft_amd_batch_destroy(h)
            L.p3dfft_amd_free(spec)

    def batch(self, idx):
        """A PoissonBatch (grouped launches) of the solves of requests `idx`."""
        return PoissonBatch(self, idx)

    def graph(self, idx, nstreams=8):
        """A PoissonGraph of the solves of requests `idx` on their fields."""
        return PoissonGraph(self, idx, nstreams)

    def run_streams(self, idx, reps=1):
        """The same batch issued eagerly: request k of `idx` on stream k mod
        slots, each with its own work area, `reps` times.  Returns seconds."""
        L = self.lib
        blks = [self.arena.alloc(self.work[i]) for i in idx]
        try:
            if any(b is None for b in blks):
                raise P3DFFTError("work budget below the batch's work areas")
            L.p3dfft_amd_synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                for k, i in enumerate(idx):
                    rc = L.p3dfft_poisson_solve(*self.shapes[i], self.field_ptr(i), self.arena.ptr(blks[k]),
                                                self.work[i], self.streams[k % self.slots])
                    if rc:
                        raise P3DFFTError(L.p3dfft_amd_last_error().decode())
            L.p3dfft_amd_synchronize()
            return time.perf_counter() - t0
        finally:
            for b in blks:
                if b is not None:
                    self.arena.release(b)

    def close(self):
        L = self.lib
        L.p3dfft_amd_synchronize()
        for s in self.streams:
            L.p3dfft_amd_stream_destroy(s)
        for e in self.events + [self.t0ev]:
            L.p3dfft_amd_event_destroy(e)
        self.arena.close()
        if self.fields:
            L.p3dfft_amd_free(self.fields)
            self.fields = None



class PoissonGraph:
    """A fixed batch of solves (requests `idx` of a PoissonServer, in place on
    their resident fields) captured once into a HIP graph by
    p3dfft_poisson_graph_create and replayed by launch(): one host call issues
    every kernel of the batch.  Work areas (one per solve) are held for the
    graph's lifetime."""

    def __init__(self, server, idx, nstreams=8):
        L = server.lib
        self.lib = L
        self.idx = list(idx)
        n = len(self.idx)
        works = [server.work[i] for i in self.idx]
        offs = np.concatenate([[0], np.cumsum([(w + 255) & ~255 for w in works])]).astype(np.int64)
        self.work_base = L.p3dfft_amd_malloc(int(offs[-1]))
        if not self.work_base:
            raise P3DFFTError("graph work allocation failed")
        shapes = (ctypes.c_int * (3 * n))(*[v for i in self.idx for v in server.shapes[i]])
        fields = (ctypes.c_void_p * n)(*[server.field_ptr(i) for i in self.idx])
        wptr = (ctypes.c_void_p * n)(*[self.work_base + int(o) for o in offs[:-1]])
        wbytes = (ctypes.c_size_t * n)(*works)
        h = ctypes.c_void_p()
        if L.p3dfft_poisson_graph_create(n, shapes, fields, wptr, wbytes, int(nstreams), ctypes.byref(h)):
            L.p3dfft_amd_free(self.work_base)
            self.work_base = None
            raise P3DFFTError(L.p3dfft_amd_last_error().decode())
        self.handle = h.value

    def launch(self, stream=None):
        if self.lib.p3dfft_poisson_graph_launch(self.handle, stream):
            raise P3DFFTError(self.lib.p3dfft_amd_last_error().decode())

    def run(self, reps=1):
        """`reps` back-to-back replays; returns seconds."""
        L = self.lib
        L.p3dfft_amd_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.launch()
        L.p3dfft_amd_synchronize()
        return time.perf_counter() - t0

    def close(self):
        L = self.lib
        L.p3dfft_amd_synchronize()
        if self.handle:
            L.p3dfft_poisson_graph_destroy(self.handle)
            self.handle = None
        if self.work_base:
            L.p3dfft_amd_free(self.work_base)
            self.work_base = None


class PoissonBatch:
    """A fixed batch of solves (requests `idx` of a PoissonServer) run by
    p3dfft_poisson_batch_*: one launch per pass and transform length for every
    solve whose lengths have specialised kernels, the rest one by one.  Work
    areas (one per solve) are held for the batch's lifetime."""

    def __init__(self, server, idx):
        L = server.lib
        self.lib = L
        self.idx = list(idx)
        n = len(self.idx)
        works = [server.work[i] for i in self.idx]
        offs = np.concatenate([[0], np.cumsum([(w + 255) & ~255 for w in works])]).astype(np.int64)
        self.work_base = L.p3dfft_amd_malloc(int(offs[-1]))
        if not self.work_base:
            raise P3DFFTError("batch work allocation failed")
        shapes = (ctypes.c_int * (3 * n))(*[v for i in self.idx for v in server.shapes[i]])
        fields = (ctypes.c_void_p * n)(*[server.field_ptr(i) for i in self.idx])
        wptr = (ctypes.c_void_p * n)(*[self.work_base + int(o) for o in offs[:-1]])
        wbytes = (ctypes.c_size_t * n)(*works)
        h = ctypes.c_void_p()
        if L.p3dfft_poisson_batch_create(n, shapes, fields, wptr, wbytes, ctypes.byref(h)):
            L.p3dfft_amd_free(self.work_base)
            self.work_base = None
            raise P3DFFTError(L.p3dfft_amd_last_error().decode())
        self.handle = h.value
        gl, ss = ctypes.c_int(), ctypes.c_int()
        L.p3dfft_poisson_batch_info(self.handle, ctypes.byref(gl), ctypes.byref(ss))
        self.grouped_launches, self.single_solves = gl.value, ss.value

    def launch(self, stream=None):
        if self.lib.p3dfft_poisson_batch_launch(self.handle, stream):
            raise P3DFFTError(self.lib.p3dfft_amd_last_error().decode())

    def run(self, reps=1):
        """`reps` back-to-back launches; returns seconds."""
        L = self.lib
        L.p3dfft_amd_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.launch()
        L.p3dfft_amd_synchronize()
        return time.perf_counter() - t0

    def capture(self):
        """Capture one launch of the batch (its grouped launches and any
        solves run one by one) into a HIP graph on a stream of its own, so
        the whole batch is one host launch (run_graph replays it)."""
        L = self.lib
        self.stream = L.p3dfft_amd_stream_create()
        g = ctypes.c_void_p()
        if L.p3dfft_amd_capture_begin(self.stream):
            raise P3DFFTError(L.p3dfft_amd_last_error().decode())
        self.launch(self.stream)
        if L.p3dfft_amd_capture_end(self.stream, ctypes.byref(g)):
            raise P3DFFTError(L.p3dfft_amd_last_error().decode())
        self.graph = g.value

    def run_graph(self, reps=1):
        """`reps` back-to-back replays of the captured batch; returns seconds."""
        L = self.lib
        L.p3dfft_amd_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            if L.p3dfft_amd_graph_launch(self.graph, self.stream):
                raise P3DFFTError(L.p3dfft_amd_last_error().decode())
        L.p3dfft_amd_stream_synchronize(self.stream)
        return time.perf_counter() - t0

    def close(self):
        L = self.lib
        L.p3dfft_amd_synchronize()
        if getattr(self, "graph", None):
            L.p3dfft_amd_graph_destroy(self.graph)
            self.graph = None
        if getattr(self, "stream", None):
            L.p3dfft_amd_stream_destroy(self.stream)
            self.stream = None
        if self.handle:
            L.p3dfft_poisson_batch_destroy(self.handle)
            self.handle = None
        if self.work_base:
            L.p3dfft_amd_free(self.work_base)
            self.work_base = None
