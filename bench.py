"""Benchmark of the forward+backward 3D FFT path on one MI355X.

One step = p3dfft_ftran_r2c ('fft') + p3dfft_btran_c2r ('tff') of a 512^3 real
field through the device-pointer C-ABI entry points, with the field, the
spectrum and the library's work buffers resident in HBM (no host traffic in
the timed region).  The headline is the FP32 build (FP32 bulk, compensated
hi+lo FP32 twiddles generated in FP64); the plain FP64 build is timed beside
it on the same field.

Prints ONE JSON line.  Besides metric, value, unit, steps and timing:
  roofline       dominant kernel: algorithmic bytes per launch / HIP-event
                 duration per launch, against the 8 TB/s HBM peak
  path           whole fwd+bwd against the compulsory-bytes roofline
                 2B = 2 (N_r s + N_c 2s) and the measured copy-kernel peak
  kernels        per-stage durations from the library's HIP-event timers
  fp64_*         the plain FP64 GPU path and the FP32 speed-up over it
  rel_l2_*       FP32 spectrum / round trip vs the FP64 path
  power_spectrum the drivers' shell sums of the 512^3 FP32 spectrum
  c1_64cubed     64^3 FP64 fwd+bwd, eager and replayed from a HIP graph
  share          one rank's stages of 2048^3 on a 16 x 16 grid
  serve          served mixed-size Poisson solves: throughput, p50/p99,
                 one-at-a-time baselines
  batched        a fixed batch of small solves: eager streams, one HIP
                 graph, grouped launches, one at a time
  cpu_baseline   the reference's distributed CPU path restated by the FP64 C
                 oracle, one rank per host core on the drivers' P1 x P2 grid
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import p3dfft_amd  # noqa: E402

HBM_PEAK = 8.0e12
METRIC = "fwd+bwd 3D FFT time & % HBM roofline at 512^3; batched/served solves/s, p99"
SLOT_NAMES = {5: "x_r2c (k_x_r2c_ct)", 7: "y_fwd (k_colio_ct)", 8: "z_fwd (k_colio_ct)",
              9: "z_bwd (k_colio_ct)", 10: "y_bwd (k_colio_ct)", 12: "x_c2r (k_x_c2r_ct)"}


def modes_field(n, seed):
    """sum of 8 random-phase Fourier modes (wavenumbers in [1,32]) + N(0,1e-3^2)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 33, size=(8, 3))
    ph = rng.random(8) * 2 * np.pi
    x = np.arange(n) / n
    f = np.zeros((n, n, n))
    for m in range(8):
        f += np.cos(2 * np.pi * (k[m, 0] * x[None, None, :] + k[m, 1] * x[None, :, None]
                                 + k[m, 2] * x[:, None, None]) + ph[m])
    f += rng.normal(0.0, 1e-3, size=f.shape)
    return f


def slot_bytes(slot, n, s):
    """algorithmic bytes per launch of the stage kernel in a timer slot"""
    nr = n ** 3
    nc = (n // 2 + 1) * n * n
    if slot in (5, 12):
        return nr * s + nc * 2 * s
    return 2 * nc * 2 * s


PMC_TRAFFIC = os.path.join(ROOT, "profiles", "pmc_traffic_512cubed_fp32.md")


def pmc_traffic(stage):
    """HBM bytes per launch of a stage kernel at 512^3 FP32, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None when absent)."""
    if not os.path.exists(PMC_TRAFFIC):
        return None
    for line in open(PMC_TRAFFIC):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == stage and cells[4].isdigit():
            return int(cells[4])
    return None


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def run_precision(precision, n, f, steps, warmup, chunk_bytes=None, timers=True):
    P = p3dfft_amd.P3DFFT(precision)
    if chunk_bytes is not None:
        P.lib.p3dfft_amd_set_chunk_bytes(chunk_bytes)
    P.lib.p3dfft_amd_set_timing(1 if timers else 0)
    P.setup((1, 1), n, n, n)
    lib = P.lib
    try:
        fin = P.alloc(P.real_bytes())
        spec = P.alloc(P.spec_bytes())
        out = P.alloc(P.real_bytes())
        fin.upload(f.astype(P.rdtype))
        for _ in range(warmup):
            P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft")
            P.btran_c2r_dev(spec.ptr, out.ptr, "tff")
        P.synchronize()
        P.set_timers()
        e0 = lib.p3dfft_amd_event_create()
        e1 = lib.p3dfft_amd_event_create()
        P.synchronize()
        t0 = time.perf_counter()
        lib.p3dfft_amd_event_record(e0, None)
        for _ in range(steps):
            P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft")
            P.btran_c2r_dev(spec.ptr, out.ptr, "tff")
        lib.p3dfft_amd_event_record(e1, None)
        P.synchronize()
        wall = time.perf_counter() - t0
        ev_ms = lib.p3dfft_amd_event_elapsed_ms(e0, e1)
        lib.p3dfft_amd_event_destroy(e0)
        lib.p3dfft_amd_event_destroy(e1)
        timers = P.get_timers()
        # one more forward for the accuracy comparison
        P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft")
        s = spec.download(P.spec_shape(), P.cdtype)
        # the drivers' power spectrum of that spectrum (HBM in, HBM out)
        kmax = int((3.0 * n * n) ** 0.5 * 0.5 + 0.5)
        ebuf = P.alloc((kmax + 1) * 8)
        for _ in range(3):
            P.power_spectrum_dev(spec.ptr, kmax, ebuf.ptr)
        P.synchronize()
        e0 = lib.p3dfft_amd_event_create()
        e1 = lib.p3dfft_amd_event_create()
        lib.p3dfft_amd_event_record(e0, None)
        for _ in range(10):
            P.power_spectrum_dev(spec.ptr, kmax, ebuf.ptr)
        lib.p3dfft_amd_event_record(e1, None)
        P.synchronize()
        spec_ms = lib.p3dfft_amd_event_elapsed_ms(e0, e1) / 10
        lib.p3dfft_amd_event_destroy(e0)
        lib.p3dfft_amd_event_destroy(e1)
        ebuf.free()
        P.btran_c2r_dev(spec.ptr, out.ptr, "tff")
        back = out.download(P.real_shape(), P.rdtype)
        fin.free(); spec.free(); out.free()
    finally:
        P.clean()
    return {"ms_per_step": ev_ms / steps, "wall_ms_per_step": wall * 1e3 / steps, "timers": timers,
            "spec": s, "back": back, "spectrum_ms": spec_ms}


def copy_peak(lib, nbytes=2 << 30, reps=10):
    a = p3dfft_amd.DeviceBuffer(lib, nbytes)
    b = p3dfft_amd.DeviceBuffer(lib, nbytes)
    lib.p3dfft_amd_memset(a.ptr, 0, nbytes)
    lib.p3dfft_amd_stream_copy(a.ptr, b.ptr, nbytes, None)
    e0 = lib.p3dfft_amd_event_create()
    e1 = lib.p3dfft_amd_event_create()
    lib.p3dfft_amd_synchronize()
    lib.p3dfft_amd_event_record(e0, None)
    for _ in range(reps):
        lib.p3dfft_amd_stream_copy(a.ptr, b.ptr, nbytes, None)
    lib.p3dfft_amd_event_record(e1, None)
    ms = lib.p3dfft_amd_event_elapsed_ms(e0, e1)
    lib.p3dfft_amd_event_destroy(e0)
    lib.p3dfft_amd_event_destroy(e1)
    a.free(); b.free()
    return 2.0 * nbytes * reps / (ms * 1e-3)


def ranks_timed(f, threads, reps):
    """The oracle's distributed path on the drivers' grid of `threads` ranks,
    halving the rank count while the grid leaves a rank without a pencil.
    Returns (p1, p2, forward times, backward times, timer slots (ranks,
    reps, 12))."""
    from oracle import oracle as O
    while True:
        p1, p2 = O.dims_create(threads)
        try:
            sl = np.zeros((p1 * p2, reps, 12))
            _, _, tf, tb = O.ranks_fwd_bwd(f, p1, p2, reps=reps, spectrum=False, back=False, timers=sl)
            return p1, p2, tf, tb, sl
        except ValueError:
            if threads <= 1:
                raise
            threads //= 2


def timer_slots(sl):
    """The 12 timer slots of the timed pairs as the reference drivers report
    them (sample/C/driver_rand.c:271-289): each rank's slots divided by the
    number of timed loops, then their average, maximum and minimum over the
    ranks (seconds per fwd+bwd pair)."""
    per_rank = sl.mean(axis=1)  # (ranks, 12)
    return {"slots": list(range(1, 13)), "avg": per_rank.mean(axis=0).tolist(),
            "max": per_rank.max(axis=0).tolist(), "min": per_rank.min(axis=0).tolist(),
            "meaning": "1-4 fcomm1/fcomm2/bcomm1/bcomm2 exchanges, 5 X r2c, 6 fcomm1 pack+unpack, 7 Y fwd, "
                       "8 Z fwd, 9 Z bwd + bcomm1 unpack, 10 Y bwd, 11 bcomm2 pack+unpack, 12 X c2r "
                       "(the fcomm2 pack is untimed, as in the reference)"}


def cpu_baseline(n, f, threads, loops=5):
    """The reference's distributed CPU path restated by the C oracle, FP64:
    `threads` ranks on the drivers' P1 x P2 grid (MPI_Dims_create, dims(1) <=
    dims(2); sample/FORTRAN/driver_rand.F90:115-137), one rank per host core
    (each rank single-threaded), the four pencil exchanges as shared-memory
    copies of the fcomm1/fcomm2/bcomm1/bcomm2 buffers; timed as the
    reference's driver loop times its transforms
    (sample/FORTRAN/driver_sine.F90:192-232: warm-up, then n x (ftran, btran),
    barrier to barrier): one warm-up pair, then `loops` timed pairs on the
    same field; the median pair is the value.  The OpenMP-over-lines 1 x 1
    oracle is timed beside it (one pair) for comparison."""
    from oracle import oracle as O
    p1, p2, tf, tb, sl = ranks_timed(f, threads, loops + 1)
    tf, tb = tf[1:], tb[1:]
    pairs = [float(a + b) for a, b in zip(tf, tb)]
    med = float(np.median(pairs))
    O.set_threads(threads)
    p = O.Plan(n, n, n)
    t0 = time.perf_counter()
    p.btran_c2r(p.ftran_r2c(f, "fft"), "tff", keep_input=False)
    t_lines = time.perf_counter() - t0
    return {"value": 1.0 / med, "unit": "fwd+bwd/s", "cores": p1 * p2, "kind": "port",
            "sample": "the reference's CPU path on a %d x %d grid of ranks (one per host core, threads of one "
                      "process, exchanges as shared-memory copies of the pack/unpack buffers) on the same "
                      "%d^3 field in FP64 by the C oracle, timed like the reference driver: one warm-up pair, "
                      "then %d x (ftran 'fft', btran 'tff'); median pair %.2f s (ftran %.2f s, btran %.2f s "
                      "median)" % (p1, p2, n, loops, med, float(np.median(tf)), float(np.median(tb))),
            "grid": [p1, p2], "pairs_s": pairs, "one_rank_openmp_lines_pair_s": t_lines,
            "timers_s": timer_slots(sl[:, 1:, :])}


def fp64_poisson(f):
    """FP64 numpy solve of -f_hat/|k|^2 (k=0 -> 0): the check of served results."""
    import scipy.fft as sf
    nz, ny, nx = f.shape
    workers = min(16, len(os.sched_getaffinity(0)))
    F = sf.rfftn(f.astype(np.float64), workers=workers)
    kx = np.arange(nx // 2 + 1)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    kz = np.fft.fftfreq(nz, 1.0 / nz)
    k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    mult = np.where(k2 > 0, -1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    return sf.irfftn(F * mult, s=(nz, ny, nx), axes=(0, 1, 2), workers=workers)


def share_leg(steps=20, warmup=3):
    """One rank of the 2048^3 transform on a 16 x 16 grid (rank (15, 3): local X
    pencil (2048, 128, 128), 65 kx in its Y/Z pencils), FP32: the per-rank
    stages of a fwd+bwd with the exchange buffers in HBM (the exchange itself
    is another rank's business).  Per stage: mean time and algorithmic GB/s."""
    n, dims, ip, jp = 2048, (16, 16), 15, 3
    S = p3dfft_amd.Share(4)
    S.setup_share(dims, ip, jp, n, n, n)
    L = S.lib
    try:
        c = 8  # bytes per FP32 complex
        nb1 = max(S.n_send1(), S.n_ypen()) * c
        nb2 = max(S.n_send2(), S.n_zpen(), S.n_spec()) * c
        inp = S.alloc(S.n_real() * 4); out = S.alloc(S.n_real() * 4)
        b1 = S.alloc(nb1); b2 = S.alloc(nb2); b3 = S.alloc(S.n_zpen() * c)
        for b in (inp, out, b1, b2, b3):
            L.p3dfft_amd_memset(b.ptr, 0, b.nbytes)
        stages = [("ftran_x", lambda: S.ftran_x(inp, b1), S.n_real() * 4 + S.n_send1() * c),
                  ("ftran_y", lambda: S.ftran_y(b1, b2), (S.n_ypen() + S.n_send2()) * c),
                  ("ftran_z", lambda: S.ftran_z(b2, b2, "fft"), (S.n_zpen() + S.n_spec()) * c),
                  ("btran_z", lambda: S.btran_z(b2, b3, "tff"), (S.n_spec() + S.n_zpen()) * c),
                  ("btran_y", lambda: S.btran_y(b3, b1), (S.n_zpen() + S.n_ypen()) * c),
                  ("btran_x", lambda: S.btran_x(b1, out), S.n_send1() * c + S.n_real() * 4)]
        for _ in range(warmup):
            for _, f, _ in stages:
                f()
        ev = [L.p3dfft_amd_event_create() for _ in range(len(stages) + 1)]
        acc = [0.0] * len(stages)
        for _ in range(steps):
            L.p3dfft_amd_event_record(ev[0], None)
            for i, (_, f, _) in enumerate(stages):
                f()
                L.p3dfft_amd_event_record(ev[i + 1], None)
            L.p3dfft_amd_synchronize()
            for i in range(len(stages)):
                acc[i] += L.p3dfft_amd_event_elapsed_ms(ev[i], ev[i + 1])
        for e in ev:
            L.p3dfft_amd_event_destroy(e)
        for b in (inp, out, b1, b2, b3):
            b.free()
    finally:
        S.clean()
    st = {name: {"ms": acc[i] / steps, "bytes": nbytes, "GBps": nbytes / (acc[i] / steps * 1e-3) / 1e9}
          for i, (name, _, nbytes) in enumerate(stages)}
    return {"config": "rank (15,3) of 2048^3 on a 16x16 grid, local X pencil (2048,128,128), FP32; "
                      "stages include the reference's pack/unpack maps, not the exchange",
            "ms_per_fwd_bwd": sum(acc) / steps, "stages": st}


def serve_leg(n_req, slots, load_frac=0.7):
    """Mixed-size spectral Poisson solves (grid sizes drawn per axis from
    {32,...,384}), FP32 with compensated twiddles, fields resident in HBM.
    Throughput at saturation (all requests queued at t=0) for the served loop
    and for one-at-a-time solves (FP32, and the plain FP64 path); p50/p99 latency of the served loop with
    exponential inter-arrival times at load_frac of its capacity."""
    from p3dfft_amd import serve
    shapes = serve.make_requests(n_req)
    S = serve.PoissonServer(shapes, precision=4, slots=slots)
    try:
        order = list(range(n_req))
        S.fill()
        S.run(order[:min(50, n_req)], np.zeros(min(50, n_req)))  # warm
        S.fill()
        # accuracy of the served results themselves, against FP64 solves of the
        # same right-hand sides: the 6 largest requests and 6 drawn at random
        by_size = sorted(order, key=lambda i: -int(np.prod(shapes[i])))
        rng = np.random.default_rng(7)
        sample = sorted(set(by_size[:6]) | set(int(i) for i in rng.choice(n_req, 6, replace=False)))
        rhs = {i: S.download(i) for i in sample}
        _, mk_served, peak = S.run(order, np.zeros(n_req))
        errs = {i: float(np.linalg.norm(S.download(i) - fp64_poisson(rhs[i])) / np.linalg.norm(fp64_poisson(rhs[i])))
                for i in sample}
        err = max(errs.values())
        err_sample = "served results of %d requests (the 6 largest, up to %s, and 6 at random) against FP64 " \
                     "solves of the same right-hand sides" % (len(sample), "x".join(map(str, shapes[by_size[0]])))
        S.fill()
        _, mk_seq, _ = S.run(order, np.zeros(n_req), sequential=True)
        thr_served, thr_seq = n_req / mk_served, n_req / mk_seq
        S.fill()
        arr = serve.exp_arrivals(n_req, 1.0 / (load_frac * thr_served))
        lat, _, _ = S.run(order, arr)
        S.fill()
        lat_q, _, _ = S.run(order, arr, sequential=True)
    finally:
        S.close()
    # the plain FP64 GPU path, one solve at a time, on the same requests:
    # (a) the fused solve entry point with a synchronise after each request,
    # (b) the reference's entry points per request (setup, ftran_r2c, the
    # multiply, btran_c2r, clean), as a caller of its single global plan
    S8 = serve.PoissonServer(shapes, precision=8, slots=1)
    try:
        S8.fill()
        S8.run(order[:min(50, n_req)], np.zeros(min(50, n_req)), sequential=True)  # warm
        S8.fill()
        _, mk_seq64, _ = S8.run(order, np.zeros(n_req), sequential=True)
        thr_seq64 = n_req / mk_seq64
        S8.fill()
        thr_plain64 = n_req / S8.run_plain(order)
        S8.fill()
        thr_api64 = n_req / S8.run_api(order)
    finally:
        S8.close()
    return {"requests": n_req, "slots": slots, "solves_per_s": thr_served, "one_at_a_time_solves_per_s": thr_seq,
            "speedup_vs_one_at_a_time": thr_served / thr_seq,
            "one_at_a_time_fp64_solves_per_s": thr_seq64,
            "speedup_vs_one_at_a_time_fp64": thr_served / thr_seq64,
            "one_at_a_time_plain_fp64_solves_per_s": thr_plain64,
            "speedup_vs_one_at_a_time_plain_fp64": thr_served / thr_plain64,
            "one_at_a_time_plain_fp64_api_solves_per_s": thr_api64,
            "speedup_vs_one_at_a_time_plain_fp64_api": thr_served / thr_api64,
            "baselines": "one_at_a_time: the same FP32 fused solve, synchronised after each request; "
                         "fp64: the FP64 fused solve likewise; plain_fp64: the plain FP64 transform path per "
                         "request with its plans prepared (forward r2c 'fft', p3dfft_poisson_multiply, backward "
                         "c2r 'tff': seven passes, no setup), synchronised; plain_fp64_api: per request p3dfft_setup, "
                         "ftran_r2c_dev 'fft', p3dfft_poisson_multiply, btran_c2r_dev 'tff', p3dfft_clean (FP64)",
            "latency_load": "exponential inter-arrival at %.0f%% of served capacity (%.0f req/s)"
                            % (100 * load_frac, load_frac * thr_served),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "one_at_a_time_p50_ms": float(np.percentile(lat_q, 50) * 1e3),
            "one_at_a_time_p99_ms": float(np.percentile(lat_q, 99) * 1e3),
            "peak_work_bytes": int(peak), "rel_l2_vs_fp64_max": err, "rel_l2_sample": err_sample,
            "mean_points_per_request": float(np.mean([np.prod(s) for s in shapes]))}


def c3_leg(steps, warmup, n=384):
    """The radix-3 precision grid (C3): 384^3 i.i.d. N(0,1), FP32 and FP64
    fwd+bwd through the device entry points, per-step time, fraction of the
    compulsory-bytes roofline and the FP32 error against the FP64 path."""
    f = np.random.default_rng(384).standard_normal((n, n, n))
    r32 = run_precision(4, n, f, steps, warmup)
    r64 = run_precision(8, n, f, steps, warmup)
    nr, nc = n ** 3, (n // 2 + 1) * n * n
    out = {"grid": [n, n, n], "field": "N(0,1), seed 384"}
    for name, r, sz in (("fp32", r32, 4), ("fp64", r64, 8)):
        t = r["ms_per_step"] * 1e-3
        B2 = 2 * (nr * sz + nc * 2 * sz)
        out[name] = {"ms_per_step": r["ms_per_step"], "compulsory_bytes": B2, "GBps": B2 / t / 1e9,
                     "frac_of_8TBps": B2 / t / HBM_PEAK,
                     "kernels_ms": {SLOT_NAMES[k]: r["timers"][k - 1] / steps * 1e3 for k in SLOT_NAMES}}
    ref = r64["spec"].astype(np.complex128)
    out["speedup_fp32_vs_fp64"] = r64["ms_per_step"] / r32["ms_per_step"]
    out["rel_l2_spectrum_fp32_vs_fp64"] = float(np.linalg.norm(r32["spec"] - ref) / np.linalg.norm(ref))
    b64 = r64["back"].astype(np.float64)
    out["rel_l2_roundtrip_fp32_vs_fp64"] = float(np.linalg.norm(r32["back"] - b64) / np.linalg.norm(b64))
    return out


def cheby_field(nx, ny, nz, Lz, seed=257):
    """Random trig modes in x, y times a random degree-64 Chebyshev polynomial in
    z (coefficients N(0,1)/(1+n)^2) on the nz Gauss-Lobatto points."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(65) / (1 + np.arange(65)) ** 2
    zz = 0.5 * Lz * np.cos(np.pi * np.arange(nz) / (nz - 1))
    g = np.polynomial.chebyshev.chebval(2.0 * zz / Lz, c)
    xs, ys = np.arange(nx) / nx, np.arange(ny) / ny
    xy = sum(np.cos(2 * np.pi * (rng.integers(1, 9) * xs[None, :] + rng.integers(1, 9) * ys[:, None])
                    + rng.random() * 2 * np.pi) for _ in range(4))
    return g[:, None, None] * xy[None]


def cheby_leg(steps=20, warmup=3, cpu_threads=1, cpu=True):
    """The Chebyshev configuration (C4): 256 x 256 x 257 FP64, p3dfft_cheby
    ('ffc' forward, scaling, derivative recurrence along z) and the backward
    btran_c2r 'cff', device entry points, fields resident.  Times of each
    direction and of the recurrence kernel alone (cheby minus the 'ffc'
    forward), against the compulsory-bytes roofline; with the CPU oracle's
    cheby + btran timed on the same field as its baseline and checker."""
    nx, ny, nz, Lz = 256, 256, 257, 2.0
    a = cheby_field(nx, ny, nz, Lz)
    P = p3dfft_amd.P3DFFT(8)
    P.setup((1, 1), nx, ny, nz)
    L = P.lib
    try:
        fin = P.alloc(P.real_bytes()); spec = P.alloc(P.spec_bytes()); out = P.alloc(P.real_bytes())
        fin.upload(a)
        e0, e1 = L.p3dfft_amd_event_create(), L.p3dfft_amd_event_create()

        def timed(fn):
            for _ in range(warmup):
                fn()
            P.synchronize()
            L.p3dfft_amd_event_record(e0, None)
            for _ in range(steps):
                fn()
            L.p3dfft_amd_event_record(e1, None)
            return L.p3dfft_amd_event_elapsed_ms(e0, e1) / steps
        t_cheby = timed(lambda: P.cheby_dev(fin.ptr, spec.ptr, Lz))
        t_ffc = timed(lambda: P.ftran_r2c_dev(fin.ptr, spec.ptr, "ffc"))
        P.cheby_dev(fin.ptr, spec.ptr, Lz)
        got = spec.download(P.spec_shape(), P.cdtype)
        # btran 'cff' on a copy of the derivative spectrum (overwrite mode
        # transforms its input in place), restored before each call
        keep = P.alloc(P.spec_bytes())
        L.p3dfft_amd_memcpy_d2d(keep.ptr, spec.ptr, P.spec_bytes())

        def bwd():
            L.p3dfft_amd_memcpy_d2d(spec.ptr, keep.ptr, P.spec_bytes())
            P.btran_c2r_dev(spec.ptr, out.ptr, "cff")
        t_b_copy = timed(lambda: L.p3dfft_amd_memcpy_d2d(spec.ptr, keep.ptr, P.spec_bytes()))
        t_bwd = timed(bwd) - t_b_copy
        L.p3dfft_amd_memcpy_d2d(spec.ptr, keep.ptr, P.spec_bytes())
        P.btran_c2r_dev(spec.ptr, out.ptr, "cff")
        back = out.download(P.real_shape(), P.rdtype)
        L.p3dfft_amd_event_destroy(e0); L.p3dfft_amd_event_destroy(e1)
        for b in (fin, spec, out, keep):
            b.free()
    finally:
        P.clean()
    nr, nc = nx * ny * nz, (nx // 2 + 1) * ny * nz
    B2 = 2 * (nr * 8 + nc * 16)
    t = (t_cheby + t_bwd) * 1e-3
    res = {"grid": [nx, ny, nz], "Lz": Lz, "cheby_ms": t_cheby, "btran_cff_ms": t_bwd,
           "recurrence_ms": t_cheby - t_ffc, "ms_per_step": t_cheby + t_bwd, "compulsory_bytes": B2,
           "GBps": B2 / t / 1e9, "frac_of_8TBps": B2 / t / HBM_PEAK,
           "recurrence_GBps": nc * 32 / ((t_cheby - t_ffc) * 1e-3) / 1e9 if t_cheby > t_ffc else None}
    if cpu:
        res["cpu_baseline"] = cpu_baseline_cheby(a, Lz, got, back, cpu_threads)
    return res


def cpu_baseline_cheby(a, Lz, got, back, threads, loops=3):
    """The FP64 CPU oracle's cheby + btran 'cff' on the same field: its time
    (median of `loops` after one warm-up) and, as the checker, the GPU
    results' relative L2 error against its outputs."""
    from oracle import oracle as O
    O.set_threads(threads)
    nz, ny, nx = a.shape
    p = O.Plan(nx, ny, nz)
    c = p.cheby(a, Lz)
    p.btran_c2r(c, "cff")
    tt = []
    for _ in range(loops):
        t0 = time.perf_counter()
        c = p.cheby(a, Lz)
        b = p.btran_c2r(c, "cff")[0]
        tt.append(time.perf_counter() - t0)
    med = float(np.median(tt))
    return {"value": 1.0 / med, "unit": "cheby+btran/s", "cores": O.get_threads(), "kind": "port",
            "sample": "the whole 256x256x257 FP64 case, one warm-up then %d x (cheby, btran 'cff'); median %.3f s"
                      % (loops, med),
            "gpu_rel_l2_cheby_vs_oracle": float(np.linalg.norm(got - c) / np.linalg.norm(c)),
            "gpu_rel_l2_btran_vs_oracle": float(np.linalg.norm(back - b) / np.linalg.norm(b))}


def cpu_baseline_small(n, threads, loops=20):
    """The CPU-runnable 64^3 FP64 case by the oracle's distributed CPU path
    (`threads` ranks on the drivers' P1 x P2 grid, as cpu_baseline), timed
    like the reference driver: one warm-up pair, then `loops` pairs; the
    median."""
    f = np.random.default_rng(64).random((n, n, n))
    p1, p2, tf, tb, sl = ranks_timed(f, threads, loops + 1)
    med = float(np.median(tf[1:] + tb[1:]))
    return {"value": 1.0 / med, "unit": "fwd+bwd/s", "cores": p1 * p2, "kind": "port",
            "sample": "64^3 U[0,1) FP64 on a %d x %d grid of ranks, one warm-up pair then %d pairs; median %.5f s"
                      % (p1, p2, loops, med), "timers_s": timer_slots(sl[:, 1:, :])}


def small_leg(n=64, reps=200):
    """The CPU-runnable 64^3 FP64 case on the GPU: fwd+bwd through the device
    entry points issued eagerly, and the same pair captured once into a HIP
    graph (p3dfft_amd_capture_*) and replayed."""
    import ctypes
    P = p3dfft_amd.P3DFFT(8)
    P.setup((1, 1), n, n, n)
    L = P.lib
    st = L.p3dfft_amd_stream_create()
    fin, spec, out = P.alloc(P.real_bytes()), P.alloc(P.spec_bytes()), P.alloc(P.real_bytes())
    g = ctypes.c_void_p()
    try:
        fin.upload(np.random.default_rng(1).random((n, n, n)))
        e0, e1 = L.p3dfft_amd_event_create(), L.p3dfft_amd_event_create()

        def timed(issue):
            for _ in range(10):
                issue()
            L.p3dfft_amd_stream_synchronize(st)
            L.p3dfft_amd_event_record(e0, st)
            for _ in range(reps):
                issue()
            L.p3dfft_amd_event_record(e1, st)
            L.p3dfft_amd_stream_synchronize(st)
            return L.p3dfft_amd_event_elapsed_ms(e0, e1) / reps

        def eager():
            P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft", st)
            P.btran_c2r_dev(spec.ptr, out.ptr, "tff", st)

        P.lib.p3dfft_amd_set_timing(0)
        t_eager = timed(eager)
        if L.p3dfft_amd_capture_begin(st):
            raise RuntimeError("capture failed")
        eager()
        if L.p3dfft_amd_capture_end(st, ctypes.byref(g)):
            raise RuntimeError("capture failed")
        t_graph = timed(lambda: L.p3dfft_amd_graph_launch(g, st))
        L.p3dfft_amd_event_destroy(e0)
        L.p3dfft_amd_event_destroy(e1)
    finally:
        if g.value:
            L.p3dfft_amd_graph_destroy(g)
        fin.free(); spec.free(); out.free()
        L.p3dfft_amd_stream_destroy(st)
        P.lib.p3dfft_amd_set_timing(1)
        P.clean()
    return {"grid": [n, n, n], "precision": "f64", "eager_ms_per_fwd_bwd": t_eager,
            "graph_ms_per_fwd_bwd": t_graph, "reps": reps}


def batch_leg(n_req=256, slots=8, reps=20):
    """A fixed batch of small mixed-size solves (axes from {32,...,128}, FP32)
    run `reps` times: issued eagerly on `slots` streams, captured once into a
    HIP graph of `slots` branches (p3dfft_poisson_graph_*), as grouped
    launches, one per pass and transform length (p3dfft_poisson_batch_*),
    and those grouped launches captured into one HIP graph (one host launch
    per batch)."""
    from p3dfft_amd import serve
    shapes = serve.make_requests(n_req, seed=2002, sizes=(32, 48, 64, 96, 128))
    S = serve.PoissonServer(shapes, precision=4, slots=slots)
    G = None
    try:
        S.arena.close()
        S.arena = serve.Arena(S.lib, sum((w + 255) & ~255 for w in S.work))
        idx = list(range(n_req))
        S.fill()
        S.run_streams(idx, reps=2)  # warm
        t_eager = min(S.run_streams(idx, reps=reps) for _ in range(3))
        G = S.graph(idx, nstreams=slots)
        G.run(reps=2)  # warm
        t_graph = min(G.run(reps=reps) for _ in range(3))
        Bt = S.batch(idx)
        launches = Bt.grouped_launches
        Bt.run(reps=2)  # warm
        t_grp = min(Bt.run(reps=reps) for _ in range(3))
        # the same grouped launches captured into one HIP graph: the whole
        # batch as one host launch
        Bt.capture()
        Bt.run_graph(reps=2)  # warm
        t_grp_graph = min(Bt.run_graph(reps=reps) for _ in range(3))
        Bt.close()
        # one at a time: each solve synchronised before the next is issued
        S.fill()
        _, mk_seq, _ = S.run(idx, np.zeros(n_req), sequential=True)
    finally:
        if G is not None:
            G.close()
        S.close()
    S8 = serve.PoissonServer(shapes, precision=8, slots=1)
    try:
        S8.fill()
        S8.run(idx[:32], np.zeros(32), sequential=True)  # warm
        _, mk_seq64, _ = S8.run(idx, np.zeros(n_req), sequential=True)
    finally:
        S8.close()
    eager, graph, grp = n_req * reps / t_eager, n_req * reps / t_graph, n_req * reps / t_grp
    grp_graph = n_req * reps / t_grp_graph
    seq, seq64 = n_req / mk_seq, n_req / mk_seq64
    return {"requests": n_req, "branches": slots, "reps": reps,
            "eager_streams_solves_per_s": eager, "graph_solves_per_s": graph,
            "speedup_graph_vs_eager": graph / eager,
            "grouped_solves_per_s": grp, "grouped_launches_per_batch": launches,
            "grouped_graph_solves_per_s": grp_graph,
            "speedup_grouped_vs_eager": grp / eager,
            "one_at_a_time_solves_per_s": seq, "speedup_grouped_vs_one_at_a_time": grp / seq,
            "one_at_a_time_fp64_solves_per_s": seq64, "speedup_grouped_vs_one_at_a_time_fp64": grp / seq64,
            "mean_points_per_request": float(np.mean([np.prod(s) for s in shapes]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--cpu-threads", type=int, default=len(os.sched_getaffinity(0)),
                    help="threads of the CPU baselines (default: every core this process may run on)")
    ap.add_argument("--no-c3", action="store_true", help="skip the 384^3 leg")
    ap.add_argument("--no-c4", action="store_true", help="skip the Chebyshev 256x256x257 leg")
    ap.add_argument("--chunk-mb", type=float, default=None,
                    help="X/Y z-chunk size in MiB of intermediate (library default 0 = no chunking)")
    ap.add_argument("--no-timers", action="store_true",
                    help="record no per-stage HIP events (whole-step time only)")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the served Poisson leg (0 = skip)")
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--batch-reps", type=int, default=20, help="repetitions of the batched leg's batch")
    ap.add_argument("--no-share", action="store_true", help="skip the share-mode (one rank of 2048^3) leg")
    ap.add_argument("--field", choices=["modes", "random"], default="modes",
                    help="modes: the 8-mode + noise field; random: N(0,1) (fast, for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU exchange is a design, "
              "not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    n = args.n
    log("field %s %d^3" % (args.field, n))
    f = modes_field(n, 2026) if args.field == "modes" else np.random.default_rng(1).standard_normal((n, n, n))
    cb = None if args.chunk_mb is None else int(args.chunk_mb * (1 << 20))
    log("fp32 leg")
    r32 = run_precision(4, n, f, args.steps, args.warmup, cb, not args.no_timers)
    log("fp32 %.4f ms/step" % r32["ms_per_step"])
    r64 = None
    if not args.no_fp64:
        r64 = run_precision(8, n, f, args.steps, args.warmup, cb, not args.no_timers)
        log("fp64 %.4f ms/step" % r64["ms_per_step"])
    lib = p3dfft_amd.load(4)
    cpk = copy_peak(lib)
    log("copy peak %.1f GB/s" % (cpk / 1e9))

    s = 4
    nr, nc = n ** 3, (n // 2 + 1) * n * n
    B = nr * s + nc * 2 * s
    t = r32["ms_per_step"] * 1e-3
    kern = {}
    for slot, name in SLOT_NAMES.items():
        avg = r32["timers"][slot - 1] / args.steps
        kern[name] = {"avg_ms": avg * 1e3, "gbps": slot_bytes(slot, n, s) / avg / 1e9 if avg > 0 else None}
    dom_slot = max(SLOT_NAMES, key=lambda k: r32["timers"][k - 1])
    dom_avg = r32["timers"][dom_slot - 1] / args.steps
    achieved = slot_bytes(dom_slot, n, s) / dom_avg
    line = {
        "metric": METRIC,
        "value": 1.0 / t,
        "unit": "fwd+bwd 512^3 transforms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 8 random-phase Fourier modes + N(0,1e-3) noise, seed 2026",
        "config": {"workload": "%d^3 r2c 'fft' + c2r 'tff', FP32 bulk, compensated twiddles, 1x1 grid, "
                               "fields resident in HBM" % n, "grid": [n, n, n], "op": ["fft", "tff"],
                   "xy_chunk_mib": 0 if args.chunk_mb is None else args.chunk_mb},
        "roofline": {"bound": "hbm", "kernel": SLOT_NAMES[dom_slot], "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": pmc_traffic(SLOT_NAMES[dom_slot].split()[0]) if n == 512 else None,
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, calibrated on a copy kernel: "
                                       "profiles/pmc_traffic_512cubed_fp32.md",
                     "bytes_per_launch": slot_bytes(dom_slot, n, s), "avg_launch_ms": dom_avg * 1e3},
        "path": {"compulsory_bytes": 2 * B, "frac_of_8TBps": 2 * B / t / HBM_PEAK,
                 "copy_kernel_GBps": cpk / 1e9, "frac_of_copy_peak": 2 * B / t / cpk,
                 "gflops_5NlogN": 2 * 5 * nr * np.log2(nr) / t / 1e9, "wall_ms_per_step": r32["wall_ms_per_step"]},
        "kernels": kern,
        "power_spectrum": {"ms": r32["spectrum_ms"], "shells": int((3.0 * n * n) ** 0.5 * 0.5 + 0.5) + 1,
                           "spectrum_read_GBps": (n // 2 + 1) * n * n * 8 / (r32["spectrum_ms"] * 1e-3) / 1e9,
                           "what": "k2-weighted shell sums of the FP32 spectrum (sample/C/driver_spec.c "
                                   "compute_spectrum), HBM in and out"},
    }
    if r64 is not None:
        ref = r64["spec"].astype(np.complex128)
        line["fp64_ms_per_step"] = r64["ms_per_step"]
        line["speedup_fp32_vs_fp64"] = r64["ms_per_step"] / r32["ms_per_step"]
        line["rel_l2_spectrum_fp32_vs_fp64"] = float(np.linalg.norm(r32["spec"] - ref) / np.linalg.norm(ref))
        b64 = r64["back"].astype(np.float64)
        line["rel_l2_roundtrip_fp32_vs_fp64"] = float(np.linalg.norm(r32["back"] - b64) / np.linalg.norm(b64))
    log("64^3 FP64 leg (eager and graph-captured)")
    line["c1_64cubed"] = small_leg()
    if not args.no_cpu:
        line["c1_64cubed"]["cpu_baseline"] = cpu_baseline_small(64, args.cpu_threads)
    if not args.no_c3:
        log("384^3 leg (FP32 and FP64)")
        line["c3_384cubed"] = c3_leg(args.steps, args.warmup)
        log("384^3 fp32 %.4f ms/step" % line["c3_384cubed"]["fp32"]["ms_per_step"])
    if not args.no_c4:
        log("Chebyshev 256x256x257 FP64 leg")
        line["c4_cheby"] = cheby_leg(args.steps, args.warmup, args.cpu_threads, cpu=not args.no_cpu)
        log("cheby %.4f ms fwd+bwd" % line["c4_cheby"]["ms_per_step"])
    if not args.no_share:
        log("share-mode leg (one rank of 2048^3 on 16x16)")
        line["share"] = share_leg()
        log("share %.3f ms per rank fwd+bwd" % line["share"]["ms_per_fwd_bwd"])
    if args.serve_requests > 0:
        log("serve leg (%d requests)" % args.serve_requests)
        line["serve"] = serve_leg(args.serve_requests, args.serve_slots)
        log("serve %.1f solves/s (one-at-a-time %.1f)" % (line["serve"]["solves_per_s"],
                                                          line["serve"]["one_at_a_time_solves_per_s"]))
        log("batched leg (graph-captured vs eager streams)")
        line["batched"] = batch_leg(slots=args.serve_slots, reps=args.batch_reps)
        log("batched %.1f solves/s grouped, %.1f graph, %.1f eager" % (
            line["batched"]["grouped_solves_per_s"], line["batched"]["graph_solves_per_s"],
            line["batched"]["eager_streams_solves_per_s"]))
    if not args.no_cpu:
        log("cpu baseline (oracle, %d threads)" % args.cpu_threads)
        line["cpu_baseline"] = cpu_baseline(n, f, args.cpu_threads)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
