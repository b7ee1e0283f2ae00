"""Served spectral Poisson solves: each result equals the same request solved
alone, the solution matches an FP64 numpy solve (the FP32 path's stated
tolerance 2e-6 relative L2 times a margin of 10 for the operator), admission
never exceeds the HBM work budget, and the sequential baseline gives the same
results."""
import os

import numpy as np
import pytest

import p3dfft_amd
from p3dfft_amd import serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert p3dfft_amd.device_count() > 0, "no GPU visible to the HIP runtime"


def fp64_solve(f):
    """FP64 solve (scipy's pocketfft, threaded): u_hat = -f_hat/|k|^2, k=0 -> 0."""
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


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision", [4, 8])
def test_served_equals_alone_and_fp64(precision):
    sizes = (32, 48, 64, 96, 128)
    shapes = serve.make_requests(40, seed=11, sizes=sizes)
    S = serve.PoissonServer(shapes, precision=precision, slots=4)
    try:
        S.fill()
        rhs = [S.download(i) for i in range(len(shapes))]
        arr = serve.exp_arrivals(len(shapes), 50e-6, seed=12)
        lat, makespan, _ = S.run(list(range(len(shapes))), arr)
        assert np.all(lat >= 0) and makespan > 0
        mixed = [S.download(i) for i in range(len(shapes))]
        S.fill()
        for i in range(len(shapes)):
            S.solve_alone(i)
            alone = S.download(i)
            assert np.array_equal(alone, mixed[i]), i
            tol = 2e-6 if precision == 4 else 1e-12
            assert rel(mixed[i], fp64_solve(rhs[i].astype(np.float64))) < tol * 10, (i, shapes[i])
    finally:
        S.close()


def test_served_full_size_mix_against_fp64():
    """40 requests drawn from the whole served size set {32, ..., 384} per axis
    (radix 3 and 5 lengths included), FP32 with compensated twiddles, served
    on 8 slots with overlapping arrivals: every served result is within
    2e-5 relative L2 of an FP64 solve of the same right-hand side, also the
    requests above 128^3 that carry most of the bytes."""
    shapes = serve.make_requests(40, seed=17)
    assert max(max(s) for s in shapes) > 128
    S = serve.PoissonServer(shapes, precision=4, slots=8)
    try:
        S.fill()
        rhs = [S.download(i) for i in range(len(shapes))]
        arr = serve.exp_arrivals(len(shapes), 20e-6, seed=18)
        S.run(list(range(len(shapes))), arr)
        worst = 0.0
        for i in range(len(shapes)):
            e = rel(S.download(i), fp64_solve(rhs[i]))
            worst = max(worst, e)
            assert e < 2e-5, (i, shapes[i], e)
        print("served full mix: worst relative L2 vs FP64 %.3e" % worst)
    finally:
        S.close()


def test_admission_respects_work_budget():
    """A budget that holds the largest request once: requests queue for memory,
    the arena never exceeds it, and results equal the unconstrained run."""
    shapes = serve.make_requests(24, seed=21, sizes=(32, 64, 96, 128, 160))
    S0 = serve.PoissonServer(shapes, precision=4, slots=4)
    try:
        S0.fill()
        S0.run(list(range(len(shapes))), np.zeros(len(shapes)))
        ref = [S0.download(i) for i in range(len(shapes))]
        work = list(S0.work)
    finally:
        S0.close()
    budget = max(work) + 4096
    S = serve.PoissonServer(shapes, precision=4, slots=4, work_budget=budget)
    try:
        S.fill()
        _, _, max_used = S.run(list(range(len(shapes))), np.zeros(len(shapes)))
        assert max_used <= budget and S.arena.peak <= budget
        for i in range(len(shapes)):
            assert np.array_equal(S.download(i), ref[i])
    finally:
        S.close()


def test_sequential_baseline_same_results():
    shapes = serve.make_requests(12, seed=31, sizes=(32, 48, 64))
    S = serve.PoissonServer(shapes, precision=4, slots=4)
    try:
        S.fill()
        S.run(list(range(len(shapes))), np.zeros(len(shapes)))
        a = [S.download(i) for i in range(len(shapes))]
        S.fill()
        S.run(list(range(len(shapes))), np.zeros(len(shapes)), sequential=True)
        for i in range(len(shapes)):
            assert np.array_equal(S.download(i), a[i])
    finally:
        S.close()


@pytest.mark.parametrize("precision", [4, 8])
def test_fused_zsolve_matches_composed(precision):
    """The fused Z forward / multiply / Z inverse kernel against the composition
    of the forward transform, the multiply pass and the backward transform
    (the same butterflies and the same multiply, so equal to a few ulps), on
    Z lengths with a specialised kernel and without one (40, 27), and both
    against the FP64 numpy solve."""
    shapes = [(32, 48, 64), (48, 40, 40), (64, 32, 27), (20, 24, 512), (96, 36, 128)]
    S = serve.PoissonServer(shapes, precision=precision, slots=2)
    L = S.lib
    try:
        S.fill()
        rhs = [S.download(i) for i in range(len(shapes))]
        fused = []
        for i in range(len(shapes)):
            S.solve_alone(i)
            fused.append(S.download(i))
        S.fill()
        L.p3dfft_poisson_set_fused(0)
        try:
            for i in range(len(shapes)):
                S.solve_alone(i)
                comp = S.download(i)
                scale = np.abs(comp).max()
                ulp = np.finfo(np.float32 if precision == 4 else np.float64).eps
                assert np.abs(fused[i] - comp).max() <= 64 * ulp * scale, shapes[i]
                tol = 2e-6 if precision == 4 else 1e-12
                assert rel(fused[i], fp64_solve(rhs[i].astype(np.float64))) < tol * 10, shapes[i]
        finally:
            L.p3dfft_poisson_set_fused(1)
    finally:
        S.close()


def test_reference_api_one_at_a_time_matches():
    """The one-at-a-time baseline through the reference's entry points (setup,
    ftran_r2c, multiply, btran_c2r, clean per request) solves the same
    problems: FP64 within 1e-12 of the numpy FP64 solve."""
    shapes = serve.make_requests(6, seed=41, sizes=(32, 48, 64))
    S = serve.PoissonServer(shapes, precision=8, slots=1)
    try:
        S.fill()
        rhs = [S.download(i) for i in range(len(shapes))]
        assert S.run_api(list(range(len(shapes)))) > 0
        for i in range(len(shapes)):
            assert rel(S.download(i), fp64_solve(rhs[i])) < 1e-12, shapes[i]
    finally:
        S.close()


@pytest.mark.parametrize("precision,fused", [(4, True), (8, True), (4, False)])
def test_graph_replays_match_eager(precision, fused):
    """A batch captured into a HIP graph (specialised, generic and prime
    lengths; fused Z-solve or the composed multiply pass) gives, on every
    replay, bit-for-bit the fields of the same solves issued eagerly on
    streams, with eager fills and solves between the replays."""
    shapes = [(32, 36, 48), (17, 32, 20), (64, 64, 64), (48, 40, 32), (36, 17, 64), (96, 32, 48), (32, 32, 32)]
    S = serve.PoissonServer(shapes, precision=precision, slots=3, work_budget=None)
    S.arena.close()
    S.arena = serve.Arena(S.lib, sum((w + 255) & ~255 for w in S.work))
    idx = list(range(len(shapes)))
    S.lib.p3dfft_poisson_set_fused(1 if fused else 0)
    G = None
    try:
        S.fill()
        rhs = [S.download(i) for i in idx]
        S.run_streams(idx, reps=1)
        once = [S.download(i) for i in idx]
        S.fill()
        S.run_streams(idx, reps=2)
        twice = [S.download(i) for i in idx]
        G = S.graph(idx, nstreams=3)
        for rep in range(3):
            S.fill()
            S.solve_alone(rep)  # eager work between replays, then a fresh right-hand side
            S.fill([rep])
            G.launch()
            S.lib.p3dfft_amd_synchronize()
            for i in idx:
                assert np.array_equal(S.download(i), once[i]), (rep, i, shapes[i])
        S.fill()
        G.run(reps=2)
        for i in idx:
            assert np.array_equal(S.download(i), twice[i]), (i, shapes[i])
        tol = 2e-6 if precision == 4 else 1e-12
        for i in idx:
            assert rel(once[i], fp64_solve(rhs[i].astype(np.float64))) < tol * 10, (i, shapes[i])
    finally:
        S.lib.p3dfft_poisson_set_fused(1)
        if G is not None:
            G.close()
        S.close()


def test_graph_create_rejects_small_work_area():
    import ctypes
    S = serve.PoissonServer([(32, 32, 32)], precision=4, slots=1)
    try:
        L = S.lib
        shapes = (ctypes.c_int * 3)(32, 32, 32)
        fields = (ctypes.c_void_p * 1)(S.field_ptr(0))
        blk = S.arena.alloc(S.work[0])
        works = (ctypes.c_void_p * 1)(S.arena.ptr(blk))
        wbytes = (ctypes.c_size_t * 1)(S.work[0] - 1)
        h = ctypes.c_void_p()
        assert L.p3dfft_poisson_graph_create(1, shapes, fields, works, wbytes, 2, ctypes.byref(h)) == -1
        assert b"work area" in L.p3dfft_amd_last_error()
        assert h.value is None
        S.arena.release(blk)
    finally:
        S.close()


def test_graph_survives_plan_measurement():
    """A plan measurement that replaces the plans of a graph's lengths retires
    their twiddle tables instead of freeing them: the graph captured before it
    still replays to the same fields, and plans reset only once no graph is
    alive."""
    shapes = [(48, 40, 36), (64, 60, 50)]
    S = serve.PoissonServer(shapes, precision=4, slots=2)
    S.arena.close()
    S.arena = serve.Arena(S.lib, sum((w + 255) & ~255 for w in S.work))
    idx = [0, 1]
    G = None
    L = S.lib
    try:
        S.fill()
        S.run_streams(idx, reps=1)
        want = [S.download(i) for i in idx]
        G = S.graph(idx, nstreams=2)
        P = p3dfft_amd.P3DFFT(4)
        L.p3dfft_amd_set_plan_mode(1)
        try:
            P.setup((1, 1), 48, 40, 36)  # measures new plans for 48, 40 and 36
        finally:
            L.p3dfft_amd_set_plan_mode(0)
            P.clean()
        assert L.p3dfft_amd_reset_plans() == -1  # a graph is alive
        S.fill()
        G.launch()
        L.p3dfft_amd_synchronize()
        for i in idx:
            assert np.array_equal(S.download(i), want[i]), i
    finally:
        if G is not None:
            G.close()
        S.close()
    assert L.p3dfft_amd_reset_plans() == 0


@pytest.mark.parametrize("precision", [4, 8])
def test_grouped_batch_bit_identical_to_single_solves(precision):
    """A batch run as grouped launches (one launch per pass and transform
    length, each workgroup finding its field in a tile table) gives every
    field bit for bit what p3dfft_poisson_solve gives it alone; shapes with
    a generic or prime length run one by one inside the same launch call."""
    shapes = [(32, 48, 64), (64, 32, 48), (48, 48, 48), (96, 64, 32), (32, 32, 32), (64, 96, 128),
              (36, 32, 40), (17, 32, 48), (128, 64, 96), (48, 96, 64)]
    S = serve.PoissonServer(shapes, precision=precision, slots=2)
    B = None
    try:
        idx = list(range(len(shapes)))
        S.fill()
        rhs = [S.download(i) for i in idx]
        for i in idx:
            S.solve_alone(i)
        want = [S.download(i) for i in idx]
        B = S.batch(idx)
        assert B.single_solves == 2  # (36, 32, 40) and (17, 32, 48)
        # distinct lengths of the eight others: 5 in x, 4 in y, 5 in z
        assert B.grouped_launches == 5 + 4 + 5 + 4 + 5
        for rep in range(2):
            S.fill()
            B.launch()
            S.lib.p3dfft_amd_synchronize()
            for i in idx:
                assert np.array_equal(S.download(i), want[i]), (rep, i, shapes[i])
        tol = 2e-6 if precision == 4 else 1e-12
        for i in idx:
            assert rel(want[i], fp64_solve(rhs[i].astype(np.float64))) < tol * 10, shapes[i]
    finally:
        if B is not None:
            B.close()
        S.close()


def test_small_requests_bypass_a_waiting_large_one():
    """With HBM for one large solve only, a large request that must wait for
    memory does not hold back the small requests behind it (some complete
    before it), it still runs after a bounded number of bypasses, the budget
    holds, and every result equals the unconstrained run."""
    shapes = [(64, 64, 64), (160, 160, 160)] + [(32, 32, 32)] * 12
    n = len(shapes)
    S0 = serve.PoissonServer(shapes, precision=4, slots=4)
    try:
        S0.fill()
        S0.run(list(range(n)), np.zeros(n))
        ref = [S0.download(i) for i in range(n)]
        work = list(S0.work)
    finally:
        S0.close()
    budget = work[1] + 4096
    S = serve.PoissonServer(shapes, precision=4, slots=4, work_budget=budget)
    try:
        S.fill()
        lat, _, max_used = S.run(list(range(n)), np.zeros(n), max_bypass=8)
        assert max_used <= budget
        assert min(lat[2:]) < lat[1]  # a small request behind the large one finished first
        assert sum(lat[2:] < lat[1]) <= 8 + 1  # bounded bypass (+ one already in flight)
        for i in range(n):
            assert np.array_equal(S.download(i), ref[i]), i
    finally:
        S.close()


def test_grouped_batch_on_the_served_size_mix():
    """40 requests drawn like the served workload (axes from {32, ..., 384},
    every specialised length, radix 3 and 5 included): the grouped batch
    equals the one-by-one solves bit for bit."""
    shapes = serve.make_requests(40, seed=71)
    S = serve.PoissonServer(shapes, precision=4, slots=2)
    B = None
    try:
        idx = list(range(len(shapes)))
        S.fill()
        for i in idx:
            S.solve_alone(i)
        want = [S.download(i) for i in idx]
        B = S.batch(idx)
        assert B.single_solves == 0
        S.fill()
        B.launch()
        S.lib.p3dfft_amd_synchronize()
        for i in idx:
            assert np.array_equal(S.download(i), want[i]), (i, shapes[i])
    finally:
        if B is not None:
            B.close()
        S.close()


def test_plain_path_baseline_same_solves():
    """The bench's plain-path baseline (PoissonServer.run_plain: forward r2c,
    the per-mode multiply and backward c2r per request, through one-grid
    plans) computes the same solves as the fused entry point: FP64 within
    1e-12 relative L2."""
    shapes = serve.make_requests(12, seed=13, sizes=(32, 48, 64, 96, 128, 160))
    S = serve.PoissonServer(shapes, precision=8, slots=1)
    try:
        order = list(range(len(shapes)))
        S.fill()
        S.run(order, np.zeros(len(shapes)), sequential=True)
        fused = [S.download(i).astype(np.float64) for i in order]
        S.fill()
        assert S.run_plain(order) > 0
        for i in order:
            assert rel(S.download(i), fused[i]) < 1e-12, (i, shapes[i])
    finally:
        S.close()


@pytest.mark.parametrize("precision", [4, 8])
def test_grouped_batch_captured_graph_bit_identical(precision):
    """The grouped launches of a batch captured into one HIP graph (the whole
    batch as one host launch) replay to the eager grouped result bit for
    bit, three times, with the right-hand sides refilled and eager solves
    between the replays."""
    shapes = [(32, 48, 64), (64, 32, 48), (48, 48, 48), (96, 64, 32), (32, 32, 32), (64, 96, 128),
              (128, 64, 96), (48, 96, 64)]
    S = serve.PoissonServer(shapes, precision=precision, slots=2)
    B = None
    try:
        idx = list(range(len(shapes)))
        B = S.batch(idx)
        S.fill()
        B.launch()
        S.lib.p3dfft_amd_synchronize()
        want = [S.download(i) for i in idx]
        B.capture()
        for rep in range(3):
            S.fill()
            S.solve_alone(0)  # eager work between the replays
            S.fill([0])
            B.run_graph(1)
            for i in idx:
                assert np.array_equal(S.download(i), want[i]), (rep, i)
    finally:
        if B is not None:
            B.close()
        S.close()
