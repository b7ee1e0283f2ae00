"""A seeded sweep of the transform options through the C-ABI against the FP64
oracle: random grid shapes (mixed radices 2/3/5/7, primes up to 31, odd and
even lengths), every Z op, random cuts (pruning) of each axis, the STRIDE1
layout and both overwrite settings, in both precisions; forward and backward
each checked against the oracle (FP64 1e-12, FP32 2e-6 relative L2).  The
draws are fixed by the seed, so a failure names a reproducible case."""
import numpy as np
import pytest

from test_gpu_parity import TOL, lib_setup, oplan, rel

pytestmark = pytest.mark.gpu

LENGTHS = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 21, 24, 25, 27, 28, 30, 31, 32,
           35, 36, 40, 42, 45, 48, 49, 50, 54, 56, 60, 63, 64, 70, 72, 75, 80, 81, 96, 100, 105, 120, 125, 128]


def draw_cases(seed, n):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n):
        nx, ny, nz = (int(rng.choice(LENGTHS)) for _ in range(3))
        while nx * ny * nz > 200000:
            nz = max(2, nz // 2)
        cut = [None, None, None]
        for ax, m in enumerate((nx, ny, nz)):
            if rng.random() < 0.3 and m > 2:
                cut[ax] = int(rng.integers(1, m))
        cases.append(dict(shape=(nx, ny, nz), cut=tuple(cut), stride1=bool(rng.random() < 0.3),
                          ow=bool(rng.random() < 0.5), zop=str(rng.choice(list("tcsn")))))
    return cases


CASES = draw_cases(2026, 24)


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("k", range(len(CASES)))
def test_sweep_against_oracle(precision, k):
    c = CASES[k]
    nx, ny, nz = c["shape"]
    nxc, nyc, nzc = c["cut"]
    zop = c["zop"]
    if zop == "c" and nz < 2:
        zop = "t"
    fop, bop = "ff" + zop, zop + "ff"
    if zop == "t":
        fop, bop = "fft", "tff"
    a = np.random.default_rng(1000 + k).standard_normal((nz, ny, nx))
    o = oplan(nx, ny, nz, nxc, nyc, nzc, ow=c["ow"], stride1=c["stride1"])
    ref_f = o.ftran_r2c(a, fop)
    ref_b = o.btran_c2r(ref_f, bop)[0]
    with lib_setup(precision, nx, ny, nz, nxc, nyc, nzc, ow=c["ow"], stride1=c["stride1"]) as P:
        s = P.ftran_r2c(a, fop)
        assert rel(s, ref_f) < TOL[precision], (c, fop)
        b = P.btran_c2r(ref_f, bop)
        assert rel(b, ref_b) < TOL[precision], (c, bop)


def test_sweep_cases_cover_the_options():
    """The drawn cases include cuts on every axis, STRIDE1, both overwrite
    settings and every Z op."""
    assert {c["zop"] for c in CASES} >= set("tcsn")
    assert any(c["stride1"] for c in CASES) and any(not c["stride1"] for c in CASES)
    assert any(c["ow"] for c in CASES) and any(not c["ow"] for c in CASES)
    for ax in range(3):
        assert any(c["cut"][ax] is not None for c in CASES), ax
