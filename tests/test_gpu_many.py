"""Multivariable transforms (build/ftran.F90:104-317, build/btran.F90:104-304)
through the C-ABI: nv variables at strides dim_in / dim_out (with padding
between variables), against the oracle applied variable by variable."""
import numpy as np
import pytest

import p3dfft_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = {8: 1e-12, 4: 2e-6}


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert p3dfft_amd.device_count() > 0, "no GPU visible to the HIP runtime"


CASES = [((16, 12, 10), (None, None, None), False, True),
         ((16, 12, 10), (12, 8, 8), False, False),
         ((24, 20, 18), (None, None, None), True, True),
         ((24, 20, 18), (20, 16, 12), True, False),
         ((32, 24, 20), (None, None, None), False, False)]


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("pad", [0, 7])
def test_many_matches_per_variable_oracle(precision, case, pad):
    (nx, ny, nz), cut, stride1, ow = case
    nv = 3
    o = O.Plan(nx, ny, nz, *(c or 0 for c in cut), ow=ow, stride1=stride1)
    nr = nx * ny * nz
    nc = int(np.prod(o.spec_shape()))
    dim_in, dim_out = nr + pad, nc + pad
    rng = np.random.default_rng(nv + pad)
    a = rng.standard_normal(nv * dim_in)
    P = p3dfft_amd.P3DFFT(precision)
    P.setup((1, 1), nx, ny, nz, *cut, overwrite=ow, stride1=stride1)
    try:
        for op in ("fft", "ffc", "ffs"):
            s = P.ftran_r2c_many(a, dim_in, dim_out, nv, op)
            for j in range(nv):
                ref = o.ftran_r2c(a[j * dim_in:j * dim_in + nr], op).reshape(-1)
                assert rel(s[j * dim_out:j * dim_out + nc], ref) < TOL[precision], (op, j)
            bop = op[2] + "ff"
            back = P.btran_c2r_many(s, dim_out, dim_in, nv, bop)
            for j in range(nv):
                ref, _ = o.btran_c2r(s[j * dim_out:j * dim_out + nc].astype(np.complex128), bop)
                assert rel(back[j * dim_in:j * dim_in + nr], ref.reshape(-1)) < TOL[precision], (bop, j)
    finally:
        P.clean()


@pytest.mark.parametrize("precision", [8, 4])
def test_cheby_many(precision):
    nx, ny, nz, Lz, nv = 16, 12, 17, 2.0, 2
    o = O.Plan(nx, ny, nz)
    rng = np.random.default_rng(3)
    a = rng.standard_normal(nv * nx * ny * nz)
    nc = (nx // 2 + 1) * ny * nz
    P = p3dfft_amd.P3DFFT(precision)
    P.setup((1, 1), nx, ny, nz)
    try:
        c = P.cheby_many(a, nx * ny * nz, nc, nv, Lz)
        tol = TOL[precision] if precision == 8 else TOL[4] * (4.0 / Lz) * nz
        for j in range(nv):
            ref = o.cheby(a[j * nx * ny * nz:(j + 1) * nx * ny * nz], Lz).reshape(-1)
            assert rel(c[j * nc:(j + 1) * nc], ref) < tol
    finally:
        P.clean()


def test_many_dims_too_small_is_reported():
    """build/ftran.F90:123-129 prints 'input array dimensions are too low'."""
    P = p3dfft_amd.P3DFFT(8)
    P.setup((1, 1), 8, 8, 8)
    try:
        with pytest.raises(p3dfft_amd.P3DFFTError, match="dimensions are too low"):
            P.ftran_r2c_many(np.zeros(2 * 511), 511, 5 * 8 * 8, 2)
    finally:
        P.clean()
