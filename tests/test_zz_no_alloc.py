"""The allocation-free power spectrum entry point against the allocating one,
through the C-ABI: identical shell sums on two calls in a row (the second on
a smaller grid), and a refusal with a message, not a fault, when the partial
sums would not fit its static scratch array."""
import numpy as np
import pytest

import p3dfft_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert p3dfft_amd.device_count() > 0, "no GPU visible to the HIP runtime"


def _kmax(nx, ny, nz):
    # the drivers' bound: 4 kmax^2 >= the squared largest stored index
    d = ((nx // 2), ny - 1, nz - 1)
    return int(np.ceil(0.5 * np.sqrt(sum(v * v for v in d)))) + 1


@pytest.mark.parametrize("precision", [8, 4])
def test_noalloc_power_spectrum_matches_allocating_entry(precision):
    L = p3dfft_amd.load(precision)
    for k, (nx, ny, nz) in enumerate(((64, 64, 64), (32, 24, 20))):
        P = p3dfft_amd.P3DFFT(precision)
        P.setup((1, 1), nx, ny, nz)
        try:
            f = np.random.default_rng(100 + k).standard_normal((nz, ny, nx))
            fin, spec = P.alloc(P.real_bytes()), P.alloc(P.spec_bytes())
            fin.upload(f.astype(P.rdtype))
            P.ftran_r2c_dev(fin.ptr, spec.ptr, "fft")
            kmax = _kmax(nx, ny, nz)
            e1, e2 = P.alloc((kmax + 1) * 8), P.alloc((kmax + 1) * 8)
            L.p3dfft_amd_memset(e2.ptr, 0xff, (kmax + 1) * 8)
            assert L.p3dfft_power_spectrum_dev(spec.ptr, kmax, e1.ptr, None) == 0
            assert L.p3dfft_power_spectrum_dev_noalloc(spec.ptr, kmax, e2.ptr, None) == 0
            P.synchronize()
            a = e1.download((kmax + 1,), np.float64)
            b = e2.download((kmax + 1,), np.float64)
            assert np.array_equal(a, b), (nx, ny, nz)
            assert a.sum() > 0
            # partials beyond the 16 MiB scratch: refused with a message
            big = 4096
            assert L.p3dfft_power_spectrum_dev_noalloc(spec.ptr, big, e2.ptr, None) == -1
            assert "scratch" in P.last_error()
            L.p3dfft_amd_clear_error()
            for buf in (fin, spec, e1, e2):
                buf.free()
        finally:
            P.clean()
