"""The Fortran drop-in: the sample driver built against module p3dfft
(p3dfft_amd/fortran/examples/driver_sine.f90, the reference's driver_sine
check) runs on the GPU and reports a correct round trip."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "p3dfft_amd", "lib")


@pytest.mark.parametrize("exe", ["driver_sine_f.x", "driver_sine_f_sp.x"])
def test_fortran_driver_sine(exe):
    path = os.path.join(LIB, exe)
    if not os.path.exists(path):
        pytest.fail("Fortran driver not built: %s" % path)
    r = subprocess.run([path, "48", "40", "36", "2"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    assert "Results are correct" in r.stdout
