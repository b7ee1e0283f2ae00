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


@pytest.mark.parametrize("exe", ["driver_sine_inplace_f.x", "driver_sine_inplace_f_sp.x"])
def test_fortran_driver_sine_inplace(exe):
    """In-place transforms from Fortran through the external ftran_r2c /
    btran_c2r (wrap.f90; the reference's build/wrap.F90:82-150), one array
    passed as input and output as sample/FORTRAN/driver_sine_inplace.F90:206,225
    does: the four modes of the sine field and the round trip."""
    path = os.path.join(LIB, exe)
    if not os.path.exists(path):
        pytest.fail("Fortran driver not built: %s" % path)
    r = subprocess.run([path, "32", "24", "20", "2"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    assert "Results are correct" in r.stdout


@pytest.mark.parametrize("exe", ["driver_utils_f.x", "driver_utils_f_sp.x"])
def test_fortran_module_utilities(exe):
    """p3dfft_get_mpi_info, rtran_x2y/y2x/x2z/z2x and print_buf of module
    p3dfft (build/module.F90:178-186 publics) on one GPU: task 0 of 1 with the
    setup's communicator, the transposes copy the whole grid with its bounds,
    print_buf lists exactly the non-zero modes."""
    path = os.path.join(LIB, exe)
    if not os.path.exists(path):
        pytest.fail("Fortran utilities driver not built: %s" % path)
    r = subprocess.run([path], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    assert "Utilities are correct" in r.stdout
    lines = [l for l in r.stdout.splitlines() if ": (" in l]
    assert len(lines) == 2
    # loop order: k outermost, then j, then i
    assert [int(v) for v in lines[0].split("(")[1].split(")")[0].split()] == [5, 1, 1]
    assert [int(v) for v in lines[1].split("(")[1].split(")")[0].split()] == [2, 3, 4]
