"""CPU checks of the drop-in boundary: the C-ABI libraries load, export every
entry point include/p3dfft.h declares, and the host-only paths behave as the
reference does (no GPU calls are made here)."""
import ctypes
import os
import re
import subprocess

import pytest

import p3dfft_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "p3dfft.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set()
    for m in re.finditer(r"^(?!static)[A-Za-z_][\w \t\*]*?\b(\w+)\s*\([^;{]*\)\s*;", src, flags=re.M):
        names.add(m.group(1))
    return sorted(names)


def test_header_declares_reference_entry_points():
    names = declared_symbols()
    for ref in ("p3dfft_setup", "p3dfft_get_dims", "p3dfft_ftran_r2c", "p3dfft_btran_c2r", "p3dfft_cheby",
                "p3dfft_clean", "get_timers", "set_timers"):
        assert ref in names


@pytest.mark.parametrize("precision", [8, 4])
def test_library_exports_every_declared_symbol(precision):
    L = p3dfft_amd.load(precision)
    for name in declared_symbols():
        assert hasattr(L, name), "%s missing from %s" % (name, p3dfft_amd.lib_path(precision))
    assert L.p3dfft_amd_precision() == precision


@pytest.mark.parametrize("precision", [8, 4])
def test_dynamic_symbol_table(precision):
    out = subprocess.run(["nm", "-D", "--defined-only", p3dfft_amd.lib_path(precision)],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared_symbols():
        assert name in exported


def test_call_before_setup_prints_and_returns():
    """build/module.F90:237-239: get_dims before setup prints an error and returns."""
    L = p3dfft_amd.load(8)
    L.p3dfft_amd_set_error_mode(1)
    L.p3dfft_amd_clear_error()
    s = (ctypes.c_int * 3)(7, 7, 7)
    e = (ctypes.c_int * 3)(7, 7, 7)
    z = (ctypes.c_int * 3)(7, 7, 7)
    c = ctypes.c_int(1)
    L.p3dfft_get_dims(s, e, z, ctypes.byref(c))
    assert "call setup before other routines" in L.p3dfft_amd_last_error().decode()
    assert list(s) == [7, 7, 7]
    L.p3dfft_amd_clear_error()


def test_invalid_dims_rejected_without_gpu():
    """build/setup.F90:125-128 and 181-184: invalid dimensions and processor
    geometry are rejected before any device work."""
    P = p3dfft_amd.P3DFFT(8)
    with pytest.raises(p3dfft_amd.P3DFFTError, match="Invalid dimensions"):
        P.setup((1, 1), 0, 8, 8)
    with pytest.raises(p3dfft_amd.P3DFFTError, match="Invalid processor geometry"):
        P.setup((2, 1), 8, 8, 8)
    with pytest.raises(p3dfft_amd.P3DFFTError, match="Invalid cut dimensions"):
        P.setup((1, 1), 8, 8, 8, nxc=9)
    # lengths above 2^24 are refused (those up to it run, above 8192 in
    # several passes)
    with pytest.raises(p3dfft_amd.P3DFFTError, match="exceeds the supported transform length 16777216"):
        P.setup((1, 1), 4, 4, (1 << 24) + 1)


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "p3dfft_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h", ".f90")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "p3dfft_oracle" not in text, f


def test_fortran_module_built():
    if not os.path.exists("/opt/rocm/bin/amdflang"):
        pytest.skip("amdflang not present")
    lib = os.path.join(ROOT, "p3dfft_amd", "lib")
    assert os.path.exists(os.path.join(lib, "mod", "p3dfft.mod"))
    assert os.path.exists(os.path.join(lib, "libp3dfft_amd_f.so"))
    assert os.path.exists(os.path.join(lib, "mod_sp", "p3dfft.mod"))


def test_poisson_graph_rejects_invalid_batches_without_gpu():
    """Batch validation (shapes, fields, work areas) precedes any device work."""
    import ctypes
    from p3dfft_amd import serve
    L = serve._bind(p3dfft_amd.load(4))
    h = ctypes.c_void_p()
    assert L.p3dfft_poisson_graph_create(0, None, None, None, None, 1, ctypes.byref(h)) == -1
    shapes = (ctypes.c_int * 3)(32, 0, 32)
    fields = (ctypes.c_void_p * 1)(16)
    works = (ctypes.c_void_p * 1)(16)
    wbytes = (ctypes.c_size_t * 1)(1 << 20)
    assert L.p3dfft_poisson_graph_create(1, shapes, fields, works, wbytes, 1, ctypes.byref(h)) == -1
    assert b"invalid grid" in L.p3dfft_amd_last_error()
    shapes = (ctypes.c_int * 3)(32, 32, 32)
    wbytes = (ctypes.c_size_t * 1)(100)
    assert L.p3dfft_poisson_graph_create(1, shapes, fields, works, wbytes, 1, ctypes.byref(h)) == -1
    assert b"work area too small" in L.p3dfft_amd_last_error()
    assert h.value is None
    assert L.p3dfft_poisson_graph_destroy(None) == 0
    L.p3dfft_amd_clear_error()


def test_plan_report_empty_before_any_measurement():
    """No measured setup yet: the plan report is empty."""
    import ctypes
    L = p3dfft_amd.load(8)
    buf = ctypes.create_string_buffer(64)
    assert L.p3dfft_amd_plan_report(buf, 64) == 0
    assert buf.value == b""


def test_serve_loop_rejects_inadmissible_requests_without_gpu():
    """The native serving loop refuses a request whose work area can never
    fit the arena (it would wait forever) before any device work."""
    import ctypes
    from p3dfft_amd import serve
    L = serve._bind(p3dfft_amd.load(4))
    lat = (ctypes.c_double * 1)()
    mk = ctypes.c_double()
    pk = ctypes.c_size_t()
    assert L.p3dfft_serve_run(0, None, None, None, None, None, 1, None, 0, 0, 16, 8, lat, ctypes.byref(mk),
                              ctypes.byref(pk)) == -1
    shapes = (ctypes.c_int * 3)(32, 32, 32)
    fields = (ctypes.c_void_p * 1)(16)
    wb = (ctypes.c_size_t * 1)(1 << 20)
    order = (ctypes.c_int * 1)(0)
    arr = (ctypes.c_double * 1)(0.0)
    assert L.p3dfft_serve_run(1, shapes, fields, wb, order, arr, 1, 4096, 1 << 19, 0, 16, 8, lat,
                              ctypes.byref(mk), ctypes.byref(pk)) == -1
    assert b"exceeds the arena" in L.p3dfft_amd_last_error()
    L.p3dfft_amd_clear_error()
