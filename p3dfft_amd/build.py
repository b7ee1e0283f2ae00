"""Build the gfx950 libraries of p3dfft_amd in-tree.

Produces, under p3dfft_amd/lib/:
  libp3dfft_amd.so     double precision (the reference's default build)
  libp3dfft_amd_sp.so  single precision (the reference's SINGLE_PREC build)
  p3dfft.mod + libp3dfft_amd_f.so (dp) / libp3dfft_amd_sp_f.so (sp)
                       the Fortran module over the C-ABI, when amdflang exists

Run:  python -m p3dfft_amd.build   (or p3dfft_amd.build.build()).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "bin", "amdflang")

SOURCES = ["engine.hip", "engine_long.hip", "share.hip", "api.cpp", "api_share.cpp", "api_poisson.cpp", "api_procs.cpp", "serve.cpp"]
# packed-FP32 instructions are disabled for all device code
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option"]

VARIANTS = {"libp3dfft_amd.so": [], "libp3dfft_amd_sp.so": ["-DP3DFFT_SINGLE"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r.stdout


def _compile_obj(src, defs, tag):
    obj = os.path.join(LIB, "obj", tag, os.path.splitext(src)[0] + ".o")
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    lang = ["-x", "hip"]
    _run([HIPCC] + lang + HIP_FLAGS + defs + ["-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_library(name, defs):
    tag = os.path.splitext(name)[0]
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile_obj(s, defs, tag), SOURCES))
    out = os.path.join(LIB, name)
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_fortran():
    """Compile the Fortran module (iso_c_binding interfaces over the C-ABI)."""
    if not os.path.exists(FLANG):
        return None
    src = os.path.join(HERE, "fortran", "p3dfft.f90")
    outs = []
    for lib, defs, modir in (("libp3dfft_amd_f.so", [], "mod"), ("libp3dfft_amd_sp_f.so", ["-DSINGLE_PREC"], "mod_sp")):
        md = os.path.join(LIB, modir)
        os.makedirs(md, exist_ok=True)
        obj = os.path.join(md, "p3dfft_mod.o")
        _run([FLANG, "-cpp", "-O2", "-fPIC"] + defs + ["-module-dir", md, "-c", src, "-o", obj])
        wobj = os.path.join(md, "wrap.o")
        _run([FLANG, "-cpp", "-O2", "-fPIC"] + defs + ["-I", md, "-module-dir", md, "-c",
              os.path.join(HERE, "fortran", "wrap.f90"), "-o", wobj])
        base = "p3dfft_amd" if not defs else "p3dfft_amd_sp"
        _run([FLANG, "-shared", "-fPIC", "-o", os.path.join(LIB, lib), obj, wobj,
              "-L" + LIB, "-l" + base, "-Wl,-rpath,$ORIGIN"])
        outs.append(os.path.join(LIB, lib))
        # the Fortran sample driver (examples/driver_sine.f90) against this module
        exe = os.path.join(LIB, "driver_sine_f.x" if not defs else "driver_sine_f_sp.x")
        _run([FLANG, "-cpp", "-O2"] + defs + ["-I", md, os.path.join(HERE, "fortran", "examples", "driver_sine.f90"),
              "-o", exe, "-L" + LIB, "-l" + os.path.splitext(lib)[0][3:], "-l" + base, "-Wl,-rpath,$ORIGIN"])
        outs.append(exe)
        # in-place use through the external wrappers (examples/driver_sine_inplace.f90)
        exe = os.path.join(LIB, "driver_sine_inplace_f.x" if not defs else "driver_sine_inplace_f_sp.x")
        _run([FLANG, "-cpp", "-O2"] + defs + ["-I", md,
              os.path.join(HERE, "fortran", "examples", "driver_sine_inplace.f90"),
              "-o", exe, "-L" + LIB, "-l" + os.path.splitext(lib)[0][3:], "-l" + base, "-Wl,-rpath,$ORIGIN"])
        outs.append(exe)
        # the utility procedures' check (examples/driver_utils.f90)
        exe = os.path.join(LIB, "driver_utils_f.x" if not defs else "driver_utils_f_sp.x")
        _run([FLANG, "-cpp", "-O2"] + defs + ["-I", md, os.path.join(HERE, "fortran", "examples", "driver_utils.f90"),
              "-o", exe, "-L" + LIB, "-l" + os.path.splitext(lib)[0][3:], "-l" + base, "-Wl,-rpath,$ORIGIN"])
        outs.append(exe)
    return outs


def build(fortran=True):
    os.makedirs(LIB, exist_ok=True)
    with ThreadPoolExecutor(max_workers=2) as ex:
        libs = list(ex.map(lambda kv: build_library(*kv), VARIANTS.items()))
    if fortran and os.path.exists(os.path.join(HERE, "fortran", "p3dfft.f90")):
        build_fortran()
    shutil.rmtree(os.path.join(LIB, "obj"), ignore_errors=True)
    return libs


if __name__ == "__main__":
    for p in build():
        print(p)
