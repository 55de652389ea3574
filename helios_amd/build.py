"""Build recipe of libhelios_ccsd.so (HIP, gfx950) and the Fortran host.

    python -m helios_amd.build            # library only
    python -m helios_amd.build --fortran  # library + Fortran driver

Objects go to helios_amd/_build/, the library to helios_amd/libhelios_ccsd.so.
Every HIP source is compiled with the packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhelios_ccsd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
ARCH = "gfx950"
HIP_SOURCES = ["hcc_gemm.hip", "hcc_ops.hip", "hcc_solver.hip", "hcc_mlp.hip", "hcc_scf.hip", "hcc_triples.hip",
               "hcc_lr.hip"]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_library(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """Fortran host (helios_amd/src) linked against libhelios_ccsd.so."""
    src = os.path.join(HERE, "src")
    order = ["modules/mod_kinds.f90", "modules/mod_ccsd_gpu.f90", "modules/mod_linalg.f90",
             "input/mod_input.f90", "input/mod_molecule.f90", "methods/mod_scf.f90",
             "calculation/mod_iterator.f90", "methods/mod_cc_gpu.f90", "main.f90"]
    files = [os.path.join(src, f) for f in order if os.path.exists(os.path.join(src, f))]
    if not files:
        return None
    fb = os.path.join(BUILD, "fortran")
    os.makedirs(fb, exist_ok=True)
    objs = []
    for f in files:   # module order matters: one file at a time
        obj = os.path.join(fb, os.path.basename(f).replace(".f90", ".o"))
        _run([FLANG, "-O2", "-c", "-module-dir", fb, "-I" + fb, f, "-o", obj])
        objs.append(obj)
    exe = os.path.join(HERE, "helios_gpu")
    _run([FLANG, "-O2", "-o", exe] + objs + ["-L" + HERE, "-lhelios_ccsd", "-Wl,-rpath,$ORIGIN"])
    return exe


if __name__ == "__main__":
    print(build_library())
    if "--fortran" in sys.argv:
        print(build_fortran())
