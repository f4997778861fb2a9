"""Build steps for the native parts (called by __graft_entry__.build()).

* quack_amd/libquack_amd.so  HIP kernels + C-ABI (hipcc, gfx950)
* quack_amd/intgen/libqkint.so  integral producer (gcc, OpenMP)
* quack_amd/fortran/  Fortran drop-in objects and the qk_gw_driver program (flang)
* oracle/_ref/libqkref.so  compiled QuAcK reference, only where the QuAcK
  sources are present (test infrastructure, built by oracle/Makefile)
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "quack_amd")
CSRC = os.path.join(PKG, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIP_SOURCES = ["qk_gemm.hip", "qk_gemm_bf16x3.hip", "qk_gw_kernels.hip", "qk_gw_api.hip", "qk_mlp.hip", "qk_sytrd.hip",
               "qk_int.hip", "qk_scf.hip", "qk_noalloc.hip", "qk_gw_unrestricted.hip", "qk_gw_batch.hip", "qk_gw_qsgw.hip",
               "qk_gw_bse.hip"]
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build_hip():
    obj = os.path.join(CSRC, "_build")
    os.makedirs(obj, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")

    def one(src):
        out = os.path.join(obj, src.replace(".hip", ".o"))
        _run([hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out])
        return out

    with ThreadPoolExecutor(len(HIP_SOURCES)) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    lib = os.path.join(PKG, "libquack_amd.so")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-o", lib + ".tmp"] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(lib + ".tmp", lib)
    return lib


def build_intgen():
    src = os.path.join(PKG, "intgen", "intgen.c")
    lib = os.path.join(PKG, "intgen", "libqkint.so")
    _run(["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-Wall", "-o", lib + ".tmp", src, "-lm"])
    os.replace(lib + ".tmp", lib)
    return lib


def build_oracle(reference="/root/reference"):
    """Compile QuAcK's own routines into oracle/_ref (only where the sources exist)."""
    if not os.path.isdir(os.path.join(reference, "src", "GW")):
        return None
    if shutil.which("flang") is None:
        return None
    _run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle"), "REF=" + reference])
    return os.path.join(ROOT, "oracle", "_ref", "libqkref.so")


def build_all():
    libs = [build_hip(), build_intgen()]
    try:
        from . import fortran_build
        libs.append(fortran_build.build())
    except ImportError:
        pass
    libs.append(build_oracle())
    return libs


if __name__ == "__main__":
    print(build_all())
