"""Builds the Fortran drop-in (flang): quack_amd/fortran/qk_gw_driver."""
import os
import shutil
import subprocess
import sys

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran")
PKG = os.path.dirname(HERE)


def build():
    fc = shutil.which("flang") or shutil.which("amdflang")
    if fc is None:
        return None
    obj = os.path.join(HERE, "_build")
    os.makedirs(obj, exist_ok=True)
    srcs = ["quack_amd_gw.f90", "RGW_hip.f90", "UGW_hip.f90", "qsRGW_hip.f90", "qk_gw_driver.f90"]
    objs = []
    for s in srcs:
        o = os.path.join(obj, s.replace(".f90", ".o"))
        r = subprocess.run([fc, "-O2", "-fPIC", "-J", obj, "-c", os.path.join(HERE, s), "-o", o],
                           capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed on " + s)
        objs.append(o)
    exe = os.path.join(HERE, "qk_gw_driver")
    r = subprocess.run([fc, "-o", exe + ".tmp"] + objs + ["-L" + PKG, "-lquack_amd", "-Wl,-rpath,$ORIGIN/.."],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of qk_gw_driver failed")
    os.replace(exe + ".tmp", exe)
    return exe
