"""Builds the Fortran ISO_C_BINDING shim (mod_titan_hip) and its driver with
flang, linked against titan_amd/libtitan_hip.so (rpath = that directory)."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)


def flang():
    for c in ("/opt/rocm/llvm/bin/flang", "/opt/rocm/bin/amdflang", shutil.which("flang") or ""):
        if c and os.path.exists(c):
            return c
    raise ImportError("no Fortran compiler (flang) found")


def build():
    fc = flang()
    obj = os.path.join(PKG, "_obj")
    os.makedirs(obj, exist_ok=True)
    exe = os.path.join(HERE, "th_shim_driver")
    subprocess.run([fc, "-O2", "-c", os.path.join(HERE, "mod_titan_hip.F90"), "-module-dir", obj,
                    "-o", os.path.join(obj, "mod_titan_hip.o")], check=True)
    subprocess.run([fc, "-O2", os.path.join(HERE, "th_shim_driver.F90"), "-I", obj, "-module-dir", obj,
                    os.path.join(obj, "mod_titan_hip.o"), "-L" + PKG, "-ltitan_hip",
                    "-Wl,-rpath," + PKG, "-o", exe], check=True)
    return exe


if __name__ == "__main__":
    print(build())
