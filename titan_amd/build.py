"""Builds titan_amd/libtitan_hip.so (gfx950) and the Fortran shim objects.

    python -m titan_amd.build          # HIP library
    python -m titan_amd.build --all    # + Fortran shim test driver

Compiles in-tree with hipcc (no JIT cache), one object per source in
parallel, then links the shared library.  Packed-FP32 instructions are
disabled for all device code.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libtitan_hip.so")
SOURCES = ["th_green.hip", "th_gemm.hip", "th_dyson.hip", "th_adjoint.hip", "th_traces.hip", "th_density.hip", "th_current.hip", "th_refine.hip", "th_abi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_hip(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        for s in SOURCES:
            o = os.path.join(OBJ, s.replace(".hip", ".o"))
            objs.append(o)
            cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", o]
            jobs.append(ex.submit(_run, cmd))
        for j in jobs:
            out = j.result()
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lrocsolver", "-lrocblas"])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """Compiles the ISO_C_BINDING shim and its test driver with flang."""
    from titan_amd.fortran import build as fb
    return fb.build()


def main(argv):
    lib = build_hip(verbose="-v" in argv)
    print("built", lib)
    if "--all" in argv:
        print("built", build_fortran())


if __name__ == "__main__":
    main(sys.argv[1:])
