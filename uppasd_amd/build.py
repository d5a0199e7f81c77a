"""Builds libuasd_hip.so in-tree for gfx950 (hipcc, no JIT cache).

The FP64 kernels and the context are compiled with -ffp-contract=off so that
every multiply and add rounds on its own, except the exchange sums, which
call fma() explicitly (the reference's release build contracts them too);
the oracle rounds the same way.  The FP32 kernels allow contraction.  Packed FP32 instructions are disabled for all
device code (-packed-fp32-ops).  The host set-up (neighbour map builder) is
plain C++ with OpenMP, built by g++.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libuasd_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_COMMON = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wno-unused-command-line-argument"]
UNITS = [
    ("uasd_kernels_f64.hip", ["-ffp-contract=off"]),
    ("uasd_kernels_f32.hip", ["-ffp-contract=fast"]),
    ("uasd_engine.hip", ["-ffp-contract=off"]),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    # hipcc reads the sources twice per unit (device, then host pass): a
    # source edited while it runs would leave host launch code for kernels
    # the device code does not have, so the build refuses to finish then
    before = _source_digest()
    # the translation units are independent: compile them side by side
    cmds, objs = [], []
    for src, extra in UNITS:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmds.append([HIPCC, *HIP_COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    setup_obj = os.path.join(objdir, "uasd_setup.o")
    cmds.append(["g++", "-O2", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off",
                 "-c", os.path.join(CSRC, "uasd_setup.cpp"), "-o", setup_obj])
    objs.append(setup_obj)
    with ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64", "-fopenmp",
          f"-Wl,-rpath,{ROCM}/lib"], verbose)
    if _source_digest() != before:
        os.remove(tmp)
        raise RuntimeError("sources changed while the library was being built; build again")
    os.replace(tmp, LIB)
    return LIB


def _source_digest() -> str:
    """SHA-256 over the sources the library is built from."""
    h = hashlib.sha256()
    files = sorted(os.listdir(CSRC)) + [os.path.join("..", "..", "include", "uasd_hip.h")]
    for name in files:
        path = os.path.normpath(os.path.join(CSRC, name))
        if os.path.isfile(path):
            h.update(name.encode())
            with open(path, "rb") as f:
                h.update(f.read())
    return h.hexdigest()


FLANG = os.path.join(ROCM, "bin", "amdflang")


def build_fortran(verbose: bool = False):
    """Compiles the Fortran binding (fortran/uasd_hip.f90) with amdflang into
    _build/fortran (object + .mod).  Returns (object, module dir), or None when
    no amdflang is installed."""
    if not os.path.exists(FLANG):
        return None
    outdir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(outdir, exist_ok=True)
    obj = os.path.join(outdir, "uasd_hip.o")
    _run([FLANG, "-O2", "-fPIC", "-module-dir", outdir, "-c",
          os.path.join(HERE, "fortran", "uasd_hip.f90"), "-o", obj], verbose)
    return obj, outdir


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
