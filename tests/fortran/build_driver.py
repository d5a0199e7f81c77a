"""Links the Fortran test driver (driver.f90) against libuasd_hip.so and the
Fortran binding module; the binary is written beside this file and finds the
library through an $ORIGIN-relative run path."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "uasd_driver")


def build(verbose: bool = False):
    sys.path.insert(0, ROOT)
    from uppasd_amd import build as B
    got = B.build_fortran(verbose)
    if got is None:
        return None
    obj, moddir = got
    cmd = [B.FLANG, "-O2", "-module-dir", moddir, "-I", moddir,
           os.path.join(HERE, "driver.f90"), obj, "-L", os.path.join(ROOT, "uppasd_amd"),
           "-luasd_hip", "-Wl,-rpath,$ORIGIN/../../uppasd_amd", "-o", BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver link failed:\n{r.stdout}\n{r.stderr}")
    return BIN


if __name__ == "__main__":
    print(build(verbose=True))
