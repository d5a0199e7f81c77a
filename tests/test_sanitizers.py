"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only,
g++/gcc): the engine's neighbour-map builder (uppasd_amd/csrc/uasd_setup.cpp)
and the oracle, driven by tests/sanitize/san_driver.cpp, which also checks
that both builders give the same lists for a periodic bcc five-shell lattice,
a triangular DM list and a stripe with vacuum boundaries."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1", "-fopenmp"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_host_code_under_asan_ubsan(tmp_path):
    objs = []
    for src in ("or_rng.c", "or_lattice.c", "or_field.c", "or_llg.c", "or_mc.c", "or_init.c",
                "or_pt.c"):
        o = str(tmp_path / (src + ".o"))
        subprocess.run(["gcc", "-std=c11", "-ffp-contract=off", *SAN, "-c",
                        os.path.join(ROOT, "oracle", src), "-o", o], check=True)
        objs.append(o)
    exe = str(tmp_path / "san_driver")
    subprocess.run(["g++", "-std=c++17", *SAN, os.path.join(ROOT, "tests", "sanitize",
                                                             "san_driver.cpp"),
                    os.path.join(ROOT, "uppasd_amd", "csrc", "uasd_setup.cpp"), *objs, "-lm",
                    "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip() == "ok"
