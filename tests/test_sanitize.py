"""The library's host-side C++ (host_graph.cpp, host_bins.cpp,
host_window.cpp) and the C oracle built with g++/gcc under AddressSanitizer,
UndefinedBehaviorSanitizer and LeakSanitizer, driven by
tests/sanitize/host_sanitize.cpp: generators, file round trips, CSR checks
on valid and malformed input, batch/bin/window builders, host coarsening,
the colouring's renumbering and full oracle runs.  No GPU code is built."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nido_amd", "csrc")
FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def _compile(args):
    r = subprocess.run(args, capture_output=True, text=True)
    return r.returncode, r.stderr[-4000:]


@pytest.mark.timeout(600)
def test_host_code_and_oracle_under_sanitizers(tmp_path):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("g++/gcc not found")
    units = [
        ["g++", "-std=c++17", *FLAGS, "-c", os.path.join(CSRC, f + ".cpp"), "-o", str(tmp_path / (f + ".o"))]
        for f in ("host_graph", "host_bins", "host_window")
    ]
    units.append(["g++", "-std=c++17", *FLAGS, "-c", os.path.join(ROOT, "tests", "sanitize", "host_sanitize.cpp"),
                  "-o", str(tmp_path / "driver.o")])
    units.append(["gcc", "-std=gnu11", *FLAGS, "-ffp-contract=off", "-c",
                  os.path.join(ROOT, "oracle", "louvain_oracle.c"), "-o", str(tmp_path / "oracle.o")])
    with ThreadPoolExecutor(max_workers=5) as ex:
        for rc, err in ex.map(_compile, units):
            assert rc == 0, err
    exe = tmp_path / "host_sanitize"
    objs = [str(tmp_path / (f + ".o")) for f in ("driver", "host_graph", "host_bins", "host_window", "oracle")]
    rc, err = _compile(["g++", *FLAGS, *objs, "-o", str(exe)])
    assert rc == 0, err
    env = dict(os.environ, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="abort_on_error=0:halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])
