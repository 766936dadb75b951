"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only, no device code):

* tests/sanitize_host.cpp with the library's host sources (MPS reader, input gate, scaling, recheck, the PDHG driver
  and the CPU backend), built with g++ -fsanitize=address,undefined, run over every MPS fixture (the reference's
  parser_* and literature files, the edge cases and gate models written here): every parse either succeeds or
  throws cleanly, and every supported model is scaled, rechecked and solved on the CPU backend;
* tests/layout_check.cpp (the kernels' host-side layout construction and their indexing replayed on the CPU) built
  with the sanitizers on its host pass only (`-Xarch_host` before each `-fsanitize=`), in all three layout modes.

Any sanitizer report makes the run exit non-zero.
"""
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "taral-lp_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def host_binary(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("san") / "sanitize_host")
    srcs = [os.path.join(ROOT, "tests", "sanitize_host.cpp")] + [os.path.join(SRC, f) for f in
                                                                 ("mps.cpp", "prep.cpp", "recheck.cpp", "driver.cpp",
                                                                  "cpu.cpp")]
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                        "-pthread", f"-I{SRC}", f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__", "-o", exe] + srcs,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def test_host_sources_under_sanitizers(host_binary):
    files = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "mps", "*.mps")))
    r = subprocess.run([host_binary] + files, capture_output=True, text=True, timeout=900, env=ENV)
    print(r.stdout)
    assert r.returncode == 0 and "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, \
        r.stderr[-4000:]
    parsed = int(r.stdout.split()[1])
    assert parsed >= 20


@pytest.mark.parametrize("mode", ["unsorted", "sorted", "slab"])
def test_layout_replay_under_sanitizers(tmp_path, mode):
    exe = str(tmp_path / "layout_check_san")
    r = subprocess.run([os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-O1", "-g", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                        "-o", exe, os.path.join(ROOT, "tests", "layout_check.cpp")],
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    env = {k: v for k, v in ENV.items() if k not in ("TPDHG_SLABS", "TPDHG_LAYOUT")}
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0 and "fails 0" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]
