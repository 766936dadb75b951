"""Host layout rules of the SpMV kernels, checked on the CPU (no GPU): tests/layout_check.cpp builds matrices of
several shapes (plain, long rows, 8 and 16 column slabs, the largest sorted blocks, a block-structured matrix that
must stay in one piece, a forced slab count), lays them out with the library's make_layout and replays every
kernel's indexing, staging and row sums on the CPU against the CSR product. Compiled host-side with hipcc (the
header it includes is the kernels' own)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


@pytest.fixture(scope="module")
def layout_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not installed")
    exe = str(tmp_path_factory.mktemp("layout") / "layout_check")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-o", exe, os.path.join(ROOT, "tests", "layout_check.cpp")],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.mark.parametrize("mode", ["sorted", "unsorted", "slab"])
def test_layouts_replayed_on_cpu(layout_check, mode):
    env = {k: v for k, v in os.environ.items() if k not in ("TPDHG_SLABS", "TPDHG_LAYOUT")}
    r = subprocess.run([layout_check, mode], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "fails 0" in r.stdout
