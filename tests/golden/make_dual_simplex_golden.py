"""Writes tests/golden/cfg0_dual_simplex.json: the reference's FP64 dual-simplex optimum of the 2,000 x 5,000
planted LP (10 nonzeros per column, seed 0).

The reference CPU engine (src/*.cpp, built by `make -C oracle ref` into oracle/_ref/taral where the reference
tree is present) is run as `taral MODEL.mps --method dual --json OUT` (src/main.cpp:425-712). Run from the repository
root: python tests/golden/make_dual_simplex_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "taral-lp_amd"))
from taral_pdhg import gen  # noqa: E402


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    md = gen.planted_lp(2000, 5000, 10, seed=0, name="CFG0")
    with tempfile.TemporaryDirectory() as d:
        mps, out = os.path.join(d, "cfg0.mps"), os.path.join(d, "cfg0.json")
        gen.write_mps(md, mps)
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "taral"), mps, "--method", "dual", "--time-limit", "1800",
                        "--json", out], check=True)
        r = json.load(open(out))
    g = {"model": "planted_lp(2000, 5000, 10, seed=0)", "solver": "reference taral --method dual (FP64 dual simplex)",
         "status": r["status"], "objective": r["objective"], "iterations": r.get("iterations"),
         "wall_s": r.get("wall_s"), "planted_objective": md.planted_objective}
    with open(os.path.join(ROOT, "tests", "golden", "cfg0_dual_simplex.json"), "w") as f:
        json.dump(g, f, indent=1)
    print(g)


if __name__ == "__main__":
    main()
