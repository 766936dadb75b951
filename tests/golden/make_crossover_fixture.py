"""Writes tests/golden/crossover/: the library's PDHG answer files for two small LPs and what the reference's
crossover consumer makes of them. Run in the container only (it needs the reference build oracle/_ref/taral, made
by `make -C oracle ref` from the reference's own sources):

    python tests/golden/make_crossover_fixture.py

For each LP the script writes the MPS text (gpu/gen_lp.py-equivalent generator), solves it with the library's `pdhg`
CLI (CPU backend, reference scheme, tol 1e-6) writing `.sol` (x, %.17g per column) and `.dual` (y = Dr z per row,
gpu/pdhg.cu:898-909), then runs the reference engine `taral MODEL.mps --method dual` cold and with
`--warm-sol X.sol --warm-dual X.dual` (src/main.cpp:661-683 -> src/crossover.cpp basis_from_point) and records
status, objective, iterations and the crossover line. The .sol/.dual files are committed beside the record (the MPS
text is not: the generator rewrites it) so the CPU test can check that the library still writes them bit for bit.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "taral-lp_amd"))
from taral_pdhg import gen  # noqa: E402

OUT = os.path.join(HERE, "crossover")
PDHG = os.path.join(ROOT, "taral-lp_amd", "build", "pdhg")
TARAL = os.path.join(ROOT, "oracle", "_ref", "taral")


def reference(mps, warm=None):
    cmd = [TARAL, mps, "--method", "dual"]
    if warm:
        cmd += ["--warm-sol", warm[0], "--warm-dual", warm[1]]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    m = re.search(r"status (\S+) objective (\S+) iterations (\d+)", r.stdout)
    cross = re.search(r"crossover: (\d+) variables farther than (\S+) from a bound", r.stderr)
    return {"exit": r.returncode, "status": m.group(1), "objective": float(m.group(2)), "iterations": int(m.group(3)),
            "crossover_interior": int(cross.group(1)) if cross else None, "stdout": r.stdout.strip()}


def main():
    if not os.path.exists(TARAL):
        sys.exit(f"{TARAL} is missing: run `make -C oracle ref` in a tree with the reference sources")
    os.makedirs(OUT, exist_ok=True)
    cases = {"transport_20x20": gen.transport_mps(20, 20)[1], "packing_2000": gen.packing_mps(2000)[1]}
    record = {"_source": __doc__.strip().splitlines()[0], "cases": []}
    tmp = tempfile.mkdtemp()
    for name, text in cases.items():
        mps = os.path.join(tmp, f"{name}.mps")  # regenerable: not kept
        with open(mps, "w") as f:
            f.write(text)
        sol, dual, js = (os.path.join(OUT, f"{name}{e}") for e in (".sol", ".dual", ".json"))
        r = subprocess.run([PDHG, mps, "--device", "cpu", "--tol", "1e-6", "--sol", sol, "--dual", dual, "--json", js],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        pd = json.load(open(js))
        os.remove(js)
        cold, warm = reference(mps), reference(mps, (sol, dual))
        record["cases"].append({"case": name, "pdhg": {k: pd[k] for k in ("status", "objective", "iterations",
                                                                         "restarts", "rel_primal", "rel_gap")},
                                "reference_cold": cold, "reference_warm": warm})
        print(name, pd["objective"], cold["iterations"], warm["iterations"], warm["status"])
    with open(os.path.join(OUT, "outcome.json"), "w") as f:
        json.dump(record, f, indent=1)


if __name__ == "__main__":
    main()
