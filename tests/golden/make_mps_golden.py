"""Writes the MPS reader fixtures: tests/golden/mps/<case>.mps and <case>.dump.

Each .dump is what the REFERENCE reader (src/mps.cpp, compiled here by `make -C oracle ref` into
oracle/_ref/ref_mps_dump) makes of the .mps file, printed in a canonical text form (see
oracle/ref_mps_dump.cpp). The cases are written here (edge cases of the MPS rules and the 13 input-gate
models of gpu/pdhg_input_tests.py:57-71), plus data files copied from the reference's own tests:
tests/repro/parser_*.mps (reader rejection fixtures) and examples/literature_lp/*.mps (small LPs with
known optima). Run from the repository root where the reference tree is present:
python tests/golden/make_mps_golden.py
"""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "mps")
DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_mps_dump")
REF = "/root/reference"


def gate_model(name, bounds=" UP B X 2\n", quadratic="", marker=False, maximize=False):
    """One-column LP in the shape the reference's input-gate test uses (gpu/pdhg_input_tests.py:34-40)."""
    sense = "OBJSENSE\n MAX\n" if maximize else ""
    start = " M0 'MARKER' 'INTORG'\n" if marker else ""
    end = " M1 'MARKER' 'INTEND'\n" if marker else ""
    return (f"NAME {name}\n{sense}ROWS\n N OBJ\n G R\nCOLUMNS\n{start} X OBJ 1 R 1\n{end}RHS\n RHS R 0.5\n"
            f"BOUNDS\n{bounds}{quadratic}ENDATA\n")


GATE = {
    "gate_continuous_lp": gate_model("LP"),
    "gate_maximize_lp": gate_model("MAXLP", maximize=True),
    "gate_zero_quadratic": gate_model("ZERO", quadratic="QUADOBJ\n X X 0\n"),
    "gate_quadobj": gate_model("QP", quadratic="QUADOBJ\n X X 2\n"),
    "gate_qmatrix": gate_model("QM", quadratic="QMATRIX\n X X 2\n"),
    "gate_qsection": gate_model("QS", quadratic="QSECTION OBJ\n X X 2\n"),
    "gate_maximize_qp": gate_model("MAXQP", quadratic="QUADOBJ\n X X -2\n", maximize=True),
    "gate_intorg": gate_model("MILP", marker=True),
    "gate_binary_bound": gate_model("BV", bounds=" BV B X\n"),
    "gate_integer_lower": gate_model("LI", bounds=" LI B X 0\n UP B X 2\n"),
    "gate_integer_upper": gate_model("UI", bounds=" UI B X 2\n"),
    "gate_fixed_integer": gate_model("FIX", bounds=" FX B X 1\n", marker=True),
    "gate_miqp": gate_model("MIQP", marker=True, quadratic="QUADOBJ\n X X 2\n"),
}

EDGE = {
    "ranges_all_kinds": """NAME RANGES
ROWS
 N  COST
 E  E1
 E  E2
 L  L1
 G  G1
 E  E3
COLUMNS
    X  COST 1  E1 1
    X  E2 2  L1 3
    X  G1 4  E3 5
    Y  COST -2  E1 1
    Y  L1 1
RHS
    RHS  E1 10  E2 20
    RHS  L1 30  G1 40
    RHS  E3 5
RANGES
    RNG  E1 4  E2 -6
    RNG  L1 -7  G1 8
ENDATA
""",
    "bounds_all_types": """NAME BOUNDS
ROWS
 N obj
 L c1
COLUMNS
 a obj 1 c1 1
 b obj 1 c1 1
 c obj 1 c1 1
 d obj 1 c1 1
 e obj 1 c1 1
 f obj 1 c1 1
 g obj 1 c1 1
 h obj 1 c1 1
 i obj 1 c1 1
 j obj 1 c1 1
RHS
 rhs c1 100
BOUNDS
 UP bnd a -3
 LO bnd b 2
 UP bnd b -1
 FR bnd c
 MI bnd d
 UP bnd d 4
 PL bnd e
 FX bnd f 2.5
 LO bnd g -1e30
 UP bnd g 1e20
 UP bnd h 1e19
 LO bnd i 1.5D2
 UP bnd i 2.5d+02
 MI bnd j
 PL bnd j
ENDATA
""",
    "duplicates_and_zeros": """NAME DUPS
ROWS
 N obj
 E r1
 E r2
 E r3
COLUMNS
 x obj 1 r1 2
 x r2 3 r1 -2
 x r3 0 r2 1.5
 x obj 4
 y r3 1 r3 1
RHS
 rhs r1 1 obj 7
 rhs r2 2
ENDATA
""",
    "sets_first_only": """NAME SETS
ROWS
 N obj
 L r1
 G r2
COLUMNS
 x obj 1 r1 1
 x r2 1
RHS
 SET1 r1 5 r2 1
 SET2 r1 9
 SET1 r2 2
RANGES
 RA r1 3
 RB r2 4
BOUNDS
 UP B1 x 7
 UP B2 x 8
 LO B1 x 1
ENDATA
""",
    "rhs_even_tokens": """NAME EVEN
ROWS
 N obj
 L r1
 L r2
COLUMNS
 x obj 1 r1 1
 x r2 1
RHS
 r1 3 r2 4
ENDATA
""",
    "objsense_forms": """NAME SENSE
OBJSENSE MAX
ROWS
 N obj
 L r
COLUMNS
 x obj 2 r 1
RHS
 rhs r 4
ENDATA
""",
    "objsense_data_line": """NAME SENSE2
OBJSENSE
    MAXIMIZE
ROWS
 N obj
 L r
COLUMNS
 x obj 2 r 1
RHS
 rhs r 4
ENDATA
""",
    "free_rows_dropped": """NAME FREE
ROWS
 N obj
 N other
 G r
COLUMNS
 x obj 1 other 5
 x r 1
RHS
 rhs r 1 other 3
ENDATA
""",
    "comments_crlf_tabs": "NAME\tCRLF\r\n* a comment line\r\nROWS\r\n N\tobj\r\n\r\n L\tr\r\nCOLUMNS\r\n\tx\tobj\t1\tr\t1\r\n"
                          "RHS\r\n\trhs\tr\t2\r\nENDATA\r\n",
    "fixed_format_spaces": (
        "NAME          FIXED\n"
        "ROWS\n"
        " N  COST\n"
        " L  LIM 1\n"
        "COLUMNS\n"
        "    X ONE     COST         1.0        LIM 1        1.0\n"
        "    X TWO     COST         2.0        LIM 1        1.0\n"
        "RHS\n"
        "    RHS       LIM 1        4.0\n"
        "BOUNDS\n"
        " UP BND       X ONE        3.0\n"
        "ENDATA\n"),
    "quadratic_triangles": """NAME QT
ROWS
 N obj
 L r
COLUMNS
 x obj 1 r 1
 y obj 1 r 1
RHS
 rhs r 1
QMATRIX
 x x 2
 x y 1
 y x 1
 y y 4
ENDATA
""",
    "marker_then_continuous": """NAME MK
ROWS
 N obj
 L r
COLUMNS
 M1 'MARKER' 'INTORG'
 i obj 1 r 1
 M2 'MARKER' 'INTEND'
 x obj 1 r 1
RHS
 rhs r 3
ENDATA
""",
    "no_columns": """NAME EMPTY
ROWS
 N obj
 E r
RHS
 rhs r 0
ENDATA
""",
    "err_unknown_row": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 q 1\nENDATA\n",
    "err_bad_number": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1e r 1\nENDATA\n",
    "err_bad_number_dot": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj . r 1\nENDATA\n",
    "err_two_blocks": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1\n y obj 1\n x r 1\nENDATA\n",
    "err_missing_endata": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\n",
    "err_bad_row_type": "NAME E\nROWS\n N obj\n Q r\nCOLUMNS\n x obj 1\nENDATA\n",
    "err_bounds_unknown_column": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\nBOUNDS\n UP b y 1\nENDATA\n",
    "err_bound_type": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\nBOUNDS\n XX b x 1\nENDATA\n",
    "err_marker": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n M 'MARKER' 'FOO'\n x obj 1 r 1\nENDATA\n",
    "err_qcmatrix": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\nQCMATRIX r\n x x 1\nENDATA\n",
    "err_section": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\nFOO\nENDATA\n",
    "err_no_objective": "NAME E\nROWS\n L r\nCOLUMNS\n x r 1\nENDATA\n",
    "err_rows_twice": "NAME E\nROWS\n N obj\nROWS\n L r\nENDATA\n",
    "err_objsense": "NAME E\nOBJSENSE UP\nROWS\n N obj\nENDATA\n",
    "err_columns_arity": "NAME E\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r\nENDATA\n",
}


def main():
    os.makedirs(OUT, exist_ok=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    cases = dict(GATE)
    cases.update(EDGE)
    for name, text in cases.items():
        with open(os.path.join(OUT, name + ".mps"), "w", newline="") as f:
            f.write(text)
    for sub, pattern in (("tests/repro", "parser_"), ("examples/literature_lp", "")):
        d = os.path.join(REF, sub)
        for fn in sorted(os.listdir(d)):
            if fn.endswith(".mps") and fn.startswith(pattern):
                shutil.copy(os.path.join(d, fn), os.path.join(OUT, ("ref_" + fn)))
    for fn in sorted(os.listdir(OUT)):
        if fn.endswith(".mps"):
            out = subprocess.run([DUMP, os.path.join(OUT, fn)], capture_output=True, text=True, check=True).stdout
            with open(os.path.join(OUT, fn[:-4] + ".dump"), "w") as f:
                f.write(out)
    print("wrote", len([f for f in os.listdir(OUT) if f.endswith(".dump")]), "dumps")


if __name__ == "__main__":
    main()
