"""MPS input contract: the library's reader against the reference reader's output on the same files.

Fixtures in tests/golden/mps: each <case>.dump is the reference reader's (src/mps.cpp) model for
<case>.mps, written by tests/golden/make_mps_golden.py. Compared field by field: row/column names and
bounds, costs, entries (values to the bit, %.17g), objective constant, sense, sentinel count, integer
flags, number of quadratic terms, the input-gate verdict (gpu/pdhg_input.hpp:7-17), and parse errors
(message text included).
"""
import glob
import math
import os

import pytest

from taral_pdhg import ffi

MPS_DIR = os.path.join(os.path.dirname(__file__), "golden", "mps")
CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(MPS_DIR, "*.mps")))


def _num(v):
    if math.isinf(v):
        return " inf" if v > 0 else " -inf"
    return " %.17g" % v


def dump_with_this_build(path):
    """The reference dump format (oracle/ref_mps_dump.cpp) from the library's reader, minus the NAME line."""
    try:
        md = ffi.Model.from_mps(path)
    except ffi.TpdhgError as e:
        assert e.code == ffi.ERR_PARSE
        return ["parse_error " + str(e).split("] ", 1)[1]]
    info = md.info()
    a = md.arrays()
    gate = md.input_error()
    lines = [f"gate {gate if gate else 'ok'}", f"maximize {info.maximize}", "obj_const" + _num(info.obj_const),
             f"sentinels {info.sentinel_bounds}", f"rows {info.m}"]
    for i in range(info.m):
        lines.append(f"r {md.name_of(True, i)}" + _num(a['row_lo'][i]) + _num(a['row_up'][i]))
    lines.append(f"cols {info.n}")
    for j in range(info.n):
        s = f"c {md.name_of(False, j)} {int(a['is_int'][j])}" + _num(a['cost'][j]) + _num(a['col_lo'][j]) + \
            _num(a['col_up'][j])
        for k in range(a['colptr'][j], a['colptr'][j + 1]):
            s += f" {a['rowidx'][k]}" + _num(a['val'][k])
        lines.append(s)
    lines.append(f"q {info.qobj_terms}")
    return lines


def reference_dump(case):
    with open(os.path.join(MPS_DIR, case + ".dump")) as f:
        lines = f.read().splitlines()
    # NAME is not part of the library's model API; quadratic terms are compared by count
    return [ln for ln in lines if not ln.startswith("name ") and not (ln.startswith("q ") and ln.count(" ") > 1)]


@pytest.mark.parametrize("case", CASES)
def test_reader_matches_reference(case):
    assert dump_with_this_build(os.path.join(MPS_DIR, case + ".mps")) == reference_dump(case)


def test_fixture_set_is_complete():
    # 13 input-gate models, the reference's parser rejection files, literature LPs and edge cases
    assert sum(c.startswith("gate_") for c in CASES) == 13
    assert sum(c.startswith("ref_parser_") for c in CASES) >= 8
    assert "ref_steel_production" in CASES and "ref_transportation" in CASES
    assert sum(c.startswith("err_") for c in CASES) >= 10


def test_generated_mps_round_trip(packing_10000, transport_100x100, tmp_path):
    """gpu/gen_lp.py-format text read back equals the generator's model bit for bit."""
    import numpy as np
    for md, text in (packing_10000, transport_100x100):
        p = tmp_path / "m.mps"
        p.write_text(text)
        a = ffi.Model.from_mps(str(p)).arrays()
        for k in ("colptr", "rowidx", "val", "cost", "col_lo", "col_up", "row_lo", "row_up"):
            assert np.array_equal(a[k], getattr(md, k)), k


def test_in_memory_text_equals_file(tmp_path):
    text = open(os.path.join(MPS_DIR, "ranges_all_kinds.mps")).read()
    import numpy as np
    a = ffi.Model.from_mps_text(text).arrays()
    b = ffi.Model.from_mps(os.path.join(MPS_DIR, "ranges_all_kinds.mps")).arrays()
    for k in a:
        assert np.array_equal(a[k], b[k])
