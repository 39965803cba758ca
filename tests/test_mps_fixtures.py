"""MPS reader parity on the reference's own parser fixtures (CPU).

The files under tests/golden/mps are the reference's datasets/linear_programming
and datasets/mixed_integer_programming MPS fixtures, copied as data.  The expected
values restate cpp/libmps_parser/tests/mps_parser_test.cpp in the free format that
cuOptReadProblem uses (cpp/src/pdlp/cuopt_c.cpp:103, input_mps_strict = false).
"""
import ctypes
import os

import numpy as np
import pytest

from cuopt_amd import c_api
from cuopt_amd.linear_programming import ParseMps

MPS = os.path.join(os.path.dirname(__file__), "golden", "mps")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
INF = np.inf


def parse(name):
    path = os.path.join(MPS, name) if os.path.exists(os.path.join(MPS, name)) else os.path.join(GOLDEN, name)
    return ParseMps(path)


def read_rc(name):
    L = c_api.lib()
    h = ctypes.c_void_p()
    rc = L.cuOptReadProblem(os.path.join(MPS, name).encode(), ctypes.byref(h))
    if h.value:
        L.cuOptDestroyProblem(ctypes.byref(h))
    return rc


def rows_of(dm):
    """The constraint matrix as {row: [(col, value), ...]} in file order."""
    off, idx, val = dm.get_constraint_matrix_offsets(), dm.get_constraint_matrix_indices(), \
        dm.get_constraint_matrix_values()
    return [list(zip(idx[off[i]:off[i + 1]].tolist(), val[off[i]:off[i + 1]].tolist()))
            for i in range(len(off) - 1)]


def check_good_1(dm, row2=((0, 2.7), (1, 10.1))):
    # mps_parser_test.cpp:59-93 (good-mps-1), :131-165 (CRLF), :249-284 (free format)
    assert dm.get_problem_name() == "good-1"
    assert list(dm.get_row_names()) == ["ROW1", "ROW2"]
    assert list(dm.get_row_types()) == ["L", "L"]
    assert dm.get_objective_name() == "COST"
    assert list(dm.get_variable_names()) == ["VAR1", "VAR2"]
    assert rows_of(dm) == [[(0, 3.0), (1, 4.0)], list(row2)]
    assert dm.get_constraint_bounds().tolist() == [5.4, 4.9]
    assert dm.get_objective_coefficients().tolist() == [0.2, 0.1]


@pytest.mark.parametrize("name", ["good-mps-1.mps", "good-mps-1-clrf.mps", "free-format-mps-1.mps"])
def test_good_mps_1(name):
    dm = parse(name)
    check_good_1(dm)
    assert not dm.get_sense()
    # mps_parser_test.cpp:337-348: no BOUNDS gives [0, inf)
    assert dm.get_variable_lower_bounds().tolist() == [0.0, 0.0]
    assert dm.get_variable_upper_bounds().tolist() == [INF, INF]


def test_comments():
    # mps_parser_test.cpp:167-199: '*' and '$' comment lines, and a '$' field
    # comments out the rest of its line ("VAR1 $ROW1 2.5 ROW2 2.6" adds nothing)
    check_good_1(parse("good-mps-1-comments.mps"), row2=((0, 2.7),))


@pytest.mark.parametrize("name", ["good-mps-fixed-no-name.mps", "good-mps-fixed-empty-name.mps"])
def test_missing_or_empty_name_parses(name):
    # mps_parser_test.cpp:201-211
    assert read_rc(name) == c_api.CUOPT_SUCCESS


def test_spaces_in_names_rejected_in_free_format():
    # mps_parser_test.cpp:286-289
    assert read_rc("good-fixed-mps-2.mps") == c_api.CUOPT_MPS_PARSE_ERROR


def test_up_low_bounds():
    # mps_parser_test.cpp:303-335
    dm = parse("lp_model_with_var_bounds.mps")
    assert dm.get_problem_name() == "lp_model_with_var_bounds"
    assert list(dm.get_row_names()) == ["con"]
    assert list(dm.get_row_types()) == ["L"]
    assert dm.get_objective_name() == "OBJ"
    assert list(dm.get_variable_names()) == ["x", "y"]
    assert rows_of(dm) == [[(0, 1.0), (1, 1.0)]]
    assert dm.get_constraint_bounds().tolist() == [3.0]
    assert dm.get_objective_coefficients().tolist() == [2.0, -1.0]
    assert dm.get_variable_lower_bounds().tolist() == [0.0, 1.0]
    assert dm.get_variable_upper_bounds().tolist() == [1.0, 2.0]


@pytest.mark.parametrize("name,lb,ub", [
    ("good-mps-some-var-bounds.mps", [-1.0, 0.0], [INF, 2.0]),       # :350-361
    ("good-mps-fixed-var.mps", [2.0, 0.0], [2.0, INF]),              # :363-374
    ("good-mps-free-var.mps", [-INF, 0.0], [INF, INF]),              # :376-387
    ("good-mps-lower-bound-inf-var.mps", [-INF, 0.0], [INF, INF]),   # :389-400
    ("good-mps-upper-bound-inf-var.mps", [0.0, 0.0], [INF, INF]),    # :410-421
])
def test_variable_bounds(name, lb, ub):
    dm = parse(name)
    assert dm.get_variable_lower_bounds().tolist() == lb
    assert dm.get_variable_upper_bounds().tolist() == ub


def test_rhs_on_objective_is_negated_offset():
    # mps_parser_test.cpp:402-408
    assert parse("good-mps-rhs-cost.mps").get_objective_offset() == -5.0


@pytest.mark.parametrize("name", ["good-mps-fixed-ranges.mps", "good-mps-free-ranges.mps"])
def test_ranges(name):
    # mps_parser_test.cpp:423-503: E rows take the range's sign, L and G rows widen
    # away from b, and E rows with R = 0 stay equalities
    dm = parse(name)
    lo = [1.2, 1.5, 7.9, 3.5, 3.9, 4.9]
    hi = [5.4, 4.9, 9.5, 6.9, 3.9, 4.9]
    np.testing.assert_allclose(dm.get_constraint_lower_bounds(), lo, atol=1e-6)
    np.testing.assert_allclose(dm.get_constraint_upper_bounds(), hi, atol=1e-6)


@pytest.mark.parametrize("name,objective", [
    ("good-mps-fixed-two-objectives.mps", "COST"),                       # :505-512 first N row
    ("good-mps-fixed-two-objectives-objname.mps", "COST6679327"),        # :514-521 OBJNAME
    ("good-mps-fixed-two-objectives-objname-next-line.mps", "COST6679327"),  # :523-530
])
def test_objective_row_choice(name, objective):
    dm = parse(name)
    assert dm.get_objective_name() == objective
    assert list(dm.get_row_names()) == ["ROW1", "ROW2"]  # further N rows are not constraints


def test_objname_row_keeps_only_its_offset(tmp_path):
    # mps_parser.cpp:851-856, 995, 1095: once OBJNAME has named the objective, every N
    # row of the ROWS section is filed as ignored, the named one included, so COLUMNS
    # entries of either N row are dropped (a zero objective); an RHS entry on the named
    # row is still read as the negated offset, and one on the other N row is an unknown
    # row name.  Kept as the reference parses it, for identical problems.
    src = open(os.path.join(MPS, "good-mps-fixed-two-objectives-objname.mps")).read()
    src = src.replace("    VAR2      COST      0.1", "    VAR2      COST      0.1\n    VAR2      COST6679327  7.5")
    f = tmp_path / "objname.mps"
    f.write_text(src.replace("    RHS1      ROW1      5.4", "    RHS1      COST6679327  2.5\n    RHS1      ROW1      5.4"))
    dm = ParseMps(str(f))
    assert dm.get_objective_name() == "COST6679327"
    assert dm.get_objective_coefficients().tolist() == [0.0, 0.0]
    assert dm.get_objective_offset() == -2.5
    g = tmp_path / "other.mps"
    g.write_text(src.replace("    RHS1      ROW1      5.4", "    RHS1      COST      2.5\n    RHS1      ROW1      5.4"))
    with pytest.raises(Exception):
        ParseMps(str(g))


BAD = ([f"bad-mps-{i}.mps" for i in (1, 2, 3, 4, 5, 6, 7, 9, 10, 11, 12, 13, 14, 15)]
       + ["bad-mps-bound-1.mps", "bad-mps-bound-2.mps",           # :556-566 bound types
          "bad-mps-fixed-objname-after-rows.mps",                 # :532-536 OBJNAME after ROWS
          "bad-mps-fixed-ranges-name.mps", "bad-mps-fixed-ranges-value.mps"])  # :544-554


@pytest.mark.parametrize("name", BAD)
def test_bad_files_are_parse_errors(name):
    # mps_parser_test.cpp:46-57 and :291-301 (bad-mps-8 does not exist): among them
    # an objective-less file (12), non-contiguous columns (13), an empty file (15)
    # and a file with only ROWS (2)
    assert read_rc(name) == c_api.CUOPT_MPS_PARSE_ERROR


@pytest.mark.parametrize("name,types,lb,ub", [
    ("good-mip-mps-1.mps", ["I", "I"], [0.0, 0.0], [10.0, 10.0]),             # :568-611
    ("good-mip-mps-1-no-mark.mps", ["I", "I"], [0.0, 0.0], [10.0, 10.0]),     # :613-656 UI/LO bounds
    ("good-mip-mps-no-bounds.mps", ["I", "C"], [0.0, 0.0], [1.0, INF]),       # :658-702 integer => [0, 1]
    ("good-mip-mps-partial-bounds.mps", ["I", "C"], [0.0, 0.0], [1.0, 10.0]),  # :704-749
])
def test_integer_columns(name, types, lb, ub):
    dm = parse(name)
    assert list(dm.get_variable_types()) == types
    assert dm.get_variable_lower_bounds().tolist() == lb
    assert dm.get_variable_upper_bounds().tolist() == ub
    assert dm.get_sense()  # OBJSENSE MAX


def test_empty_matrix():
    # one column, no constraints, MI and UP bounds
    dm = parse("empty_matrix.mps")
    assert len(dm.get_constraint_matrix_offsets()) == 1
    assert dm.get_objective_coefficients().tolist() == [-3.0]
    assert dm.get_variable_lower_bounds().tolist() == [-INF]
    assert dm.get_variable_upper_bounds().tolist() == [2.0]


def _counts(L, h):
    m, n = ctypes.c_int32(), ctypes.c_int32()
    assert L.cuOptGetNumConstraints(h, ctypes.byref(m)) == c_api.CUOPT_SUCCESS
    assert L.cuOptGetNumVariables(h, ctypes.byref(n)) == c_api.CUOPT_SUCCESS
    return m.value, n.value


def test_mutated_files_parse_or_fail_cleanly(tmp_path):
    # Seeded mutations of the fixtures (a changed byte, a truncated file, a cleared
    # or repeated line) either parse or give MPS_PARSE_ERROR; a file that parses
    # writes out and reads back with the same dimensions.
    rng = np.random.default_rng(7)
    L = c_api.lib()
    src_files = sorted(f for f in os.listdir(MPS) if f.endswith(".mps"))
    path, out = tmp_path / "fuzz.mps", tmp_path / "out.mps"
    parsed = 0
    for name in src_files:
        src = open(os.path.join(MPS, name), "rb").read()
        for _ in range(25):
            s = bytearray(src)
            if not s:
                break
            pos = int(rng.integers(len(s)))
            kind = int(rng.integers(4))
            end = s.find(b"\n", pos)
            if kind == 0:
                s[pos] = b" \n$*-+.eE0123456789ABCXYZ"[int(rng.integers(25))]
            elif kind == 1:
                del s[pos:]
            elif kind == 2 and end >= 0:
                del s[pos:end]
            elif end >= 0:
                s[pos:pos] = s[pos:end + 1]
            path.write_bytes(bytes(s))
            h = ctypes.c_void_p()
            rc = L.cuOptReadProblem(str(path).encode(), ctypes.byref(h))
            assert rc in (c_api.CUOPT_SUCCESS, c_api.CUOPT_MPS_PARSE_ERROR), (name, bytes(s))
            if rc != c_api.CUOPT_SUCCESS:
                assert h.value is None
                continue
            parsed += 1
            try:
                assert L.cuOptWriteProblem(h, str(out).encode(), 0) == c_api.CUOPT_SUCCESS
                h2 = ctypes.c_void_p()
                assert L.cuOptReadProblem(str(out).encode(), ctypes.byref(h2)) == c_api.CUOPT_SUCCESS
                assert _counts(L, h2) == _counts(L, h), (name, bytes(s))
                L.cuOptDestroyProblem(ctypes.byref(h2))
            finally:
                L.cuOptDestroyProblem(ctypes.byref(h))
    assert parsed > 100
