"""The MPS reader (bo_mps_read: MpsParser of include/bharatopt/mps_parser.hpp),
host only, no GPU.

* Every file under tests/golden/mps/ (the reference's conformance set, its
  example1 and refinery demo) reads to the same model as the reference's own
  MpsParser: sizes and a SHA-256 of the flattened arrays recorded from the
  reference build in tests/golden/mps_golden.json, and, where oracle/_ref is
  built, array by array against it live.
* The recorded LP optima are pinned by published figures: afiro -464.75314286
  (docs/validation/pre_phase_26/reference-results.csv:2), the refinery demo
  3870 (docs/validation/refinery_demo_final_validation.md:45), example1 36
  (tests/test_mps_example1_correctness.cpp:54) and the Netlib optima of
  adlittle, blend, kb2, sc50a and share2b.
* Hand-written texts exercise each rule and error status, compared with
  values worked out by hand and, live, with the reference's parser.
"""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MPS_DIR = os.path.join(HERE, "golden", "mps")
GOLD = json.load(open(os.path.join(HERE, "golden", "mps_golden.json")))
KEYS = ("lb", "ub", "obj", "var_type", "con_sense", "rhs", "range_upper", "term_ptr", "term_var", "term_val")


def flat_sha(fm):
    h = hashlib.sha256()
    for k in KEYS:
        dt = np.int64 if k in ("var_type", "con_sense", "term_ptr", "term_var") else np.float64
        h.update(np.ascontiguousarray(np.asarray(getattr(fm, k)), dt).tobytes())
    h.update(np.array([fm.sense], np.int64).tobytes())
    h.update(np.array([fm.obj_offset], np.float64).tobytes())
    return h.hexdigest()


def same_as_reference(fm, r):
    assert (fm.sense, fm.obj_offset) == (r["sense"], r["obj_offset"])
    for k in KEYS:
        a, b = np.asarray(getattr(fm, k)), np.asarray(r[k])
        assert a.shape == b.shape and np.array_equal(a, b), k


@pytest.mark.parametrize("fname", sorted(GOLD))
def test_mps_files_read_like_the_reference(pkg, po, fname):
    text = open(os.path.join(MPS_DIR, fname)).read()
    info, fm = pkg.read_mps(text=text)
    g = GOLD[fname]
    assert info["status"] == g["status"] == 0
    assert (info["num_vars"], info["num_cons"], int(fm.term_ptr[-1])) == (g["num_vars"], g["num_cons"], g["terms"])
    assert info["integer_vars"] + info["binary_vars"] == g["integer_vars"]
    assert flat_sha(fm) == g["model_sha"]
    r = po.ref_mps(text)
    if r is not None:
        same_as_reference(fm, r)


def test_file_reading_and_search_path(pkg, tmp_path, monkeypatch):
    info, fm = pkg.read_mps(path=os.path.join(MPS_DIR, "afiro.free.mps"))
    assert info["status"] == 0 and fm.num_vars == 32
    # a bare name is also looked for under benchmarks/corpus/ (src/mps_parser.cpp:47-74)
    (tmp_path / "benchmarks" / "corpus").mkdir(parents=True)
    (tmp_path / "benchmarks" / "corpus" / "tiny.mps").write_text(open(os.path.join(MPS_DIR, "example1.mps")).read())
    monkeypatch.chdir(tmp_path)
    info, fm = pkg.read_mps(path="tiny.mps")
    assert info["status"] == 0 and fm.num_vars == 2
    info, fm = pkg.read_mps(path="no_such_model.mps")
    assert info["status_name"] == "FILE_NOT_FOUND" and fm is None


def test_published_optima_pin_the_fixture():
    published = {"afiro": -464.75314286, "adlittle": 225494.96316, "blend": -30.812149846, "kb2": -1749.9001299,
                 "sc50a": -64.575077059, "share2b": -415.73224074}
    for name, f in published.items():
        for fmt in ("fixed", "free"):
            g = GOLD[f"{name}.{fmt}.mps"]["highs_objective"]
            assert abs(g - f) <= 1e-8 * (1 + abs(f)), (name, g, f)
    assert GOLD["bharatopt_refinery_demo.mps"]["highs_objective"] == pytest.approx(3870.0, abs=1e-6)
    assert GOLD["example1.mps"]["highs_objective"] == pytest.approx(36.0, abs=1e-9)


RULES = """NAME          RULES
OBJSENSE
    MAXIMIZE
ROWS
 N  PROFIT
 L  CAP       $ trailing comment
 G  DEMAND
 E  BAL
 L  RL
 G  RG
 E  RE
 N  EXTRA
COLUMNS
    X         PROFIT    3.0        CAP       1.0
    X         DEMAND    .5         BAL       1D0
    X         NOSUCHROW 9.0
    MARKER    'MARKER'  'INTORG'
    Y         PROFIT    2          CAP       -.25
    Y         RL        1          RG        1
    Y         PROFIT    7          EXTRA     4
    MARKER    'MARKER'  'INTEND'
    Z         RE        1          PROFIT    -1
RHS
    RHS1      CAP       10         DEMAND    1
    RHS1      BAL       2          PROFIT    5
    RHS1      RL        8          RG        3
    RHS1      RE        4
    RHS2      CAP       99
RANGES
    RNG       RL        2          RG        -3
    RNG       RE        -1
    RNG2      RL        50
BOUNDS
 UP BND       X         4
 MI BND       Y
 UP BND       Y         6
 FR BND       Z
 LO OTHER     X         -7
ENDATA
"""


def test_each_rule(pkg, po):
    info, fm = pkg.read_mps(text=RULES)
    assert info["status"] == 0 and info["problem_name"] == "RULES"
    assert fm.sense == 1 and fm.obj_offset == -5.0          # OBJSENSE; objective-row RHS
    assert fm.obj.tolist() == [3.0, 2.0, -1.0]              # first objective entry per column
    assert fm.lb.tolist() == [0.0, -np.inf, -np.inf]        # UP keeps lb; MI; FR
    assert fm.ub.tolist() == [4.0, 6.0, np.inf]             # second bound vector ignored
    assert fm.var_type.tolist() == [0, 1, 0]                # MARKER block
    # rows: CAP DEMAND BAL RL RG RE and the second N row as a <= row
    assert fm.con_sense.tolist() == [0, 2, 1, 3, 3, 3, 0]
    assert fm.rhs.tolist() == [10.0, 1.0, 2.0, 6.0, 0.0, 3.0, 0.0]   # RHS2 and RNG2 ignored
    assert fm.range_upper.tolist() == [0.0, 0.0, 0.0, 8.0, 3.0, 4.0, 0.0]
    rows = [list(zip(fm.term_var[fm.term_ptr[i]:fm.term_ptr[i + 1]].tolist(),
                     fm.term_val[fm.term_ptr[i]:fm.term_ptr[i + 1]].tolist())) for i in range(fm.num_cons)]
    assert rows == [[(0, 1.0), (1, -0.25)], [(0, 0.5)], [(0, 1.0)], [(1, 1.0)], [(1, 1.0)], [(2, 1.0)], [(1, 4.0)]]
    r = po.ref_mps(RULES)
    if r is not None:
        same_as_reference(fm, r)


BAD = {
    "MISSING_ENDATA": "NAME A\nROWS\n N obj\nCOLUMNS\n    X obj 1\n",
    "MALFORMED_ROW": "NAME A\nROWS\n Q obj\nENDATA\n",
    "MALFORMED_COLUMN": "NAME A\nROWS\n N obj\nCOLUMNS\n    X obj\nENDATA\n",
    "MALFORMED_BOUND": "NAME A\nROWS\n N obj\nCOLUMNS\n    X obj 1\nBOUNDS\n XX B X 1\nENDATA\n",
    "UNCLOSED_INTEGER_MARKER": "NAME A\nROWS\n N obj\nCOLUMNS\n    M 'MARKER' 'INTORG'\n    X obj 1\nENDATA\n",
    "INVALID_NUMERIC_VALUE": "NAME A\nROWS\n N obj\nCOLUMNS\n    X obj abc\nENDATA\n",
}


@pytest.mark.parametrize("status", sorted(BAD))
def test_error_statuses(pkg, po, status):
    info, fm = pkg.read_mps(text=BAD[status])
    assert info["status_name"] == status and fm is None and info["message"]
    r = po.ref_mps(BAD[status])
    if r is not None:
        assert r["status"] == info["status"]


def test_bad_bound_value_reads_as_zero(pkg, po):
    text = "NAME A\nROWS\n N obj\n L c\nCOLUMNS\n    X obj 1 c 1\nBOUNDS\n UP B X zz\nENDATA\n"
    info, fm = pkg.read_mps(text=text)
    assert info["status"] == 0 and fm.ub.tolist() == [0.0]
    r = po.ref_mps(text)
    if r is not None:
        same_as_reference(fm, r)
