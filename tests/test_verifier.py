"""SolutionVerifier through the C ABI (bo_verify_solution), host FP64, no
GPU: every field of its result equals the reference verifier's
(src/solution_verifier.cpp:42-215) on candidates that break each rule --
lower and upper bounds, integer and binary integrality, NaN and +-Inf
entries, LE / GE / EQ rows and both sides of a RANGED row, several rules at
once, short / long / empty candidates, a reported objective within and
beyond the objective tolerance (and with the check off), loosened
tolerances -- and on 40 seeded random models with mixed senses and types.
The expected values are tests/golden/verifier_kat.json, written by the
reference build (tests/golden/make_golden.py verifier); live against the
reference build too where it is present."""
import json
import math
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "verifier_kat.json")))
sys.path.insert(0, HERE)


@pytest.fixture(scope="module")
def cases():
    from golden import make_golden
    return make_golden.verifier_cases()


def _same(a, b):
    if isinstance(b, str):  # float.hex of the expected value
        b = float.fromhex(b)
    if isinstance(a, float) or isinstance(b, float):
        return (math.isnan(a) and math.isnan(b)) or a == b
    return a == b


@pytest.mark.parametrize("name", sorted(KAT))
def test_verifier_matches_reference_fixture(pkg, cases, name):
    fm, x, kw = cases[name]
    got = pkg.verify_solution(fm, x, **kw)
    want = KAT[name]
    assert set(got) == set(want)
    for k in want:
        assert _same(got[k], want[k]), (name, k, got[k], want[k])


def test_every_rule_is_exercised():
    statuses = {v["status"] for v in KAT.values()}
    assert statuses == {0, 1, 2, 3}
    assert any(v["violated_integrality_count"] > 0 for v in KAT.values())
    assert any(v["violated_bound_count"] > 0 and v["violated_constraint_count"] == 0 for v in KAT.values())
    assert any(float.fromhex(v["maximum_constraint_violation"]) == math.inf for v in KAT.values())


def test_verifier_matches_reference_live(pkg, ref, cases):
    for name, (fm, x, kw) in cases.items():
        a = pkg.verify_solution(fm, x, **kw)
        b = ref.ref_verify_full(fm, x, **kw)
        for k in b:
            assert _same(a[k], b[k]), (name, k, a[k], b[k])
