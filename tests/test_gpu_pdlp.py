"""GPU parity of the PDLP engine (FP32 iterates, FP64 checks) through the C ABI.

The engine's answer is judged by the oracle, never by itself:
* the oracle's FP64 KKT evaluator (tests/test_oracle.py pins it) recomputes
  rel-KKT of the returned (x, y) on the unscaled problem: must be <= 1e-6,
  the tolerance of the benchmark metric;
* the objective must be within 1e-6 (relative, |f - f*| / (1 + |f*|)) of two
  exact FP64 solvers recorded in tests/golden/lp_optima.json: the
  reference's simplex chain and HiGHS;
* the engine's own FP64 KKT report must agree with the oracle's;
* warm starts at the optimum stop at the first check; the FP64-iterate
  mode reaches the same optimum.
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
OPT = json.load(open(os.path.join(HERE, "golden", "lp_optima.json")))
sys.path.insert(0, HERE)
TOL = 1e-6


@pytest.fixture(scope="module")
def instances():
    from golden import make_golden
    return make_golden.pdlp_instances()


def check_solution(po, lp, x, y, stats, name):
    k = po.kkt(lp, x, y)
    print(name, "iters", stats["iterations"], "restarts", stats["restarts"], "engine rel-KKT %.3e" % stats["rel_kkt"],
          "oracle rel-KKT %.3e" % k["rel_kkt"], flush=True)
    assert stats["status"] == 0
    assert k["rel_kkt"] <= TOL, k
    assert k["max_bound_viol"] == 0.0
    # the engine's FP64 report describes the returned point
    assert abs(k["rel_kkt"] - stats["rel_kkt"]) <= 1e-9 + 1e-3 * k["rel_kkt"]
    f_star = OPT[name]["highs"]
    assert abs(k["primal_obj"] - f_star) / (1 + abs(f_star)) <= TOL
    if "reference_simplex" in OPT[name]:
        fr = OPT[name]["reference_simplex"]
        assert abs(k["primal_obj"] - fr) / (1 + abs(fr)) <= TOL
    return k


@pytest.mark.parametrize("name", sorted(OPT))
def test_pdlp_reaches_tolerance_and_optimum(pkg, po, instances, name):
    lp = instances[name]
    print("case:", name, lp.m, lp.n, lp.nnz, flush=True)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, max_iterations=200000)
    s = h.solve()
    x, y = h.solution()
    check_solution(po, lp, x, y, s, name)
    h.close()


def test_pdlp_fp64_iterates(pkg, po, instances):
    name = "random_300x1000_s42"
    lp = instances[name]
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, fp64_iterates=1)
    s = h.solve()
    x, y = h.solution()
    check_solution(po, lp, x, y, s, name)


def test_pdlp_warm_start_at_optimum(pkg, po, instances):
    lp = instances["random_300x1000_s43"]
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL)
    s = h.solve()
    x, y = h.solution()
    assert s["status"] == 0
    h.set_warm_start(x, y)
    s2 = h.solve()
    # the starting point is checked before any step: it is already optimal
    assert s2["status"] == 0 and s2["iterations"] == 0, s2
    assert abs(s2["rel_kkt"] - s["rel_kkt"]) <= 1e-9
    x2, y2 = h.solution()
    assert po.kkt(lp, x2, y2)["rel_kkt"] <= 1e-6


def test_pdlp_no_graph_same_answer(pkg, instances):
    lp = instances["random_60x200_s7"]
    a = pkg.Pdlp(lp, use_graph=1)
    sa = a.solve()
    b = pkg.Pdlp(lp, use_graph=0, lanes_rows=sa["lanes_rows"], lanes_cols=sa["lanes_cols"])
    sb = b.solve()
    assert sa["iterations"] == sb["iterations"]
    xa, ya = a.solution()
    xb, yb = b.solution()
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb)  # deterministic reductions


def test_pdlp_deterministic_repeat(pkg, instances):
    lp = instances["zipf_2000x5000_s9"]
    h = pkg.Pdlp(lp)
    s1 = h.solve()
    x1, y1 = h.solution()
    s2 = h.solve()
    x2, y2 = h.solution()
    assert s1["iterations"] == s2["iterations"]
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)


def test_pdlp_max_iterations_status(pkg, instances):
    lp = instances["random_300x1000_s44"]
    h = pkg.Pdlp(lp, max_iterations=128)
    s = h.solve()
    assert s["status"] == 1 and s["iterations"] == 128


def test_dropin_pdlp_engine(pkg, po, lpdata):
    # the whole drop-in with engine PDLP on the hand LP (max 3x + 5y)
    mb = lpdata.ModelBuilder(sense=1)
    x = mb.add_variable(0.0, np.inf, 3.0)
    y = mb.add_variable(0.0, np.inf, 5.0)
    mb.add_constraint([(x, 2.0), (y, 1.0)], lpdata.LE, 8.0)
    mb.add_constraint([(x, 1.0), (y, 2.0)], lpdata.LE, 8.0)
    r = pkg.first_order_solve(mb.build(), engine=1)
    assert r["status"] == 0
    assert abs(r["objective_value"] - 64 / 3) < 1e-5
    np.testing.assert_allclose(r["y"], [1 / 3, 7 / 3], atol=1e-5)


def test_pdlp_rejects_bad_csr(pkg, lpdata):
    lp = lpdata.random_lp(20, 50, 3, 1)
    bad = lpdata.StdLP(lp.m, lp.n, lp.row_ptr, lp.col_idx.copy(), lp.val, lp.b, lp.c, lp.lb, lp.ub, lp.is_eq)
    bad.col_idx[3] = lp.n + 5  # column out of range: rejected on the host, no launch
    with pytest.raises(pkg.BoError) as e:
        pkg.Pdlp(bad)
    assert e.value.status == 4
