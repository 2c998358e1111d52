"""MPS models through the whole drop-in on the GPU (bo_mps_read ->
bo_first_order_solve): the reference's conformance set, example1 and the
refinery demo.

* PDLP engine, run as the benchmark runs it: rel-KKT 1e-6 plus the
  objective-gap stop at 5e-7.  rel-KKT 1e-6 alone leaves objective errors of
  1-2e-6 on adlittle, blend, p0033 and share2b (its gap term is relative to
  1 + |c'x| + |dual objective|); the gap stop at 1e-6 still leaves 1.26e-6 on
  sc50a in the FP32 batch, where x's infeasibility (within the KKT tolerance)
  moves c'x past the optimum by more than the gap; at 5e-7 the worst is
  5.1e-7, for 0.5% more iterations.  Every model HiGHS solves reaches OPTIMAL
  with its objective within 1e-6 (relative) of the LP optimum in
  tests/golden/mps_golden.json
  (pinned by published figures in tests/test_mps.py); models HiGHS finds
  infeasible or unbounded never come back OPTIMAL.
* Reference engine: bit-equal to the reference's own FirstOrderLPSolver on
  six of these models through the golden KATs (tests/test_gpu_reference_path.py).
* Batched engine: the models as one batch, the badly scaled ones finishing
  in its FP64 polish.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MPS_DIR = os.path.join(HERE, "golden", "mps")
GOLD = json.load(open(os.path.join(HERE, "golden", "mps_golden.json")))
FREE = sorted(k for k in GOLD if not k.endswith(".fixed.mps"))  # the fixed-format twins read identically


def read(pkg, fname):
    info, fm = pkg.read_mps(text=open(os.path.join(MPS_DIR, fname)).read())
    assert info["status"] == 0
    return fm


@pytest.mark.parametrize("fname", FREE)
def test_mps_pdlp(pkg, fname):
    g = GOLD[fname]
    fm = read(pkg, fname)
    print("case:", fname, g["highs_status"], g.get("highs_objective"), flush=True)
    if g["highs_status"] == "optimal":
        r = pkg.first_order_solve(fm, engine=1, max_iterations=200000, eps_rel_kkt=1e-6, eps_objective=5e-7)
        f = g["highs_objective"]
        print("status", r["status"], "iterations", r["iterations"], "objective", r["objective_value"], flush=True)
        assert r["status"] == 0, r["message"]
        err = abs(r["objective_value"] - f) / (1 + abs(f))
        print("objective relative error %.2e" % err, flush=True)
        assert err <= 1e-6
        assert len(r["x"]) == fm.num_vars
    else:
        r = pkg.first_order_solve(fm, engine=1, max_iterations=20000, eps_rel_kkt=1e-6)
        print("status", r["status"], r["message"], flush=True)
        # a model that presolve empties of rows takes the reference's
        # bound-only closed form (src/first_order_solver.cpp:42-66), which
        # reports OPTIMAL whatever the bounds; the iteration never does
        assert r["status"] != 0 or "bound optimization" in r["message"]


def test_mps_models_as_one_batch(pkg, po):
    # The LP-solvable models as one block-diagonal batch (bo_pdlp_batch_*):
    # the badly scaled Netlib models stall with FP32 iterates (within a
    # 4,096-iteration stall window) and finish in the FP64 polish; every LP
    # must reach rel-KKT 1e-6 (by the oracle's evaluator) and its LP optimum.
    # (With the default 16,384 window blend.free.mps instead meets both
    # stopping tests in FP32, 1.4e-6 off its optimum: on badly scaled models
    # FP32 iterates that pass rel-KKT 1e-6 and the gap test can still miss
    # the objective by more than 1e-6.)
    names = [f for f in FREE if GOLD[f]["highs_status"] == "optimal"]
    fms = [read(pkg, f) for f in names]
    lps = [pkg.lpdata.flat_to_std(fm) for fm in fms]
    b = pkg.PdlpBatch(lps, eps_rel_kkt=1e-6, eps_objective=5e-7, max_iterations=200000, stall_iterations=4096)
    per, tot = b.solve()
    print("batch", tot, flush=True)
    assert tot["solved"] == len(lps) and tot["polished"] >= 1
    for k, f in enumerate(names):
        x, y = b.solution(k)
        kk = po.kkt(lps[k], x, y)
        sm = -1.0 if fms[k].sense == 1 else 1.0
        obj = sm * kk["primal_obj"] + fms[k].obj_offset
        fs = GOLD[f]["highs_objective"]
        assert per[k]["status"] == 0 and kk["rel_kkt"] <= 1e-6, (f, per[k], kk["rel_kkt"])
        print(f, "objective relative error %.2e" % (abs(obj - fs) / (1 + abs(fs))), flush=True)
        assert abs(obj - fs) <= 1e-6 * (1 + abs(fs)), (f, obj, fs)
    b.close()
