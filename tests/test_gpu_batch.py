"""GPU parity of the batched PDLP engine (bo_pdlp_batch_*): many LPs of uneven
size laid out block-diagonally and solved together.

Every LP of a batch is checked on its own, by the oracle, as in
tests/test_gpu_pdlp.py: FP64 rel-KKT of the returned point <= 1e-6 on the
unscaled problem, objective within 1e-6 of the exact optima recorded in
tests/golden/lp_optima.json (reference simplex chain and HiGHS), and the
engine's own FP64 report agreeing with the oracle's.
"""
import json
import os
import sys
import time

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


def check_lp(po, lp, x, y, s, f_star=None, name=""):
    k = po.kkt(lp, x, y)
    assert s["status"] == 0, (name, s)
    assert k["rel_kkt"] <= TOL, (name, k)
    assert k["max_bound_viol"] == 0.0
    assert abs(k["rel_kkt"] - s["rel_kkt"]) <= 1e-9 + 1e-3 * k["rel_kkt"], (name, k["rel_kkt"], s["rel_kkt"])
    if f_star is not None:
        assert abs(k["primal_obj"] - f_star) / (1 + abs(f_star)) <= TOL, (name, k["primal_obj"], f_star)
    return k


@pytest.mark.parametrize("fp64", [0, 1])
def test_batch_of_golden_instances(pkg, po, instances, fp64):
    names = sorted(OPT)
    lps = [instances[nm] for nm in names]
    print("case: batch of", len(lps), "golden LPs, fp64", fp64, flush=True)
    b = pkg.PdlpBatch(lps, eps_rel_kkt=TOL, max_iterations=200000, fp64_iterates=fp64)
    per, tot = b.solve()
    print("batch", tot, flush=True)
    assert tot["count"] == len(lps) and tot["solved"] == len(lps)
    for k, nm in enumerate(names):
        x, y = b.solution(k)
        check_lp(po, lps[k], x, y, per[k], OPT[nm]["highs"], nm)
        if "reference_simplex" in OPT[nm]:
            fr = OPT[nm]["reference_simplex"]
            assert abs(po.kkt(lps[k], x, y)["primal_obj"] - fr) / (1 + abs(fr)) <= TOL


def test_batch_ragged_random(pkg, po, lpdata):
    # uneven sizes (a single-row LP among them).  Every LP must be solved --
    # LPs whose FP32 iterates stall finish in the FP64 polish -- pass the
    # oracle, and agree with the single-LP engine's objective to 1e-6.
    rng = np.random.default_rng(3)
    sizes = [1, 7] + [int(v) for v in rng.integers(20, 900, size=22)]
    lps = [lpdata.random_lp_sampled(mi, 3 * mi, min(5, mi), 100 + i) for i, mi in enumerate(sizes)]
    print("case: ragged batch", sizes, flush=True)
    b = pkg.PdlpBatch(lps, eps_rel_kkt=TOL, max_iterations=200000)
    per, tot = b.solve()
    print("batch", tot, flush=True)
    print("per LP (status, iterations, fp64 polish)",
          [(p["status"], p["iterations"], p["fp64_polish_iterations"]) for p in per], flush=True)
    assert tot["count"] == len(lps) and tot["solved"] == len(lps)
    for k, lp in enumerate(lps):
        x, y = b.solution(k)
        kb = check_lp(po, lp, x, y, per[k], None, "lp%d" % k)
        h = pkg.Pdlp(lp, eps_rel_kkt=TOL, max_iterations=200000, fp64_iterates=1)
        s = h.solve()
        assert s["status"] == 0, ("single-LP engine", k, s)
        xs, ys = h.solution()
        ks = po.kkt(lp, xs, ys)
        assert ks["rel_kkt"] <= TOL
        assert abs(kb["primal_obj"] - ks["primal_obj"]) / (1 + abs(ks["primal_obj"])) <= TOL, k
        h.close()


def test_batch_time_limit(pkg, po, lpdata):
    # A wall-clock limit far below what the batch needs: every LP ends solved
    # (0) or stopped by the limit (3) -- never cut into a tail round or
    # polished after it -- and the call returns within the limit plus one
    # check interval and the last round's set-up.
    lps = [lpdata.random_lp_sampled(mi, 3 * mi, 5, 500 + mi) for mi in
           (40, 900, 2000, 3500, 60, 1200, 2500, 300, 800, 3000, 1500, 700, 2200, 90, 1800, 2700)]
    limit = 0.05
    print("case: batch time limit", limit, flush=True)
    b = pkg.PdlpBatch(lps, eps_rel_kkt=1e-9, max_iterations=10**7, time_limit_s=limit)
    t0 = time.perf_counter()
    per, tot = b.solve()
    wall = time.perf_counter() - t0
    print("batch", tot, "wall %.3f s" % wall, [(p["status"], p["iterations"]) for p in per], flush=True)
    assert all(p["status"] in (0, 3) for p in per)
    assert any(p["status"] == 3 for p in per)
    assert wall <= limit + 1.0
    for k, p in enumerate(per):
        assert p["iterations"] < 10**7
        x, y = b.solution(k)
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(y))
    b.close()


def test_batch_deterministic_and_warm_start(pkg, po, lpdata):
    lps = [lpdata.random_lp_sampled(mi, 3 * mi, 5, 7 + mi) for mi in (150, 400, 90, 260)]
    b = pkg.PdlpBatch(lps, eps_rel_kkt=TOL)
    per1, tot1 = b.solve()
    sol1 = [b.solution(k) for k in range(len(lps))]
    per2, tot2 = b.solve()
    sol2 = [b.solution(k) for k in range(len(lps))]
    assert tot1["solved"] == len(lps)
    for k in range(len(lps)):
        assert per1[k]["iterations"] == per2[k]["iterations"]
        assert np.array_equal(sol1[k][0], sol2[k][0]) and np.array_equal(sol1[k][1], sol2[k][1])
    # warm start at the optimum: every LP stops at its first check
    b.set_warm_start([s[0] for s in sol1], [s[1] for s in sol1])
    per3, tot3 = b.solve()
    assert tot3["solved"] == len(lps) and tot3["max_iterations"] == 0, tot3
    for k in range(len(lps)):
        x, y = b.solution(k)
        assert po.kkt(lps[k], x, y)["rel_kkt"] <= TOL
    b.set_warm_start(None, None)
    per4, _ = b.solve()
    assert [p["iterations"] for p in per4] == [p["iterations"] for p in per1]


def test_batch_rejects_bad_lp(pkg, lpdata):
    good = lpdata.random_lp(20, 50, 3, 1)
    bad = lpdata.StdLP(good.m, good.n, good.row_ptr, good.col_idx.copy(), good.val, good.b, good.c, good.lb,
                       good.ub, good.is_eq)
    bad.col_idx[2] = good.n + 1
    with pytest.raises(pkg.BoError) as e:
        pkg.PdlpBatch([good, bad])
    assert e.value.status == 4


def test_batch_fp64_polish_of_stalled_lps(pkg, po):
    # Draws of a staircase family whose FP32 iterates stop short of 1e-6
    # (their rel-KKT stops halving within 4096 iterations, the window these
    # draws are known to stall in): with the polish off they end STALLED
    # (status 4); with it on, the FP64 engine finishes them from their FP32
    # points within the same iteration budget, and the oracle accepts them.
    fam = pkg.lpdata.StaircaseFamily(periods=12, blk_rows=100, blk_cols=250, inv_per_link=10, seed=7)
    lps = [fam.draw(k) for k in (31, 143, 185, 1, 2)]
    print("case: fp64 polish, staircase family draws 31 143 185 1 2", flush=True)
    off = pkg.PdlpBatch(lps, eps_rel_kkt=TOL, max_iterations=40000, fp64_polish=0, stall_iterations=4096)
    per0, tot0 = off.solve()
    print("polish off", [(p["status"], p["iterations"]) for p in per0], flush=True)
    assert any(p["status"] == 4 for p in per0)
    assert all(p["status"] in (0, 1, 4) for p in per0)
    off.close()
    on = pkg.PdlpBatch(lps, eps_rel_kkt=TOL, max_iterations=40000, stall_iterations=4096)
    per, tot = on.solve()
    print("polish on", [(p["status"], p["iterations"], p["fp64_polish_iterations"]) for p in per], flush=True)
    assert tot["solved"] == len(lps)
    for k, lp in enumerate(lps):
        assert per[k]["iterations"] <= 40000
        assert (per[k]["fp64_polish_iterations"] > 0) == (per0[k]["status"] == 4)
        if per0[k]["status"] == 1:  # the cap, not a stall: the same run with the polish on
            assert per[k]["iterations"] == 40000
        if per0[k]["status"] == 0:
            assert per[k]["iterations"] == per0[k]["iterations"]
        x, y = on.solution(k)
        check_lp(po, lp, x, y, per[k], None, "draw%d" % k)
    on.close()
