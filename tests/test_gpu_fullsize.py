"""The benchmark workloads at their full sizes on the GPU, each LP checked
against an FP64 oracle (tests/golden/objectives.json, written in the build
container: HiGHS where it is tractable, the FP64 restatement
oracle/pdlp_omp.c to rel-KKT 1e-9 beyond):

* C0  the reference's CPU-sized random LP (300 x 1,000), single-LP engine;
* C1  the 200,000 x 500,000 Zipf LP (17.1M nonzeros), single-LP engine;
* C2  the 24-period staircase LP, single-LP engine;
* C3  8 LPs of the 512-LP ragged batch (the largest among them), batched engine;
* C4  16 held-out draws of the 12-period staircase family, family engine.

Each returned point must reach rel-KKT <= 1e-6 by the oracle's FP64
evaluator and an objective within 1e-6 of the recorded optimum
(|f - f*| / (1 + |f*|)), the rule LPs/s counts by.  The instance's SHA-256
is checked first, so a changed generator fails loudly instead of comparing
against another LP's optimum."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "objectives.json")))
TOL = 1e-6


def instance_sha(lp):
    h = hashlib.sha256()
    for a, dt in ((lp.row_ptr, np.int32), (lp.col_idx, np.int32), (lp.val, np.float64), (lp.b, np.float64),
                  (lp.c, np.float64), (lp.lb, np.float64), (lp.ub, np.float64), (lp.is_eq, np.uint8)):
        h.update(np.ascontiguousarray(a, dt).tobytes())
    return h.hexdigest()


def check(po, lp, x, y, entry, name):
    assert instance_sha(lp) == entry["instance_sha"], name + ": generator output changed"
    k = po.kkt(lp, x, y)
    fstar = entry["objective"]
    err = abs(k["primal_obj"] - fstar) / (1 + abs(fstar))
    print(f"{name}: rel-KKT {k['rel_kkt']:.3e}, objective {k['primal_obj']:.10g} vs {fstar:.10g}, "
          f"relative error {err:.2e}", flush=True)
    assert k["rel_kkt"] <= TOL, (name, k["rel_kkt"])
    assert k["max_bound_viol"] == 0.0, name
    assert err <= TOL, (name, err)


SINGLE = {
    "c0": lambda lpd: lpd.random_lp(300, 1000, 5, 42),
    "c1": lambda lpd: lpd.zipf_lp(200_000, 500_000, 1.5, 2000, seed=42),
    "c2": lambda lpd: lpd.staircase_lp(24, 2000, 5000, 200, seed=42),
}


@pytest.mark.parametrize("workload", sorted(SINGLE))
def test_single_lp_workloads_full_size(pkg, po, lpdata, workload):
    if workload not in FX:
        pytest.skip(f"no objective fixture for {workload}")
    lp = SINGLE[workload](lpdata)
    print("case:", workload, lp.m, lp.n, lp.nnz, flush=True)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, max_iterations=200000)
    s = h.solve()
    x, y = h.solution()
    h.close()
    print(workload, "status", s["status"], "iterations", s["iterations"], "engine rel-KKT %.3e" % s["rel_kkt"],
          flush=True)
    assert s["status"] == 0
    check(po, lp, x, y, FX[workload], workload)


def test_batched_workload_full_size_lps(pkg, po, lpdata):
    fx = FX["c3"]["lps"]
    sizes = lpdata.batch_sizes(512, 200, 20000, 42)
    ids = sorted(int(k) for k in fx)
    lps = [lpdata.random_lp_sampled(int(sizes[i]), 3 * int(sizes[i]), 5, 42 + 1 + i) for i in ids]
    print("case: c3 LPs", ids, [lp.m for lp in lps], flush=True)
    b = pkg.PdlpBatch(lps, eps_rel_kkt=TOL, max_iterations=200000)
    per, tot = b.solve()
    print("batch", tot, flush=True)
    assert tot["solved"] == len(lps)
    for k, i in enumerate(ids):
        x, y = b.solution(k)
        check(po, lps[k], x, y, fx[str(i)], f"c3 LP {i}")
    b.close()


def test_family_workload_full_size_draws(pkg, po, lpdata):
    fx = FX["c4"]["draws"]
    fam = lpdata.StaircaseFamily(periods=12, blk_rows=2000, blk_cols=5000, inv_per_link=200, seed=42)
    ids = sorted(int(k) for k in fx)
    draws = [fam.draw(k) for k in ids]
    print("case: c4 draws", ids[0], "..", ids[-1], flush=True)
    f = pkg.PdlpFamily(fam.base, capacity=64, eps_rel_kkt=TOL, max_iterations=200000)
    f.set_lps([d.c for d in draws], [d.b for d in draws])
    per, tot = f.solve()
    print("family", tot, flush=True)
    assert tot["solved"] == len(draws)
    X, Y = f.solutions()
    for i, k in enumerate(ids):
        check(po, draws[i], X[i], Y[i], fx[str(k)], f"c4 draw {k}")
    f.close()


def test_zipf_lp_fp64_iterates_over_fp32_values_full_size(pkg, po, lpdata):
    # the 200,000 x 500,000 Zipf LP with FP64 iterates over FP32 matrix
    # values: the FP64 engine's iteration count at 6 bytes per stored entry
    if "c1" not in FX:
        pytest.skip("no objective fixture for c1")
    lp = SINGLE["c1"](lpdata)
    print("case: c1, fp64_iterates = 2", flush=True)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, max_iterations=200000, fp64_iterates=2)
    s = h.solve()
    x, y = h.solution()
    h.close()
    print("status", s["status"], "iterations", s["iterations"], "engine rel-KKT %.3e" % s["rel_kkt"], flush=True)
    assert s["status"] == 0
    check(po, lp, x, y, FX["c1"], "c1 fp64 iterates over fp32 values")
