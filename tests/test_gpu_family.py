"""GPU parity of the parametric-family engine (bo_pdlp_family_*): up to 64
LPs over one matrix, vectors LP-minor, one LP per lane.

Every LP is checked by the oracle's FP64 KKT evaluator (rel-KKT <= 1e-6 on the
unscaled LP, bounds exact), its objective must match the block-diagonal
batched engine's on the same draw, and the engine's own FP64 report must
agree with the oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-6


@pytest.fixture(scope="module")
def family(pkg):
    return pkg.lpdata.StaircaseFamily(periods=12, blk_rows=100, blk_cols=250, inv_per_link=10, seed=7)


def _check(po, lp, x, y, s, name):
    k = po.kkt(lp, x, y)
    assert s["status"] == 0, (name, s)
    assert k["rel_kkt"] <= TOL, (name, k)
    assert k["max_bound_viol"] == 0.0
    assert abs(k["rel_kkt"] - s["rel_kkt"]) <= 1e-9 + 1e-3 * k["rel_kkt"], (name, k["rel_kkt"], s["rel_kkt"])
    return k


@pytest.mark.parametrize("fp64", [0, 1])
def test_family_matches_batch_engine(pkg, po, family, fp64):
    ids = list(range(1, 41))  # 40 LPs: lanes 40..63 idle
    draws = [family.draw(k) for k in ids]
    print("case: family of", len(ids), "draws, fp64", fp64, flush=True)
    f = pkg.PdlpFamily(family.base, capacity=64, eps_rel_kkt=TOL, max_iterations=60000, fp64_iterates=fp64)
    f.set_lps([d.c for d in draws], [d.b for d in draws])
    per, tot = f.solve()
    print("family", tot, flush=True)
    X, Y = f.solutions()
    b = pkg.PdlpBatch(draws, eps_rel_kkt=TOL, max_iterations=60000, fp64_iterates=fp64)
    perb, totb = b.solve()
    print("batch", totb, flush=True)
    assert tot["solved"] == len(ids)
    for i, d in enumerate(draws):
        k = _check(po, d, X[i], Y[i], per[i], "draw%d" % ids[i])
        if perb[i]["status"] == 0:
            xb, yb = b.solution(i)
            kb = po.kkt(d, xb, yb)
            assert abs(k["primal_obj"] - kb["primal_obj"]) <= TOL * (1 + abs(kb["primal_obj"]))
    its_f = np.array([p["iterations"] for p in per])
    its_b = np.array([p["iterations"] for p in perb])
    print("iterations family median %d batch median %d" % (np.median(its_f), np.median(its_b)), flush=True)
    b.close()
    f.close()


def test_family_reuse_warm_start_and_count(pkg, po, family):
    f = pkg.PdlpFamily(family.base, capacity=8, eps_rel_kkt=TOL, max_iterations=60000)
    d1 = [family.draw(k) for k in (101, 102, 103)]
    f.set_lps([d.c for d in d1], [d.b for d in d1])
    per1, tot1 = f.solve()
    X1, Y1 = f.solutions()
    per1b, _ = f.solve()  # same LPs again: same run
    X1b, Y1b = f.solutions()
    assert [p["iterations"] for p in per1] == [p["iterations"] for p in per1b]
    assert np.array_equal(X1, X1b) and np.array_equal(Y1, Y1b)
    for i, d in enumerate(d1):
        _check(po, d, X1[i], Y1[i], per1[i], "reuse%d" % i)
    # warm start at the optimum: every LP stops at its first check
    f.set_warm_start(X1, Y1)
    per2, tot2 = f.solve()
    assert tot2["solved"] == 3 and tot2["max_iterations"] == 0, tot2
    # a different count of different LPs in the same handle
    d2 = [family.draw(k) for k in range(201, 209)]
    f.set_lps([d.c for d in d2], [d.b for d in d2])
    per3, tot3 = f.solve()
    X3, Y3 = f.solutions()
    assert len(per3) == 8
    for i, d in enumerate(d2):  # every draw solved (the FP64 polish takes any that stall)
        _check(po, d, X3[i], Y3[i], per3[i], "count8_%d" % i)
    assert tot3["solved"] == 8
    f.close()


def test_family_fp64_polish(pkg, po, family):
    # draws whose FP32 iterates stall short of 1e-6 (see the batched engine's
    # polish test): the family engine hands them to its FP64 engine
    ids = [31, 143, 185, 1, 2]
    draws = [family.draw(k) for k in ids]
    f = pkg.PdlpFamily(family.base, capacity=len(ids), eps_rel_kkt=TOL, max_iterations=40000,
                       stall_iterations=4096)
    f.set_lps([d.c for d in draws], [d.b for d in draws])
    per, tot = f.solve()
    print("polish", [(p["status"], p["iterations"], p["fp64_polish_iterations"]) for p in per], flush=True)
    X, Y = f.solutions()
    assert tot["solved"] == len(ids)
    for i, d in enumerate(draws):
        assert per[i]["iterations"] <= 40000
        _check(po, d, X[i], Y[i], per[i], "polish%d" % ids[i])
    f.close()


def test_family_rejects_bad_input(pkg, family):
    with pytest.raises(pkg.BoError):
        pkg.PdlpFamily(family.base, capacity=65)
    f = pkg.PdlpFamily(family.base, capacity=4)
    d = family.draw(5)
    with pytest.raises(pkg.BoError):
        f.set_lps([d.c] * 5, [d.b] * 5)  # above capacity
    c = d.c.copy()
    c[3] = np.nan
    with pytest.raises(pkg.BoError):
        f.set_lps([c], [d.b])
    f.close()


def test_family_primal_weights(pkg, po, family):
    # starting primal weights: each LP solved again from the weight its cold
    # solve ended with (the learned warm start's second output); still every
    # LP at rel-KKT 1e-6, and bad weights are refused before anything runs
    ids = list(range(101, 133))
    draws = [family.draw(k) for k in ids]
    f = pkg.PdlpFamily(family.base, capacity=64, eps_rel_kkt=TOL, max_iterations=60000)
    f.set_lps([d.c for d in draws], [d.b for d in draws])
    per, _ = f.solve()
    w = np.array([p["primal_weight"] for p in per])
    assert np.all(w > 0)
    f.set_primal_weights(w)
    per2, _ = f.solve()
    X, Y = f.solutions()
    for k, d in enumerate(draws):
        _check(po, d, X[k], Y[k], per2[k], f"lp {k}")
    print("cold iters", int(np.max([p["iterations"] for p in per])), "from final weights",
          int(np.max([p["iterations"] for p in per2])), flush=True)
    bad = w.copy()
    bad[3] = -1.0
    with pytest.raises(pkg.BoError):
        f.set_primal_weights(bad)
    bad[3] = np.nan
    with pytest.raises(pkg.BoError):
        f.set_primal_weights(bad)
    f.set_primal_weights(None)
    per3, _ = f.solve()
    assert [p["iterations"] for p in per3] == [p["iterations"] for p in per]  # cleared: the cold run again
    f.close()


def test_family_time_limit(pkg, family):
    # A wall-clock limit far below a solve to 1e-9: every LP ends solved (0)
    # or stopped by the limit (3), the FP64 polish does not start after it,
    # and the call returns within the limit plus one check interval.
    import time
    ids = list(range(300, 340))
    draws = [family.draw(k) for k in ids]
    f = pkg.PdlpFamily(family.base, capacity=64, eps_rel_kkt=1e-9, max_iterations=10**7, time_limit_s=0.05)
    f.set_lps([d.c for d in draws], [d.b for d in draws])
    t0 = time.perf_counter()
    per, tot = f.solve()
    wall = time.perf_counter() - t0
    print("family", tot, "wall %.3f s" % wall, flush=True)
    assert all(p["status"] in (0, 3) for p in per)
    assert any(p["status"] == 3 for p in per)
    assert wall <= 0.05 + 1.0
    f.close()
