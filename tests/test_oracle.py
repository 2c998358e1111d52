"""The CPU oracle, pinned before it is trusted (no GPU needed).

* the reference's published Benchmark-10 figures (docs/phase-12-report.md:92):
  24,700 iterations, OPTIMAL, max violation 8.3e-8;
* bit-for-bit agreement with the reference's own build (golden SHA-256 of
  the FP64 result vectors, tests/golden/ref_pdhg_kat.json, and live against
  oracle/_ref when present);
* the sparse KATs of tests/test_sparse_matrix.cpp / test_gpu_foundation.cpp;
* the FP64 KKT evaluator: closed forms, HiGHS optimal primal-dual pairs, weak
  duality, and planted mistakes that it must flag;
* the OpenMP restatement of the restarted Halpern PDHG (oracle/pdlp_omp.c):
  its own stop test agrees with the evaluator, and its objectives match the
  exact optima of tests/golden/lp_optima.json.
"""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "ref_pdhg_kat.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def kat_models():
    from golden import make_golden  # noqa: F401  (sys.path set below)
    return make_golden.kat_models()


@pytest.fixture(scope="module")
def models():
    import sys
    sys.path.insert(0, HERE)
    from golden import make_golden
    return make_golden.kat_models()


OPT_KW = {"max_iterations": "max_iterations", "adaptive_step_size": "adaptive",
          "check_frequency": "check_frequency"}


def oracle_kwargs(kw):
    return {OPT_KW[k]: v for k, v in kw.items()}


# ---------------------------------------------------------------- sparse KATs
def test_csr_from_coo_sorts_sums_prunes(po):
    # tests/test_sparse_matrix.cpp:197-253 semantics: (row, col) order,
    # duplicates summed, explicit zeros removed after summing
    ri = [1, 0, 1, 0, 2, 2]
    ci = [2, 1, 2, 0, 1, 1]
    v = [1.5, 2.0, 2.5, -1.0, 3.0, -3.0]
    rp, c, val = po.csr_from_coo(3, 3, ri, ci, v)
    assert rp.tolist() == [0, 2, 3, 3]
    assert c.tolist() == [0, 1, 2]
    assert val.tolist() == [-1.0, 2.0, 4.0]


def test_csr_from_coo_out_of_range(po):
    with pytest.raises(IndexError):
        po.csr_from_coo(2, 2, [0, 2], [0, 0], [1.0, 1.0])


@pytest.mark.parametrize("shape", [(3, 3), (4, 2), (2, 3), (40, 25)])
def test_spmv_matches_dense(po, shape):
    rng = np.random.default_rng(sum(shape))
    m, n = shape
    dense = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.4)
    ri, ci = np.nonzero(dense)
    rp, c, v = po.csr_from_coo(m, n, ri, ci, dense[ri, ci])
    x = rng.standard_normal(n)
    y = rng.standard_normal(m)
    np.testing.assert_allclose(po.spmv(rp, c, v, x), dense @ x, rtol=0, atol=1e-12)
    np.testing.assert_allclose(po.spmv_t(rp, c, v, y, n), dense.T @ y, rtol=0, atol=1e-12)


def test_spmv_known_answers(po):
    # tests/test_gpu_foundation.cpp:91-130: 1*4 + 2*6 = 16, 3*5 = 15; rectangular 4x2
    rp, c, v = po.csr_from_coo(2, 3, [0, 0, 1], [0, 2, 1], [1.0, 2.0, 3.0])
    assert po.spmv(rp, c, v, [4.0, 5.0, 6.0]).tolist() == [16.0, 15.0]
    rp, c, v = po.csr_from_coo(4, 2, [0, 1, 2, 3], [0, 1, 0, 1], [1.0, 2.0, 3.0, 4.0])
    assert po.spmv(rp, c, v, [10.0, 20.0]).tolist() == [10.0, 40.0, 30.0, 80.0]
    rp, c, v = po.csr_from_coo(1, 1, [0], [0], [1e-8])
    assert abs(po.spmv(rp, c, v, [1e-8])[0] - 1e-16) <= 1e-18


def test_spmv_bitequal_reference(ref):
    rng = np.random.default_rng(3)
    m, n, z = 50, 70, 600
    ri, ci, v = rng.integers(0, m, z), rng.integers(0, n, z), rng.standard_normal(z)
    x, xt = rng.standard_normal(n), rng.standard_normal(m)
    nnz, y, yt = ref.ref_spmv(m, n, ri, ci, v, x, xt)
    rp, c, val = ref.csr_from_coo(m, n, ri, ci, v)
    assert nnz == len(val)
    assert np.array_equal(ref.spmv(rp, c, val, x), y)
    assert np.array_equal(ref.spmv_t(rp, c, val, xt, n), yt)


# ---------------------------------------------------------------- reference PDHG
def test_benchmark10_published_figures(po, lpdata):
    lp = lpdata.flat_to_std(lpdata.benchmark10_model())
    r = po.pdhg_ref(lp)
    assert r.status == 0 and r.iterations == 24700
    assert abs(r.constraint_violation - 8.3e-8) < 0.05e-8
    assert not r.used_average


@pytest.mark.parametrize("name", sorted(KAT))
def test_oracle_matches_reference_golden(po, lpdata, models, name):
    fm, kw = models[name]
    g = KAT[name]
    import sys
    sys.path.insert(0, HERE)
    from golden.make_golden import model_sha
    assert model_sha(fm) == g["model_sha"], "generator changed: regenerate the fixture"
    pres = po.ref_presolve_dims(fm) if po.ref_lib() is not None else None
    if pres is not None and pres["status"] != 3:  # presolve changed the model
        pytest.skip("presolve reduces this model; covered by the GPU drop-in test")
    lp = lpdata.flat_to_std(fm)
    r = po.pdhg_ref(lp, **oracle_kwargs(kw))
    assert r.status == g["status"]
    assert r.iterations == g["iterations"]
    assert float.hex(r.constraint_violation) == g["constraint_violation"]
    assert sha(r.x) == g["x_sha"] and sha(r.y) == g["y_sha"]


def test_oracle_matches_reference_live(ref, lpdata):
    for seed in (90, 91):
        lp = lpdata.random_lp(80, 240, 4, seed)
        a = ref.ref_fo_solve(lp.to_flat(), max_iterations=1234, check_frequency=37)
        b = ref.pdhg_ref(lp, max_iterations=1234, check_frequency=37)
        assert (a["status"], a["iterations"]) == (b.status, b.iterations)
        assert np.array_equal(a["x"], b.x) and np.array_equal(a["y"], b.y)


def test_iteration_limit_kat(po, lpdata, models):
    # tests/test_first_order_solver.cpp:361-375: max_iterations = 5 -> MAX_ITERATIONS, 5
    fm, _ = models["iteration_limit_5"]
    r = po.pdhg_ref(lpdata.flat_to_std(fm), max_iterations=5)
    assert r.status == 1 and r.iterations == 5


def test_hand_lp_optimum(po, lpdata, models):
    fm, _ = models["hand_lp_64_3"]
    lp = lpdata.flat_to_std(fm)
    r = po.pdhg_ref(lp)
    assert r.status == 0
    np.testing.assert_allclose(r.x, [8 / 3, 8 / 3], atol=1e-2)
    assert abs(-(lp.c @ r.x) - 64 / 3) < 1e-2


# ---------------------------------------------------------------- KKT evaluator
def _hand_lp(lpdata):
    mb = lpdata.ModelBuilder(sense=0)  # min -3x - 5y, 2x + y <= 8, x + 2y <= 8
    x = mb.add_variable(0.0, np.inf, -3.0)
    y = mb.add_variable(0.0, np.inf, -5.0)
    mb.add_constraint([(x, 2.0), (y, 1.0)], lpdata.LE, 8.0)
    mb.add_constraint([(x, 1.0), (y, 2.0)], lpdata.LE, 8.0)
    return lpdata.flat_to_std(mb.build())


def test_kkt_closed_form_optimum(po, lpdata):
    lp = _hand_lp(lpdata)
    # optimal pair: x = (8/3, 8/3), y = (1/3, 7/3): c + A'y = 0, complementary
    k = po.kkt(lp, [8 / 3, 8 / 3], [1 / 3, 7 / 3])
    assert k["rel_kkt"] < 1e-15
    assert abs(k["primal_obj"] + 64 / 3) < 1e-12 and abs(k["dual_obj"] + 64 / 3) < 1e-12


def test_kkt_flags_planted_mistakes(po, lpdata):
    lp = _hand_lp(lpdata)
    # infeasible primal: row 1 violated by exactly 1
    k = po.kkt(lp, [3.0, 3.0], [1 / 3, 7 / 3])
    assert abs(k["r_primal"] - np.hypot(1.0, 1.0)) < 1e-12
    # wrong dual sign: y1 negative -> counted in the dual residual
    k = po.kkt(lp, [8 / 3, 8 / 3], [-1 / 3, 7 / 3])
    assert k["r_dual"] > 0.3
    # transposed dual (y swapped): gap and dual residual both non-zero
    k = po.kkt(lp, [8 / 3, 8 / 3], [7 / 3, 1 / 3])
    assert k["rel_kkt"] > 1e-2


def test_kkt_against_highs_pairs(po, lpdata):
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix
    for seed in (1, 2):
        lp = lpdata.random_lp(60, 200, 5, seed)
        A = csr_matrix((lp.val, lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n))
        res = linprog(lp.c, A_ub=A, b_ub=lp.b, bounds=list(zip(lp.lb, lp.ub)), method="highs")
        assert res.status == 0
        y = -res.ineqlin.marginals  # HiGHS: d obj / d b <= 0 for <= rows; here y >= 0
        k = po.kkt(lp, res.x, y)
        assert k["rel_kkt"] < 1e-9, k


def test_kkt_weak_duality(po, lpdata):
    rng = np.random.default_rng(4)
    lp = lpdata.random_lp(60, 200, 5, 4)
    for _ in range(5):
        y = rng.random(lp.m)
        k = po.kkt(lp, lp.x0, y)
        assert k["dual_obj"] <= k["primal_obj"] + 1e-9
        assert k["r_primal"] == 0.0  # x0 is feasible by construction


def test_kkt_equality_rows_and_free_columns(po):
    class L:  # min x1 + x2, x1 + x2 = 2, x1 free, x2 in [0, 5]
        m, n = 1, 2
        row_ptr = np.array([0, 2], np.int32)
        col_idx = np.array([0, 1], np.int32)
        val = np.array([1.0, 1.0])
        b = np.array([2.0])
        c = np.array([1.0, 1.0])
        lb = np.array([-np.inf, 0.0])
        ub = np.array([np.inf, 5.0])
        is_eq = np.array([1], np.uint8)
    # y = -1: lambda = c + A'y = 0 everywhere; any feasible x is optimal
    k = po.kkt(L, [1.5, 0.5], [-1.0])
    assert k["rel_kkt"] < 1e-15
    # free column with nonzero reduced cost -> dual residual
    k = po.kkt(L, [1.5, 0.5], [0.0])
    assert abs(k["r_dual"] - 1.0) < 1e-15


# ------------------------------------------- the new algorithm on host cores
OPT = json.load(open(os.path.join(HERE, "golden", "lp_optima.json")))


@pytest.mark.parametrize("name", sorted(OPT))
def test_pdlp_omp_reaches_tolerance_and_optimum(po, name):
    """oracle/pdlp_omp.c (the multi-core CPU baseline of the restarted
    Halpern PDHG): it must stop at rel-KKT <= 1e-6 by its own FP64 check, the
    independent evaluator must agree on the returned point, and the objective
    must match the exact optima (HiGHS, and the reference's simplex chain
    where it ran) to 1e-6 relative."""
    import sys
    sys.path.insert(0, HERE)
    from golden import make_golden
    lp = make_golden.pdlp_instances()[name]
    st, x, y = po.pdlp_omp(lp, threads=2)
    assert st["status"] == 0 and st["rel_kkt"] <= 1e-6, st
    k = po.kkt(lp, x, y)
    assert abs(k["rel_kkt"] - st["rel_kkt"]) <= 1e-9 + 1e-3 * k["rel_kkt"]
    assert k["max_bound_viol"] == 0.0
    for key in ("highs", "reference_simplex"):
        if key in OPT[name]:
            f = OPT[name][key]
            assert abs(k["primal_obj"] - f) / (1 + abs(f)) <= 1e-6, (key, k["primal_obj"], f)


def test_pdlp_omp_iteration_limit_and_threads(po, lpdata):
    lp = lpdata.random_lp(300, 1000, 5, 42)
    st, _, _ = po.pdlp_omp(lp, max_iterations=100, threads=1)
    assert st["status"] == 1 and st["iterations"] == 100  # 64 + 36: the last interval is cut
    a, xa, _ = po.pdlp_omp(lp, threads=1)
    b, xb, _ = po.pdlp_omp(lp, threads=4)
    assert a["status"] == b["status"] == 0 and b["threads"] == 4
    # reductions combine in another order with more threads; the path barely moves
    assert abs(a["iterations"] - b["iterations"]) <= 256
    assert np.max(np.abs(xa - xb)) <= 1e-3 * (1 + np.max(np.abs(xa)))


def test_pdlp_omp_objective_gap_stop(po, lpdata):
    """The objective-gap stop (eps_objective): a run that stops on rel-KKT
    1e-6 alone can leave |f - f*| / (1 + |f*|) above 1e-6 (its gap term is
    relative to 1 + |c'x| + |dual obj|); with eps_objective = 1e-6 the stop
    also needs |c'x - dual obj| <= 1e-6 (1 + |c'x|), and the objective is then
    within 1e-6 of the exact optimum on every golden LP."""
    import sys
    sys.path.insert(0, HERE)
    from golden import make_golden
    inst = make_golden.pdlp_instances()
    for name in sorted(OPT):
        st, x, y = po.pdlp_omp(inst[name], threads=2, eps_objective=1e-6)
        assert st["status"] == 0, (name, st)
        k = po.kkt(inst[name], x, y)
        assert k["rel_kkt"] <= 1e-6
        assert abs(k["primal_obj"] - k["dual_obj"]) <= 1e-6 * (1 + abs(k["primal_obj"])) * (1 + 1e-3), name
        f = OPT[name]["highs"]
        assert abs(k["primal_obj"] - f) / (1 + abs(f)) <= 1e-6, (name, k["primal_obj"], f)


def test_pdlp_omp_warm_start_and_checkpoint_resume(po, lpdata, tmp_path):
    """The FP64 restatement's optional starting point and checkpoints (what
    tests/golden/make_objectives.py uses to survive interruptions of its
    multi-hour runs): started at its own optimum it stops at the first check
    (0 iterations); a run cut into 30 calls of growing iteration caps, each
    continuing the checkpoint, ends on the uninterrupted run's iteration
    count and objective, and a checkpoint of another LP's dimensions is
    ignored."""
    lp = lpdata.staircase_lp(3, 200, 500, 20, seed=3)
    st0, x0, y0 = po.pdlp_omp(lp, max_iterations=500000, eps_rel_kkt=1e-9, eps_objective=1e-9, threads=1)
    assert st0["status"] == 0
    st1, _, _ = po.pdlp_omp(lp, max_iterations=500000, eps_rel_kkt=1e-9, eps_objective=1e-9, threads=1,
                            x_start=x0, y_start=y0, primal_weight_start=st0["primal_weight"])
    assert st1["status"] == 0 and st1["iterations"] == 0
    ck = str(tmp_path / "run.ckpt")
    cap, calls = 1024, 0
    while True:
        st, x, y = po.pdlp_omp(lp, max_iterations=cap, eps_rel_kkt=1e-9, eps_objective=1e-9, threads=1,
                               checkpoint_path=ck, checkpoint_every=512)
        calls += 1
        if st["status"] == 0 or cap >= 200000:
            break
        cap += 1024
    assert calls > 10 and st["status"] == 0
    assert st["iterations"] == st0["iterations"] and st["primal_obj"] == st0["primal_obj"]
    other = lpdata.random_lp(60, 200, 3, 7)  # other dimensions: the file is not used
    st2, _, _ = po.pdlp_omp(other, max_iterations=100000, eps_rel_kkt=1e-6, threads=1, checkpoint_path=ck,
                            checkpoint_every=0)
    ref, _, _ = po.pdlp_omp(other, max_iterations=100000, eps_rel_kkt=1e-6, threads=1)
    assert st2["status"] == 0 and st2["iterations"] == ref["iterations"]
