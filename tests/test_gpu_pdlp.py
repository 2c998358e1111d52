"""GPU parity of the PDLP engine (FP32 iterates, FP64 checks) through the C ABI.

The engine's answer is checked by the oracle, never by itself:
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


@pytest.mark.parametrize("mode", [1, 2])
def test_pdlp_fp64_iterates(pkg, po, instances, mode):
    # mode 1: FP64 iterates and matrix values; mode 2: FP64 iterates over
    # FP32 matrix values (the multi-kernel loop only)
    name = "random_300x1000_s42"
    lp = instances[name]
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, fp64_iterates=mode)
    s = h.solve()
    x, y = h.solution()
    check_solution(po, lp, x, y, s, name)


@pytest.mark.parametrize("name", ["random_300x1000_s42", "random_300x1000_s44", "staircase_3x200x500_s3",
                                  "zipf_2000x5000_s9"])
@pytest.mark.parametrize("one_wg", [0, 1])
def test_pdlp_fp64_follows_the_cpu_restatement(pkg, po, instances, name, one_wg):
    # The engine in FP64 against oracle/pdlp_omp.c, an independent FP64
    # restatement of the same algorithm (scaling, power iteration, steps,
    # checks, restart rule, primal weight): only summation orders differ, so
    # the runs restart alike and stop within a few check intervals of each
    # other, at the same optimum.
    lp = instances[name]
    print("case:", name, "one workgroup", one_wg, flush=True)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, fp64_iterates=1, check_every=64, single_workgroup=one_wg)
    s = h.solve()
    x, y = h.solution()
    h.close()
    c, xc, yc = po.pdlp_omp(lp, eps_rel_kkt=TOL, check_every=64, threads=2)
    print("gpu", s["iterations"], s["restarts"], "cpu", c["iterations"], c["restarts"], flush=True)
    assert s["status"] == 0 and c["status"] == 0
    assert abs(s["iterations"] - c["iterations"]) <= max(256, 0.1 * c["iterations"])
    # the same 40 power-iteration rounds; the engine seeds them in its storage
    # order (columns sorted by length), so the estimates of ||A_hat||_2 agree
    # to the iteration's convergence (1e-4 on the Zipf LP), not to rounding
    assert abs(s["step_size"] - c["step_size"]) <= 1e-3 * c["step_size"]
    kg, kc = po.kkt(lp, x, y), po.kkt(lp, xc, yc)
    assert abs(kg["primal_obj"] - kc["primal_obj"]) / (1 + abs(kc["primal_obj"])) <= 2e-6


@pytest.mark.parametrize("seed,K", [(42, 256), (43, 128)])
def test_pdlp_fp64_any_check_cadence(pkg, po, lpdata, seed, K):
    # With FP64 iterates the engine follows the CPU restatement at any check
    # cadence (with FP32 iterates these two cadences do not converge on
    # these LPs within 50,000 iterations, which is why the cadence tuner
    # never leaves 64 for a one-workgroup solve)
    lp = lpdata.random_lp(300, 1000, 5, seed)
    c, _, _ = po.pdlp_omp(lp, eps_rel_kkt=TOL, check_every=K, threads=2)
    for sw in (0, 1):
        h = pkg.Pdlp(lp, eps_rel_kkt=TOL, fp64_iterates=1, check_every=K, single_workgroup=sw, max_iterations=50000)
        s = h.solve()
        x, y = h.solution()
        h.close()
        print("seed", seed, "K", K, "sw", sw, "gpu", s["iterations"], s["restarts"], "cpu", c["iterations"], c["restarts"],
              flush=True)
        assert s["status"] == 0 and c["status"] == 0
        assert abs(s["iterations"] - c["iterations"]) <= max(256, 0.1 * c["iterations"])
        assert po.kkt(lp, x, y)["rel_kkt"] <= TOL


def test_pdlp_staircase_against_fp64_solve_to_1e9(pkg, po, lpdata):
    # A staircase LP of the C2 generator at 6 periods (12,000 x 31,000), too
    # large for the exact solvers of the golden fixtures: the FP32 engine's
    # objective at rel-KKT 1e-6 against the FP64 restatement solved to rel-KKT
    # 1e-9 (certified by the evaluator), the oracle the survey names for the
    # large configurations.  (At the full 24 periods the two agree to 3.1e-8.)
    lp = lpdata.staircase_lp(6, 2000, 5000, 200, seed=42)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL)
    s = h.solve()
    x, y = h.solution()
    h.close()
    kg = po.kkt(lp, x, y)
    c, xc, yc = po.pdlp_omp(lp, eps_rel_kkt=1e-9, max_iterations=400000, threads=16)
    kc = po.kkt(lp, xc, yc)
    print("gpu", s["iterations"], kg["rel_kkt"], kg["primal_obj"], "fp64", c["iterations"], kc["rel_kkt"],
          kc["primal_obj"], flush=True)
    assert s["status"] == 0 and kg["rel_kkt"] <= TOL
    assert c["status"] == 0 and kc["rel_kkt"] <= 1.01e-9
    assert abs(kg["primal_obj"] - kc["primal_obj"]) <= 1e-6 * (1 + abs(kc["primal_obj"]))


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


@pytest.mark.parametrize("hot", [1000, 4096])
def test_pdlp_lds_hot_columns_bitequal(pkg, instances, hot):
    # the persistent CSR pass reading the hottest x+ entries from LDS gathers the
    # same values in the same order as the plain pass: identical iterates
    lp = instances["zipf_2000x5000_s9"]
    a = pkg.Pdlp(lp, lanes_rows=4, lanes_cols=4, hot_columns=0)
    b = pkg.Pdlp(lp, lanes_rows=4, lanes_cols=4, hot_columns=hot)
    sa, sb = a.solve(), b.solve()
    assert sb["hot_columns"] == hot and sa["hot_columns"] == 0
    assert sa["iterations"] == sb["iterations"]
    xa, ya = a.solution()
    xb, yb = b.solution()
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb)


@pytest.mark.parametrize("cap", [1024, 2048, 4096])
def test_pdlp_stream_schedule(pkg, po, instances, cap):
    # both passes on the stream kernel (LDS-staged products); the zipf case has
    # columns longer than every budget, so the one-segment blocks run too
    name = "zipf_2000x5000_s9"
    lp = instances[name]
    cl = np.bincount(lp.col_idx, minlength=lp.n)
    rl = np.diff(lp.row_ptr)
    print("case: stream cap", cap, "max col", cl.max(), "max row", rl.max(), flush=True)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, lanes_rows=-cap, lanes_cols=-cap, hot_columns=0)
    s = h.solve()
    assert s["lanes_rows"] == -cap and s["lanes_cols"] == -cap
    x, y = h.solution()
    check_solution(po, lp, x, y, s, name)
    s2 = h.solve()  # deterministic
    x2, y2 = h.solution()
    assert s2["iterations"] == s["iterations"] and np.array_equal(x, x2) and np.array_equal(y, y2)


def dense_line_lp(lpdata, m=1500, n=5000, k=4, seed=11):
    """random_lp entries plus one dense row (n > 4096 nonzeros) and one dense
    column (m nonzeros, between 1024 and 2048): segments longer than every
    stream budget, reduced by a workgroup of their own."""
    rng = np.random.default_rng(seed)
    rows = np.argsort(rng.random((n, m)), axis=1)[:, :k]
    cols = np.repeat(np.arange(n), k)
    rows = rows.reshape(-1)
    keep = (rows != 0) & (cols != 0)
    rows = np.concatenate([rows[keep], np.zeros(n, np.int64), np.arange(1, m)])
    cols = np.concatenate([cols[keep], np.arange(n), np.zeros(m - 1, np.int64)])
    vals = rng.standard_normal(rows.size)
    rp, ci, v = lpdata._csr_from_cols(m, n, cols, rows, vals)
    return lpdata._finish(m, n, rp, ci, v, rng, rng.uniform(0.0, 10.0, n))


@pytest.mark.parametrize("cap", [1024, 4096])
def test_pdlp_stream_long_segments(pkg, po, lpdata, cap):
    lp = dense_line_lp(lpdata)
    assert np.diff(lp.row_ptr).max() > 4096 and 1024 < np.bincount(lp.col_idx).max() < 2048
    print("case: dense-line LP, stream cap", cap, flush=True)
    seg = pkg.Pdlp(lp, eps_rel_kkt=TOL, lanes_rows=4, lanes_cols=4, hot_columns=0)
    sa = seg.solve()
    xa, ya = seg.solution()
    stm = pkg.Pdlp(lp, eps_rel_kkt=TOL, lanes_rows=-cap, lanes_cols=-cap, hot_columns=0)
    sb = stm.solve()
    xb, yb = stm.solution()
    ka, kb = po.kkt(lp, xa, ya), po.kkt(lp, xb, yb)
    print("segment iters", sa["iterations"], "rel-KKT %.3e" % ka["rel_kkt"], "stream iters", sb["iterations"],
          "rel-KKT %.3e" % kb["rel_kkt"], flush=True)
    assert sa["status"] == 0 and sb["status"] == 0
    assert ka["rel_kkt"] <= TOL and kb["rel_kkt"] <= TOL
    fa, fb = ka["primal_obj"], kb["primal_obj"]
    assert abs(fa - fb) / (1 + abs(fa)) <= 1e-5


@pytest.mark.parametrize("k", [1, 2, 4])
def test_pdlp_tiled_schedule(pkg, po, instances, k):
    # both passes on the tiled gather (vector block in LDS, SELL-64 slices,
    # FP64 partials) + reduction kernels
    name = "zipf_2000x5000_s9"
    lp = instances[name]
    print("case: tiled", k, "per CU", flush=True)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, lanes_rows=-k, lanes_cols=-k, hot_columns=0)
    s = h.solve()
    assert s["lanes_rows"] == -k and s["lanes_cols"] == -k
    for p in (0, 1):
        info = h.pass_info(p)
        assert info["kernels"] == 2 and info["tiles"] >= 1 and info["padded"] >= lp.nnz, info
    x, y = h.solution()
    check_solution(po, lp, x, y, s, name)
    s2 = h.solve()
    x2, y2 = h.solution()
    assert s2["iterations"] == s["iterations"] and np.array_equal(x, x2) and np.array_equal(y, y2)


@pytest.mark.parametrize("fp64", [0, 1, 2])
def test_pdlp_tiled_many_blocks(pkg, po, lpdata, fp64):
    # n = 90000 gathered entries do not fit one LDS block: several blocks,
    # partials added by the reduction kernel; compared with the segment schedule
    lp = lpdata.random_lp(3000, 90000, 3, 5)
    print("case: tiled many blocks, fp64", fp64, flush=True)
    seg = pkg.Pdlp(lp, eps_rel_kkt=TOL, lanes_rows=4, lanes_cols=4, hot_columns=0, fp64_iterates=fp64)
    sa = seg.solve()
    xa, ya = seg.solution()
    til = pkg.Pdlp(lp, eps_rel_kkt=TOL, lanes_rows=-2, lanes_cols=-2, hot_columns=0, fp64_iterates=fp64)
    sb = til.solve()
    xb, yb = til.solution()
    info = til.pass_info(0)
    print("segment iters", sa["iterations"], "tiled iters", sb["iterations"], "csr pass", info, flush=True)
    assert info["kernels"] == 2 and info["blocks"] >= 3
    ka, kb = po.kkt(lp, xa, ya), po.kkt(lp, xb, yb)
    assert sa["status"] == 0 and sb["status"] == 0
    assert ka["rel_kkt"] <= TOL and kb["rel_kkt"] <= TOL
    assert abs(ka["primal_obj"] - kb["primal_obj"]) / (1 + abs(ka["primal_obj"])) <= 1e-5


def test_pdlp_tiled_long_segments(pkg, po, lpdata):
    # a dense row and a dense column: pairs of 128+ nonzeros take wide slices
    lp = dense_line_lp(lpdata)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, lanes_rows=-1, lanes_cols=-1, hot_columns=0)
    s = h.solve()
    x, y = h.solution()
    k = po.kkt(lp, x, y)
    print("tiled dense-line iters", s["iterations"], "rel-KKT %.3e" % k["rel_kkt"], flush=True)
    assert s["status"] == 0 and k["rel_kkt"] <= TOL and s["lanes_rows"] == -1


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


@pytest.mark.parametrize("cap", [100, 128, 200])
def test_pdlp_interval_queue_stops_at_max_iterations(pkg, po, instances, cap):
    # the multi-kernel loop keeps two check intervals queued (64 steps each)
    # while far from the stopping test: a cap inside the second interval, at
    # its end, or after a partial third one stops exactly there, with the
    # state of the last step taken
    lp = instances["zipf_2000x5000_s9"]
    h = pkg.Pdlp(lp, eps_rel_kkt=1e-12, max_iterations=cap, single_workgroup=0)
    s = h.solve()
    x, y = h.solution()
    h.close()
    k = po.kkt(lp, x, y)
    print("cap", cap, "iterations", s["iterations"], "engine rel-KKT %.3e oracle %.3e" % (s["rel_kkt"], k["rel_kkt"]),
          flush=True)
    assert s["status"] == 1 and s["iterations"] == cap
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(y))
    assert abs(k["rel_kkt"] - s["rel_kkt"]) <= 1e-9 + 1e-3 * k["rel_kkt"]  # the engine's report of the same point


def test_pdlp_time_limit_stops_the_interval_queue(pkg, lpdata):
    # a time limit ends the multi-kernel loop with status 3 (TIME_LIMIT) at
    # most one queued interval late
    import time
    lp = lpdata.zipf_lp(20000, 50000, 1.5, 500, seed=5)
    h = pkg.Pdlp(lp, eps_rel_kkt=1e-12, max_iterations=10**7, time_limit_s=0.05, single_workgroup=0)
    t0 = time.perf_counter()
    s = h.solve()
    dt = time.perf_counter() - t0
    h.close()
    print("time-limited solve: status", s["status"], "iterations", s["iterations"], "%.3f s" % dt, flush=True)
    assert s["status"] == 3 and 0 < s["iterations"] < 10**7 and dt < 5.0


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


@pytest.mark.parametrize("fp64", [0, 1])
@pytest.mark.parametrize("name", sorted(OPT))
def test_pdlp_single_workgroup(pkg, po, instances, name, fp64):
    # the whole solve in one workgroup (k_small): same stopping rule, same
    # optimum as the multi-kernel loop
    lp = instances[name]
    print("case: one workgroup", name, "fp64", fp64, flush=True)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, max_iterations=200000, single_workgroup=1, fp64_iterates=fp64)
    s = h.solve()
    # 2: the scaled matrix (both orientations) fits the LDS beside the vectors
    resident = (lp.m <= 2048 and lp.n <= 2048 and
                (8 if fp64 else 4) * (2 * lp.nnz + lp.m + lp.n) + 4 * (lp.m + lp.n + 2 + 2 * lp.nnz) + 4096 <= 160 * 1024)
    assert s["single_workgroup"] == (2 if resident else 1)
    x, y = h.solution()
    k = check_solution(po, lp, x, y, s, name)
    h.close()
    g = pkg.Pdlp(lp, eps_rel_kkt=TOL, max_iterations=200000, single_workgroup=0, fp64_iterates=fp64)
    sg = g.solve()
    assert sg["single_workgroup"] == 0 and sg["status"] == 0
    xg, yg = g.solution()
    kg = po.kkt(lp, xg, yg)
    assert abs(k["primal_obj"] - kg["primal_obj"]) <= 2e-6 * (1 + abs(kg["primal_obj"]))
    g.close()


def test_pdlp_single_workgroup_limits_and_repeat(pkg, po, instances):
    name = sorted(OPT)[0]
    lp = instances[name]
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, max_iterations=200000, single_workgroup=1)
    s1 = h.solve()
    x1, y1 = h.solution()
    s2 = h.solve()
    x2, y2 = h.solution()
    assert s1["iterations"] == s2["iterations"]
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    h.set_warm_start(x1, y1)  # at the optimum: stops at the first check
    s3 = h.solve()
    assert s3["status"] == 0 and s3["iterations"] == 0
    h.close()
    c = pkg.Pdlp(lp, eps_rel_kkt=1e-12, max_iterations=100, single_workgroup=1)
    s4 = c.solve()
    assert s4["status"] == 1 and s4["iterations"] == 100
    c.close()


def test_pdlp_rejects_bad_csr(pkg, lpdata):
    lp = lpdata.random_lp(20, 50, 3, 1)
    bad = lpdata.StdLP(lp.m, lp.n, lp.row_ptr, lp.col_idx.copy(), lp.val, lp.b, lp.c, lp.lb, lp.ub, lp.is_eq)
    bad.col_idx[3] = lp.n + 5  # column out of range: rejected on the host, no launch
    with pytest.raises(pkg.BoError) as e:
        pkg.Pdlp(bad)
    assert e.value.status == 4


@pytest.mark.parametrize("fname", ["adlittle.free.mps", "blend.free.mps", "p0033.free.mps"])
def test_pdlp_fp64_iterates_over_fp32_values_on_badly_scaled_models(pkg, po, fname):
    # Netlib models on which FP32 iterates do not converge (100,000
    # iterations: rel-KKT 3.5e-2, 2.7e-3 and 1.7e-2): FP64 iterates over the
    # FP32-rounded scaled matrix solve them in the FP64 engine's iteration
    # counts.  (kb2 and share2b need FP64 matrix values as well: with FP32
    # values they end at rel-KKT 2.7e-4 and 1.0e-6 after 100,000 iterations.)
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    gold = json.load(open(os.path.join(here, "golden", "mps_golden.json")))[fname]
    info, fm = pkg.read_mps(text=open(os.path.join(here, "golden", "mps", fname)).read())
    lp = pkg.lpdata.flat_to_std(fm)
    print("case:", fname, lp.m, lp.n, lp.nnz, flush=True)
    h = pkg.Pdlp(lp, eps_rel_kkt=TOL, eps_objective=5e-7, max_iterations=100000, fp64_iterates=2)
    s = h.solve()
    x, y = h.solution()
    h.close()
    k = po.kkt(lp, x, y)
    obj = (-1.0 if fm.sense == 1 else 1.0) * k["primal_obj"] + fm.obj_offset
    f = gold["highs_objective"]
    print(fname, "status", s["status"], "iterations", s["iterations"], "rel-KKT %.2e" % k["rel_kkt"],
          "objective error %.2e" % (abs(obj - f) / (1 + abs(f))), flush=True)
    assert s["status"] == 0 and k["rel_kkt"] <= TOL
    assert abs(obj - f) <= 1e-6 * (1 + abs(f))
