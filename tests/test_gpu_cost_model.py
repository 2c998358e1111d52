"""The routing cost model against live solves on LPs outside its
calibration table: the predicted wall time (creation + set-up + measured
iteration count x per-iteration cost) must be within a factor of 1.6 of the
measured one, for both engines."""
import time

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def warm_device(pkg, lpdata):
    """One-time start-up (code objects, first allocations) is not a solve's
    cost: a small solve of each engine first."""
    lp = lpdata.random_lp_sampled(50, 150, 5, 1)
    h = pkg.Pdlp(lp, max_iterations=64)
    h.solve()
    h.close()
    pkg.pdhg_ref_solve(lp, max_iterations=10)


@pytest.mark.parametrize("m", [2000, 20000, 60000])
def test_pdlp_cost_prediction(pkg, lpdata, m):
    lp = lpdata.random_lp_sampled(m, 3 * m, 5, 77 + m)
    t0 = time.perf_counter()
    h = pkg.Pdlp(lp, eps_rel_kkt=1e-6, max_iterations=100000)
    s = h.solve()
    x, y = h.solution()
    wall_ms = (time.perf_counter() - t0) * 1e3
    h.close()
    est = pkg.cost_estimate(lp.m, lp.n, lp.nnz, engine=1, iterations=s["iterations"])
    print("case: random", m, "nnz", lp.nnz, "iterations", s["iterations"], "wall %.1f ms predicted %.1f ms"
          % (wall_ms, est["total_ms"]), flush=True)
    assert est["in_domain"] == 1
    assert wall_ms / 1.6 <= est["total_ms"] <= wall_ms * 1.6


@pytest.mark.parametrize("m", [400, 20000])
def test_reference_engine_cost_prediction(pkg, lpdata, m):
    lp = lpdata.random_lp_sampled(m, 3 * m, 5, 91 + m)
    t0 = time.perf_counter()
    r = pkg.pdhg_ref_solve(lp, max_iterations=3000)
    wall_ms = (time.perf_counter() - t0) * 1e3
    est = pkg.cost_estimate(lp.m, lp.n, lp.nnz, engine=0, iterations=r.iterations)
    print("case: random", m, "loop", r.loop, "iterations", r.iterations, "wall %.1f ms predicted %.1f ms"
          % (wall_ms, est["total_ms"]), flush=True)
    assert est["loop"] == r.loop
    assert wall_ms / 1.6 <= est["total_ms"] <= wall_ms * 1.6
