"""bo_pdhg_cost_estimate, the MI355X-measured cost of a GPU first-order
solve that replaces the reference router's analytic GPU estimate
(src/cost_estimator.cpp:188-197: 24.5 + 0.0008 nnz ms).  Host arithmetic:
these checks need no GPU; tests/test_gpu_cost_model.py compares the
predictions with live solves."""
import pytest


def test_estimate_reproduces_measured_points(pkg):
    # the calibration's own points come back exactly (C1: 908 ms creation,
    # 0.0808 ms per iteration; C2's staircase LP: 0.0171 ms per iteration)
    c1 = pkg.cost_estimate(200000, 500000, 17069815, engine=1, iterations=12032)
    assert abs(c1["create_ms"] - 908.066) < 1e-3 and abs(c1["per_iteration_ms"] - 0.08081) < 1e-9
    assert abs(c1["total_ms"] - (c1["create_ms"] + c1["fixed_ms"] + 12032 * c1["per_iteration_ms"])) < 1e-9
    c2 = pkg.cost_estimate(48000, 124600, 1385701, engine=1, iterations=9024)
    assert abs(c2["per_iteration_ms"] - 0.017122) < 1e-9
    # between points: log-log interpolation stays between the neighbours
    mid = pkg.cost_estimate(100000, 300000, 3000000, engine=1)["per_iteration_ms"]
    assert 0.017122 < mid < 0.063617


def test_iteration_defaults_and_engine_tiers(pkg):
    assert pkg.cost_estimate(300, 1000, 5000, engine=1)["iterations"] == 10000
    assert pkg.cost_estimate(300, 1000, 5000, engine=0)["iterations"] == 50000
    # the reference engine's loop tier follows bo_pdhg_ref_solve's rule
    assert pkg.cost_estimate(300, 1000, 5000, engine=0)["loop"] == 2        # LP resident in LDS
    assert pkg.cost_estimate(1000, 3000, 15000, engine=0)["loop"] == 1      # one workgroup over memory
    assert pkg.cost_estimate(200000, 500000, 17069815, engine=0)["loop"] == 0
    assert pkg.cost_estimate(300, 1000, 5000, engine=1)["loop"] == -1


def test_domain_flag_and_monotone_large_sizes(pkg):
    assert pkg.cost_estimate(10, 10, 10)["in_domain"] == 0
    assert pkg.cost_estimate(10**6, 10**6, 10**8)["in_domain"] == 0
    assert pkg.cost_estimate(300, 1000, 5000)["in_domain"] == 1
    prev = 0.0
    for nnz in (450000, 10**6, 3 * 10**6, 10**7, 17069815):
        t = pkg.cost_estimate(nnz // 50, nnz // 20, nnz)["per_iteration_ms"]
        assert t > prev
        prev = t


def test_invalid_arguments(pkg):
    for args in ((-1, 1, 1, 1, 0), (1, 1, -5, 1, 0), (1, 1, 1, 2, 0), (1, 1, 1, 1, -1)):
        with pytest.raises(pkg.BoError) as e:
            pkg.cost_estimate(*args)
        assert e.value.status == 4
