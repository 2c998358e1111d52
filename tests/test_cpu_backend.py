"""CPUFirstOrderSolver, the default backend of FirstOrderSolverOptions
(include/bharatopt/first_order_solver.hpp:36): the reference's PDHG loop
(src/first_order_solver.cpp:14-290) restated on the host inside libbo_hip.so.
It needs no GPU, so these tests run in the CPU suite.

* the golden KATs written from the reference build (status, iteration count,
  objective and violation as hex, SHA-256 of x and y) are matched bit for bit;
* live models with mixed row senses and objective senses against the
  reference build, where it is present;
* the C++ drop-in program's default-option pass: the reference's
  FirstOrder_* scenarios with a default-constructed FirstOrderLPSolver.
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "ref_pdhg_kat.json")))
sys.path.insert(0, HERE)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def models():
    from golden import make_golden
    return make_golden.kat_models()


@pytest.mark.parametrize("name", sorted(KAT))
def test_cpu_backend_matches_reference_golden(pkg, models, name):
    fm, kw = models[name]
    g = KAT[name]
    r = pkg.first_order_solve(fm, backend="cpu", **kw)
    assert r["capi_status"] == 0
    assert r["status"] == g["status"]
    assert r["iterations"] == g["iterations"]
    assert float.hex(r["objective_value"]) == g["objective"]
    assert float.hex(r["constraint_violation"]) == g["constraint_violation"]
    assert (len(r["x"]), len(r["y"])) == (g["x_len"], g["y_len"])
    assert sha(r["x"]) == g["x_sha"] and sha(r["y"]) == g["y_sha"]


def test_cpu_backend_benchmark10_published_iterations(pkg, lpdata):
    # docs/phase-12-report.md:92: Benchmark 10 converges in 24,700 iterations
    r = pkg.first_order_solve(lpdata.benchmark10_model(), backend="cpu")
    assert (r["status"], r["iterations"]) == (0, 24700)
    assert r["constraint_violation"] <= 1e-7


def test_cpu_backend_matches_reference_live(pkg, ref, lpdata):
    rng = np.random.default_rng(11)
    for seed in range(8):
        lp = lpdata.random_lp(25, 70, 3, 300 + seed)
        fm = lp.to_flat()
        fm.con_sense = rng.choice([0, 1, 2], size=fm.num_cons).astype(np.int32)
        fm.sense = int(seed % 2)
        kw = dict(max_iterations=int(rng.integers(1, 4000)), check_frequency=int(rng.integers(1, 150)),
                  adaptive_step_size=bool(seed % 3))
        a = ref.ref_fo_solve(fm, **kw)
        b = pkg.first_order_solve(fm, backend="cpu", **kw)
        assert (a["status"], a["iterations"]) == (b["status"], b["iterations"]), (seed, kw)
        assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"]), seed
        assert a["objective"] == b["objective_value"]


def test_cpu_backend_stage_outcomes(pkg, lpdata):
    mb = lpdata.ModelBuilder()
    mb.add_variable(5.0, 2.0, 1.0)  # lb > ub: validation failure
    r = pkg.first_order_solve(mb.build(), backend="cpu")
    assert r["status"] == 2 and "validation" in r["message"] and len(r["x"]) == 0
    mb = lpdata.ModelBuilder()
    x = mb.add_variable(0.0, 1.0, 1.0)
    mb.add_constraint([(x, 1.0)], lpdata.GE, 3.0)  # singleton row -> lb 3 > ub 1
    assert pkg.first_order_solve(mb.build(), backend="cpu")["status"] == 3
    mb = lpdata.ModelBuilder(sense=1)  # bound-only: max 4x, 0 <= x <= 7
    mb.add_variable(0.0, 7.0, 4.0)
    r = pkg.first_order_solve(mb.build(), backend="cpu")
    assert r["status"] == 0 and r["x"][0] == 7.0 and r["objective_value"] == 28.0


def test_cpp_dropin_default_options_pass():
    exe = os.path.join(HERE, "cpp", "bin", "test_dropin")
    assert os.path.exists(exe), "tests/cpp not built"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(p.stdout[-4000:])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    passed = [ln for ln in p.stdout.splitlines() if ln.startswith("PASS [default options] FirstOrder_")]
    assert len(passed) >= 24


def test_dropin_results_carry_the_verifiers_judgement(pkg, lpdata, models):
    """Every returned solution is judged by SolutionVerifier against the
    original model with the reported objective (verify_status: 0 verified,
    1 failed, 2 objective mismatch; -1 no solution); the verdict equals the
    verifier's own on the same point."""
    for name in ("hand_lp_64_3", "benchmark10", "presolve_mix", "iteration_limit_5", "mps_afiro.free.mps"):
        fm, kw = models[name]
        r = pkg.first_order_solve(fm, backend="cpu", **kw)
        v = pkg.verify_solution(fm, r["x"], reported_objective=r["objective_value"])
        assert r["verify_status"] == v["status"], name
        assert r["verify_max_violation"] == max(v["maximum_constraint_violation"], v["maximum_lower_bound_violation"],
                                                 v["maximum_upper_bound_violation"])
        if r["status"] == 0 and name != "iteration_limit_5":
            assert r["verify_status"] == 0, (name, r["verify_max_violation"])
    mb = lpdata.ModelBuilder()
    mb.add_variable(5.0, 2.0, 1.0)  # validation failure: no solution, nothing to verify
    assert pkg.first_order_solve(mb.build(), backend="cpu")["verify_status"] == -1
