"""GPU parity of the reference-semantics path (engine REFERENCE_PDHG):
bit-for-bit against the oracle and the reference build, through the C ABI.

* GpuBackend FP64 primitives: KATs of tests/test_gpu_foundation.cpp plus
  bit-equality of the ordered SpMV with the reference's CSR loop;
* the PDHG loop (bo_pdhg_ref_solve): status, iteration count, violation and
  every bit of x and y equal to the oracle on seeded models, including the
  24,700-iteration Benchmark-10 case and edge cases (empty rows/columns,
  free and fixed-width columns, equality rows, non-adaptive steps, odd check
  frequencies, iteration limits 0/1/5);
* the whole drop-in (bo_first_order_solve: validate, presolve, standard form,
  GPU loop, postsolve) equal to the reference's FirstOrderLPSolver on the
  golden models (tests/golden/ref_pdhg_kat.json, written from the reference
  build: seeded random and staircase models, mixed row and objective senses,
  option variants, six MPS models), and the C++ drop-in test program.
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
KAT = json.load(open(os.path.join(HERE, "golden", "ref_pdhg_kat.json")))
sys.path.insert(0, HERE)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module", autouse=True)
def gpu_present(pkg):
    st, info = pkg.device_info()
    assert st == 0, "no GPU: " + info.status_message.decode()
    print("device:", info.device_name.decode(), flush=True)


@pytest.fixture(scope="module")
def models():
    from golden import make_golden
    return make_golden.kat_models()


# ---------------------------------------------------------------- primitives
def test_spmv_f64_bitequal(pkg, po, lpdata):
    for seed, (m, n) in enumerate([(1, 1), (5, 3), (300, 1000), (2000, 700)]):
        lp = lpdata.random_lp(m, n, min(3, m), seed)
        x = np.random.default_rng(seed).standard_normal(n)
        y, t = pkg.spmv_f64(lp, x)
        assert np.array_equal(y, po.spmv(lp.row_ptr, lp.col_idx, lp.val, x))
        assert t.kernel_time_ms >= 0.0


def test_spmv_f64_empty_rows(pkg, po):
    lp = pkg.lpdata.StdLP(4, 3, np.array([0, 0, 2, 2, 3], np.int32), np.array([0, 2, 1], np.int32),
                          np.array([1.5, -2.0, 4.0]), np.zeros(4), np.zeros(3), np.zeros(3),
                          np.ones(3), np.zeros(4, np.uint8))
    y, _ = pkg.spmv_f64(lp, [1.0, 2.0, 3.0])
    assert y.tolist() == [0.0, 1.5 - 6.0, 0.0, 8.0]


def test_vector_primitives(pkg):
    lib = pkg.lib()
    a = pkg.DeviceBuffer.from_host(np.array([1.0, 2.0, 3.0]))
    b = pkg.DeviceBuffer.from_host(np.array([4.0, 5.0, 6.0]))
    r = pkg.DeviceBuffer(3)
    assert lib.bo_vec_add_f64(a.ptr, b.ptr, r.ptr, 3, None) == 0
    assert r.download().tolist() == [5.0, 7.0, 9.0]
    assert lib.bo_vec_sub_f64(b.ptr, a.ptr, r.ptr, 3, None) == 0
    assert r.download().tolist() == [3.0, 3.0, 3.0]
    assert lib.bo_vec_scale_f64(a.ptr, 2.5, r.ptr, 3, None) == 0
    assert r.download().tolist() == [2.5, 5.0, 7.5]
    y = pkg.DeviceBuffer.from_host(np.array([10.0, 20.0, 30.0]))
    assert lib.bo_axpy_f64(3.0, a.ptr, y.ptr, r.ptr, 3, None) == 0
    assert r.download().tolist() == [13.0, 26.0, 39.0]
    import ctypes as C
    out = C.c_double()
    assert lib.bo_dot_f64(a.ptr, b.ptr, C.byref(out), 3, None) == 0 and out.value == 32.0
    v = pkg.DeviceBuffer.from_host(np.array([3.0, 4.0, 0.0]))
    assert lib.bo_norm2_f64(v.ptr, C.byref(out), 3, None) == 0 and out.value == 5.0
    big = np.random.default_rng(0).standard_normal(1_000_003)
    d = pkg.DeviceBuffer.from_host(big)
    assert lib.bo_dot_f64(d.ptr, d.ptr, C.byref(out), big.size, None) == 0
    assert abs(out.value - big @ big) <= 1e-9 * (big @ big)
    assert lib.bo_dot_f64(None, None, C.byref(out), 10, None) == 4  # INVALID_ARGUMENT


# ---------------------------------------------------------------- PDHG loop
def _same(g, o):
    assert (g.status, g.iterations) == (o.status, o.iterations)
    assert g.constraint_violation == o.constraint_violation
    assert g.bound_violation == o.bound_violation
    assert g.used_average == o.used_average
    assert np.array_equal(g.x, o.x), np.max(np.abs(g.x - o.x))
    assert np.array_equal(g.y, o.y), np.max(np.abs(g.y - o.y))


def test_benchmark10_on_gpu(pkg, po, lpdata):
    lp = lpdata.flat_to_std(lpdata.benchmark10_model())
    print("case: benchmark10", flush=True)
    g = pkg.pdhg_ref_solve(lp)
    assert g.status == 0 and g.iterations == 24700
    assert abs(g.constraint_violation - 8.3e-8) < 0.05e-8
    _same(g, po.pdhg_ref(lp))


@pytest.mark.parametrize("case", [
    # loop 2: the whole LP resident in one workgroup's LDS (one or two
    # columns / rows per thread); loop 1: one workgroup over the matrix in
    # memory (n + m <= 7680, nnz <= 65536); loop 0: a kernel per pass
    dict(m=300, n=1000, k=5, seed=42, it=2000, loop=2),
    dict(m=60, n=200, k=5, seed=7, it=50000, loop=2),
    dict(m=1, n=4, k=1, seed=3, it=300, loop=2),
    dict(m=500, n=120, k=7, seed=8, it=700, loop=2),
    dict(m=200, n=1800, k=2, seed=19, it=600, loop=2),
    dict(m=2000, n=5000, k=3, seed=17, it=500, loop=1),
    dict(m=3000, n=6000, k=3, seed=11, it=400, loop=0),
    dict(m=400, n=1200, k=60, seed=13, it=300, loop=0),
])
def test_pdhg_ref_bitequal_random(pkg, po, lpdata, case):
    print("case:", case, flush=True)
    lp = lpdata.random_lp(case["m"], case["n"], case["k"], case["seed"])
    g = pkg.pdhg_ref_solve(lp, max_iterations=case["it"])
    assert g.loop == case["loop"]
    _same(g, po.pdhg_ref(lp, max_iterations=case["it"]))


def test_pdhg_ref_bitequal_staircase(pkg, po, lpdata):
    lp = lpdata.staircase_lp(2, 60, 150, 10, seed=5)
    _same(pkg.pdhg_ref_solve(lp, max_iterations=3000), po.pdhg_ref(lp, max_iterations=3000))


@pytest.mark.parametrize("kw", [dict(max_iterations=0), dict(max_iterations=1), dict(max_iterations=5),
                                dict(max_iterations=777, check_frequency=1),
                                dict(max_iterations=900, check_frequency=250),
                                dict(max_iterations=800, adaptive_step_size=False)])
def test_pdhg_ref_options(pkg, po, lpdata, kw):
    lp = lpdata.random_lp(40, 120, 4, 12)
    okw = {("adaptive" if k == "adaptive_step_size" else k): v for k, v in kw.items()}
    _same(pkg.pdhg_ref_solve(lp, **kw), po.pdhg_ref(lp, **okw))


def test_pdhg_ref_edge_columns(pkg, po):
    # empty column, free column, infinite upper bound, equality row, empty row
    class L:
        pass
    lp = pkg.lpdata.StdLP(
        4, 5, np.array([0, 2, 4, 4, 6], np.int32), np.array([0, 1, 1, 3, 0, 3], np.int32),
        np.array([1.0, 2.0, -1.0, 1.0, 3.0, -2.0]), np.array([4.0, 1.0, 0.0, 2.0]),
        np.array([1.0, -1.0, 0.5, 2.0, -0.3]), np.array([0.0, -np.inf, 0.0, -np.inf, 1.0]),
        np.array([10.0, np.inf, 3.0, np.inf, 1.0]), np.array([0, 1, 0, 0], np.uint8))
    _same(pkg.pdhg_ref_solve(lp, max_iterations=2500), po.pdhg_ref(lp, max_iterations=2500))


@pytest.mark.parametrize("step", [1.0, 10.0])
def test_pdhg_ref_numerical_failure(pkg, po, step):
    # fixed steps too long for the free columns: the iterates blow up and the
    # first check reports NUMERICAL_FAILURE (iterations 0, loop stopped at 100)
    lp = pkg.lpdata.StdLP(
        4, 5, np.array([0, 2, 4, 4, 6], np.int32), np.array([0, 1, 1, 3, 0, 3], np.int32),
        np.array([1.0, 2.0, -1.0, 1.0, 3.0, -2.0]), np.array([4.0, 1.0, 0.0, 2.0]),
        np.array([1.0, -1.0, 0.5, 2.0, -0.3]), np.array([0.0, -np.inf, 0.0, -np.inf, 1.0]),
        np.array([10.0, np.inf, 3.0, np.inf, 1.0]), np.array([0, 1, 0, 0], np.uint8))
    g = pkg.pdhg_ref_solve(lp, max_iterations=3000, adaptive_step_size=False, primal_step_size=step,
                           dual_step_size=step)
    o = po.pdhg_ref(lp, max_iterations=3000, adaptive=False, primal_step=step, dual_step=step)
    assert g.loop == 2
    assert (g.status, g.iterations, g.loop_iterations) == (o.status, o.iterations, o.loop_iterations) == (2, 0, 100)
    assert g.constraint_violation == o.constraint_violation


# ---------------------------------------------------------------- whole drop-in
OPT = {"max_iterations", "adaptive_step_size", "check_frequency"}


@pytest.mark.parametrize("name", sorted(KAT))
def test_dropin_matches_reference_golden(pkg, models, name):
    fm, kw = models[name]
    g = KAT[name]
    print("case:", name, flush=True)
    r = pkg.first_order_solve(fm, engine=0, **kw)
    assert r["status"] == g["status"]
    assert r["iterations"] == g["iterations"]
    assert float.hex(r["objective_value"]) == g["objective"]
    assert float.hex(r["constraint_violation"]) == g["constraint_violation"]
    assert sha(r["x"]) == g["x_sha"] and sha(r["y"]) == g["y_sha"]
    # the drop-in's own verifier verdict on the returned point
    v = pkg.verify_solution(fm, r["x"], reported_objective=r["objective_value"])
    assert r["verify_status"] == v["status"]


def test_dropin_validation_and_infeasible_presolve(pkg, lpdata):
    mb = lpdata.ModelBuilder()
    mb.add_variable(5.0, 2.0, 1.0)  # lb > ub: validation failure
    r = pkg.first_order_solve(mb.build())
    assert r["status"] == 2 and "validation" in r["message"]
    mb = lpdata.ModelBuilder()
    x = mb.add_variable(0.0, 1.0, 1.0)
    mb.add_constraint([(x, 1.0)], lpdata.GE, 3.0)  # singleton row -> lb 3 > ub 1
    r = pkg.first_order_solve(mb.build())
    assert r["status"] == 3


def test_cpp_dropin_program():
    exe = os.path.join(HERE, "cpp", "bin", "test_dropin")
    assert os.path.exists(exe), "tests/cpp not built"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(p.stdout[-4000:])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
