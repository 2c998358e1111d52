"""Host-side checks of libbo_hip.so that need no GPU: it loads, exports every
symbol include/bo_capi.h declares, its host presolve reproduces the
reference's PresolveEngine, and without a GPU every compute entry point
fails loudly (BO_GPU_UNAVAILABLE) instead of computing on the CPU."""
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "include", "bo_capi.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(bo_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_list_agree(pkg):
    assert declared_symbols() == sorted(pkg.EXPORTS)


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_status_ordinals_match_gpustatus(pkg):
    # GPUStatus order (include/bharatopt/gpu_backend.hpp:15-22)
    assert pkg.STATUS_NAMES == ["SUCCESS", "GPU_UNAVAILABLE", "GPU_UNINITIALIZED",
                                "OUT_OF_MEMORY", "INVALID_ARGUMENT", "EXECUTION_FAILED"]


def _no_gpu(pkg):
    st, _ = pkg.device_info()
    return st != 0


def test_compute_entry_points_fail_without_gpu(pkg, lpdata):
    if not _no_gpu(pkg):
        pytest.skip("a GPU is visible")
    lp = lpdata.random_lp(10, 30, 3, 1)
    with pytest.raises(pkg.BoError) as e:
        pkg.pdhg_ref_solve(lp, max_iterations=10)
    assert e.value.status == 1
    with pytest.raises(pkg.BoError) as e:
        pkg.Pdlp(lp)
    assert e.value.status == 1
    with pytest.raises(pkg.BoError):
        pkg.first_order_solve(lp.to_flat())
    with pytest.raises(pkg.BoError) as e:
        pkg.DeviceBuffer(8)
    assert e.value.status == 1


def _random_presolve_model(lpdata, rng):
    """Models that make every presolve pass fire."""
    mb = lpdata.ModelBuilder(sense=int(rng.integers(0, 2)))
    n = int(rng.integers(5, 14))
    for j in range(n):
        kind = rng.integers(0, 6)
        if kind == 0:
            v = float(rng.integers(-2, 3))
            mb.add_variable(v, v, float(rng.standard_normal()))                 # fixed
        elif kind == 1:
            mb.add_variable(0.0, 1.0, float(rng.standard_normal()), vtype=2)    # binary
        elif kind == 2:
            mb.add_variable(-np.inf, np.inf, float(rng.standard_normal()))      # free
        else:
            mb.add_variable(0.0, float(rng.integers(1, 20)), float(rng.standard_normal()))
    for i in range(int(rng.integers(2, 10))):
        k = int(rng.integers(0, 4))
        cols = rng.choice(n, size=min(k, n), replace=False) if k else []
        terms = [(int(j), float(np.round(rng.standard_normal() * 4, 2)) or 1.0) for j in cols]
        sense = int(rng.integers(0, 4))
        rhs = float(rng.integers(-3, 15))
        mb.add_constraint(terms, sense, rhs, rhs + float(rng.integers(0, 5)))
    return mb.build()


def test_presolve_matches_reference(pkg, ref, lpdata):
    rng = np.random.default_rng(2024)
    seen = set()
    for _ in range(400):
        fm = _random_presolve_model(lpdata, rng)
        a = pkg.presolve_stats(fm)
        b = ref.ref_presolve_dims(fm)
        assert a["status"] == b["status"]
        if b["status"] in (0, 3):  # SUCCESS / NO_CHANGE: the reduced model must agree
            assert (a["vars"], a["cons"], a["nnz"]) == (b["vars"], b["cons"], b["nnz"])
            assert a["offset"] == b["offset"]
        seen.add(b["status"])
    assert seen == {0, 1, 2, 3}  # every outcome exercised


def _malformed_csr_cases(lpdata):
    """(label, StdLP, nnz override) with a CSR no entry point may read."""
    base = lpdata.random_lp(6, 9, 2, 3)
    out = []
    bad = base.row_ptr.copy()
    bad[3], bad[4] = bad[4], bad[3] - 1      # decreasing offsets
    out.append(("non-monotone row_ptr", bad, None))
    bad = base.row_ptr.copy()
    bad[0] = -2                              # negative start
    out.append(("negative row_ptr[0]", bad, None))
    bad = base.row_ptr.copy()
    bad[0] = 1                               # does not start at 0
    out.append(("row_ptr[0] != 0", bad, None))
    out.append(("nnz larger than row_ptr[m]", base.row_ptr.copy(), base.nnz + 5))
    out.append(("nnz smaller than row_ptr[m]", base.row_ptr.copy(), base.nnz - 1))
    out.append(("negative nnz", base.row_ptr.copy(), -1))
    return base, out


def test_malformed_csr_rejected_before_any_device_work(pkg, lpdata):
    """Every entry point that takes a bo_lp validates its CSR first and
    returns BO_INVALID_ARGUMENT (4) -- with or without a GPU -- instead of
    reading col_idx / val through bad offsets."""
    import ctypes as C
    L = pkg.lib()
    base, cases = _malformed_csr_cases(lpdata)
    for label, rp, nnz in cases:
        lp = pkg.lpdata.StdLP(base.m, base.n, rp, base.col_idx, base.val, base.b, base.c, base.lb, base.ub,
                              base.is_eq)
        h = pkg._LpHolder(lp)
        h.s.nnz = base.nnz if nnz is None else nnz
        o = pkg.RefOptions(100, 1e-7, 0.01, 0.01, 1, 100)
        r = pkg.RefResult()
        x = np.zeros(base.n)
        y = np.zeros(base.m)
        assert L.bo_pdhg_ref_solve(C.byref(h.s), C.byref(o), pkg._p(x), pkg._p(y), C.byref(r)) == 4, label
        p = pkg.default_pdlp_params()
        out = C.c_void_p()
        assert L.bo_pdlp_create(C.byref(h.s), C.byref(p), C.byref(out)) == 4, label
        assert not out.value
        arr = (pkg.BoLp * 1)(h.s)
        assert L.bo_pdlp_batch_create(arr, 1, C.byref(p), C.byref(out)) == 4, label
        assert L.bo_pdlp_family_create(C.byref(h.s), 4, C.byref(p), C.byref(out)) == 4, label
    # a column index outside [0, n)
    bad_ci = base.col_idx.copy()
    bad_ci[2] = base.n
    lp = pkg.lpdata.StdLP(base.m, base.n, base.row_ptr, bad_ci, base.val, base.b, base.c, base.lb, base.ub,
                          base.is_eq)
    h = pkg._LpHolder(lp)
    r = pkg.RefResult()
    o = pkg.RefOptions(100, 1e-7, 0.01, 0.01, 1, 100)
    assert L.bo_pdhg_ref_solve(C.byref(h.s), C.byref(o), pkg._p(np.zeros(base.n)), pkg._p(np.zeros(base.m)),
                               C.byref(r)) == 4


def test_noalloc_variant_refuses_oversized_lp_before_device_work(pkg, lpdata):
    """bo_pdhg_ref_solve_noalloc checks its fixed limits (m + n <= 7680,
    nnz <= 65536) before anything touches a device, with or without a GPU."""
    lp = lpdata.random_lp_sampled(2000, 6000, 5, 11)
    with pytest.raises(pkg.BoError) as e:
        pkg.pdhg_ref_solve(lp, max_iterations=10, noalloc=True)
    assert e.value.status == 4 and "m + n <= 7680" in str(e.value)
