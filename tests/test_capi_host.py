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
