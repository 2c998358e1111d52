"""bo_pdhg_ref_solve_noalloc, the allocation-free variant of bo_pdhg_ref_solve
(every device array from one fixed per-GPU array, the host-side CSC in fixed
static host arrays): called through the C ABI, its results must equal
bo_pdhg_ref_solve's bit for bit -- status, iteration counts, violations, the
iterate choice, the loop tier, every bit of x and y -- on calls in a row with
shrinking LPs (the array keeps the previous call's contents), on both
one-workgroup tiers, and an LP beyond its limit must be refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_present(pkg):
    st, info = pkg.device_info()
    assert st == 0, "no GPU: " + info.status_message.decode()


def _same(a, b):
    assert (a.status, a.iterations, a.loop_iterations, a.used_average, a.loop) == \
           (b.status, b.iterations, b.loop_iterations, b.used_average, b.loop)
    assert a.constraint_violation == b.constraint_violation or (np.isnan(a.constraint_violation) and
                                                                 np.isnan(b.constraint_violation))
    assert a.bound_violation == b.bound_violation
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_noalloc_equals_allocating_entry_on_shrinking_calls(pkg, lpdata):
    # matrix in memory (m + n = 6000 > the LDS tier), then two LDS-resident
    # LPs, each smaller than the one before, then the first again
    cases = [("random 1500x4500", lpdata.random_lp(1500, 4500, 5, 3), 3000),
             ("random 300x1000", lpdata.random_lp(300, 1000, 5, 42), 2000),
             ("random 60x200", lpdata.random_lp(60, 200, 5, 7), 50000),
             ("random 1500x4500 again", lpdata.random_lp(1500, 4500, 5, 3), 3000)]
    tiers = set()
    for name, lp, its in cases:
        print("case:", name, "nnz", lp.nnz, "max_iterations", its, flush=True)
        a = pkg.pdhg_ref_solve(lp, max_iterations=its, noalloc=True)
        b = pkg.pdhg_ref_solve(lp, max_iterations=its)
        _same(a, b)
        tiers.add(a.loop)
        print("  status", a.status, "iterations", a.iterations, "loop", a.loop, flush=True)
    assert tiers == {1, 2}


def test_noalloc_option_variants_and_equality_rows(pkg, lpdata):
    lp = lpdata.staircase_lp(2, 60, 150, 10, seed=5)  # equality rows, scaled columns
    for kw in (dict(max_iterations=1), dict(max_iterations=7, check_frequency=3),
               dict(max_iterations=1500, adaptive_step_size=False), dict(max_iterations=3000)):
        print("case: staircase 120x310", kw, flush=True)
        _same(pkg.pdhg_ref_solve(lp, noalloc=True, **kw), pkg.pdhg_ref_solve(lp, **kw))


def test_noalloc_refuses_lp_beyond_its_limit(pkg, lpdata):
    lp = lpdata.random_lp_sampled(2000, 6000, 5, 11)  # m + n = 8000 > 7680
    with pytest.raises(pkg.BoError) as e:
        pkg.pdhg_ref_solve(lp, max_iterations=10, noalloc=True)
    assert e.value.status == 4 and "m + n <= 7680" in str(e.value)
