// test_dropin.cpp -- the scenarios of the reference's tests/test_first_order_solver.cpp
// (FirstOrder_01..25) and tests/test_gpu_foundation.cpp (GpuFoundation_01..25),
// run through the drop-in classes.  Models and expected values are the
// reference's.  Every FirstOrder_* scenario runs twice: first exactly as the
// reference's tests run it -- a default-constructed FirstOrderLPSolver, whose
// default backend CPU_FIRST_ORDER is the host loop -- and then, when a GPU is
// visible, with backend GPU_FIRST_ORDER forced.  Without a GPU only the first
// pass and the host-only cases run.
#include <bharatopt/first_order_solver.hpp>
#include <bharatopt/gpu_backend.hpp>
#include <bharatopt/solution_verifier.hpp>
#include <bharatopt/sparse_matrix.hpp>

#include <cmath>
#include <cstring>
#include <vector>

#include "harness.hpp"

using namespace bharatopt;

namespace {
bool g_gpu_pass = false;  // the second pass: GPU_FIRST_ORDER forced

FirstOrderSolverOptions gpu_opts(FirstOrderEngine e = FirstOrderEngine::REFERENCE_PDHG) {
    FirstOrderSolverOptions o;
    o.backend_type = FirstOrderBackendType::GPU_FIRST_ORDER;
    o.engine = e;
    return o;
}
// The options a scenario starts from: the defaults, or the GPU backend.
FirstOrderSolverOptions pass_opts() { return g_gpu_pass ? gpu_opts() : FirstOrderSolverOptions{}; }
FirstOrderSolverResult solve(const LPModel& m) {
    if (!g_gpu_pass) {
        FirstOrderLPSolver solver;  // as the reference's tests construct it
        return solver.solve(m);
    }
    return FirstOrderLPSolver(gpu_opts()).solve(m);
}
FirstOrderSolverResult solve(const LPModel& m, FirstOrderSolverOptions o) { return FirstOrderLPSolver(o).solve(m); }
const double INF = BHARATOPT_INFINITY;
// primal value j, NaN when the solve returned no solution
double xs(const FirstOrderSolverResult& r, size_t j) {
    return j < r.primal_solution.size() ? r.primal_solution[j] : std::nan("");
}
bool same_bits(const std::vector<real_t>& a, const std::vector<real_t>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(real_t)) == 0);
}
}  // namespace

// ---------------------------------------------------------------- FirstOrder_*
TEST_CASE(FirstOrder_01_OneVariableBounded) {
    LPModel m("one_var_bounded");
    index_t x = m.add_variable("x", 2.0, 5.0, 3.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), 2.0, 1e-3);
    EXPECT_NEAR(r.objective_value, 6.0, 1e-3);
}

TEST_CASE(FirstOrder_02_TwoVariableSimple) {
    LPModel m("two_var_simple");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 10.0, 2.0);
    index_t y = m.add_variable("y", 0.0, 10.0, 3.0);
    m.add_constraint("c1", {{x, 1.0}, {y, 1.0}}, ConstraintSense::LESS_EQUAL, 6.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 18.0, 1e-2);
}

TEST_CASE(FirstOrder_03_EqualityConstrained) {
    LPModel m("equality_constrained");
    index_t x = m.add_variable("x", 0.0, 10.0, 1.0);
    index_t y = m.add_variable("y", 0.0, 10.0, 2.0);
    m.add_constraint("c1", {{x, 1.0}, {y, 1.0}}, ConstraintSense::EQUAL, 4.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x) + xs(r, y), 4.0, 1e-3);
    EXPECT_NEAR(r.objective_value, 4.0, 1e-3);
}

TEST_CASE(FirstOrder_04_GreaterThanConstraint) {
    LPModel m("greater_than");
    index_t x = m.add_variable("x", 0.0, 10.0, 4.0);
    m.add_constraint("c1", {{x, 1.0}}, ConstraintSense::GREATER_EQUAL, 3.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), 3.0, 1e-3);
    EXPECT_NEAR(r.objective_value, 12.0, 1e-3);
}

TEST_CASE(FirstOrder_05_MixedConstraints) {
    LPModel m("mixed_constraints");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 10.0, 3.0);
    index_t y = m.add_variable("y", 0.0, 10.0, 5.0);
    m.add_constraint("c1", {{x, 1.0}, {y, 1.0}}, ConstraintSense::LESS_EQUAL, 8.0);
    m.add_constraint("c2", {{x, 1.0}}, ConstraintSense::GREATER_EQUAL, 1.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 38.0, 1e-2);
}

TEST_CASE(FirstOrder_06_LowerBoundVariables) {
    LPModel m("lb_vars");
    index_t x = m.add_variable("x", 3.0, INF, 2.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), 3.0, 1e-3);
}

TEST_CASE(FirstOrder_07_UpperBoundVariables) {
    LPModel m("ub_vars");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 7.0, 5.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), 7.0, 1e-3);
}

TEST_CASE(FirstOrder_08_DoublyBoundedVariables) {
    LPModel m("doubly_bounded");
    index_t x = m.add_variable("x", -2.0, 4.0, 3.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), -2.0, 1e-3);
}

TEST_CASE(FirstOrder_09_FreeVariable) {
    LPModel m("free_variable");
    index_t x = m.add_variable("x", -INF, INF, 1.0);
    m.add_constraint("c1", {{x, 1.0}}, ConstraintSense::GREATER_EQUAL, -5.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), -5.0, 1e-2);
}

TEST_CASE(FirstOrder_10_FixedVariable) {
    LPModel m("fixed_var");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 4.0, 4.0, 5.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), 4.0, 1e-3);
    EXPECT_NEAR(r.objective_value, 20.0, 1e-3);
}

TEST_CASE(FirstOrder_11_ZeroObjective) {
    LPModel m("zero_obj");
    index_t x = m.add_variable("x", 0.0, 5.0, 0.0);
    m.add_constraint("c1", {{x, 1.0}}, ConstraintSense::GREATER_EQUAL, 2.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 0.0, 1e-3);
}

TEST_CASE(FirstOrder_12_RedundantConstraint) {
    LPModel m("redundant_cons");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 10.0, 4.0);
    m.add_constraint("c1", {{x, 1.0}}, ConstraintSense::LESS_EQUAL, 5.0);
    m.add_constraint("c2", {{x, 1.0}}, ConstraintSense::LESS_EQUAL, 8.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), 5.0, 1e-3);
}

TEST_CASE(FirstOrder_13_DegenerateCase) {
    LPModel m("degenerate_case");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 10.0, 2.0);
    index_t y = m.add_variable("y", 0.0, 10.0, 3.0);
    m.add_constraint("c1", {{x, 1.0}, {y, 1.0}}, ConstraintSense::LESS_EQUAL, 0.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 0.0, 1e-3);
}

TEST_CASE(FirstOrder_14_SmallCoefficients) {
    LPModel m("small_coeffs");
    index_t x = m.add_variable("x", 0.0, 10.0, 1e-4);
    m.add_constraint("c1", {{x, 1e-4}}, ConstraintSense::GREATER_EQUAL, 1e-3);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x), 10.0, 1e-2);
}

TEST_CASE(FirstOrder_15_MixedCoefficientScales) {
    LPModel m("mixed_scales");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 10.0, 100.0);
    index_t y = m.add_variable("y", 0.0, 10.0, 1.0);
    m.add_constraint("c1", {{x, 10.0}, {y, 0.1}}, ConstraintSense::LESS_EQUAL, 50.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 500.0, 1e-2);
}

TEST_CASE(FirstOrder_16_PresolveReducedModel) {
    LPModel m("presolve_pipeline");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x1 = m.add_variable("x1", 3.0, 3.0, 2.0);
    index_t x2 = m.add_variable("x2", 0.0, 10.0, 4.0);
    m.add_constraint("c1", {{x1, 1.0}, {x2, 1.0}}, ConstraintSense::LESS_EQUAL, 8.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, x1), 3.0, 1e-3);
    EXPECT_NEAR(xs(r, x2), 5.0, 1e-2);
}

TEST_CASE(FirstOrder_17_RectangularSparseModel) {
    LPModel m("rectangular_sparse");
    index_t x1 = m.add_variable("x1", 0.0, 10.0, 1.0);
    index_t x2 = m.add_variable("x2", 0.0, 10.0, 2.0);
    m.add_variable("x3", 0.0, 10.0, 3.0);
    m.add_constraint("c1", {{x1, 1.0}, {x2, 1.0}}, ConstraintSense::GREATER_EQUAL, 4.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 4.0, 1e-2);
}

TEST_CASE(FirstOrder_18_LargerSparseModel) {
    LPModel m("larger_sparse");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    std::vector<index_t> v;
    for (int j = 0; j < 20; ++j) v.push_back(m.add_variable("x_" + std::to_string(j), 0.0, 5.0, 1.0));
    for (int i = 0; i < 10; ++i)
        m.add_constraint("c_" + std::to_string(i), {{v[i], 1.0}, {v[i + 10], 1.0}}, ConstraintSense::LESS_EQUAL, 6.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 60.0, 1e-1);
}

static LPModel hand_lp() {
    // max 3x + 5y s.t. 2x + y <= 8, x + 2y <= 8; optimum x = y = 8/3, 64/3
    LPModel m("primary_hand_derived_pdhg");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, INF, 3.0);
    index_t y = m.add_variable("y", 0.0, INF, 5.0);
    m.add_constraint("c1", {{x, 2.0}, {y, 1.0}}, ConstraintSense::LESS_EQUAL, 8.0);
    m.add_constraint("c2", {{x, 1.0}, {y, 2.0}}, ConstraintSense::LESS_EQUAL, 8.0);
    return m;
}

TEST_CASE(FirstOrder_19_MandatoryHandLP) {
    LPModel m = hand_lp();
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(xs(r, 0), 8.0 / 3.0, 1e-2);
    EXPECT_NEAR(xs(r, 1), 8.0 / 3.0, 1e-2);
    EXPECT_NEAR(r.objective_value, 64.0 / 3.0, 1e-2);
}

TEST_CASE_GPU(FirstOrder_19b_HandLP_PDLP) {
    LPModel m = hand_lp();
    auto r = solve(m, gpu_opts(FirstOrderEngine::PDLP));
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 64.0 / 3.0, 1e-4);
    EXPECT_TRUE(r.stats.relative_kkt <= 1e-6);
    EXPECT_EQ(r.dual_solution.size(), static_cast<size_t>(2));
    EXPECT_NEAR(r.dual_solution[0], 1.0 / 3.0, 1e-4);  // shadow prices of max 3x+5y in <=-form min
    EXPECT_NEAR(r.dual_solution[1], 7.0 / 3.0, 1e-4);
}

TEST_CASE(FirstOrder_20_InfeasibilityTest) {
    LPModel m("infeasible_pdhg");
    m.add_variable("x", 5.0, 2.0, 1.0);
    auto r = solve(m);
    EXPECT_TRUE(r.status != FirstOrderSolverStatus::OPTIMAL);
}

TEST_CASE(FirstOrder_21_UnboundedTest) {
    LPModel m("unbounded_pdhg");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, INF, 10.0);
    auto r = solve(m);
    EXPECT_TRUE(r.status != FirstOrderSolverStatus::OPTIMAL || xs(r, x) >= 0.0);
}

TEST_CASE(FirstOrder_22_IndependentVerification) {
    LPModel m("verifier_check_pdhg");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 10.0, 4.0);
    m.add_constraint("c1", {{x, 1.0}}, ConstraintSense::LESS_EQUAL, 6.0);
    auto r = solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_TRUE(SolutionVerifier::verify_solution_feasibility(m, r.primal_solution));
}

TEST_CASE_GPU(FirstOrder_23b_ReferenceVsPdlpEngines) {
    LPModel m("cpu_vs_gpu_pdhg");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 10.0, 2.0);
    index_t y = m.add_variable("y", 0.0, 10.0, 3.0);
    m.add_constraint("c1", {{x, 1.0}, {y, 1.0}}, ConstraintSense::LESS_EQUAL, 5.0);
    auto a = solve(m, gpu_opts(FirstOrderEngine::REFERENCE_PDHG));
    auto b = solve(m, gpu_opts(FirstOrderEngine::PDLP));
    EXPECT_EQ(a.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_EQ(b.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(a.objective_value, b.objective_value, 1e-2);
}

TEST_CASE(FirstOrder_24_NumericalStability) {
    LPModel m("stability_pdhg");
    index_t x = m.add_variable("x", 0.0, 10.0, 1.0);
    m.add_constraint("c1", {{x, 1.0}}, ConstraintSense::GREATER_EQUAL, 2.0);
    FirstOrderSolverOptions o = pass_opts();
    o.max_iterations = 1000;
    auto r = solve(m, o);
    EXPECT_TRUE(r.status == FirstOrderSolverStatus::OPTIMAL || r.status == FirstOrderSolverStatus::MAX_ITERATIONS);
}

TEST_CASE(FirstOrder_25_IterationLimitTest) {
    LPModel m("iter_limit_pdhg");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 100.0, 5.0);
    index_t y = m.add_variable("y", 0.0, 100.0, 6.0);
    m.add_constraint("c1", {{x, 1.0}, {y, 1.0}}, ConstraintSense::LESS_EQUAL, 50.0);
    FirstOrderSolverOptions o = pass_opts();
    o.max_iterations = 5;
    auto r = solve(m, o);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::MAX_ITERATIONS);
    EXPECT_EQ(r.stats.iterations, static_cast<size_t>(5));
}

TEST_CASE_GPU(FirstOrder_26_RangedRowHonoured) {
    // 2 <= x + y <= 4, min x + 2y: the reference treats the row as x + y <= 2
    // (lower value); this library keeps both sides, optimum x = 2, y = 0.
    LPModel m("ranged");
    index_t x = m.add_variable("x", 0.0, 10.0, 1.0);
    index_t y = m.add_variable("y", 0.0, 10.0, 2.0);
    m.add_constraint("r", {{x, 1.0}, {y, 1.0}}, ConstraintSense::RANGED, 2.0, 4.0);
    auto r = solve(m, gpu_opts(FirstOrderEngine::PDLP));
    EXPECT_EQ(r.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(r.objective_value, 2.0, 1e-4);
    EXPECT_TRUE(SolutionVerifier::verify_solution_feasibility(m, r.primal_solution, 1e-5));
}

TEST_CASE_GPU(FirstOrder_23_CPUvsGPUComparison) {
    // The reference's test compares the objectives to 1e-2; the GPU engine
    // runs the same loop bit for bit, so the whole results are compared.
    LPModel m("cpu_vs_gpu_pdhg");
    m.set_sense(ObjectiveSense::MAXIMIZE);
    index_t x = m.add_variable("x", 0.0, 10.0, 2.0);
    index_t y = m.add_variable("y", 0.0, 10.0, 3.0);
    m.add_constraint("c1", {{x, 1.0}, {y, 1.0}}, ConstraintSense::LESS_EQUAL, 5.0);
    FirstOrderSolverOptions oc;
    oc.backend_type = FirstOrderBackendType::CPU_FIRST_ORDER;
    auto a = FirstOrderLPSolver(oc).solve(m);
    auto b = solve(m, gpu_opts());
    EXPECT_EQ(a.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_EQ(b.status, FirstOrderSolverStatus::OPTIMAL);
    EXPECT_NEAR(a.objective_value, b.objective_value, 1e-2);
    EXPECT_EQ(a.stats.iterations, b.stats.iterations);
    EXPECT_TRUE(same_bits(a.primal_solution, b.primal_solution));
    EXPECT_TRUE(same_bits(a.dual_solution, b.dual_solution));
}

TEST_CASE_GPU(FirstOrder_27_HostAndGpuLoopsBitEqual) {
    // A 60 x 150 LP whose run ends on the iteration limit with the ergodic
    // average chosen or not: the host loop and the GPU loop agree bit for bit
    // in status, iterations, violations, x and y.
    LPModel m("bit_equal");
    std::vector<index_t> v;
    for (int j = 0; j < 150; ++j)
        v.push_back(m.add_variable("x" + std::to_string(j), 0.0, 20.0, std::sin(0.7 * j + 0.3)));
    for (int i = 0; i < 60; ++i) {
        std::vector<std::pair<index_t, real_t>> t;
        for (int k = 0; k < 5; ++k) t.emplace_back(v[(7 * i + 11 * k) % 150], std::cos(1.3 * i + 0.9 * k));
        m.add_constraint("r" + std::to_string(i), t, i % 3 == 0 ? ConstraintSense::GREATER_EQUAL
                                                                 : ConstraintSense::LESS_EQUAL, 3.0 + (i % 7));
    }
    for (size_t iters : {size_t(1), size_t(37), size_t(2000)}) {
        FirstOrderSolverOptions oc;
        oc.max_iterations = iters;
        oc.check_frequency = 13;
        FirstOrderSolverOptions og = oc;
        og.backend_type = FirstOrderBackendType::GPU_FIRST_ORDER;
        auto a = FirstOrderLPSolver(oc).solve(m);
        auto b = FirstOrderLPSolver(og).solve(m);
        EXPECT_EQ(a.status, b.status);
        EXPECT_EQ(a.stats.iterations, b.stats.iterations);
        EXPECT_TRUE(a.stats.constraint_violation == b.stats.constraint_violation);
        EXPECT_TRUE(a.stats.bound_violation == b.stats.bound_violation);
        EXPECT_TRUE(a.objective_value == b.objective_value);
        EXPECT_TRUE(same_bits(a.primal_solution, b.primal_solution));
        EXPECT_TRUE(same_bits(a.dual_solution, b.dual_solution));
    }
}

TEST_CASE_HOST(FirstOrder_28_GpuSolverWithoutGpuFailsLoudly) {
    // GPUFirstOrderSolver never falls back to the host loop: without a GPU it
    // returns NUMERICAL_FAILURE and no solution.
    if (GpuBackend::instance().initialize() == GPUStatus::SUCCESS) return;  // only meaningful without a GPU
    LPModel m("no_gpu");
    index_t x = m.add_variable("x", 0.0, 10.0, 1.0);
    m.add_constraint("c1", {{x, 1.0}}, ConstraintSense::GREATER_EQUAL, 2.0);
    auto r = FirstOrderLPSolver(gpu_opts()).solve(m);
    EXPECT_EQ(r.status, FirstOrderSolverStatus::NUMERICAL_FAILURE);
    EXPECT_TRUE(r.primal_solution.empty());
}

// ---------------------------------------------------------------- GpuFoundation_*
TEST_CASE_GPU(GpuFoundation_01_GpuAvailability) {
    GPUStatus s = GpuBackend::instance().initialize();
    EXPECT_TRUE(s == GPUStatus::SUCCESS);
    EXPECT_TRUE(GpuBackend::instance().is_available());
}

TEST_CASE_GPU(GpuFoundation_03_DeviceMetadata) {
    GpuBackend::instance().initialize();
    const auto& info = GpuBackend::instance().device_info();
    EXPECT_TRUE(!info.device_name.empty());
    EXPECT_TRUE(info.total_memory_bytes > 0);
}

TEST_CASE_GPU(GpuFoundation_04_DeviceAllocation) {
    GpuBuffer<real_t> b;
    EXPECT_EQ(b.size(), static_cast<size_t>(0));
    EXPECT_EQ(b.allocate(100), GPUStatus::SUCCESS);
    EXPECT_EQ(b.size(), static_cast<size_t>(100));
    EXPECT_EQ(b.bytes(), 100 * sizeof(real_t));
    b.deallocate();
    EXPECT_EQ(b.size(), static_cast<size_t>(0));
}

TEST_CASE_GPU(GpuFoundation_06_RoundTrip) {
    std::vector<real_t> in = {10.0, 20.0, 30.0}, out(3, 0.0);
    GpuBuffer<real_t> d;
    EXPECT_EQ(d.copy_host_to_device(in.data(), 3), GPUStatus::SUCCESS);
    EXPECT_EQ(d.size(), static_cast<size_t>(3));
    EXPECT_EQ(d.copy_device_to_host(out.data(), 3), GPUStatus::SUCCESS);
    EXPECT_NEAR(out[0], 10.0, 1e-12);
    EXPECT_NEAR(out[2], 30.0, 1e-12);
    EXPECT_EQ(d.copy_device_to_host(out.data(), 4), GPUStatus::INVALID_ARGUMENT);
}

static void check_spmv(int& th_fail, COOMatrix coo, std::vector<real_t> x, std::vector<real_t> expect) {
    CSRMatrix a = CSRMatrix::from_coo(coo);
    std::vector<real_t> y;
    real_t err = 1.0;
    GpuTimingResult t;
    EXPECT_EQ(GpuBackend::instance().spmv_with_cpu_comparison(a, x, y, err, &t), GPUStatus::SUCCESS);
    EXPECT_NEAR(err, 0.0, 1e-12);
    EXPECT_EQ(y.size(), expect.size());
    for (size_t i = 0; i < expect.size() && i < y.size(); ++i) EXPECT_NEAR(y[i], expect[i], 1e-12);
    EXPECT_TRUE(t.h2d_time_ms >= 0.0 && t.d2h_time_ms >= 0.0 && t.total_gpu_time_ms >= 0.0);
}

TEST_CASE_GPU(GpuFoundation_07_IdentitySpMV) {
    COOMatrix c(3, 3);
    for (int i = 0; i < 3; ++i) c.add_entry(i, i, 1.0);
    check_spmv(th_fail, c, {5.0, -3.0, 7.0}, {5.0, -3.0, 7.0});
}

TEST_CASE_GPU(GpuFoundation_09_SmallSparseSpMV) {
    COOMatrix c(2, 3);
    c.add_entry(0, 0, 1.0);
    c.add_entry(0, 2, 2.0);
    c.add_entry(1, 1, 3.0);
    check_spmv(th_fail, c, {4.0, 5.0, 6.0}, {16.0, 15.0});
}

TEST_CASE_GPU(GpuFoundation_10_RectangularSpMV) {
    COOMatrix c(4, 2);
    c.add_entry(0, 0, 1.0);
    c.add_entry(1, 1, 2.0);
    c.add_entry(2, 0, 3.0);
    c.add_entry(3, 1, 4.0);
    check_spmv(th_fail, c, {10.0, 20.0}, {10.0, 40.0, 30.0, 80.0});
}

TEST_CASE_GPU(GpuFoundation_12_LargeSparseSpMV) {
    COOMatrix c(1000, 1000);
    for (int i = 0; i < 1000; ++i) c.add_entry(i, i, 2.0);
    check_spmv(th_fail, c, std::vector<real_t>(1000, 3.0), std::vector<real_t>(1000, 6.0));
}

TEST_CASE_GPU(GpuFoundation_19_NumericalTolerance) {
    COOMatrix c(1, 1);
    c.add_entry(0, 0, 1e-8);
    CSRMatrix a = CSRMatrix::from_coo(c);
    std::vector<real_t> y;
    real_t err = 1.0;
    EXPECT_EQ(GpuBackend::instance().spmv_with_cpu_comparison(a, {1e-8}, y, err), GPUStatus::SUCCESS);
    EXPECT_NEAR(y[0], 1e-16, 1e-18);
}

TEST_CASE_GPU(GpuFoundation_13_VectorOps) {
    GpuBuffer<real_t> a(3), b(3), r(3);
    std::vector<real_t> ha = {1.0, 2.0, 3.0}, hb = {4.0, 5.0, 6.0}, hr(3);
    a.copy_host_to_device(ha.data(), 3);
    b.copy_host_to_device(hb.data(), 3);
    auto& g = GpuBackend::instance();
    EXPECT_EQ(g.vector_add(a, b, r, 3), GPUStatus::SUCCESS);
    r.copy_device_to_host(hr.data(), 3);
    EXPECT_NEAR(hr[0], 5.0, 1e-12);
    EXPECT_NEAR(hr[2], 9.0, 1e-12);
    EXPECT_EQ(g.vector_sub(b, a, r, 3), GPUStatus::SUCCESS);
    r.copy_device_to_host(hr.data(), 3);
    EXPECT_NEAR(hr[1], 3.0, 1e-12);
    EXPECT_EQ(g.vector_scale(a, 2.5, r, 3), GPUStatus::SUCCESS);
    r.copy_device_to_host(hr.data(), 3);
    EXPECT_NEAR(hr[2], 7.5, 1e-12);
    std::vector<real_t> hy = {10.0, 20.0, 30.0};
    GpuBuffer<real_t> y(3), out;
    y.copy_host_to_device(hy.data(), 3);
    EXPECT_EQ(g.axpy(3.0, a, y, out, 3), GPUStatus::SUCCESS);  // out auto-allocated
    out.copy_device_to_host(hr.data(), 3);
    EXPECT_NEAR(hr[0], 13.0, 1e-12);
    EXPECT_NEAR(hr[2], 39.0, 1e-12);
    real_t dot = 0.0, nrm = 0.0;
    EXPECT_EQ(g.dot_product(a, b, dot, 3), GPUStatus::SUCCESS);
    EXPECT_NEAR(dot, 32.0, 1e-12);
    std::vector<real_t> hv = {3.0, 4.0, 0.0};
    GpuBuffer<real_t> v(3);
    v.copy_host_to_device(hv.data(), 3);
    EXPECT_EQ(g.norm2(v, nrm, 3), GPUStatus::SUCCESS);
    EXPECT_NEAR(nrm, 5.0, 1e-12);
}

TEST_CASE_GPU(GpuFoundation_20_InvalidArguments) {
    GpuBuffer<real_t> empty;
    GpuMatrixCSR none;
    EXPECT_EQ(GpuBackend::instance().spmv(none, empty, empty), GPUStatus::INVALID_ARGUMENT);
    real_t dot = 0.0;
    EXPECT_EQ(GpuBackend::instance().dot_product(empty, empty, dot, 10), GPUStatus::INVALID_ARGUMENT);
}

TEST_CASE_GPU(GpuFoundation_22_KernelTiming) {
    GpuBuffer<real_t> a(10), b(10), r(10);
    std::vector<real_t> h(10, 1.0);
    a.copy_host_to_device(h.data(), 10);
    b.copy_host_to_device(h.data(), 10);
    GpuTimingResult t;
    EXPECT_EQ(GpuBackend::instance().vector_add(a, b, r, 10, &t), GPUStatus::SUCCESS);
    EXPECT_TRUE(t.kernel_time_ms >= 0.0);
}

int main(int argc, char** argv) {
    const char* filter = argc > 1 ? argv[1] : nullptr;
    int run = 0;
    int failed = th_run(filter, {th::SCENARIO, th::HOST_ONLY}, "default options", run);
    if (GpuBackend::instance().initialize() == GPUStatus::SUCCESS) {
        g_gpu_pass = true;
        failed += th_run(filter, {th::SCENARIO, th::GPU_ONLY}, "GPU_FIRST_ORDER", run);
    } else {
        std::printf("no GPU (%s): GPU pass skipped\n", GpuBackend::instance().device_info().status_message.c_str());
    }
    std::printf("%d/%d passed\n", run - failed, run);
    return failed ? 1 : 0;
}
