// first_order_solver.hpp -- the drop-in boundary: the reference's
// include/bharatopt/first_order_solver.hpp:15-113 with GPUFirstOrderSolver
// running on the MI355X.  Added fields are appended with defaults, so code
// written against the original header compiles and behaves unchanged.
#ifndef BHARATOPT_FIRST_ORDER_SOLVER_HPP
#define BHARATOPT_FIRST_ORDER_SOLVER_HPP

#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include <bharatopt/config.hpp>
#include <bharatopt/gpu_backend.hpp>
#include <bharatopt/lp_model.hpp>
#include <bharatopt/solution_verifier.hpp>
#include <bharatopt/sparse_matrix.hpp>

namespace bharatopt {

enum class FirstOrderSolverStatus { OPTIMAL, MAX_ITERATIONS, NUMERICAL_FAILURE, INFEASIBLE_DETECTED, UNBOUNDED_DETECTED, INTERRUPTED };

enum class FirstOrderBackendType { CPU_FIRST_ORDER, GPU_FIRST_ORDER };

// Which GPU algorithm GPUFirstOrderSolver runs (new; default keeps the
// reference's algorithm and results).
enum class FirstOrderEngine {
    REFERENCE_PDHG,  // the reference's diagonal PDHG and stopping rule, FP64, bit-equal results
    PDLP             // restarted PDHG, FP32 iterates, FP64 checks, stops at relative KKT tolerance
};

struct FirstOrderSolverOptions {
    size_t max_iterations{50000};
    real_t feasibility_tolerance{DEFAULT_FEASIBILITY_TOLERANCE};
    real_t optimality_tolerance{DEFAULT_OPTIMALITY_TOLERANCE};
    real_t primal_step_size{0.01};
    real_t dual_step_size{0.01};
    bool adaptive_step_size{true};
    FirstOrderBackendType backend_type{FirstOrderBackendType::CPU_FIRST_ORDER};
    size_t check_frequency{100};
    // --- additions ---
    FirstOrderEngine engine{FirstOrderEngine::REFERENCE_PDHG};
    real_t relative_kkt_tolerance{1e-6};  // PDLP engine stopping tolerance
    real_t objective_gap_tolerance{0.0};  // PDLP engine: > 0 also requires |c'x - dual obj| <= this (1 + |c'x|)
    bool verify_solution{true};  // run SolutionVerifier on every returned primal solution (result.verification)
};

struct FirstOrderSolverStats {
    size_t iterations{0};
    real_t primal_objective{0.0};
    real_t dual_objective{0.0};
    real_t primal_residual{0.0};
    real_t dual_residual{0.0};
    real_t constraint_violation{0.0};
    real_t bound_violation{0.0};
    double solve_time_ms{0.0};
    double transfer_time_ms{0.0};
    double kernel_time_ms{0.0};
    // --- additions (PDLP engine) ---
    real_t relative_kkt{0.0};
    size_t restarts{0};
};

struct FirstOrderSolverResult {
    FirstOrderSolverStatus status{FirstOrderSolverStatus::NUMERICAL_FAILURE};
    std::vector<real_t> primal_solution;  // original space
    std::vector<real_t> dual_solution;    // presolved <=-form constraints (reference convention)
    real_t objective_value{0.0};
    FirstOrderSolverStats stats;
    std::string message;
    // --- additions ---
    // SolutionVerifier::verify of primal_solution against the original model
    // with its reported objective (feasibility tolerance of the options);
    // informational: it does not change status.  variable_count == 0 when no
    // solution was returned or verify_solution is off.
    VerificationResult verification;
};

class IFirstOrderLPSolver {
public:
    virtual ~IFirstOrderLPSolver() = default;
    virtual FirstOrderSolverResult solve(const LPModel& model) = 0;
};

// The reference's PDHG loop on the host, FP64, bit-equal to the reference
// (src/first_order_solver.cpp:14-290); the default backend.
class CPUFirstOrderSolver : public IFirstOrderLPSolver {
public:
    explicit CPUFirstOrderSolver(FirstOrderSolverOptions options = {});
    FirstOrderSolverResult solve(const LPModel& model) override;
    const FirstOrderSolverOptions& options() const { return options_; }

private:
    FirstOrderSolverOptions options_;
};

// PDHG on the MI355X: validate -> presolve -> standard form -> GPU -> postsolve.
// Without a GPU it reports NUMERICAL_FAILURE (no host fallback).
class GPUFirstOrderSolver : public IFirstOrderLPSolver {
public:
    explicit GPUFirstOrderSolver(FirstOrderSolverOptions options = {});
    FirstOrderSolverResult solve(const LPModel& model) override;
    const FirstOrderSolverOptions& options() const { return options_; }

private:
    FirstOrderSolverOptions options_;
};

class FirstOrderLPSolver {
public:
    explicit FirstOrderLPSolver(FirstOrderSolverOptions options = {});
    FirstOrderSolverResult solve(const LPModel& model);
    const FirstOrderSolverOptions& options() const { return options_; }
    void set_options(const FirstOrderSolverOptions& options) { options_ = options; }

private:
    FirstOrderSolverOptions options_;
};

}  // namespace bharatopt

#endif
