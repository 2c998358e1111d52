// first_order_solver.cpp -- the reference's first-order solver classes
// (src/first_order_solver.cpp:14-319) with two back ends:
//
//   * CPUFirstOrderSolver runs the reference's PDHG loop on the host, in
//     FP64, in the reference's operation order: its results are bit-equal to
//     the reference's for the same model (the default backend of
//     FirstOrderSolverOptions, so default-constructed solvers keep the
//     reference's behaviour);
//   * GPUFirstOrderSolver runs the loop on the MI355X through the C ABI
//     (bo_pdhg_ref_solve, bit-equal to the same loop, or bo_pdlp_*).
//
// Both share the stages around the loop, in the reference's order and with
// its messages and result fields: validation gate, presolve, bound-only
// closed form when no rows or columns remain, standard form (<= rows, GE
// rows negated, EQ rows flagged), the loop, postsolve, objective on the
// original model.
//
// Divergences, all deliberate:
//   * GPUFirstOrderSolver without a GPU returns NUMERICAL_FAILURE with a
//     message; it never falls back to the host loop;
//   * RANGED rows become two <= rows (the reference keeps only "<= rhs");
//     their dual is reported as y_upper - y_lower;
//   * engine PDLP stops on the FP64 relative KKT error instead of the
//     reference's stagnation test.
#include <bharatopt/first_order_solver.hpp>
#include <bharatopt/model_validator.hpp>
#include <bharatopt/presolve.hpp>

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>

namespace bharatopt {

namespace {

using clk = std::chrono::high_resolution_clock;

real_t original_objective(const LPModel& model, const std::vector<real_t>& x) {
    real_t obj = model.obj_offset();
    for (size_t j = 0; j < model.num_variables(); ++j) obj += model.variables()[j].obj_coeff * x[j];
    return obj;
}

// min c'x, rows <= b (or == b), l <= x <= u; ranged rows split in two.
struct StandardForm {
    int m_model = 0;  // rows of the (reduced) model
    int m = 0;        // rows after splitting ranged rows
    int n = 0;
    CSRMatrix A;
    std::vector<real_t> b, c, lb, ub;
    std::vector<unsigned char> is_eq;
    std::vector<int> lower_of;  // extra row -> model row (ranged lower sides)
};

StandardForm build_standard_form(const LPModel& model) {
    StandardForm sf;
    const real_t sense_mult = model.sense() == ObjectiveSense::MINIMIZE ? 1.0 : -1.0;
    sf.n = static_cast<int>(model.num_variables());
    sf.m_model = static_cast<int>(model.num_constraints());
    int extra = 0;
    for (const auto& c : model.constraints()) extra += c.sense == ConstraintSense::RANGED;
    sf.m = sf.m_model + extra;
    sf.c.resize(sf.n);
    sf.lb.resize(sf.n);
    sf.ub.resize(sf.n);
    for (int j = 0; j < sf.n; ++j) {
        const Variable& v = model.variables()[static_cast<size_t>(j)];
        sf.c[j] = v.obj_coeff * sense_mult;
        sf.lb[j] = v.lower_bound;
        sf.ub[j] = v.upper_bound;
    }
    sf.b.assign(sf.m, 0.0);
    sf.is_eq.assign(sf.m, 0);
    // CSR straight from the rows, with CSRMatrix::from_coo's rule per row:
    // columns ascending (stable, so repeated columns add in insertion order),
    // repeated columns summed, exact zeros dropped.  The ranged rows' extra
    // "-a'x <= -lower" rows follow the model's rows in constraint order.
    std::vector<index_t> ptr(static_cast<size_t>(sf.m) + 1, 0), idx;
    std::vector<real_t> val;
    size_t total = 0;
    for (const auto& c : model.constraints())
        total += c.terms.size() * (c.sense == ConstraintSense::RANGED ? 2 : 1);
    idx.reserve(total);
    val.reserve(total);
    std::vector<std::pair<index_t, real_t>> row;
    auto append_row = [&](const Constraint& con, real_t mult, int r) {
        row.clear();
        for (const auto& t : con.terms) {
            if (t.first < 0 || t.first >= sf.n)
                throw std::out_of_range("COO triplet index out of range during compression.");
            row.emplace_back(t.first, t.second * mult);
        }
        auto by_col = [](const std::pair<index_t, real_t>& a, const std::pair<index_t, real_t>& b) {
            return a.first < b.first;
        };
        if (!std::is_sorted(row.begin(), row.end(), by_col)) std::stable_sort(row.begin(), row.end(), by_col);
        for (size_t k = 0; k < row.size();) {
            real_t sum = row[k].second;
            size_t q = k + 1;
            for (; q < row.size() && row[q].first == row[k].first; ++q) sum += row[q].second;
            if (sum != 0.0) {
                idx.push_back(row[k].first);
                val.push_back(sum);
            }
            k = q;
        }
        ptr[static_cast<size_t>(r) + 1] = static_cast<index_t>(idx.size());
    };
    for (int i = 0; i < sf.m_model; ++i) {
        const Constraint& con = model.constraints()[static_cast<size_t>(i)];
        real_t mult = 1.0;
        real_t rhs = con.rhs;
        if (con.sense == ConstraintSense::GREATER_EQUAL) {
            mult = -1.0;
            rhs = -rhs;
        } else if (con.sense == ConstraintSense::EQUAL) {
            sf.is_eq[i] = 1;
        } else if (con.sense == ConstraintSense::RANGED) {
            rhs = con.range_upper;  // a'x <= upper here, -a'x <= -lower in the extra row
            sf.b[sf.m_model + static_cast<int>(sf.lower_of.size())] = -con.rhs;
            sf.lower_of.push_back(i);
        }
        sf.b[i] = rhs;
        append_row(con, mult, i);
    }
    for (size_t e = 0; e < sf.lower_of.size(); ++e)
        append_row(model.constraints()[static_cast<size_t>(sf.lower_of[e])], -1.0, sf.m_model + static_cast<int>(e));
    sf.A = CSRMatrix(static_cast<size_t>(sf.m), static_cast<size_t>(sf.n), std::move(ptr), std::move(idx),
                     std::move(val));
    return sf;
}

std::vector<real_t> fold_duals(const StandardForm& sf, const std::vector<real_t>& y) {
    std::vector<real_t> out(y.begin(), y.begin() + sf.m_model);
    for (size_t k = 0; k < sf.lower_of.size(); ++k) out[sf.lower_of[k]] -= y[sf.m_model + k];
    return out;
}

void max_violation(const StandardForm& sf, const std::vector<real_t>& x, real_t& row_viol, real_t& bound_viol) {
    std::vector<real_t> ax = sf.A.multiply(x);
    row_viol = 0.0;
    for (int i = 0; i < sf.m; ++i) {
        const real_t r = ax[i] - sf.b[i];
        const real_t v = sf.is_eq[i] ? std::abs(r) : std::max(0.0, r);
        row_viol = std::max(row_viol, v);
    }
    bound_viol = 0.0;
    for (int j = 0; j < sf.n; ++j) {
        if (x[j] < sf.lb[j] - 1e-12) bound_viol = std::max(bound_viol, sf.lb[j] - x[j]);
        if (x[j] > sf.ub[j] + 1e-12) bound_viol = std::max(bound_viol, x[j] - sf.ub[j]);
    }
}

bo_lp as_bo_lp(const StandardForm& sf) {
    bo_lp lp{};
    lp.m = sf.m;
    lp.n = sf.n;
    lp.nnz = static_cast<long long>(sf.A.nnz());
    lp.row_ptr = sf.A.row_offsets().data();
    lp.col_idx = sf.A.col_indices().data();
    lp.val = sf.A.values().data();
    lp.b = sf.b.data();
    lp.c = sf.c.data();
    lp.lb = sf.lb.data();
    lp.ub = sf.ub.data();
    lp.is_eq = sf.is_eq.data();
    return lp;
}

std::string device_error(bo_status s) {
    const char* e = bo_last_error();
    return std::string("GPU execution failed (status ") + std::to_string(static_cast<int>(s)) + "): " + (e ? e : "");
}

// What the loop stage hands back to the shared pipeline.
enum class LoopEnd { CONVERGED, NOT_CONVERGED, EXPLODED, FAILED };

// The reference's PDHG loop (src/first_order_solver.cpp:103-279) on the host:
// diagonal Pock-Chambolle steps, A'y as a row-order scatter, the primal step
// with the ergodic sum, 2x - x_prev, A x_hat as row dot products, the dual
// step, and every check_freq iterations the violation of x and of the
// clamped average, the explosion test and the feasibility + stagnation stop.
// Every product and sum is one FP64 operation in the reference's order (the
// file is built without contraction), so the results are the reference's
// bits.  x_out receives the chosen iterate (x, or the average when it
// violates less), y the last dual iterate.
LoopEnd host_reference_loop(const StandardForm& sf, const FirstOrderSolverOptions& o, std::vector<real_t>& x_out,
                            std::vector<real_t>& y, FirstOrderSolverResult& result) {
    const size_t m = static_cast<size_t>(sf.m), n = static_cast<size_t>(sf.n);
    const std::vector<index_t>& rp = sf.A.row_offsets();
    const std::vector<index_t>& ci = sf.A.col_indices();
    const std::vector<real_t>& av = sf.A.values();

    std::vector<real_t> col_abs(n, 0.0), row_abs(m, 0.0);
    for (size_t i = 0; i < m; ++i)
        for (index_t k = rp[i]; k < rp[i + 1]; ++k) {
            const real_t a = std::abs(av[static_cast<size_t>(k)]);
            row_abs[i] += a;
            col_abs[static_cast<size_t>(ci[static_cast<size_t>(k)])] += a;
        }
    std::vector<real_t> tau(n, o.primal_step_size), sigma(m, o.dual_step_size);
    if (o.adaptive_step_size) {
        for (size_t j = 0; j < n; ++j) tau[j] = col_abs[j] > 1e-12 ? 0.95 / col_abs[j] : 1.0;
        for (size_t i = 0; i < m; ++i) sigma[i] = row_abs[i] > 1e-12 ? 0.95 / row_abs[i] : 1.0;
    }

    std::vector<real_t> x(n);
    for (size_t j = 0; j < n; ++j) {
        const real_t l = sf.lb[j], u = sf.ub[j];
        real_t start = 0.0;
        if (l > -BHARATOPT_INFINITY && u < BHARATOPT_INFINITY) start = 0.5 * (l + u);
        else if (l > -BHARATOPT_INFINITY) start = l;
        else if (u < BHARATOPT_INFINITY) start = u;
        x[j] = clamp(start, l, u);
    }
    y.assign(m, 0.0);
    std::vector<real_t> x_prev = x, x_hat(n, 0.0), aty(n, 0.0), ax(m, 0.0), x_sum(n, 0.0), x_avg(n, 0.0);

    // max over rows of (a_i x - b_i)+ (|a_i x - b_i| on equality rows)
    auto row_violation = [&](const std::vector<real_t>& v) {
        sf.A.multiply(v.data(), ax.data());
        real_t worst = 0.0;
        for (size_t i = 0; i < m; ++i) {
            const real_t r = ax[i] - sf.b[i];
            worst = std::max(worst, sf.is_eq[i] ? std::abs(r) : std::max(0.0, r));
        }
        return worst;
    };
    auto average = [&](size_t count) {
        for (size_t j = 0; j < n; ++j)
            x_avg[j] = clamp(x_sum[j] / static_cast<real_t>(count), sf.lb[j], sf.ub[j]);
    };

    const size_t check_freq = std::max<size_t>(1, std::min<size_t>(o.check_frequency, 100));
    size_t iter = 0, count = 0;
    bool converged = false;
    while (iter < o.max_iterations) {
        ++iter;
        sf.A.multiply_transpose(y.data(), aty.data());
        real_t dx = 0.0;
        for (size_t j = 0; j < n; ++j) {
            const real_t moved = clamp(x[j] - tau[j] * (sf.c[j] + aty[j]), sf.lb[j], sf.ub[j]);
            dx = std::max(dx, std::abs(moved - x[j]));
            x[j] = moved;
            x_sum[j] += moved;
        }
        ++count;
        for (size_t j = 0; j < n; ++j) x_hat[j] = 2.0 * x[j] - x_prev[j];
        sf.A.multiply(x_hat.data(), ax.data());
        real_t dy = 0.0;
        for (size_t i = 0; i < m; ++i) {
            const real_t step = y[i] + sigma[i] * (ax[i] - sf.b[i]);
            const real_t moved = sf.is_eq[i] ? step : std::max(0.0, step);
            dy = std::max(dy, std::abs(moved - y[i]));
            y[i] = moved;
        }
        x_prev = x;

        if (iter % check_freq != 0 && iter != o.max_iterations) continue;
        const real_t viol_x = row_violation(x);
        real_t viol_bounds = 0.0;
        for (size_t j = 0; j < n; ++j) {
            if (x[j] < sf.lb[j] - 1e-12) viol_bounds = std::max(viol_bounds, sf.lb[j] - x[j]);
            if (x[j] > sf.ub[j] + 1e-12) viol_bounds = std::max(viol_bounds, x[j] - sf.ub[j]);
        }
        average(count);
        const real_t best = std::min(viol_x, row_violation(x_avg));
        result.stats.constraint_violation = best;
        result.stats.bound_violation = viol_bounds;
        if (std::isnan(best) || best > 1e15) return LoopEnd::EXPLODED;
        if (best <= o.feasibility_tolerance && viol_bounds <= o.feasibility_tolerance && dx <= 1e-7 && dy <= 1e-7 &&
            iter >= 100) {
            converged = true;
            break;
        }
    }
    result.stats.iterations = iter;

    if (count > 0) average(count);
    else x_avg = x;
    const bool use_average = !converged && row_violation(x_avg) <= result.stats.constraint_violation;
    x_out = use_average ? x_avg : x;
    return converged ? LoopEnd::CONVERGED : LoopEnd::NOT_CONVERGED;
}

// A loop stage fills x, y and the result's statistics.  A stage that has a
// dual objective (PDLP) stores it in the <=-form minimisation sense in
// `dual_min`; the pipeline converts it to the model's sense.
using LoopStage = std::function<LoopEnd(const StandardForm&, std::vector<real_t>& x, std::vector<real_t>& y,
                                        FirstOrderSolverResult&, real_t& dual_min, std::string& failure)>;

// validate -> presolve -> bound-only form or standard form -> loop ->
// postsolve -> objective, as CPUFirstOrderSolver::solve orders them
// (src/first_order_solver.cpp:14-290).  solve_time_ms covers everything up
// to the end of the loop (the reference stops its clock before postsolve).
// The verifier's judgement of a returned solution (see FirstOrderSolverResult::verification).
void verify_result(const LPModel& model, const FirstOrderSolverOptions& o, FirstOrderSolverResult& r) {
    if (!o.verify_solution || r.primal_solution.empty()) return;
    VerificationOptions vo;
    vo.feasibility_tolerance = o.feasibility_tolerance;
    r.verification = SolutionVerifier(vo).verify(model, r.primal_solution, r.objective_value, vo);
}

FirstOrderSolverResult run_pipeline(const LPModel& model, const FirstOrderSolverOptions& o, clk::time_point t0,
                                    const LoopStage& loop) {
    FirstOrderSolverResult result;
    if (!ModelValidator().validate(model).is_valid()) {
        result.status = FirstOrderSolverStatus::NUMERICAL_FAILURE;
        result.message = "Model validation failed before First-Order LP Solver.";
        return result;
    }

    PresolveResult pre = PresolveEngine().presolve(model);
    if (pre.status == PresolveStatus::INFEASIBLE) {
        result.status = FirstOrderSolverStatus::INFEASIBLE_DETECTED;
        result.message = "Presolve detected infeasibility.";
        return result;
    }
    const LPModel& work = pre.reduced_model;
    const size_t m0 = work.num_constraints(), n0 = work.num_variables();
    const real_t sense_mult = work.sense() == ObjectiveSense::MINIMIZE ? 1.0 : -1.0;

    if (m0 == 0 || n0 == 0) {  // bound-only problem: each variable at its cheaper bound
        std::vector<real_t> x(n0, 0.0);
        for (size_t j = 0; j < n0; ++j) {
            const Variable& v = work.variables()[j];
            const real_t cj = v.obj_coeff * sense_mult;
            const real_t l = v.lower_bound, u = v.upper_bound;
            if (cj > 1e-12) x[j] = (l > -BHARATOPT_INFINITY) ? l : 0.0;
            else if (cj < -1e-12) x[j] = (u < BHARATOPT_INFINITY) ? u : 0.0;
            else x[j] = clamp(0.0, (l > -BHARATOPT_INFINITY) ? l : -1e9, (u < BHARATOPT_INFINITY) ? u : 1e9);
        }
        result.primal_solution = pre.postsolve.recover_solution(x);
        result.status = FirstOrderSolverStatus::OPTIMAL;
        result.objective_value = original_objective(model, result.primal_solution);
        result.stats.primal_objective = result.objective_value;
        result.message = "LP model solved via bound optimization & presolve postsolve.";
        verify_result(model, o, result);
        return result;
    }

    const StandardForm sf = build_standard_form(work);
    std::vector<real_t> x(static_cast<size_t>(sf.n)), y(static_cast<size_t>(sf.m));
    std::string failure;
    real_t dual_min = std::numeric_limits<real_t>::quiet_NaN();
    const LoopEnd end = loop(sf, x, y, result, dual_min, failure);
    if (end == LoopEnd::EXPLODED || end == LoopEnd::FAILED) {
        result.status = FirstOrderSolverStatus::NUMERICAL_FAILURE;
        result.message = end == LoopEnd::EXPLODED
                             ? "Numerical failure: exploded iterate detected during PDHG iterations."
                             : failure;
        return result;
    }
    if (!std::isnan(dual_min)) result.stats.dual_objective = dual_min * sense_mult + work.obj_offset();
    result.stats.solve_time_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    if (end == LoopEnd::CONVERGED) {
        result.status = FirstOrderSolverStatus::OPTIMAL;
        result.message = "PDHG/PDLP-style first-order solver converged to OPTIMAL solution.";
    } else {
        result.status = FirstOrderSolverStatus::MAX_ITERATIONS;
        result.message = "PDHG/PDLP-style first-order solver reached maximum iteration limit.";
    }
    result.primal_solution = pre.postsolve.recover_solution(x);
    result.dual_solution = fold_duals(sf, y);
    result.objective_value = original_objective(model, result.primal_solution);
    result.stats.primal_objective = result.objective_value;
    verify_result(model, o, result);
    return result;
}

// The GPU loop stage: the reference's algorithm (bit-equal) or PDLP.
LoopEnd gpu_loop(const FirstOrderSolverOptions& options, const StandardForm& sf, std::vector<real_t>& x,
                 std::vector<real_t>& y, FirstOrderSolverResult& result, real_t& dual_min, std::string& failure) {
    const bo_lp lp = as_bo_lp(sf);
    if (options.engine == FirstOrderEngine::REFERENCE_PDHG) {
        bo_ref_options o{};
        o.max_iterations = static_cast<long long>(options.max_iterations);
        o.feasibility_tolerance = options.feasibility_tolerance;
        o.primal_step_size = options.primal_step_size;
        o.dual_step_size = options.dual_step_size;
        o.adaptive_step_size = options.adaptive_step_size ? 1 : 0;
        o.check_frequency = static_cast<long long>(options.check_frequency);
        bo_ref_result r{};
        const bo_status s = bo_pdhg_ref_solve(&lp, &o, x.data(), y.data(), &r);
        if (s != BO_SUCCESS) {
            failure = device_error(s);
            return LoopEnd::FAILED;
        }
        result.stats.constraint_violation = r.constraint_violation;
        result.stats.bound_violation = r.bound_violation;
        result.stats.kernel_time_ms = r.kernel_time_ms;
        result.stats.transfer_time_ms = r.transfer_time_ms;
        if (r.status == 2) return LoopEnd::EXPLODED;
        result.stats.iterations = static_cast<size_t>(r.iterations);
        return r.status == 0 ? LoopEnd::CONVERGED : LoopEnd::NOT_CONVERGED;
    }
    // Precision: LPs up to 2^18 nonzeros run with FP64 iterates (they run in
    // one workgroup, where FP64 costs about the same, and badly scaled models
    // such as Netlib's kb2 or share2b do not converge with FP32 iterates);
    // larger ones run FP32 iterates for at most 50,000 iterations and, if that
    // does not converge, go on in FP64 from where it stopped, for the
    // iterations left.
    bo_pdlp_params p;
    bo_pdlp_default_params(&p);
    p.max_iterations = static_cast<long long>(options.max_iterations);
    p.eps_rel_kkt = options.relative_kkt_tolerance;
    p.eps_objective = options.objective_gap_tolerance;
    p.fp64_iterates = lp.nnz <= (1LL << 18) ? 1 : 0;
    const long long budget = p.max_iterations;
    constexpr long long kFp32Iterations = 50000;
    if (!p.fp64_iterates && budget > kFp32Iterations) p.max_iterations = kFp32Iterations;
    bo_pdlp* h = nullptr;
    bo_status s = bo_pdlp_create(&lp, &p, &h);
    bo_pdlp_stats st{};
    if (s == BO_SUCCESS) s = bo_pdlp_solve(h, &st);
    if (s == BO_SUCCESS) s = bo_pdlp_get_solution(h, x.data(), y.data());
    bo_pdlp_destroy(h);
    if (s == BO_SUCCESS && !p.fp64_iterates && (st.status == 1 || st.status == 4) && st.iterations < budget) {
        const long long used = st.iterations;
        p.fp64_iterates = 1;
        p.max_iterations = budget - used;
        h = nullptr;
        s = bo_pdlp_create(&lp, &p, &h);
        if (s == BO_SUCCESS) s = bo_pdlp_set_warm_start(h, x.data(), y.data());
        if (s == BO_SUCCESS) s = bo_pdlp_solve(h, &st);
        if (s == BO_SUCCESS) s = bo_pdlp_get_solution(h, x.data(), y.data());
        bo_pdlp_destroy(h);
        st.iterations += used;
    }
    if (s != BO_SUCCESS) {
        failure = device_error(s);
        return LoopEnd::FAILED;
    }
    if (st.status == 2) {
        failure = "Numerical failure: non-finite KKT error during PDLP iterations.";
        return LoopEnd::FAILED;
    }
    result.stats.iterations = static_cast<size_t>(st.iterations);
    result.stats.relative_kkt = st.rel_kkt;
    result.stats.restarts = static_cast<size_t>(st.restarts);
    dual_min = st.dual_obj;
    result.stats.primal_residual = st.rel_primal;
    result.stats.dual_residual = st.rel_dual;
    result.stats.kernel_time_ms = st.solve_ms;
    real_t rv = 0.0, bv = 0.0;
    max_violation(sf, x, rv, bv);
    result.stats.constraint_violation = rv;
    result.stats.bound_violation = bv;
    return st.status == 0 ? LoopEnd::CONVERGED : LoopEnd::NOT_CONVERGED;
}

}  // namespace

CPUFirstOrderSolver::CPUFirstOrderSolver(FirstOrderSolverOptions options) : options_(options) {}

FirstOrderSolverResult CPUFirstOrderSolver::solve(const LPModel& model) {
    const FirstOrderSolverOptions& o = options_;
    return run_pipeline(model, o, clk::now(),
                        [&o](const StandardForm& sf, std::vector<real_t>& x, std::vector<real_t>& y,
                             FirstOrderSolverResult& r, real_t&, std::string&) {
                            return host_reference_loop(sf, o, x, y, r);
                        });
}

GPUFirstOrderSolver::GPUFirstOrderSolver(FirstOrderSolverOptions options) : options_(options) {}

FirstOrderSolverResult GPUFirstOrderSolver::solve(const LPModel& model) {
    const auto t0 = clk::now();
    GpuBackend& gpu = GpuBackend::instance();
    gpu.initialize();
    if (!gpu.is_available()) {
        FirstOrderSolverResult result;
        result.status = FirstOrderSolverStatus::NUMERICAL_FAILURE;
        result.message = "GPU unavailable: " + gpu.device_info().status_message;
        return result;
    }
    const FirstOrderSolverOptions& o = options_;
    return run_pipeline(model, o, t0,
                        [&o](const StandardForm& sf, std::vector<real_t>& x, std::vector<real_t>& y,
                             FirstOrderSolverResult& r, real_t& dual_min, std::string& failure) {
                            return gpu_loop(o, sf, x, y, r, dual_min, failure);
                        });
}

FirstOrderLPSolver::FirstOrderLPSolver(FirstOrderSolverOptions options) : options_(options) {}

FirstOrderSolverResult FirstOrderLPSolver::solve(const LPModel& model) {
    if (options_.backend_type == FirstOrderBackendType::GPU_FIRST_ORDER) return GPUFirstOrderSolver(options_).solve(model);
    return CPUFirstOrderSolver(options_).solve(model);
}

}  // namespace bharatopt
