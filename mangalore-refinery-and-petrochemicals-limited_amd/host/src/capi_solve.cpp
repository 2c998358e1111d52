// capi_solve.cpp -- bo_first_order_solve: the whole drop-in solve behind a C
// signature (flattened LPModel in, FirstOrderSolverResult fields out), for
// callers that bind C rather than C++ (ctypes, cgo, JNI).
#include <bharatopt/first_order_solver.hpp>
#include <bharatopt/presolve.hpp>
#include <bharatopt/solution_verifier.hpp>

#include <algorithm>
#include <cstring>
#include <exception>
#include <string>

using namespace bharatopt;

namespace bo {
void set_last_error(const std::string& msg);  // bo_last_error() text (backend.hip)
}

namespace {

LPModel model_from_flat(const bo_model* fm) {
    LPModel model("flat_model");
    model.set_sense(fm->sense == 1 ? ObjectiveSense::MAXIMIZE : ObjectiveSense::MINIMIZE);
    model.set_obj_offset(fm->obj_offset);
    for (int j = 0; j < fm->num_vars; ++j) {
        VariableType t = VariableType::CONTINUOUS;
        if (fm->var_type) t = fm->var_type[j] == 2 ? VariableType::BINARY
                            : (fm->var_type[j] == 1 ? VariableType::INTEGER : VariableType::CONTINUOUS);
        model.add_variable("x" + std::to_string(j), fm->lb[j], fm->ub[j], fm->obj[j], t);
    }
    for (int i = 0; i < fm->num_cons; ++i) {
        std::vector<std::pair<index_t, real_t>> terms;
        terms.reserve(static_cast<size_t>(fm->term_ptr[i + 1] - fm->term_ptr[i]));
        for (long long k = fm->term_ptr[i]; k < fm->term_ptr[i + 1]; ++k) terms.emplace_back(fm->term_var[k], fm->term_val[k]);
        static const ConstraintSense senses[] = {ConstraintSense::LESS_EQUAL, ConstraintSense::EQUAL,
                                                 ConstraintSense::GREATER_EQUAL, ConstraintSense::RANGED};
        const int s = fm->con_sense[i];
        model.add_constraint("c" + std::to_string(i), terms, senses[(s >= 0 && s <= 3) ? s : 0], fm->rhs[i],
                             fm->range_upper ? fm->range_upper[i] : 0.0);
    }
    return model;
}

// 0 verified, 1 failed, 2 objective mismatch, 3 dimension mismatch.
int verification_status(const VerificationResult& vr) {
    if (vr.status_message.find("Dimension mismatch") != std::string::npos) return 3;
    if (vr.status_message.find("objective mismatch") != std::string::npos) return 2;
    return vr.verified ? 0 : 1;
}

void set_message(bo_fo_result* r, const std::string& m) {
    std::strncpy(r->message, m.c_str(), sizeof(r->message) - 1);
    r->message[sizeof(r->message) - 1] = 0;
}

}  // namespace

extern "C" bo_status bo_presolve_stats(const bo_model* fm, int* status, int* reduced_vars, int* reduced_cons,
                                       double* obj_offset, long long* reduced_nnz) {
    if (!fm || fm->num_vars < 0 || fm->num_cons < 0) return BO_INVALID_ARGUMENT;
    try {
        PresolveResult pr = PresolveEngine().presolve(model_from_flat(fm));
        if (status) *status = static_cast<int>(pr.status);
        if (reduced_vars) *reduced_vars = static_cast<int>(pr.reduced_model.num_variables());
        if (reduced_cons) *reduced_cons = static_cast<int>(pr.reduced_model.num_constraints());
        if (obj_offset) *obj_offset = pr.reduced_model.obj_offset();
        if (reduced_nnz) *reduced_nnz = static_cast<long long>(pr.reduced_model.num_nonzeros());
        return BO_SUCCESS;
    } catch (...) {
        return BO_INVALID_ARGUMENT;
    }
}

extern "C" bo_status bo_verify_solution(const bo_model* fm, const double* x, long long x_len,
                                        const double* reported, const bo_verify_options* opts,
                                        bo_verify_result* out) {
    if (!fm || !out || fm->num_vars < 0 || fm->num_cons < 0 || x_len < 0 || (x_len > 0 && !x)) return BO_INVALID_ARGUMENT;
    if ((fm->num_vars > 0 && (!fm->lb || !fm->ub || !fm->obj)) ||
        (fm->num_cons > 0 && (!fm->con_sense || !fm->rhs || !fm->term_ptr)))
        return BO_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    try {
        const LPModel model = model_from_flat(fm);
        VerificationOptions vo;
        if (opts) {
            vo.feasibility_tolerance = opts->feasibility_tolerance;
            vo.integrality_tolerance = opts->integrality_tolerance;
            vo.objective_tolerance = opts->objective_tolerance;
            vo.check_objective = opts->check_objective != 0;
        }
        const std::vector<real_t> cand(x, x + x_len);
        const SolutionVerifier sv(vo);
        const VerificationResult vr = reported ? sv.verify(model, cand, *reported, vo) : sv.verify(model, cand, vo);
        out->verified = vr.verified ? 1 : 0;
        out->has_reported_objective = vr.has_reported_objective ? 1 : 0;
        out->objective_reported = vr.objective_reported;
        out->objective_recomputed = vr.objective_recomputed;
        out->objective_difference = vr.objective_difference;
        out->maximum_constraint_violation = vr.maximum_constraint_violation;
        out->maximum_lower_bound_violation = vr.maximum_lower_bound_violation;
        out->maximum_upper_bound_violation = vr.maximum_upper_bound_violation;
        out->maximum_integrality_violation = vr.maximum_integrality_violation;
        out->violated_constraint_count = static_cast<long long>(vr.violated_constraint_count);
        out->violated_bound_count = static_cast<long long>(vr.violated_bound_count);
        out->violated_integrality_count = static_cast<long long>(vr.violated_integrality_count);
        out->variable_count = static_cast<long long>(vr.variable_count);
        out->constraint_count = static_cast<long long>(vr.constraint_count);
        out->status = verification_status(vr);
        out->diagnostics = static_cast<int>(vr.diagnostic_messages.size());
        return BO_SUCCESS;
    } catch (const std::exception& e) {
        bo::set_last_error(std::string("bo_verify_solution: ") + e.what());
        return BO_INVALID_ARGUMENT;
    }
}

extern "C" bo_status bo_first_order_solve(const bo_model* fm, const bo_fo_options* o, bo_fo_result* r,
                                          double* primal_out, double* dual_out) {
    if (!fm || !o || !r || fm->num_vars < 0 || fm->num_cons < 0) return BO_INVALID_ARGUMENT;
    if ((fm->num_vars > 0 && (!fm->lb || !fm->ub || !fm->obj)) ||
        (fm->num_cons > 0 && (!fm->con_sense || !fm->rhs || !fm->term_ptr)))
        return BO_INVALID_ARGUMENT;
    std::memset(r, 0, sizeof(*r));
    try {
        LPModel model = model_from_flat(fm);
        FirstOrderSolverOptions opts;
        opts.max_iterations = static_cast<size_t>(o->max_iterations);
        opts.feasibility_tolerance = o->feasibility_tolerance;
        opts.optimality_tolerance = o->optimality_tolerance;
        opts.primal_step_size = o->primal_step_size;
        opts.dual_step_size = o->dual_step_size;
        opts.adaptive_step_size = o->adaptive_step_size != 0;
        opts.check_frequency = static_cast<size_t>(o->check_frequency);
        opts.backend_type = o->cpu_backend == 1 ? FirstOrderBackendType::CPU_FIRST_ORDER
                                                : FirstOrderBackendType::GPU_FIRST_ORDER;
        opts.engine = o->engine == 1 ? FirstOrderEngine::PDLP : FirstOrderEngine::REFERENCE_PDHG;
        if (o->eps_rel_kkt > 0.0) opts.relative_kkt_tolerance = o->eps_rel_kkt;
        if (o->eps_objective > 0.0) opts.objective_gap_tolerance = o->eps_objective;
        FirstOrderSolverResult res = FirstOrderLPSolver(opts).solve(model);
        r->status = static_cast<int>(res.status);
        r->iterations = static_cast<long long>(res.stats.iterations);
        r->objective_value = res.objective_value;
        r->primal_objective = res.stats.primal_objective;
        r->dual_objective = res.stats.dual_objective;
        r->primal_residual = res.stats.primal_residual;
        r->dual_residual = res.stats.dual_residual;
        r->constraint_violation = res.stats.constraint_violation;
        r->bound_violation = res.stats.bound_violation;
        r->solve_time_ms = res.stats.solve_time_ms;
        r->transfer_time_ms = res.stats.transfer_time_ms;
        r->kernel_time_ms = res.stats.kernel_time_ms;
        r->primal_len = static_cast<long long>(res.primal_solution.size());
        r->dual_len = static_cast<long long>(res.dual_solution.size());
        set_message(r, res.message);
        r->verify_status = res.verification.variable_count > 0 || res.verification.status_message.size()
                               ? verification_status(res.verification)
                               : -1;
        r->verify_max_violation = std::max({res.verification.maximum_constraint_violation,
                                            res.verification.maximum_lower_bound_violation,
                                            res.verification.maximum_upper_bound_violation});
        if (primal_out && !res.primal_solution.empty())
            std::memcpy(primal_out, res.primal_solution.data(), res.primal_solution.size() * sizeof(double));
        if (dual_out && !res.dual_solution.empty())
            std::memcpy(dual_out, res.dual_solution.data(), res.dual_solution.size() * sizeof(double));
        if (opts.backend_type == FirstOrderBackendType::GPU_FIRST_ORDER && !GpuBackend::instance().is_available())
            return BO_GPU_UNAVAILABLE;
        return BO_SUCCESS;
    } catch (const std::exception& e) {
        set_message(r, e.what());
        return BO_INVALID_ARGUMENT;
    }
}
