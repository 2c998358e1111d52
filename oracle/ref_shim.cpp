// ref_shim.cpp -- extern "C" entry points over the reference's own sources,
// compiled by oracle/Makefile into oracle/_ref/libbharatopt_ref.so.
// TEST INFRASTRUCTURE ONLY: this is the checker (and the timed CPU baseline),
// never the product path.  The shim only builds LPModel objects from flat
// arrays and forwards to the reference classes; all arithmetic is theirs:
//   FirstOrderLPSolver::solve        include/bharatopt/first_order_solver.hpp:103-113
//   PresolveEngine::presolve          include/bharatopt/presolve.hpp:96-102
//   DualRevisedSimplex / RevisedSimplex (oracle chain of src/conformance_runner.cpp:81-101)
//   SolutionVerifier::verify          include/bharatopt/solution_verifier.hpp:46-65
//   CSRMatrix::from_coo / multiply / multiply_transpose  src/sparse_matrix.cpp
//   MpsParser::parse_string           include/bharatopt/mps_parser.hpp:40-49
#include <bharatopt/dual_revised_simplex.hpp>
#include <bharatopt/mps_parser.hpp>
#include <bharatopt/first_order_solver.hpp>
#include <bharatopt/presolve.hpp>
#include <bharatopt/revised_simplex.hpp>
#include <bharatopt/solution_verifier.hpp>
#include <bharatopt/sparse_matrix.hpp>

#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace bharatopt;

extern "C" {

// Flat LPModel: row-wise terms in insertion order (same layout as bo_model).
struct ref_model {
    int num_vars;
    int num_cons;
    int sense;  // 0 minimize, 1 maximize
    double obj_offset;
    const double* lb;
    const double* ub;
    const double* obj;
    const int* var_type;      // may be NULL (all continuous)
    const int* con_sense;     // 0 LE, 1 EQ, 2 GE, 3 RANGED
    const double* rhs;
    const double* range_upper;  // may be NULL
    const long long* term_ptr;  // num_cons + 1
    const int* term_var;
    const double* term_val;
};

struct ref_fo_options {
    long long max_iterations;
    double feasibility_tolerance;
    double optimality_tolerance;
    double primal_step_size;
    double dual_step_size;
    int adaptive_step_size;
    int backend_gpu;
    long long check_frequency;
};

struct ref_fo_result {
    int status;
    long long iterations;
    double objective_value;
    double constraint_violation;
    double bound_violation;
    double solve_time_ms;
    long long primal_len;  // filled; arrays must hold num_vars / num_cons
    long long dual_len;
    char message[256];
};

}  // extern "C"

static LPModel build_model(const ref_model* fm) {
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
        for (long long k = fm->term_ptr[i]; k < fm->term_ptr[i + 1]; ++k)
            terms.push_back({fm->term_var[k], fm->term_val[k]});
        ConstraintSense s = ConstraintSense::LESS_EQUAL;
        switch (fm->con_sense[i]) {
            case 1: s = ConstraintSense::EQUAL; break;
            case 2: s = ConstraintSense::GREATER_EQUAL; break;
            case 3: s = ConstraintSense::RANGED; break;
            default: break;
        }
        model.add_constraint("c" + std::to_string(i), terms, s, fm->rhs[i],
                             fm->range_upper ? fm->range_upper[i] : 0.0);
    }
    return model;
}

extern "C" {

int ref_fo_solve(const ref_model* fm, const ref_fo_options* o, ref_fo_result* r,
                 double* primal_out, double* dual_out) {
    try {
        LPModel model = build_model(fm);
        FirstOrderSolverOptions opts;
        opts.max_iterations = static_cast<size_t>(o->max_iterations);
        opts.feasibility_tolerance = o->feasibility_tolerance;
        opts.optimality_tolerance = o->optimality_tolerance;
        opts.primal_step_size = o->primal_step_size;
        opts.dual_step_size = o->dual_step_size;
        opts.adaptive_step_size = o->adaptive_step_size != 0;
        opts.backend_type = o->backend_gpu ? FirstOrderBackendType::GPU_FIRST_ORDER
                                           : FirstOrderBackendType::CPU_FIRST_ORDER;
        opts.check_frequency = static_cast<size_t>(o->check_frequency);
        FirstOrderLPSolver solver(opts);
        FirstOrderSolverResult res = solver.solve(model);
        r->status = static_cast<int>(res.status);
        r->iterations = static_cast<long long>(res.stats.iterations);
        r->objective_value = res.objective_value;
        r->constraint_violation = res.stats.constraint_violation;
        r->bound_violation = res.stats.bound_violation;
        r->solve_time_ms = res.stats.solve_time_ms;
        r->primal_len = static_cast<long long>(res.primal_solution.size());
        r->dual_len = static_cast<long long>(res.dual_solution.size());
        std::strncpy(r->message, res.message.c_str(), sizeof(r->message) - 1);
        r->message[sizeof(r->message) - 1] = 0;
        if (primal_out && !res.primal_solution.empty())
            std::memcpy(primal_out, res.primal_solution.data(), res.primal_solution.size() * sizeof(double));
        if (dual_out && !res.dual_solution.empty())
            std::memcpy(dual_out, res.dual_solution.data(), res.dual_solution.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        std::strncpy(r->message, e.what(), sizeof(r->message) - 1);
        r->message[sizeof(r->message) - 1] = 0;
        return -1;
    }
}

// Presolve: returns status and reduced dimensions.
int ref_presolve_dims(const ref_model* fm, int* status, int* red_vars, int* red_cons,
                      double* red_offset, long long* red_nnz) {
    try {
        LPModel model = build_model(fm);
        PresolveEngine pe;
        PresolveResult pr = pe.presolve(model);
        *status = static_cast<int>(pr.status);
        *red_vars = static_cast<int>(pr.reduced_model.num_variables());
        *red_cons = static_cast<int>(pr.reduced_model.num_constraints());
        *red_offset = pr.reduced_model.obj_offset();
        *red_nnz = static_cast<long long>(pr.reduced_model.num_nonzeros());
        return 0;
    } catch (...) {
        return -1;
    }
}

// FP64 simplex oracle chain: DualRevisedSimplex, then RevisedSimplex (Dantzig),
// then RevisedSimplex (Bland), as src/conformance_runner.cpp:81-101 and
// tests/test_conformance_suite.cpp:83-135 do.  Returns which stage answered
// (1 dual, 2 primal Dantzig, 3 primal Bland) via *stage; status 0 = OPTIMAL.
int ref_simplex_solve(const ref_model* fm, double* x_out, double* obj_out, int* stage,
                      long long* iterations) {
    try {
        LPModel model = build_model(fm);
        DualRevisedSimplex dual;
        DualRevisedSimplexResult d = dual.solve(model);
        if (d.status == DualRevisedSimplexStatus::OPTIMAL) {
            *stage = 1;
            *obj_out = d.objective_value;
            *iterations = static_cast<long long>(d.iterations);
            std::memcpy(x_out, d.primal_solution.data(), d.primal_solution.size() * sizeof(double));
            return 0;
        }
        if (d.status == DualRevisedSimplexStatus::INFEASIBLE) { *stage = 1; return 2; }
        if (d.status == DualRevisedSimplexStatus::UNBOUNDED) { *stage = 1; return 1; }
        for (int pass = 0; pass < 2; ++pass) {
            RevisedSimplexOptions ro;
            ro.max_iterations = 20000;
            ro.entering_rule = pass == 0 ? EnteringRule::MOST_NEGATIVE : EnteringRule::BLANDS_RULE;
            RevisedSimplex rs(ro);
            RevisedSimplexResult p = rs.solve(model);
            if (p.status == RevisedSimplexStatus::OPTIMAL) {
                *stage = 2 + pass;
                *obj_out = p.objective_value;
                *iterations = static_cast<long long>(p.iterations);
                std::memcpy(x_out, p.primal_solution.data(), p.primal_solution.size() * sizeof(double));
                return 0;
            }
            if (p.status == RevisedSimplexStatus::INFEASIBLE) { *stage = 2 + pass; return 2; }
            if (p.status == RevisedSimplexStatus::UNBOUNDED) { *stage = 2 + pass; return 1; }
        }
        *stage = 3;
        return 3;
    } catch (...) {
        return -1;
    }
}

// SolutionVerifier::verify (FP64 verifier): returns verified flag and max violations.
int ref_verify(const ref_model* fm, const double* x, double tol, double* max_cons_viol,
               double* max_lb_viol, double* max_ub_viol, double* obj_recomputed) {
    try {
        LPModel model = build_model(fm);
        VerificationOptions vo;
        vo.feasibility_tolerance = tol;
        SolutionVerifier sv(vo);
        std::vector<real_t> xs(x, x + fm->num_vars);
        VerificationResult vr = sv.verify(model, xs, vo);
        *max_cons_viol = vr.maximum_constraint_violation;
        *max_lb_viol = vr.maximum_lower_bound_violation;
        *max_ub_viol = vr.maximum_upper_bound_violation;
        *obj_recomputed = vr.objective_recomputed;
        return vr.verified ? 1 : 0;
    } catch (...) {
        return -1;
    }
}

// SolutionVerifier::verify with every field of VerificationResult; the
// candidate has x_len entries (a length other than num_vars exercises the
// dimension check); reported != NULL uses the overload with a reported
// objective.  status: 0 verified, 1 failed verification, 2 objective
// mismatch, 3 dimension mismatch (from status_message).
struct ref_verify_out {
    int verified, has_reported_objective;
    double objective_reported, objective_recomputed, objective_difference;
    double max_constraint_violation, max_lower_bound_violation, max_upper_bound_violation,
        max_integrality_violation;
    long long violated_constraint_count, violated_bound_count, violated_integrality_count, variable_count,
        constraint_count;
    int status, diagnostics;
};

int ref_verify_full(const ref_model* fm, const double* x, long long x_len, const double* reported, double feas_tol,
                    double int_tol, double obj_tol, int check_objective, ref_verify_out* o) {
    try {
        LPModel model = build_model(fm);
        VerificationOptions vo;
        vo.feasibility_tolerance = feas_tol;
        vo.integrality_tolerance = int_tol;
        vo.objective_tolerance = obj_tol;
        vo.check_objective = check_objective != 0;
        SolutionVerifier sv(vo);
        std::vector<real_t> xs(x, x + x_len);
        VerificationResult vr = reported ? sv.verify(model, xs, *reported, vo) : sv.verify(model, xs, vo);
        o->verified = vr.verified ? 1 : 0;
        o->has_reported_objective = vr.has_reported_objective ? 1 : 0;
        o->objective_reported = vr.objective_reported;
        o->objective_recomputed = vr.objective_recomputed;
        o->objective_difference = vr.objective_difference;
        o->max_constraint_violation = vr.maximum_constraint_violation;
        o->max_lower_bound_violation = vr.maximum_lower_bound_violation;
        o->max_upper_bound_violation = vr.maximum_upper_bound_violation;
        o->max_integrality_violation = vr.maximum_integrality_violation;
        o->violated_constraint_count = static_cast<long long>(vr.violated_constraint_count);
        o->violated_bound_count = static_cast<long long>(vr.violated_bound_count);
        o->violated_integrality_count = static_cast<long long>(vr.violated_integrality_count);
        o->variable_count = static_cast<long long>(vr.variable_count);
        o->constraint_count = static_cast<long long>(vr.constraint_count);
        const std::string& m = vr.status_message;
        o->status = m.find("Dimension mismatch") != std::string::npos ? 3
                    : m.find("objective mismatch") != std::string::npos ? 2
                    : vr.verified ? 0 : 1;
        o->diagnostics = static_cast<int>(vr.diagnostic_messages.size());
        return 0;
    } catch (...) {
        return -1;
    }
}

// CSRMatrix::from_coo + multiply / multiply_transpose.  Returns nnz after
// cleaning, -1 on exceptions.  y has rows entries, yt has cols entries.
long long ref_csr_spmv(int rows, int cols, long long nnz, const int* ri, const int* ci,
                       const double* v, const double* x, const double* xt, double* y,
                       double* yt) {
    try {
        COOMatrix coo(static_cast<size_t>(rows), static_cast<size_t>(cols));
        for (long long k = 0; k < nnz; ++k) coo.add_entry(ri[k], ci[k], v[k]);
        CSRMatrix a = CSRMatrix::from_coo(coo);
        if (x && y) a.multiply(x, y);
        if (xt && yt) a.multiply_transpose(xt, yt);
        return static_cast<long long>(a.nnz());
    } catch (...) {
        return -1;
    }
}

// MpsParser::parse_string, the model flattened like ref_model into caller
// buffers.  dims[0..4] = status, vars, cons, terms (row-wise, insertion
// order), sense.  Call with cap_terms < 0 for the sizes only.  Returns 0, or
// -1 on an exception or buffers too small.
int ref_mps_parse(const char* text, long long* dims, double* offset, long long cap_terms, double* lb, double* ub,
                  double* obj, int* vtype, int* csense, double* rhs, double* rup, long long* tptr, int* tvar,
                  double* tval) {
    try {
        MpsParseResult r = MpsParser().parse_string(text);
        const LPModel& m = r.model;
        long long terms = 0;
        for (size_t i = 0; i < m.num_constraints(); ++i)
            terms += static_cast<long long>(m.get_constraint(static_cast<index_t>(i)).terms.size());
        dims[0] = static_cast<long long>(r.status);
        dims[1] = static_cast<long long>(m.num_variables());
        dims[2] = static_cast<long long>(m.num_constraints());
        dims[3] = terms;
        dims[4] = m.sense() == ObjectiveSense::MAXIMIZE ? 1 : 0;
        *offset = m.obj_offset();
        if (cap_terms < 0) return 0;
        if (cap_terms < terms) return -1;
        for (size_t j = 0; j < m.num_variables(); ++j) {
            const Variable& v = m.get_variable(static_cast<index_t>(j));
            lb[j] = v.lower_bound;
            ub[j] = v.upper_bound;
            obj[j] = v.obj_coeff;
            vtype[j] = v.type == VariableType::BINARY ? 2 : (v.type == VariableType::INTEGER ? 1 : 0);
        }
        long long t = 0;
        tptr[0] = 0;
        for (size_t i = 0; i < m.num_constraints(); ++i) {
            const Constraint& c = m.get_constraint(static_cast<index_t>(i));
            csense[i] = c.sense == ConstraintSense::EQUAL ? 1
                        : c.sense == ConstraintSense::GREATER_EQUAL ? 2
                        : c.sense == ConstraintSense::RANGED ? 3 : 0;
            rhs[i] = c.rhs;
            rup[i] = c.range_upper;
            for (const auto& term : c.terms) {
                tvar[t] = static_cast<int>(term.first);
                tval[t] = term.second;
                ++t;
            }
            tptr[i + 1] = t;
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
