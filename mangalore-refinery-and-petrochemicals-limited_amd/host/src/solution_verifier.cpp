// solution_verifier.cpp -- bounds, integrality, row activities and objective of
// a candidate, all in FP64 on the original model; the acceptance rules are the
// reference's (src/solution_verifier.cpp:42-215).
#include <bharatopt/solution_verifier.hpp>

#include <algorithm>
#include <cmath>
#include <sstream>

namespace bharatopt {

std::string VerificationResult::to_string() const {
    std::ostringstream os;
    os << "Verification: " << (verified ? "VERIFIED" : "FAILED") << " - " << status_message << "\n"
       << "  max constraint violation " << maximum_constraint_violation << ", max bound violation "
       << std::max(maximum_lower_bound_violation, maximum_upper_bound_violation) << ", objective "
       << objective_recomputed << "\n";
    for (const auto& d : diagnostic_messages) os << "  " << d << "\n";
    return os.str();
}

SolutionVerifier::SolutionVerifier(VerificationOptions options) : default_options_(options) {}

VerificationResult SolutionVerifier::verify(const LPModel& model, const std::vector<real_t>& x,
                                            const VerificationOptions& o) const {
    VerificationResult r;
    r.variable_count = model.num_variables();
    r.constraint_count = model.num_constraints();
    r.verified = true;
    if (x.size() != r.variable_count) {
        r.verified = false;
        r.status_message = "Dimension mismatch between model variables and candidate solution.";
        r.diagnostic_messages.push_back("Candidate size (" + std::to_string(x.size()) + ") != Model variable count (" +
                                        std::to_string(r.variable_count) + ")");
        return r;
    }
    const auto& vars = model.variables();
    for (size_t j = 0; j < vars.size(); ++j) {
        const real_t v = x[j];
        if (!std::isfinite(v)) {
            r.verified = false;
            r.violated_bound_count++;
            r.maximum_lower_bound_violation = r.maximum_upper_bound_violation = BHARATOPT_INFINITY;
            r.diagnostic_messages.push_back("Variable '" + vars[j].name + "' is non-finite (NaN/Inf).");
            continue;
        }
        if (v < vars[j].lower_bound - o.feasibility_tolerance) {
            r.maximum_lower_bound_violation = std::max(r.maximum_lower_bound_violation, vars[j].lower_bound - v);
            r.violated_bound_count++;
            r.verified = false;
            r.diagnostic_messages.push_back("Variable '" + vars[j].name + "' violates lower bound.");
        }
        if (v > vars[j].upper_bound + o.feasibility_tolerance) {
            r.maximum_upper_bound_violation = std::max(r.maximum_upper_bound_violation, v - vars[j].upper_bound);
            r.violated_bound_count++;
            r.verified = false;
            r.diagnostic_messages.push_back("Variable '" + vars[j].name + "' violates upper bound.");
        }
        if (vars[j].type != VariableType::CONTINUOUS) {
            const real_t d = std::abs(v - std::round(v));
            r.maximum_integrality_violation = std::max(r.maximum_integrality_violation, d);
            if (d > o.integrality_tolerance) {
                r.violated_integrality_count++;
                r.verified = false;
                r.diagnostic_messages.push_back("Variable '" + vars[j].name + "' violates integrality.");
            }
        }
    }
    for (const Constraint& c : model.constraints()) {
        real_t act = 0.0;
        for (const auto& t : c.terms)
            if (t.first >= 0 && static_cast<size_t>(t.first) < x.size()) act += t.second * x[static_cast<size_t>(t.first)];
        real_t viol = 0.0;
        const real_t tol = o.feasibility_tolerance;
        switch (c.sense) {
            case ConstraintSense::LESS_EQUAL: if (act > c.rhs + tol) viol = act - c.rhs; break;
            case ConstraintSense::GREATER_EQUAL: if (act < c.rhs - tol) viol = c.rhs - act; break;
            case ConstraintSense::EQUAL: if (std::abs(act - c.rhs) > tol) viol = std::abs(act - c.rhs); break;
            case ConstraintSense::RANGED:
                if (act < c.rhs - tol) viol = c.rhs - act;
                else if (act > c.range_upper + tol) viol = act - c.range_upper;
                break;
        }
        if (viol > 0.0) {
            r.maximum_constraint_violation = std::max(r.maximum_constraint_violation, viol);
            r.violated_constraint_count++;
            r.verified = false;
            r.diagnostic_messages.push_back("Constraint '" + c.name + "' violated by " + std::to_string(viol));
        }
    }
    real_t obj = model.obj_offset();
    for (size_t j = 0; j < vars.size(); ++j) obj += vars[j].obj_coeff * x[j];
    r.objective_recomputed = obj;
    r.status_message = r.verified ? "Candidate solution is mathematically valid and verified against original model."
                                  : "Candidate solution failed independent mathematical verification.";
    return r;
}

VerificationResult SolutionVerifier::verify(const LPModel& model, const std::vector<real_t>& x, real_t reported,
                                            const VerificationOptions& o) const {
    VerificationResult r = verify(model, x, o);
    r.has_reported_objective = true;
    r.objective_reported = reported;
    r.objective_difference = std::abs(r.objective_recomputed - reported);
    if (o.check_objective && r.objective_difference > o.objective_tolerance) {
        r.verified = false;
        r.diagnostic_messages.push_back("Reported objective differs from recomputed objective.");
        r.status_message = "Candidate solution failed verification: objective mismatch.";
    }
    return r;
}

bool SolutionVerifier::verify_solution_feasibility(const LPModel& model, const std::vector<real_t>& x, real_t tol) {
    VerificationOptions o;
    o.feasibility_tolerance = tol;
    o.check_objective = false;
    return SolutionVerifier().verify(model, x, o).verified;
}

}  // namespace bharatopt
