// model_validator.cpp -- which models the solver refuses.  Same error/warning
// classification as the reference's ModelValidator::validate
// (src/model_validator.cpp:70-205): errors for NaN/Inf data, lb > ub, bad binary
// bounds, duplicate names, bad variable references and inconsistent ranges;
// warnings for empty names, explicit zeros, repeated terms and non-integral
// integer bounds.
#include <bharatopt/model_validator.hpp>

#include <cmath>
#include <sstream>
#include <unordered_set>

namespace bharatopt {

std::string ValidationIssue::to_string() const {
    std::ostringstream os;
    os << (severity == IssueSeverity::ERROR ? "[ERROR] " : "[WARNING] ") << category << ": " << message;
    if (!entity_name.empty()) os << " (" << entity_name << ")";
    return os.str();
}

std::vector<ValidationIssue> ValidationResult::errors() const {
    std::vector<ValidationIssue> out;
    for (const auto& i : issues_)
        if (i.severity == IssueSeverity::ERROR) out.push_back(i);
    return out;
}

std::vector<ValidationIssue> ValidationResult::warnings() const {
    std::vector<ValidationIssue> out;
    for (const auto& i : issues_)
        if (i.severity == IssueSeverity::WARNING) out.push_back(i);
    return out;
}

void ValidationResult::add_issue(IssueSeverity severity, const std::string& category, const std::string& message,
                                 const std::string& entity_name, index_t entity_index) {
    issues_.push_back({severity, category, message, entity_name, entity_index});
    if (severity == IssueSeverity::ERROR) ++errors_;
    else ++warnings_;
}

std::string ValidationResult::to_string() const {
    std::ostringstream os;
    os << "Validation Status: " << (is_valid() ? "VALID" : "INVALID") << " (" << errors_ << " errors, " << warnings_
       << " warnings)\n";
    for (const auto& i : issues_) os << "  " << i.to_string() << "\n";
    return os.str();
}

ValidationResult ModelValidator::validate(const LPModel& model) const {
    ValidationResult r;
    if (model.name().empty()) r.add_warning("STRUCTURE", "Model name is empty.");
    const real_t off = model.obj_offset();
    if (std::isnan(off)) r.add_error("OBJECTIVE", "Objective constant offset is NaN.");
    else if (std::isinf(off)) r.add_error("OBJECTIVE", "Objective constant offset is infinite.");

    const auto& vars = model.variables();
    std::unordered_set<std::string> names;
    names.reserve(vars.size());
    for (size_t j = 0; j < vars.size(); ++j) {
        const Variable& v = vars[j];
        const index_t id = static_cast<index_t>(j);
        if (!names.insert(v.name).second)
            r.add_error("VARIABLE", "Duplicate variable name '" + v.name + "'.", v.name, id);
        const bool lnan = std::isnan(v.lower_bound), unan = std::isnan(v.upper_bound);
        if (lnan) r.add_error("VARIABLE", "Variable lower bound is NaN.", v.name, id);
        if (unan) r.add_error("VARIABLE", "Variable upper bound is NaN.", v.name, id);
        if (!lnan && !unan && v.lower_bound > v.upper_bound)
            r.add_error("VARIABLE", "Lower bound > Upper bound.", v.name, id);
        if (std::isnan(v.obj_coeff)) r.add_error("OBJECTIVE", "Objective coefficient is NaN.", v.name, id);
        else if (std::isinf(v.obj_coeff)) r.add_error("OBJECTIVE", "Objective coefficient is infinite.", v.name, id);
        if (v.type == VariableType::BINARY) {
            if (v.lower_bound < 0.0 || v.upper_bound > 1.0)
                r.add_error("VARIABLE", "Binary variable bounds must be within [0, 1].", v.name, id);
            else if (v.lower_bound != 0.0 || v.upper_bound != 1.0)
                r.add_warning("VARIABLE", "Binary variable has non-standard restricted bounds.", v.name, id);
        } else if (v.type == VariableType::INTEGER) {
            if (!std::isinf(v.lower_bound) && v.lower_bound != std::floor(v.lower_bound))
                r.add_warning("VARIABLE", "Integer variable has non-integer lower bound.", v.name, id);
            if (!std::isinf(v.upper_bound) && v.upper_bound != std::ceil(v.upper_bound))
                r.add_warning("VARIABLE", "Integer variable has non-integer upper bound.", v.name, id);
        }
    }

    const auto& cons = model.constraints();
    // repeated terms: the row that last used each variable (one array for all
    // rows instead of a set per row)
    std::vector<index_t> last_row(vars.size(), -1);
    std::unordered_set<std::string> cnames;
    cnames.reserve(cons.size());
    for (size_t i = 0; i < cons.size(); ++i) {
        const Constraint& c = cons[i];
        const index_t id = static_cast<index_t>(i);
        if (!cnames.insert(c.name).second)
            r.add_error("CONSTRAINT", "Duplicate constraint name '" + c.name + "'.", c.name, id);
        if (std::isnan(c.rhs)) r.add_error("CONSTRAINT", "Constraint RHS is NaN.", c.name, id);
        else if (std::isinf(c.rhs)) r.add_error("CONSTRAINT", "Constraint RHS is infinite.", c.name, id);
        if (c.sense == ConstraintSense::RANGED) {
            if (std::isnan(c.range_upper))
                r.add_error("CONSTRAINT", "Ranged constraint upper bound is NaN.", c.name, id);
            else if (std::isinf(c.range_upper))
                r.add_error("CONSTRAINT", "Ranged constraint upper bound is infinite.", c.name, id);
            else if (!std::isnan(c.rhs) && c.rhs > c.range_upper)
                r.add_error("CONSTRAINT", "Ranged constraint lower bound > upper range.", c.name, id);
        }
        for (const auto& t : c.terms) {
            if (t.first < 0 || static_cast<size_t>(t.first) >= vars.size()) {
                r.add_error("COEFFICIENT",
                            "Constraint references invalid variable index #" + std::to_string(t.first) + ".", c.name,
                            id);
                continue;
            }
            if (std::isnan(t.second)) r.add_error("COEFFICIENT", "Constraint term coefficient is NaN.", c.name, id);
            else if (std::isinf(t.second))
                r.add_error("COEFFICIENT", "Constraint term coefficient is infinite.", c.name, id);
            else if (t.second == 0.0)
                r.add_warning("COEFFICIENT", "Constraint contains explicit zero coefficient.", c.name, id);
            index_t& last = last_row[static_cast<size_t>(t.first)];
            const bool repeated = last == id;
            last = id;
            if (repeated)
                r.add_warning("COEFFICIENT",
                              "Constraint contains duplicate terms for variable #" + std::to_string(t.first) + ".",
                              c.name, id);
        }
    }
    return r;
}

}  // namespace bharatopt
