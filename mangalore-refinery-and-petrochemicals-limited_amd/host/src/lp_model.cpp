// lp_model.cpp -- LPModel storage.  Behaviour of the reference's
// src/lp_model.cpp (duplicate names and bad indices throw, set_coeff
// overwrites or appends a non-zero term).
#include <bharatopt/lp_model.hpp>

namespace bharatopt {

namespace {
template <class Vec>
auto& checked_at(Vec& v, index_t i, const char* what) {
    if (i < 0 || static_cast<size_t>(i) >= v.size())
        throw std::out_of_range(std::string(what) + " index out of range: " + std::to_string(i));
    return v[static_cast<size_t>(i)];
}
}  // namespace

LPModel::LPModel(std::string name) : name_(std::move(name)) {}

index_t LPModel::add_variable(const std::string& name, real_t lower_bound, real_t upper_bound,
                              real_t obj_coeff, VariableType type) {
    if (var_lookup_.count(name))
        throw std::invalid_argument("Variable with name '" + name + "' already exists in LPModel.");
    const index_t idx = static_cast<index_t>(vars_.size());
    vars_.push_back(Variable{name, lower_bound, upper_bound, obj_coeff, type, idx});
    var_lookup_.emplace(name, idx);
    return idx;
}

const Variable& LPModel::get_variable(index_t index) const { return checked_at(vars_, index, "Variable"); }
Variable& LPModel::get_variable(index_t index) { return checked_at(vars_, index, "Variable"); }

const Variable& LPModel::get_variable(const std::string& name) const {
    auto it = var_lookup_.find(name);
    if (it == var_lookup_.end()) throw std::invalid_argument("Variable name not found: " + name);
    return vars_[static_cast<size_t>(it->second)];
}

bool LPModel::has_variable(const std::string& name) const { return var_lookup_.count(name) != 0; }

index_t LPModel::add_constraint(const std::string& name, ConstraintSense sense, real_t rhs, real_t range_upper) {
    return add_constraint(name, {}, sense, rhs, range_upper);
}

index_t LPModel::add_constraint(const std::string& name, const std::vector<std::pair<index_t, real_t>>& terms,
                                ConstraintSense sense, real_t rhs, real_t range_upper) {
    if (con_lookup_.count(name))
        throw std::invalid_argument("Constraint with name '" + name + "' already exists in LPModel.");
    for (const auto& t : terms)
        if (t.first < 0 || static_cast<size_t>(t.first) >= vars_.size())
            throw std::out_of_range("Invalid variable index in constraint terms: " + std::to_string(t.first));
    const index_t idx = static_cast<index_t>(cons_.size());
    Constraint c;
    c.name = name;
    c.sense = sense;
    c.rhs = rhs;
    c.range_upper = range_upper;
    c.index = idx;
    c.terms = terms;
    cons_.push_back(std::move(c));
    con_lookup_.emplace(name, idx);
    return idx;
}

const Constraint& LPModel::get_constraint(index_t index) const { return checked_at(cons_, index, "Constraint"); }
Constraint& LPModel::get_constraint(index_t index) { return checked_at(cons_, index, "Constraint"); }

const Constraint& LPModel::get_constraint(const std::string& name) const {
    auto it = con_lookup_.find(name);
    if (it == con_lookup_.end()) throw std::invalid_argument("Constraint name not found: " + name);
    return cons_[static_cast<size_t>(it->second)];
}

bool LPModel::has_constraint(const std::string& name) const { return con_lookup_.count(name) != 0; }

void LPModel::set_coeff(index_t constraint_idx, index_t var_idx, real_t value) {
    Constraint& c = checked_at(cons_, constraint_idx, "Constraint");
    checked_at(vars_, var_idx, "Variable");
    for (auto& t : c.terms)
        if (t.first == var_idx) {
            t.second = value;
            return;
        }
    if (value != 0.0) c.terms.emplace_back(var_idx, value);
}

real_t LPModel::get_coeff(index_t constraint_idx, index_t var_idx) const {
    const Constraint& c = checked_at(cons_, constraint_idx, "Constraint");
    checked_at(vars_, var_idx, "Variable");
    for (const auto& t : c.terms)
        if (t.first == var_idx) return t.second;
    return 0.0;
}

size_t LPModel::num_nonzeros() const {
    size_t count = 0;
    for (const auto& c : cons_)
        for (const auto& t : c.terms) count += (t.second != 0.0);
    return count;
}

void LPModel::clear() {
    *this = LPModel();
}

}  // namespace bharatopt
