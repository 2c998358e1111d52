// lp_model.hpp -- the LPModel input type of the solve() boundary, with the
// public interface of include/bharatopt/lp_model.hpp:14-127 of the reference.
#ifndef BHARATOPT_LP_MODEL_HPP
#define BHARATOPT_LP_MODEL_HPP

#include <cmath>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include <bharatopt/config.hpp>

namespace bharatopt {

enum class ObjectiveSense { MINIMIZE, MAXIMIZE };

enum class ConstraintSense {
    LESS_EQUAL,     // a'x <= rhs
    EQUAL,          // a'x == rhs
    GREATER_EQUAL,  // a'x >= rhs
    RANGED          // rhs <= a'x <= range_upper
};

enum class VariableType { CONTINUOUS, INTEGER, BINARY };

struct Variable {
    std::string name;
    real_t lower_bound{0.0};
    real_t upper_bound{BHARATOPT_INFINITY};
    real_t obj_coeff{0.0};
    VariableType type{VariableType::CONTINUOUS};
    index_t index{-1};

    bool is_free() const { return std::isinf(lower_bound) && std::isinf(upper_bound); }
    bool is_fixed() const { return lower_bound == upper_bound; }
};

struct Constraint {
    std::string name;
    ConstraintSense sense{ConstraintSense::LESS_EQUAL};
    real_t rhs{0.0};
    real_t range_upper{0.0};
    index_t index{-1};
    std::vector<std::pair<index_t, real_t>> terms;  // (variable index, coefficient)
};

class LPModel {
public:
    explicit LPModel(std::string name = "unnamed_model");

    const std::string& name() const { return name_; }
    void set_name(const std::string& name) { name_ = name; }
    ObjectiveSense sense() const { return sense_; }
    void set_sense(ObjectiveSense sense) { sense_ = sense; }
    real_t obj_offset() const { return obj_offset_; }
    void set_obj_offset(real_t offset) { obj_offset_ = offset; }

    index_t add_variable(const std::string& name, real_t lower_bound = 0.0,
                         real_t upper_bound = BHARATOPT_INFINITY, real_t obj_coeff = 0.0,
                         VariableType type = VariableType::CONTINUOUS);
    size_t num_variables() const { return vars_.size(); }
    const Variable& get_variable(index_t index) const;
    Variable& get_variable(index_t index);
    const Variable& get_variable(const std::string& name) const;
    bool has_variable(const std::string& name) const;

    index_t add_constraint(const std::string& name, ConstraintSense sense, real_t rhs,
                           real_t range_upper = 0.0);
    index_t add_constraint(const std::string& name, const std::vector<std::pair<index_t, real_t>>& terms,
                           ConstraintSense sense, real_t rhs, real_t range_upper = 0.0);
    size_t num_constraints() const { return cons_.size(); }
    const Constraint& get_constraint(index_t index) const;
    Constraint& get_constraint(index_t index);
    const Constraint& get_constraint(const std::string& name) const;
    bool has_constraint(const std::string& name) const;

    void set_coeff(index_t constraint_idx, index_t var_idx, real_t value);
    real_t get_coeff(index_t constraint_idx, index_t var_idx) const;

    size_t num_nonzeros() const;
    void clear();

    const std::vector<Variable>& variables() const { return vars_; }
    const std::vector<Constraint>& constraints() const { return cons_; }

private:
    friend class PresolveEngine;  // hands over its working copy when it removed nothing

    std::string name_;
    ObjectiveSense sense_{ObjectiveSense::MINIMIZE};
    real_t obj_offset_{0.0};
    std::vector<Variable> vars_;
    std::vector<Constraint> cons_;
    std::unordered_map<std::string, index_t> var_lookup_;
    std::unordered_map<std::string, index_t> con_lookup_;
};

}  // namespace bharatopt

#endif
