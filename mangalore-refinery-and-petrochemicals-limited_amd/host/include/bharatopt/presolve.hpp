// presolve.hpp -- host-side presolve / postsolve stage of the solve path, with
// the interface of the reference's include/bharatopt/presolve.hpp:12-102.
#ifndef BHARATOPT_PRESOLVE_HPP
#define BHARATOPT_PRESOLVE_HPP

#include <string>
#include <unordered_map>
#include <vector>

#include <bharatopt/config.hpp>
#include <bharatopt/lp_model.hpp>

namespace bharatopt {

enum class PresolveStatus { SUCCESS, INFEASIBLE, UNBOUNDED, NO_CHANGE };

enum class ReductionType {
    FIXED_VARIABLE,
    EMPTY_ROW_REDUNDANT,
    EMPTY_ROW_INFEASIBLE,
    EMPTY_COL_FIXED,
    EMPTY_COL_UNBOUNDED,
    SINGLETON_ROW,
    BOUND_TIGHTENING,
    REDUNDANT_CONSTRAINT
};

struct TransformationRecord {
    ReductionType type;
    std::string description;
    std::string entity_name;
    index_t orig_var_idx{-1};
    index_t orig_cons_idx{-1};
    real_t value{0.0};
};

class Postsolve {
public:
    Postsolve() = default;
    void initialize(const LPModel& orig_model);
    void record_fixed_var(index_t orig_var_idx, real_t val);
    void record_var_map(index_t orig_var_idx, index_t reduced_var_idx);
    void record_tightened_bounds(index_t orig_var_idx, real_t lb, real_t ub);
    std::vector<real_t> recover_solution(const std::vector<real_t>& reduced_x) const;
    bool verify_original_feasibility(const std::vector<real_t>& orig_x,
                                     real_t tol = DEFAULT_FEASIBILITY_TOLERANCE) const;
    real_t compute_original_objective(const std::vector<real_t>& orig_x) const;
    const LPModel& original_model() const { return orig_; }

private:
    size_t n_{0};
    LPModel orig_;
    std::vector<real_t> fixed_value_;
    std::vector<bool> fixed_;
    std::vector<index_t> to_reduced_;
    std::vector<real_t> lb_;
    std::vector<real_t> ub_;
};

struct PresolveStatistics {
    size_t original_vars{0};
    size_t original_cons{0};
    size_t reduced_vars{0};
    size_t reduced_cons{0};
    size_t vars_removed{0};
    size_t cons_removed{0};
    double presolve_time_ms{0.0};
    std::vector<TransformationRecord> transformations;
};

struct PresolveOptions {
    bool enable_fixed_variable{true};
    bool enable_empty_row{true};
    bool enable_empty_col{true};
    bool enable_singleton_row{true};
    bool enable_bound_tightening{true};
    bool enable_scaling{true};  // accepted for compatibility; scaling happens on the GPU
};

struct PresolveResult {
    PresolveStatus status{PresolveStatus::NO_CHANGE};
    LPModel reduced_model;
    Postsolve postsolve;
    PresolveStatistics stats;
    std::string to_string() const;
};

class PresolveEngine {
public:
    PresolveEngine() = default;
    PresolveResult presolve(const LPModel& original_model, const PresolveOptions& options = {}) const;
};

}  // namespace bharatopt

#endif
