// presolve.cpp -- reductions applied on the host before the GPU solve.
//
// The reductions, their order, tolerances and floating-point updates are the
// reference's (src/presolve.cpp:157-517), because the drop-in must hand the
// GPU exactly the reduced model the reference hands its CPU loop:
//   A fixed variables (|l-u| <= 1e-12) substituted out, rhs -= a*v in
//     ascending variable order;   B empty rows dropped or proven infeasible;
//   C empty columns fixed at the cost-preferred finite bound (UNBOUNDED stops
//     presolve);  D singleton rows turned into bounds;  E big-M coefficients of
//     binaries in <= rows tightened;  repeated while anything changed, at most
//     100 rounds;  F the reduced model keeps active variables and non-empty
//     active rows in original order.
// The substitution of pass A walks each row once per round instead of once
// per fixed variable; subtractions are applied in ascending variable order,
// so every rhs is rounded exactly as in the reference.
#include <bharatopt/presolve.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <sstream>

namespace bharatopt {

// ---------------------------------------------------------------- postsolve
void Postsolve::initialize(const LPModel& orig_model) {
    orig_ = orig_model;
    n_ = orig_model.num_variables();
    fixed_value_.assign(n_, 0.0);
    fixed_.assign(n_, false);
    to_reduced_.assign(n_, -1);
    lb_.resize(n_);
    ub_.resize(n_);
    for (size_t j = 0; j < n_; ++j) {
        lb_[j] = orig_model.variables()[j].lower_bound;
        ub_[j] = orig_model.variables()[j].upper_bound;
    }
}

void Postsolve::record_fixed_var(index_t j, real_t val) {
    if (j < 0 || static_cast<size_t>(j) >= n_) return;
    fixed_[static_cast<size_t>(j)] = true;
    fixed_value_[static_cast<size_t>(j)] = val;
    to_reduced_[static_cast<size_t>(j)] = -1;
}

void Postsolve::record_var_map(index_t j, index_t r) {
    if (j >= 0 && static_cast<size_t>(j) < n_) to_reduced_[static_cast<size_t>(j)] = r;
}

void Postsolve::record_tightened_bounds(index_t j, real_t lb, real_t ub) {
    if (j < 0 || static_cast<size_t>(j) >= n_) return;
    lb_[static_cast<size_t>(j)] = lb;
    ub_[static_cast<size_t>(j)] = ub;
}

std::vector<real_t> Postsolve::recover_solution(const std::vector<real_t>& reduced_x) const {
    std::vector<real_t> x(n_, 0.0);
    const bool minimize = orig_.sense() == ObjectiveSense::MINIMIZE;
    for (size_t j = 0; j < n_; ++j) {
        if (fixed_[j]) {
            x[j] = fixed_value_[j];
            continue;
        }
        const index_t r = to_reduced_[j];
        if (r >= 0 && static_cast<size_t>(r) < reduced_x.size()) {
            x[j] = reduced_x[static_cast<size_t>(r)];
            continue;
        }
        // not in the reduced model: the bound its cost prefers, else any finite bound
        const real_t cost = minimize ? orig_.variables()[j].obj_coeff : -orig_.variables()[j].obj_coeff;
        const bool lf = !std::isinf(lb_[j]), uf = !std::isinf(ub_[j]);
        if (cost > 0.0) x[j] = lf ? lb_[j] : (uf ? ub_[j] : 0.0);
        else if (cost < 0.0) x[j] = uf ? ub_[j] : (lf ? lb_[j] : 0.0);
        else x[j] = lf ? lb_[j] : (uf ? ub_[j] : 0.0);
    }
    return x;
}

bool Postsolve::verify_original_feasibility(const std::vector<real_t>& x, real_t tol) const {
    if (x.size() != n_) return false;
    for (size_t j = 0; j < n_; ++j) {
        const Variable& v = orig_.variables()[j];
        if (x[j] < v.lower_bound - tol || x[j] > v.upper_bound + tol) return false;
    }
    for (const Constraint& c : orig_.constraints()) {
        real_t lhs = 0.0;
        for (const auto& t : c.terms)
            if (t.first >= 0 && static_cast<size_t>(t.first) < x.size()) lhs += t.second * x[static_cast<size_t>(t.first)];
        switch (c.sense) {
            case ConstraintSense::LESS_EQUAL: if (lhs > c.rhs + tol) return false; break;
            case ConstraintSense::GREATER_EQUAL: if (lhs < c.rhs - tol) return false; break;
            case ConstraintSense::EQUAL: if (std::abs(lhs - c.rhs) > tol) return false; break;
            case ConstraintSense::RANGED: if (lhs < c.rhs - tol || lhs > c.range_upper + tol) return false; break;
        }
    }
    return true;
}

real_t Postsolve::compute_original_objective(const std::vector<real_t>& x) const {
    real_t obj = orig_.obj_offset();
    for (size_t j = 0; j < n_ && j < x.size(); ++j) obj += orig_.variables()[j].obj_coeff * x[j];
    return obj;
}

std::string PresolveResult::to_string() const {
    static const char* names[] = {"SUCCESS", "INFEASIBLE", "UNBOUNDED", "NO_CHANGE"};
    std::ostringstream os;
    os << "Presolve Status: " << names[static_cast<int>(status)] << "\n"
       << "  Original Model: " << stats.original_vars << " Vars, " << stats.original_cons << " Cons\n"
       << "  Reduced Model:  " << stats.reduced_vars << " Vars, " << stats.reduced_cons << " Cons\n"
       << "  Reductions:     " << stats.vars_removed << " Vars Removed, " << stats.cons_removed << " Cons Removed\n"
       << "  Execution Time: " << stats.presolve_time_ms << " ms\n";
    return os.str();
}

// ---------------------------------------------------------------- presolve
namespace {

bool infinite_bound(real_t v) { return std::isinf(v) || std::abs(v) >= 1e19; }

struct Work {
    std::vector<Variable> vars;
    std::vector<Constraint> cons;
    std::vector<char> var_on, con_on;
    real_t offset{0.0};
};

}  // namespace

PresolveResult PresolveEngine::presolve(const LPModel& orig, const PresolveOptions& opt) const {
    const auto t0 = std::chrono::high_resolution_clock::now();
    PresolveResult res;
    res.postsolve.initialize(orig);
    res.stats.original_vars = orig.num_variables();
    res.stats.original_cons = orig.num_constraints();
    // every return goes through here; res is not used afterwards, so it is
    // moved out rather than copied (two models and the postsolve record)
    auto finish = [&](PresolveStatus s) -> PresolveResult {
        res.status = s;
        res.stats.presolve_time_ms =
            std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
        return std::move(res);
    };

    const size_t nv = orig.num_variables(), nc = orig.num_constraints();
    Work w;
    w.vars = orig.variables();
    w.cons = orig.constraints();
    w.var_on.assign(nv, 1);
    w.con_on.assign(nc, 1);
    w.offset = orig.obj_offset();
    const ObjectiveSense sense = orig.sense();

    bool changed = true;
    for (int round = 0; changed && round < 100; ++round) {
        changed = false;

        if (opt.enable_fixed_variable) {  // ---- A
            std::vector<char> fix_now(nv, 0);
            bool any = false;
            for (size_t j = 0; j < nv; ++j) {
                if (!w.var_on[j]) continue;
                const real_t lb = w.vars[j].lower_bound, ub = w.vars[j].upper_bound;
                if (lb > ub + 1e-12) return finish(PresolveStatus::INFEASIBLE);
                if (std::abs(lb - ub) <= 1e-12) {
                    fix_now[j] = 1;
                    any = true;
                    w.var_on[j] = 0;
                    res.postsolve.record_fixed_var(static_cast<index_t>(j), lb);
                    w.offset += w.vars[j].obj_coeff * lb;
                    res.stats.transformations.push_back({ReductionType::FIXED_VARIABLE,
                                                         "Fixed variable " + w.vars[j].name + " at value " +
                                                             std::to_string(lb),
                                                         w.vars[j].name, static_cast<index_t>(j), -1, lb});
                }
            }
            if (any) {
                changed = true;
                for (size_t i = 0; i < nc; ++i) {
                    if (!w.con_on[i]) continue;
                    Constraint& c = w.cons[i];
                    // positions of terms on variables fixed now, in ascending variable order
                    std::vector<std::pair<index_t, size_t>> hits;
                    for (size_t k = 0; k < c.terms.size(); ++k)
                        if (fix_now[static_cast<size_t>(c.terms[k].first)]) hits.emplace_back(c.terms[k].first, k);
                    if (hits.empty()) continue;
                    std::stable_sort(hits.begin(), hits.end(),
                                     [](const auto& a, const auto& b) { return a.first < b.first; });
                    for (const auto& h : hits) {
                        const real_t v = w.vars[static_cast<size_t>(h.first)].lower_bound;
                        c.rhs -= c.terms[h.second].second * v;
                        if (c.sense == ConstraintSense::RANGED) c.range_upper -= c.terms[h.second].second * v;
                    }
                    std::vector<std::pair<index_t, real_t>> kept;
                    kept.reserve(c.terms.size() - hits.size());
                    for (const auto& t : c.terms)
                        if (!fix_now[static_cast<size_t>(t.first)]) kept.push_back(t);
                    c.terms.swap(kept);
                }
            }
        }

        if (opt.enable_empty_row) {  // ---- B
            for (size_t i = 0; i < nc; ++i) {
                if (!w.con_on[i]) continue;
                Constraint& c = w.cons[i];
                c.terms.erase(std::remove_if(c.terms.begin(), c.terms.end(),
                                             [&](const std::pair<index_t, real_t>& t) {
                                                 return !w.var_on[static_cast<size_t>(t.first)] || t.second == 0.0;
                                             }),
                              c.terms.end());
                if (!c.terms.empty()) continue;
                bool infeasible = false;
                switch (c.sense) {
                    case ConstraintSense::LESS_EQUAL: infeasible = c.rhs < -1e-12; break;
                    case ConstraintSense::GREATER_EQUAL: infeasible = c.rhs > 1e-12; break;
                    case ConstraintSense::EQUAL: infeasible = std::abs(c.rhs) > 1e-12; break;
                    case ConstraintSense::RANGED: infeasible = c.rhs > 1e-12 || c.range_upper < -1e-12; break;
                }
                if (infeasible) return finish(PresolveStatus::INFEASIBLE);
                w.con_on[i] = 0;
                changed = true;
                res.stats.transformations.push_back({ReductionType::EMPTY_ROW_REDUNDANT,
                                                     "Removed empty redundant constraint " + c.name, c.name, -1,
                                                     static_cast<index_t>(i), 0.0});
            }
        }

        if (opt.enable_empty_col) {  // ---- C
            std::vector<size_t> count(nv, 0);
            for (size_t i = 0; i < nc; ++i) {
                if (!w.con_on[i]) continue;
                for (const auto& t : w.cons[i].terms)
                    if (w.var_on[static_cast<size_t>(t.first)]) count[static_cast<size_t>(t.first)]++;
            }
            for (size_t j = 0; j < nv; ++j) {
                if (!w.var_on[j] || count[j] != 0) continue;
                const real_t cj = w.vars[j].obj_coeff, lb = w.vars[j].lower_bound, ub = w.vars[j].upper_bound;
                // bound the objective prefers: +1 lower, -1 upper, 0 no preference
                int want = 0;
                if (cj != 0.0) {
                    const bool up = (sense == ObjectiveSense::MINIMIZE) ? (cj < 0.0) : (cj > 0.0);
                    want = up ? -1 : 1;
                }
                real_t v = 0.0;
                if (want == 0) {
                    v = !infinite_bound(lb) ? lb : (!infinite_bound(ub) ? ub : 0.0);
                } else if (want == 1) {
                    if (infinite_bound(lb)) return finish(PresolveStatus::UNBOUNDED);
                    v = lb;
                } else {
                    if (infinite_bound(ub)) return finish(PresolveStatus::UNBOUNDED);
                    v = ub;
                }
                w.var_on[j] = 0;
                res.postsolve.record_fixed_var(static_cast<index_t>(j), v);
                w.offset += cj * v;
                changed = true;
                res.stats.transformations.push_back({ReductionType::EMPTY_COL_FIXED,
                                                     "Fixed empty column " + w.vars[j].name + " at value " +
                                                         std::to_string(v),
                                                     w.vars[j].name, static_cast<index_t>(j), -1, v});
            }
        }

        if (opt.enable_singleton_row) {  // ---- D
            for (size_t i = 0; i < nc; ++i) {
                if (!w.con_on[i]) continue;
                const Constraint& c = w.cons[i];
                if (c.terms.size() != 1) continue;
                const size_t j = static_cast<size_t>(c.terms[0].first);
                if (!w.var_on[j]) continue;
                const real_t a = c.terms[0].second;
                if (a == 0.0) continue;
                const real_t v = c.rhs / a;
                real_t lo = -BHARATOPT_INFINITY, hi = BHARATOPT_INFINITY;
                switch (c.sense) {
                    case ConstraintSense::LESS_EQUAL: (a > 0.0 ? hi : lo) = v; break;
                    case ConstraintSense::GREATER_EQUAL: (a > 0.0 ? lo : hi) = v; break;
                    case ConstraintSense::EQUAL: lo = hi = v; break;
                    case ConstraintSense::RANGED: {
                        const real_t vu = c.range_upper / a;
                        lo = a > 0.0 ? v : vu;
                        hi = a > 0.0 ? vu : v;
                        break;
                    }
                }
                Variable& var = w.vars[j];
                if (lo > var.lower_bound) var.lower_bound = lo;
                if (hi < var.upper_bound) var.upper_bound = hi;
                res.postsolve.record_tightened_bounds(static_cast<index_t>(j), var.lower_bound, var.upper_bound);
                if (var.lower_bound > var.upper_bound + 1e-12) return finish(PresolveStatus::INFEASIBLE);
                w.con_on[i] = 0;
                changed = true;
                res.stats.transformations.push_back({ReductionType::SINGLETON_ROW,
                                                     "Singleton row " + c.name + " tightened variable " + var.name,
                                                     var.name, static_cast<index_t>(j), static_cast<index_t>(i), v});
            }
        }

        if (opt.enable_bound_tightening) {  // ---- E
            for (size_t i = 0; i < nc; ++i) {
                if (!w.con_on[i] || w.cons[i].sense != ConstraintSense::LESS_EQUAL) continue;
                Constraint& c = w.cons[i];
                for (auto& term : c.terms) {
                    const size_t j = static_cast<size_t>(term.first);
                    if (!w.var_on[j] || w.vars[j].type != VariableType::BINARY || !(term.second < -1e-5)) continue;
                    const real_t coeff = term.second;
                    real_t worst = 0.0;  // largest activity of the other terms
                    bool bounded = true;
                    for (const auto& o : c.terms) {
                        const size_t k = static_cast<size_t>(o.first);
                        if (o.first == term.first || !w.var_on[k]) continue;
                        if (o.second > 0.0) {
                            if (std::isinf(w.vars[k].upper_bound)) {
                                bounded = false;
                                break;
                            }
                            worst += o.second * w.vars[k].upper_bound;
                        } else if (o.second < 0.0 && !std::isinf(w.vars[k].lower_bound)) {
                            worst += o.second * w.vars[k].lower_bound;
                        }
                    }
                    if (!bounded || !(worst > 0.0)) continue;
                    const real_t big_m = worst - c.rhs;
                    if (big_m > 0.0 && -coeff > big_m + 1e-4) {
                        term.second = -big_m;
                        changed = true;
                        res.stats.transformations.push_back(
                            {ReductionType::BOUND_TIGHTENING,
                             "Tightened Big-M coefficient on " + c.name + " for " + w.vars[j].name, w.vars[j].name,
                             static_cast<index_t>(j), static_cast<index_t>(i), -big_m});
                    }
                }
            }
        }
    }

    // ---- F: reduced model
    LPModel red("reduced_" + orig.name());
    red.set_sense(sense);
    red.set_obj_offset(w.offset);
    // Nothing removed (every variable and row active, every row non-empty,
    // every term on a valid variable): the working copy is the reduced model,
    // handed over without rebuilding it -- the same variables, rows and terms
    // in the same order as the general path below, whose add_variable /
    // add_constraint calls it replaces (names re-registered, so a repeated
    // name still throws).
    bool kept_all = std::all_of(w.var_on.begin(), w.var_on.end(), [](char on) { return on != 0; }) &&
                    std::all_of(w.con_on.begin(), w.con_on.end(), [](char on) { return on != 0; });
    for (size_t i = 0; kept_all && i < nc; ++i) {
        kept_all = !w.cons[i].terms.empty();
        for (const auto& t : w.cons[i].terms) kept_all = kept_all && t.first >= 0 && static_cast<size_t>(t.first) < nv;
    }
    if (kept_all) {
        red.var_lookup_.reserve(nv);
        for (size_t j = 0; j < nv; ++j) {
            Variable& v = w.vars[j];
            v.index = static_cast<index_t>(j);
            if (!red.var_lookup_.emplace(v.name, v.index).second)
                throw std::invalid_argument("Variable with name '" + v.name + "' already exists in LPModel.");
            res.postsolve.record_var_map(v.index, v.index);
        }
        red.con_lookup_.reserve(nc);
        for (size_t i = 0; i < nc; ++i) {
            Constraint& c = w.cons[i];
            c.index = static_cast<index_t>(i);
            if (!red.con_lookup_.emplace(c.name, c.index).second)
                throw std::invalid_argument("Constraint with name '" + c.name + "' already exists in LPModel.");
        }
        red.vars_ = std::move(w.vars);
        red.cons_ = std::move(w.cons);
    }
    std::vector<index_t> to_red(kept_all ? 0 : nv, -1);
    for (size_t j = 0; !kept_all && j < nv; ++j) {
        if (!w.var_on[j]) continue;
        const Variable& v = w.vars[j];
        to_red[j] = red.add_variable(v.name, v.lower_bound, v.upper_bound, v.obj_coeff, v.type);
        res.postsolve.record_var_map(static_cast<index_t>(j), to_red[j]);
    }
    for (size_t i = 0; !kept_all && i < nc; ++i) {
        if (!w.con_on[i]) continue;
        const Constraint& c = w.cons[i];
        std::vector<std::pair<index_t, real_t>> terms;
        terms.reserve(c.terms.size());
        for (const auto& t : c.terms)
            if (t.first >= 0 && static_cast<size_t>(t.first) < nv && w.var_on[static_cast<size_t>(t.first)])
                terms.emplace_back(to_red[static_cast<size_t>(t.first)], t.second);
        if (!terms.empty()) red.add_constraint(c.name, terms, c.sense, c.rhs, c.range_upper);
    }
    res.stats.reduced_vars = red.num_variables();
    res.stats.reduced_cons = red.num_constraints();
    res.stats.vars_removed = res.stats.original_vars - res.stats.reduced_vars;
    res.stats.cons_removed = res.stats.original_cons - res.stats.reduced_cons;
    const bool any_change = res.stats.vars_removed > 0 || res.stats.cons_removed > 0 ||
                            std::abs(w.offset - orig.obj_offset()) > 1e-12;
    res.reduced_model = std::move(red);
    return finish(any_change ? PresolveStatus::SUCCESS : PresolveStatus::NO_CHANGE);
}

}  // namespace bharatopt
