// capi_mps.cpp -- the MPS reader behind a C signature: a handle that owns
// the parsed model flattened into the bo_model layout (row-wise terms in
// insertion order), for callers that bind C (ctypes, cgo, JNI).
#include <bharatopt/mps_parser.hpp>

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "bo_capi.h"

using namespace bharatopt;

struct bo_mps {
    int num_vars = 0, num_cons = 0, sense = 0;
    double obj_offset = 0.0;
    std::vector<double> lb, ub, obj, rhs, range_upper, term_val;
    std::vector<int> var_type, con_sense, term_var;
    std::vector<long long> term_ptr;
};

namespace {

void copy_text(char* dst, size_t cap, const std::string& s) {
    std::strncpy(dst, s.c_str(), cap - 1);
    dst[cap - 1] = 0;
}

void flatten(const LPModel& m, bo_mps* h) {
    h->num_vars = static_cast<int>(m.num_variables());
    h->num_cons = static_cast<int>(m.num_constraints());
    h->sense = m.sense() == ObjectiveSense::MAXIMIZE ? 1 : 0;
    h->obj_offset = m.obj_offset();
    for (const Variable& v : m.variables()) {
        h->lb.push_back(v.lower_bound);
        h->ub.push_back(v.upper_bound);
        h->obj.push_back(v.obj_coeff);
        h->var_type.push_back(v.type == VariableType::BINARY ? 2 : (v.type == VariableType::INTEGER ? 1 : 0));
    }
    h->term_ptr.push_back(0);
    for (const Constraint& c : m.constraints()) {
        h->con_sense.push_back(c.sense == ConstraintSense::EQUAL           ? 1
                               : c.sense == ConstraintSense::GREATER_EQUAL ? 2
                               : c.sense == ConstraintSense::RANGED        ? 3
                                                                           : 0);
        h->rhs.push_back(c.rhs);
        h->range_upper.push_back(c.range_upper);
        for (const auto& t : c.terms) {
            h->term_var.push_back(static_cast<int>(t.first));
            h->term_val.push_back(t.second);
        }
        h->term_ptr.push_back(static_cast<long long>(h->term_var.size()));
    }
}

}  // namespace

extern "C" bo_status bo_mps_read(const char* text, const char* path, bo_mps_info* info, bo_mps** model) {
    if (!info || !model || (!text && !path)) return BO_INVALID_ARGUMENT;
    *model = nullptr;
    std::memset(info, 0, sizeof(*info));
    try {
        const MpsParser parser;
        MpsParseResult r = text ? parser.parse_string(text) : parser.parse_file(path);
        info->status = static_cast<int>(r.status);
        info->num_vars = static_cast<int>(r.cols_parsed);
        info->num_cons = static_cast<int>(r.rows_parsed);
        info->nonzeros = static_cast<long long>(r.nonzeros_parsed);
        info->integer_vars = static_cast<int>(r.integer_vars_count);
        info->binary_vars = static_cast<int>(r.binary_vars_count);
        info->parse_time_ms = r.parse_time_ms;
        copy_text(info->problem_name, sizeof(info->problem_name), r.problem_name);
        copy_text(info->message, sizeof(info->message), r.error_message);
        if (r.status != MpsParseStatus::SUCCESS) return BO_SUCCESS;
        bo_mps* h = new bo_mps;
        flatten(r.model, h);
        *model = h;
        return BO_SUCCESS;
    } catch (const std::bad_alloc&) {
        return BO_OUT_OF_MEMORY;
    } catch (const std::exception& e) {  // LPModel rejects e.g. a repeated row name
        copy_text(info->message, sizeof(info->message), e.what());
        return BO_INVALID_ARGUMENT;
    }
}

extern "C" bo_status bo_mps_model(const bo_mps* h, bo_model* v) {
    if (!h || !v) return BO_INVALID_ARGUMENT;
    v->num_vars = h->num_vars;
    v->num_cons = h->num_cons;
    v->sense = h->sense;
    v->obj_offset = h->obj_offset;
    v->lb = h->lb.data();
    v->ub = h->ub.data();
    v->obj = h->obj.data();
    v->var_type = h->var_type.data();
    v->con_sense = h->con_sense.data();
    v->rhs = h->rhs.data();
    v->range_upper = h->range_upper.data();
    v->term_ptr = h->term_ptr.data();
    v->term_var = h->term_var.data();
    v->term_val = h->term_val.data();
    return BO_SUCCESS;
}

extern "C" void bo_mps_destroy(bo_mps* h) { delete h; }
