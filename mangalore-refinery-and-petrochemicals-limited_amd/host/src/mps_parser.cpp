// mps_parser.cpp -- free-format MPS into an LPModel, with the reference
// reader's semantics (src/mps_parser.cpp:95-529), which the drop-in keeps so
// a model reads the same under both:
//   * a line whose first character is not blank (or any line before the first
//     section) and whose first field is a section name starts that section;
//     '*' lines are comments and '$' starts a trailing comment;
//   * the first N row is the objective; further N rows become <= rows with
//     right-hand side 0 (the reference keeps them as constraints);
//   * RHS / RANGES / BOUNDS: a first field that is not a row name names the
//     vector; only the first vector named is used;
//   * the right-hand side of the objective row is the objective offset's
//     negation;
//   * a RANGES entry R turns the row into a ranged row: L rows [b - R, b] for
//     R > 0 and [b, b - R] otherwise; G and E rows [b, b + R] for R > 0 and
//     [b + R, b] otherwise;
//   * bounds: LO, UP, FX, FR, MI, PL, BV, LI, UI.  UP never touches the lower
//     bound; a malformed bound value reads as 0; a column given both infinite
//     bounds is free;
//   * 'MARKER' lines with INTORG / INTEND bracket integer columns;
//   * coefficients on rows that ROWS does not list are dropped; the objective
//     coefficient of a column is the first one given for the objective row.
#include <bharatopt/mps_parser.hpp>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <fstream>
#include <iterator>
#include <sstream>
#include <unordered_map>

namespace bharatopt {

namespace {

std::string upper(std::string s) {
    for (char& ch : s) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
    return s;
}

// Fortran-style numbers as MPS files write them: ".5", "-.5", "1D3".
// Throws std::invalid_argument / std::out_of_range like std::stod.
real_t parse_number(std::string s) {
    const size_t a = s.find_first_not_of(" \t");
    if (a == std::string::npos) return 0.0;
    s = s.substr(a, s.find_last_not_of(" \t") - a + 1);
    if (s[0] == '.') s.insert(0, "0");
    else if (s.size() > 1 && (s[0] == '-' || s[0] == '+') && s[1] == '.') s.insert(1, "0");
    std::replace_if(s.begin(), s.end(), [](char ch) { return ch == 'D' || ch == 'd'; }, 'E');
    return std::stod(s);
}

enum class Section { kNone, kName, kObjSense, kRows, kColumns, kRhs, kRanges, kBounds, kEnd };

struct BoundSpec {
    real_t lo{0.0};
    real_t hi{BHARATOPT_INFINITY};
    VariableType type{VariableType::CONTINUOUS};
    bool free_col{false};
};

struct Entry {
    std::string row;
    real_t value;
};

// Everything the sections collect before the model is built.
struct Collected {
    std::string objective_row;
    bool sense_given{false};
    ObjectiveSense sense{ObjectiveSense::MINIMIZE};
    std::unordered_map<std::string, char> row_type;
    std::vector<std::string> rows;  // in file order, repeats kept
    std::unordered_map<std::string, std::vector<Entry>> column_entries;
    std::unordered_map<std::string, VariableType> column_type;
    std::vector<std::string> columns;  // first appearance order
    std::unordered_map<std::string, real_t> rhs, ranges;
    std::unordered_map<std::string, BoundSpec> bounds;
    std::string rhs_vector, range_vector, bound_vector;
    bool integer_block{false};
};

bool sense_word(const std::string& w, ObjectiveSense* s) {
    const std::string u = upper(w);
    if (u == "MAX" || u == "MAXIMIZE") *s = ObjectiveSense::MAXIMIZE;
    else if (u == "MIN" || u == "MINIMIZE") *s = ObjectiveSense::MINIMIZE;
    else return false;
    return true;
}

// RHS and RANGES lines: [vector] row value [row value]
bool read_pairs(const std::vector<std::string>& t, const Collected& c, std::string* vector,
                std::unordered_map<std::string, real_t>* into) {
    size_t k = 0;
    if (c.row_type.count(t[0]) == 0 && t.size() > 1) {
        if (vector->empty()) *vector = t[0];
        if (t[0] != *vector) return true;
        k = 1;
    }
    for (; k + 1 < t.size(); k += 2) {
        try {
            (*into)[t[k]] = parse_number(t[k + 1]);
        } catch (...) {
            return false;
        }
    }
    return true;
}

}  // namespace

std::vector<std::string> MpsParser::tokenize_line(const std::string& line) const {
    std::vector<std::string> out;
    if (line.empty() || line[0] == '*') return out;
    std::istringstream in(line.substr(0, line.find('$')));
    for (std::string w; in >> w;) out.push_back(w);
    return out;
}

MpsParseResult MpsParser::parse_file(const std::string& filepath) const {
    const size_t slash = filepath.find_last_of("/\\");
    const std::string base = slash == std::string::npos ? filepath : filepath.substr(slash + 1);
    std::vector<std::string> candidates = {filepath};
    for (const char* up : {"", "../", "../../", "../../../"}) {
        const std::string u(up);
        candidates.push_back(u + filepath);
        for (const char* dir : {"benchmarks/corpus/", "benchmarks/netlib/", "benchmarks/miplib/"})
            candidates.push_back(u + dir + base);
        candidates.push_back(u + base);
    }
    for (const std::string& p : candidates) {
        std::ifstream f(p);
        if (f) return parse_string(std::string(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>()));
    }
    MpsParseResult r;
    r.status = MpsParseStatus::FILE_NOT_FOUND;
    r.error_message = "Could not open MPS file: " + filepath;
    return r;
}

MpsParseResult MpsParser::parse_string(const std::string& content, const std::string& problem_name_override) const {
    const auto t0 = std::chrono::high_resolution_clock::now();
    MpsParseResult r;
    auto fail = [&r](MpsParseStatus s, const std::string& msg) {
        r.status = s;
        r.error_message = msg;
        return r;
    };
    Collected c;
    Section sec = Section::kNone;
    bool ended = false;
    std::istringstream in(content);
    std::string line;
    size_t lineno = 0;
    static const char* kHeaders[] = {"NAME", "OBJSENSE", "ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS", "ENDATA"};
    while (std::getline(in, line)) {
        ++lineno;
        const std::string at = " at line " + std::to_string(lineno);
        const std::vector<std::string> t = tokenize_line(line);
        if (t.empty()) continue;
        const std::string key = upper(t[0]);
        const bool starts_line = line[0] != ' ' && line[0] != '\t';
        if ((starts_line || sec == Section::kNone) &&
            std::find(std::begin(kHeaders), std::end(kHeaders), key) != std::end(kHeaders)) {
            if (key == "NAME") {
                sec = Section::kName;
                r.problem_name = t.size() > 1 ? t[1] : "unnamed_mps";
                if (!problem_name_override.empty()) r.problem_name = problem_name_override;
            } else if (key == "OBJSENSE") {
                sec = Section::kObjSense;
                if (t.size() > 1 && sense_word(t[1], &c.sense)) c.sense_given = true;
            } else if (key == "ROWS") {
                sec = Section::kRows;
            } else if (key == "COLUMNS") {
                sec = Section::kColumns;
            } else if (key == "RHS") {
                sec = Section::kRhs;
            } else if (key == "RANGES") {
                sec = Section::kRanges;
            } else if (key == "BOUNDS") {
                sec = Section::kBounds;
            } else {
                ended = true;
                break;
            }
            continue;
        }
        switch (sec) {
            case Section::kObjSense:
                if (sense_word(t[0], &c.sense)) c.sense_given = true;
                break;
            case Section::kRows: {
                if (t.size() < 2) return fail(MpsParseStatus::MALFORMED_ROW, "Malformed ROWS line" + at);
                const char type = static_cast<char>(std::toupper(static_cast<unsigned char>(t[0][0])));
                if (type != 'N' && type != 'E' && type != 'L' && type != 'G')
                    return fail(MpsParseStatus::MALFORMED_ROW, "Invalid row type '" + std::string(1, type) + "'" + at);
                if (type == 'N' && c.objective_row.empty()) c.objective_row = t[1];
                c.row_type[t[1]] = type;
                c.rows.push_back(t[1]);
                break;
            }
            case Section::kColumns: {
                bool marker = false;
                for (const std::string& w : t) marker = marker || upper(w).find("MARKER") != std::string::npos;
                if (marker) {
                    for (const std::string& w : t) {
                        const std::string u = upper(w);
                        if (u.find("INTORG") != std::string::npos) c.integer_block = true;
                        else if (u.find("INTEND") != std::string::npos) c.integer_block = false;
                    }
                    break;
                }
                if (t.size() < 3) return fail(MpsParseStatus::MALFORMED_COLUMN, "Malformed COLUMNS line" + at);
                auto ins = c.column_entries.emplace(t[0], std::vector<Entry>{});
                if (ins.second) {
                    c.columns.push_back(t[0]);
                    c.column_type[t[0]] = c.integer_block ? VariableType::INTEGER : VariableType::CONTINUOUS;
                }
                for (size_t k = 1; k + 1 < t.size(); k += 2) {
                    try {
                        ins.first->second.push_back({t[k], parse_number(t[k + 1])});
                    } catch (...) {
                        return fail(MpsParseStatus::INVALID_NUMERIC_VALUE,
                                    "Invalid coefficient '" + t[k + 1] + "'" + at);
                    }
                }
                break;
            }
            case Section::kRhs:
                if (!read_pairs(t, c, &c.rhs_vector, &c.rhs))
                    return fail(MpsParseStatus::INVALID_NUMERIC_VALUE, "Invalid RHS value" + at);
                break;
            case Section::kRanges:
                if (!read_pairs(t, c, &c.range_vector, &c.ranges))
                    return fail(MpsParseStatus::INVALID_NUMERIC_VALUE, "Invalid RANGES value" + at);
                break;
            case Section::kBounds: {
                if (t.size() < 3) return fail(MpsParseStatus::MALFORMED_BOUND, "Malformed BOUNDS line" + at);
                const std::string kind = upper(t[0]);
                if (c.bound_vector.empty()) c.bound_vector = t[1];
                if (t[1] != c.bound_vector) break;
                BoundSpec& b = c.bounds[t[2]];
                real_t v = 0.0;
                if (t.size() > 3) {
                    try {
                        v = parse_number(t[3]);
                    } catch (...) {
                        v = 0.0;
                    }
                }
                if (kind == "LO") {
                    b.lo = v;
                } else if (kind == "UP") {
                    b.hi = v;
                } else if (kind == "FX") {
                    b.lo = b.hi = v;
                } else if (kind == "FR") {
                    b.free_col = true;
                    b.lo = -BHARATOPT_INFINITY;
                    b.hi = BHARATOPT_INFINITY;
                } else if (kind == "MI") {
                    b.lo = -BHARATOPT_INFINITY;
                } else if (kind == "PL") {
                    b.hi = BHARATOPT_INFINITY;
                } else if (kind == "BV") {
                    b.lo = 0.0;
                    b.hi = 1.0;
                    b.type = VariableType::BINARY;
                } else if (kind == "LI") {
                    b.lo = v;
                    b.type = VariableType::INTEGER;
                } else if (kind == "UI") {
                    b.hi = v;
                    b.type = VariableType::INTEGER;
                } else {
                    return fail(MpsParseStatus::MALFORMED_BOUND, "Unknown bound type '" + kind + "'" + at);
                }
                break;
            }
            default:
                break;  // NAME continuation lines and text before the first section
        }
    }
    if (!ended) return fail(MpsParseStatus::MISSING_ENDATA, "MPS file missing ENDATA card");
    if (c.integer_block)
        return fail(MpsParseStatus::UNCLOSED_INTEGER_MARKER,
                    "MPS file contains unclosed integer marker section (missing INTEND)");

    LPModel model(r.problem_name.empty() ? "mps_model" : r.problem_name);
    if (c.sense_given) model.set_sense(c.sense);
    if (auto it = c.rhs.find(c.objective_row); it != c.rhs.end()) model.set_obj_offset(-it->second);

    std::unordered_map<std::string, index_t> col_index;
    for (const std::string& name : c.columns) {
        VariableType type = c.column_type[name];
        real_t lo = 0.0, hi = BHARATOPT_INFINITY;
        if (auto it = c.bounds.find(name); it != c.bounds.end()) {
            const BoundSpec& b = it->second;
            const bool free_col = b.free_col || (b.lo <= -BHARATOPT_INFINITY && b.hi >= BHARATOPT_INFINITY);
            lo = free_col ? -BHARATOPT_INFINITY : b.lo;
            hi = free_col ? BHARATOPT_INFINITY : b.hi;
            if (b.type == VariableType::BINARY) type = VariableType::BINARY;
            else if (b.type == VariableType::INTEGER && type != VariableType::BINARY) type = VariableType::INTEGER;
        }
        if (type == VariableType::BINARY) r.binary_vars_count++;
        else if (type == VariableType::INTEGER) r.integer_vars_count++;
        real_t cost = 0.0;
        for (const Entry& e : c.column_entries[name])
            if (e.row == c.objective_row) {
                cost = e.value;
                break;
            }
        col_index[name] = model.add_variable(name, lo, hi, cost, type);
    }

    std::unordered_map<std::string, index_t> row_index;
    for (const std::string& name : c.rows) {
        if (name == c.objective_row) continue;
        const char type = c.row_type[name];
        real_t rhs = 0.0, upper_side = 0.0;
        if (auto it = c.rhs.find(name); it != c.rhs.end()) rhs = it->second;
        ConstraintSense sense = type == 'E' ? ConstraintSense::EQUAL
                                : type == 'G' ? ConstraintSense::GREATER_EQUAL
                                              : ConstraintSense::LESS_EQUAL;
        if (auto it = c.ranges.find(name); it != c.ranges.end()) {
            const real_t R = it->second;
            real_t lo = rhs, hi = rhs;
            if (type == 'L') {
                lo = R > 0.0 ? rhs - R : rhs;
                hi = R > 0.0 ? rhs : rhs - R;
            } else if (type == 'G' || type == 'E') {
                lo = R > 0.0 ? rhs : rhs + R;
                hi = R > 0.0 ? rhs + R : rhs;
            }
            sense = ConstraintSense::RANGED;
            rhs = lo;
            upper_side = hi;
        }
        row_index[name] = model.add_constraint(name, sense, rhs, upper_side);
    }

    size_t coefficients = 0;
    for (const std::string& name : c.columns) {
        const index_t j = col_index[name];
        for (const Entry& e : c.column_entries[name]) {
            if (e.row == c.objective_row) continue;
            auto it = row_index.find(e.row);
            if (it == row_index.end()) continue;
            model.set_coeff(it->second, j, e.value);
            ++coefficients;
        }
    }

    r.status = MpsParseStatus::SUCCESS;
    r.model = model;
    r.rows_parsed = model.num_constraints();
    r.cols_parsed = model.num_variables();
    r.nonzeros_parsed = coefficients;
    r.parse_time_ms =
        std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
    return r;
}

}  // namespace bharatopt
