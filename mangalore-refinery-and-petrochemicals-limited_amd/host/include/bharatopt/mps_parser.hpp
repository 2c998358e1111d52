// mps_parser.hpp -- MPS reader producing the LPModel the first-order solver
// takes, with the public interface of include/bharatopt/mps_parser.hpp:14-55
// of the reference (status enum, result struct, parser class), so callers of
// the reference's MpsParser compile unchanged.
#ifndef BHARATOPT_MPS_PARSER_HPP
#define BHARATOPT_MPS_PARSER_HPP

#include <cstddef>
#include <string>
#include <vector>

#include <bharatopt/config.hpp>
#include <bharatopt/lp_model.hpp>

namespace bharatopt {

enum class MpsParseStatus {
    SUCCESS,
    FILE_NOT_FOUND,
    MISSING_ENDATA,
    UNKNOWN_SECTION,
    MALFORMED_ROW,
    MALFORMED_COLUMN,
    MALFORMED_BOUND,
    UNCLOSED_INTEGER_MARKER,
    INVALID_NUMERIC_VALUE
};

struct MpsParseResult {
    MpsParseStatus status{MpsParseStatus::FILE_NOT_FOUND};
    std::string error_message;
    std::string problem_name;
    LPModel model;

    size_t rows_parsed{0};
    size_t cols_parsed{0};
    size_t nonzeros_parsed{0};
    size_t integer_vars_count{0};
    size_t binary_vars_count{0};
    double parse_time_ms{0.0};
};

class MpsParser {
public:
    MpsParser() = default;

    // Reads the file at `filepath`; when it does not open, the same name is
    // tried under ../, ../../, ../../../ and their benchmarks/{corpus,netlib,
    // miplib}/ folders, as the reference does.
    MpsParseResult parse_file(const std::string& filepath) const;

    // Parses MPS text (free format: whitespace-separated fields).
    MpsParseResult parse_string(const std::string& content, const std::string& problem_name_override = "") const;

private:
    std::vector<std::string> tokenize_line(const std::string& line) const;
};

}  // namespace bharatopt

#endif
