// config.hpp -- scalar types and tolerances of the BharatOpt API
// (include/bharatopt/config.hpp:20-36 of the reference), unchanged so that
// callers compile against this library as they do against the original.
#ifndef BHARATOPT_CONFIG_HPP
#define BHARATOPT_CONFIG_HPP

#include <cstddef>
#include <limits>

namespace bharatopt {

using real_t = double;
using index_t = int;
using size_t = std::size_t;

// Same semantics as the reference helper: lo if v < lo, hi if hi < v, else v.
template <typename T>
constexpr const T& clamp(const T& v, const T& lo, const T& hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

constexpr real_t DEFAULT_FEASIBILITY_TOLERANCE = 1e-7;
constexpr real_t DEFAULT_OPTIMALITY_TOLERANCE = 1e-7;
constexpr real_t DEFAULT_INTEGRALITY_TOLERANCE = 1e-5;
constexpr real_t DEFAULT_PIVOT_TOLERANCE = 1e-10;
constexpr real_t DEFAULT_ZERO_TOLERANCE = 1e-12;
constexpr real_t BHARATOPT_INFINITY = std::numeric_limits<real_t>::infinity();

}  // namespace bharatopt

#endif
