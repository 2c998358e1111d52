// cost_model.cpp -- bo_pdhg_cost_estimate: the predicted wall time of a GPU
// first-order solve on the MI355X, the estimate the reference's router
// prices its GPU branch with (AnalyticalCostEstimator::estimate,
// src/cost_estimator.cpp:188-197, where it is the analytic 24.5 + 0.0008 nnz
// ms of a device the reference never ran on).
//
// The model is a table of costs measured on an MI355X (gfx950) by
// calibrate_cost.py -- random LPs of 1,500 to 4.5M nonzeros plus the
// benchmark's staircase (1.39M nonzeros) and Zipf (17.1M) LPs -- interpolated
// linearly in log(nnz) /
// log(ms) between measured sizes and extrapolated along the end segments:
//   * handle creation (upload, tile layout, schedule tuning), PDLP engine;
//   * per-solve device set-up (scaling, power iteration), PDLP engine;
//   * per-iteration cost, PDLP engine (the schedule the tuner picks);
//   * per-iteration cost of the reference engine, per loop tier: the LP
//     resident in one workgroup's LDS, one workgroup over memory, or a
//     kernel per pass;
//   * host-side preparation and upload of the reference engine.
// Iterations are the caller's (e.g. from an earlier solve of a similar LP);
// 0 assumes 10,000 for the PDLP engine (the median of the measured
// workloads: 2,240 to 12,032) and 50,000 (the default iteration cap, which
// the reference loop usually reaches on large LPs) for the reference engine.
#include <bo_capi.h>

#include <cmath>
#include <cstring>

namespace {

struct Point {
    double nnz, ms;
};

// log-log interpolation over points sorted by nnz; extrapolates the end
// segments (flat when a table has a single point)
double interp(const Point* p, int count, double nnz) {
    if (count == 1) return p[0].ms;
    const double x = std::log(nnz);
    int i = 0;
    while (i + 2 < count && nnz > p[i + 1].nnz) ++i;
    const double x0 = std::log(p[i].nnz), x1 = std::log(p[i + 1].nnz);
    const double y0 = std::log(p[i].ms), y1 = std::log(p[i + 1].ms);
    const double t = (x - x0) / (x1 - x0);
    return std::exp(y0 + t * (y1 - y0));
}

// MI355X measurements (ms), calibrate_cost.py: random LPs (m rows, 3m
// columns, 5 nonzeros per column) and the benchmark's 24-period staircase
// LP (1,385,701 nonzeros) and 200,000 x 500,000 Zipf LP (17,069,815); the
// random LP of 1.5M nonzeros is left out beside the staircase LP, and the
// reference engine's first (start-up) preparation.
const Point kPdlpIter[] = {
    {1500, 0.001469}, {4500, 0.002123}, {15000, 0.00773}, {45000, 0.007837}, {150000, 0.008017},
    {450000, 0.012409}, {1385701, 0.017122}, {4500000, 0.063617}, {17069815, 0.08081}};
const Point kPdlpCreate[] = {
    {1500, 7.1952}, {4500, 15.0288}, {15000, 10.1233}, {45000, 16.5648}, {150000, 66.8471},
    {450000, 26.7681}, {1385701, 71.9719}, {4500000, 759.139}, {17069815, 908.066}};
const Point kPdlpFixed[] = {
    {1500, 0.367842}, {4500, 0.406039}, {15000, 0.945271}, {45000, 1.37488}, {150000, 0.844171},
    {450000, 1.19684}, {1385701, 1.25743}, {4500000, 5.3733}, {17069815, 12.9928}};
const Point kRefCreate[] = {
    {4500, 0.150275}, {15000, 0.282436}, {45000, 0.687374}, {150000, 2.58976}, {450000, 3.17237},
    {1385701, 5.51259}, {4500000, 42.0964}, {17069815, 166.511}};
const Point kRefIterLds[] = {
    {1500, 0.001487}, {4500, 0.002286}};
const Point kRefIterSmall[] = {
    {15000, 0.004501}};
const Point kRefIterKernels[] = {
    {45000, 0.01574}, {150000, 0.016471}, {450000, 0.019681}, {1385701, 0.04068}, {4500000, 0.212743},
    {17069815, 1.75564}};
constexpr double kMinNnz = 1500, kMaxNnz = 17069815;

// The reference engine's loop tier for an LP of this size (the rule of
// bo_pdhg_ref_solve with the MI355X's 160 KB of LDS per workgroup).
int ref_tier(long long m, long long n, long long nnz) {
    const long long big = m > n ? m : n;
    const long long per_thread = (big + 1023) / 1024;
    const long long lds = 8 * (m + n + 2 * nnz) + 4 * (m + n + 2 + 2 * nnz);
    if (per_thread <= 2 && lds + 1024 <= 160 * 1024) return 2;
    if (m + n <= 7680 && nnz <= 65536) return 1;
    return 0;
}

}  // namespace

extern "C" bo_status bo_pdhg_cost_estimate(int m, int n, long long nnz, int engine, long long iterations,
                                           bo_cost_estimate* out) {
    if (!out || m < 0 || n < 0 || nnz < 0 || (engine != 0 && engine != 1) || iterations < 0)
        return BO_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    const double z = nnz > 1 ? static_cast<double>(nnz) : 1.0;
    out->in_domain = (z >= kMinNnz && z <= kMaxNnz) ? 1 : 0;
    if (engine == 1) {
        out->iterations = iterations > 0 ? iterations : 10000;
        out->create_ms = interp(kPdlpCreate, sizeof(kPdlpCreate) / sizeof(Point), z);
        out->fixed_ms = interp(kPdlpFixed, sizeof(kPdlpFixed) / sizeof(Point), z);
        out->per_iteration_ms = interp(kPdlpIter, sizeof(kPdlpIter) / sizeof(Point), z);
        out->loop = -1;
    } else {
        out->iterations = iterations > 0 ? iterations : 50000;
        const int tier = ref_tier(m, n, nnz);
        out->loop = tier;
        const Point* tab = tier == 2 ? kRefIterLds : (tier == 1 ? kRefIterSmall : kRefIterKernels);
        const int count = static_cast<int>(tier == 2 ? sizeof(kRefIterLds) / sizeof(Point)
                                                     : (tier == 1 ? sizeof(kRefIterSmall) / sizeof(Point)
                                                                  : sizeof(kRefIterKernels) / sizeof(Point)));
        out->per_iteration_ms = interp(tab, count, z);
        // host-side CSC build and upload (bo_ref_result.transfer_time_ms)
        out->create_ms = interp(kRefCreate, sizeof(kRefCreate) / sizeof(Point), z);
        out->fixed_ms = 0.1;  // step sizes and the final iterate choice (measured 0.1-0.4 ms)
    }
    out->total_ms = out->create_ms + out->fixed_ms + static_cast<double>(out->iterations) * out->per_iteration_ms;
    return BO_SUCCESS;
}
