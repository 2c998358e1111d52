// router_check.cpp -- checker of the MI355X routing adapter
// (host/integration/measured_cost_estimator.cpp) inside the reference's own
// router; test infrastructure, built by oracle/Makefile (target router) from
// the reference's sources where they lie, output only in oracle/_ref/.
//
// For random LPs of several sizes (5 nonzeros per column, fixed seeds) it
// runs the reference's ExecutionRouter::analyse (include/bharatopt/
// execution_router.hpp; src/execution_router.cpp:26-28), then with the
// device flags of the features set off and on: estimate_all of the adapter
// and of the reference's AnalyticalCostEstimator (src/cost_estimator.cpp:
// 213-220), the router's decide with the adapter (src/execution_router.cpp:
// 34-161), and bo_pdhg_cost_estimate called directly on the same sizes.
// One JSON line per case on stdout.
//
// bo_pdhg_cost_estimate is forwarded to libbo_hip.so (argv[1]) loaded with
// RTLD_LOCAL, so the library's C++ symbols stay apart from the reference's
// in this program.
#include <bharatopt/execution_router.hpp>
#include <bharatopt/lp_model.hpp>

#include <bo_capi.h>
#include <dlfcn.h>

#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "measured_cost_estimator.hpp"

using namespace bharatopt;

namespace {
void* g_lib = nullptr;
}

extern "C" bo_status bo_pdhg_cost_estimate(int m, int n, long long nnz, int engine, long long iterations,
                                           bo_cost_estimate* out) {
    using F = bo_status (*)(int, int, long long, int, long long, bo_cost_estimate*);
    static F f = g_lib ? reinterpret_cast<F>(dlsym(g_lib, "bo_pdhg_cost_estimate")) : nullptr;
    if (!f) return BO_EXECUTION_FAILED;
    return f(m, n, nnz, engine, iterations, out);
}

static LPModel random_model(int m, int n, unsigned seed) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_int_distribution<int> row(0, m - 1);
    std::vector<std::vector<std::pair<index_t, real_t>>> rows(m);
    LPModel lp("router_check");
    for (int j = 0; j < n; ++j) {
        const index_t v = lp.add_variable("x" + std::to_string(j), 0.0, 20.0, nd(rng));
        for (int k = 0; k < 5; ++k) rows[row(rng)].push_back({v, nd(rng)});
    }
    for (int i = 0; i < m; ++i) {
        // merge repeated columns of a row
        std::vector<std::pair<index_t, real_t>> t;
        for (auto& p : rows[i]) {
            bool hit = false;
            for (auto& q : t)
                if (q.first == p.first) {
                    q.second += p.second;
                    hit = true;
                }
            if (!hit) t.push_back(p);
        }
        lp.add_constraint("r" + std::to_string(i), t, ConstraintSense::LESS_EQUAL, 10.0 + 5.0 * (i % 7));
    }
    return lp;
}

static void print_est(const char* key, const std::vector<CostEstimate>& v) {
    std::printf("\"%s\": [", key);
    for (size_t k = 0; k < v.size(); ++k) {
        const CostEstimate& e = v[k];
        std::printf("%s{\"solver\": \"%s\", \"variant\": \"%s\", \"total\": %.17g, \"compute\": %.17g, "
                    "\"transfer\": %.17g, \"valid\": %d, \"gpu_status\": %d, \"extrapolation\": %d, "
                    "\"confidence\": %d, \"model\": \"%s\"}",
                    k ? ", " : "", e.solver_name.c_str(), e.solver_variant.c_str(), e.predicted_total_time_ms,
                    e.predicted_compute_time_ms, e.predicted_transfer_time_ms, e.prediction_valid ? 1 : 0,
                    static_cast<int>(e.gpu_execution_status), static_cast<int>(e.extrapolation_status),
                    static_cast<int>(e.confidence_status), e.model_version.c_str());
    }
    std::printf("]");
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: router_check <path to libbo_hip.so>\n");
        return 2;
    }
    g_lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!g_lib) {
        std::fprintf(stderr, "dlopen: %s\n", dlerror());
        return 2;
    }
    const int sizes[][2] = {{30, 90}, {400, 1200}, {3000, 9000}, {20000, 60000}};
    for (const auto& sz : sizes) {
        const LPModel lp = random_model(sz[0], sz[1], 1000u + static_cast<unsigned>(sz[0]));
        auto measured = std::make_shared<MeasuredGpuCostEstimator>();
        ExecutionRouter router(RoutingConfiguration{}, measured);
        const WorkloadFeatures base = router.analyse(lp);
        for (int gpu = 0; gpu < 2; ++gpu) {
            WorkloadFeatures f = base;
            f.gpu_available = gpu != 0;
            f.cuda_available = gpu != 0;
            f.gpu_name = gpu ? "AMD Instinct MI355X" : "N/A";
            const std::vector<CostEstimate> mine = router.estimate(f);
            const std::vector<CostEstimate> analytic = AnalyticalCostEstimator().estimate_all(f);
            const RoutingDecision d = router.decide(lp, f, mine);
            bo_cost_estimate c{};
            const int st = bo_pdhg_cost_estimate(static_cast<int>(f.m), static_cast<int>(f.n),
                                                  static_cast<long long>(f.nnz), 0, 0, &c);
            std::printf("{\"m\": %zu, \"n\": %zu, \"nnz\": %zu, \"orig_n\": %zu, \"orig_nnz\": %zu, \"gpu\": %d, ",
                        f.m, f.n, f.nnz, f.orig_n, f.orig_nnz, gpu);
            print_est("measured", mine);
            std::printf(", ");
            print_est("analytic", analytic);
            std::printf(", \"direct\": {\"status\": %d, \"total\": %.17g, \"create\": %.17g, \"in_domain\": %d}, ",
                        st, c.total_ms, c.create_ms, c.in_domain);
            std::printf("\"decision\": {\"solver\": %d, \"variant\": \"%s\", \"device\": \"%s\", "
                        "\"gpu_cost\": %.17g, \"selected_cost\": %.17g}}\n",
                        static_cast<int>(d.selected_solver), d.solver_variant.c_str(), d.execution_device.c_str(),
                        d.predicted_gpu_cost_ms, d.selected_predicted_cost_ms);
        }
    }
    return 0;
}
