// measured_cost_estimator.hpp -- an ICostEstimator for BharatOpt's
// ExecutionRouter whose GPU first-order cost is the MI355X measurement
// behind bo_pdhg_cost_estimate.
//
// It replaces, in the router's adaptive path, the analytic GPU estimate
// 24.5 + 0.0008 nnz ms of AnalyticalCostEstimator::estimate
// (src/cost_estimator.cpp:186-197), through the router's own constructor
// argument (include/bharatopt/execution_router.hpp: ExecutionRouter(
// RoutingConfiguration, std::shared_ptr<ICostEstimator>)).  Every other
// solver's estimate, the order of estimate_all and the CPU-fallback
// semantics when no GPU is present are those of the estimator it wraps.
//
// It is compiled in a BharatOpt tree (it includes the tree's
// bharatopt/cost_estimator.hpp) linked against libbo_hip.so; it is not part
// of libbo_hip.so itself.
#pragma once

#include <bharatopt/cost_estimator.hpp>

#include <memory>
#include <vector>

namespace bharatopt {

class MeasuredGpuCostEstimator : public ICostEstimator {
public:
    // engine: 0 the reference-semantics loop (what GPU_FIRST_ORDER with
    // default FirstOrderSolverOptions runs), 1 the PDLP engine.
    // iterations: expected iterations, 0 for the model's default (50,000
    // for engine 0, the default iteration cap; 10,000 for engine 1).
    // cpu: estimator of every other solver (AnalyticalCostEstimator when null).
    explicit MeasuredGpuCostEstimator(int engine = 0, long long iterations = 0,
                                      std::shared_ptr<ICostEstimator> cpu = nullptr);

    CostEstimate estimate(const WorkloadFeatures& features, BenchmarkSolverType solver) const override;
    std::vector<CostEstimate> estimate_all(const WorkloadFeatures& features) const override;

private:
    int engine_;
    long long iterations_;
    std::shared_ptr<ICostEstimator> cpu_;
};

}  // namespace bharatopt
