// measured_cost_estimator.cpp -- see measured_cost_estimator.hpp.
#include "measured_cost_estimator.hpp"

#include <bo_capi.h>

#include <climits>

namespace bharatopt {

MeasuredGpuCostEstimator::MeasuredGpuCostEstimator(int engine, long long iterations,
                                                   std::shared_ptr<ICostEstimator> cpu)
    : engine_(engine), iterations_(iterations),
      cpu_(cpu ? std::move(cpu) : std::make_shared<AnalyticalCostEstimator>()) {}

CostEstimate MeasuredGpuCostEstimator::estimate(const WorkloadFeatures& f, BenchmarkSolverType solver) const {
    CostEstimate e = cpu_->estimate(f, solver);
    // without a device the wrapped estimator's CPU-fallback estimate stands
    if (solver != BenchmarkSolverType::GPU_FIRST_ORDER || !f.gpu_available) return e;
    bo_cost_estimate c{};
    if (f.m > static_cast<size_t>(INT_MAX) || f.n > static_cast<size_t>(INT_MAX) ||
        f.nnz > static_cast<size_t>(LLONG_MAX) ||
        bo_pdhg_cost_estimate(static_cast<int>(f.m), static_cast<int>(f.n), static_cast<long long>(f.nnz), engine_,
                              iterations_, &c) != BO_SUCCESS) {
        e.prediction_valid = false;
        e.validity_status = ValidityStatus::INVALID;
        e.gpu_execution_status = GpuExecutionStatus::INVALID;
        return e;
    }
    e.predicted_transfer_time_ms = c.create_ms;
    e.predicted_compute_time_ms = c.fixed_ms + static_cast<double>(c.iterations) * c.per_iteration_ms;
    e.predicted_total_time_ms = c.total_ms;
    e.prediction_valid = true;
    e.validity_status = ValidityStatus::VALID;
    e.gpu_execution_status = GpuExecutionStatus::NATIVE_GPU;
    e.model_version = "mi355x-measured-v1";
    e.hardware_profile = f.gpu_name;
    if (c.in_domain) {
        e.extrapolation_status = ExtrapolationStatus::IN_DOMAIN;
        e.confidence_status = ConfidenceLevel::HIGH_DATA_SUPPORT;
        e.extrapolation_message = "In the measured range (1,500 to 17.1M nonzeros).";
    } else {
        e.extrapolation_status = ExtrapolationStatus::EXTRAPOLATION_WARNING;
        e.confidence_status = ConfidenceLevel::EXTRAPOLATION_WARNING;
        e.extrapolation_message = "Outside the measured range (1,500 to 17.1M nonzeros): extrapolated.";
    }
    return e;
}

std::vector<CostEstimate> MeasuredGpuCostEstimator::estimate_all(const WorkloadFeatures& f) const {
    std::vector<CostEstimate> all = cpu_->estimate_all(f);
    for (CostEstimate& e : all)
        if (e.solver_name == "FirstOrderLP" && e.solver_variant == "GPU_PDHG")
            e = estimate(f, BenchmarkSolverType::GPU_FIRST_ORDER);
    return all;
}

}  // namespace bharatopt
