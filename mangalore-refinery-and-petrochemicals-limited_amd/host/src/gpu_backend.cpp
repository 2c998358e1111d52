// gpu_backend.cpp -- GpuBackend over the C ABI.  Argument checks, automatic
// (re)allocation of undersized outputs and status codes follow the reference
// (src/gpu_backend.cpp:34-327); the work runs on the MI355X.
#include <bharatopt/gpu_backend.hpp>

#include <chrono>
#include <cmath>
#include <mutex>

namespace bharatopt {

namespace {
using clk = std::chrono::high_resolution_clock;
double ms_since(clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); }

void copy_timing(const bo_timing& t, GpuTimingResult* out) {
    if (!out) return;
    out->kernel_time_ms = t.kernel_time_ms;
    out->sync_time_ms = t.sync_time_ms;
    out->total_gpu_time_ms = t.total_gpu_time_ms;
}
}  // namespace

GpuBackend& GpuBackend::instance() {
    static GpuBackend backend;
    return backend;
}

GPUStatus GpuBackend::initialize() {
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (initialized_) return init_status_;
    bo_device_info d{};
    bo_status s = bo_device_init(&d);
    info_.is_available = d.is_available != 0;
    info_.device_name = d.device_name[0] ? d.device_name : "None";
    info_.total_memory_bytes = d.total_memory_bytes;
    info_.free_memory_bytes = d.free_memory_bytes;
    info_.compute_capability_major = d.compute_capability_major;
    info_.compute_capability_minor = d.compute_capability_minor;
    info_.status_message = d.status_message;
    init_status_ = to_gpu_status(s);
    initialized_ = true;
    return init_status_;
}

GPUStatus GpuBackend::spmv(const GpuMatrixCSR& A, const GpuBuffer<real_t>& x, GpuBuffer<real_t>& y,
                           GpuTimingResult* timing) {
    if (!A.is_valid() || !x.is_allocated() || x.size() < A.cols) return GPUStatus::INVALID_ARGUMENT;
    if (y.size() < A.rows) {
        GPUStatus s = y.allocate(A.rows);
        if (s != GPUStatus::SUCCESS) return s;
    }
    bo_timing t{};
    bo_status s = bo_spmv_csr_f64(A.rows, A.cols, A.nnz, A.row_offsets.data(), A.col_indices.data(), A.values.data(),
                                  x.data(), y.data(), timing ? &t : nullptr);
    if (timing) {
        timing->h2d_time_ms = 0.0;
        timing->d2h_time_ms = 0.0;
        copy_timing(t, timing);
    }
    return to_gpu_status(s);
}

GPUStatus GpuBackend::spmv_with_cpu_comparison(const CSRMatrix& cpu_A, const std::vector<real_t>& cpu_x,
                                               std::vector<real_t>& gpu_y, real_t& max_abs_error,
                                               GpuTimingResult* timing) {
    if (!cpu_A.is_valid() || cpu_x.size() != cpu_A.cols()) return GPUStatus::INVALID_ARGUMENT;
    const auto t_cpu = clk::now();
    const std::vector<real_t> cpu_y = cpu_A.multiply(cpu_x);
    const double cpu_ms = ms_since(t_cpu);

    const auto t_all = clk::now();
    GpuMatrixCSR dA;
    GPUStatus s = dA.upload_from_cpu(cpu_A);
    if (s != GPUStatus::SUCCESS) return s;
    GpuBuffer<real_t> dx(cpu_x.size());
    s = dx.copy_host_to_device(cpu_x.data(), cpu_x.size());
    if (s != GPUStatus::SUCCESS) return s;
    const double h2d_ms = ms_since(t_all);
    GpuBuffer<real_t> dy(cpu_A.rows());
    GpuTimingResult k{};
    s = spmv(dA, dx, dy, &k);
    if (s != GPUStatus::SUCCESS) return s;
    const auto t_d2h = clk::now();
    gpu_y.resize(cpu_A.rows());
    s = dy.copy_device_to_host(gpu_y.data(), cpu_A.rows());
    if (s != GPUStatus::SUCCESS) return s;
    const double d2h_ms = ms_since(t_d2h);

    max_abs_error = 0.0;
    for (size_t i = 0; i < cpu_A.rows(); ++i) {
        const real_t d = std::abs(gpu_y[i] - cpu_y[i]);
        if (d > max_abs_error) max_abs_error = d;
    }
    if (timing) {
        timing->h2d_time_ms = h2d_ms;
        timing->kernel_time_ms = k.kernel_time_ms;
        timing->d2h_time_ms = d2h_ms;
        timing->sync_time_ms = k.sync_time_ms;
        timing->total_gpu_time_ms = ms_since(t_all);
        timing->cpu_reference_time_ms = cpu_ms;
    }
    return GPUStatus::SUCCESS;
}

namespace {
template <class Fn>
GPUStatus binary_op(const GpuBuffer<real_t>& a, const GpuBuffer<real_t>* b, GpuBuffer<real_t>& out, size_t n,
                    GpuTimingResult* timing, Fn&& fn) {
    if (!a.is_allocated() || a.size() < n || (b && (!b->is_allocated() || b->size() < n)))
        return GPUStatus::INVALID_ARGUMENT;
    if (out.size() < n) {
        GPUStatus s = out.allocate(n);
        if (s != GPUStatus::SUCCESS) return s;
    }
    bo_timing t{};
    bo_status s = fn(timing ? &t : nullptr);
    copy_timing(t, timing);
    return to_gpu_status(s);
}
}  // namespace

GPUStatus GpuBackend::vector_add(const GpuBuffer<real_t>& a, const GpuBuffer<real_t>& b, GpuBuffer<real_t>& r,
                                 size_t n, GpuTimingResult* timing) {
    return binary_op(a, &b, r, n, timing, [&](bo_timing* t) { return bo_vec_add_f64(a.data(), b.data(), r.data(), n, t); });
}

GPUStatus GpuBackend::vector_sub(const GpuBuffer<real_t>& a, const GpuBuffer<real_t>& b, GpuBuffer<real_t>& r,
                                 size_t n, GpuTimingResult* timing) {
    return binary_op(a, &b, r, n, timing, [&](bo_timing* t) { return bo_vec_sub_f64(a.data(), b.data(), r.data(), n, t); });
}

GPUStatus GpuBackend::vector_scale(const GpuBuffer<real_t>& in, real_t alpha, GpuBuffer<real_t>& r, size_t n,
                                   GpuTimingResult* timing) {
    return binary_op(in, nullptr, r, n, timing,
                     [&](bo_timing* t) { return bo_vec_scale_f64(in.data(), alpha, r.data(), n, t); });
}

GPUStatus GpuBackend::axpy(real_t alpha, const GpuBuffer<real_t>& x, const GpuBuffer<real_t>& y_in,
                           GpuBuffer<real_t>& y_out, size_t n, GpuTimingResult* timing) {
    return binary_op(x, &y_in, y_out, n, timing,
                     [&](bo_timing* t) { return bo_axpy_f64(alpha, x.data(), y_in.data(), y_out.data(), n, t); });
}

GPUStatus GpuBackend::dot_product(const GpuBuffer<real_t>& a, const GpuBuffer<real_t>& b, real_t& result, size_t n,
                                  GpuTimingResult* timing) {
    if (!a.is_allocated() || !b.is_allocated() || a.size() < n || b.size() < n) return GPUStatus::INVALID_ARGUMENT;
    bo_timing t{};
    bo_status s = bo_dot_f64(a.data(), b.data(), &result, n, timing ? &t : nullptr);
    copy_timing(t, timing);
    return to_gpu_status(s);
}

GPUStatus GpuBackend::norm2(const GpuBuffer<real_t>& v, real_t& result, size_t n, GpuTimingResult* timing) {
    if (!v.is_allocated() || v.size() < n) return GPUStatus::INVALID_ARGUMENT;
    bo_timing t{};
    bo_status s = bo_norm2_f64(v.data(), &result, n, timing ? &t : nullptr);
    copy_timing(t, timing);
    return to_gpu_status(s);
}

}  // namespace bharatopt
