// backend.hip -- device management, buffers and the FP64 primitives behind
// bharatopt::GpuBackend (include/bharatopt/gpu_backend.hpp:171-233).
// Built with -ffp-contract=off: the ordered SpMV and axpy round every
// multiply and add separately, as the reference's host loops do.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"

namespace bo {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

std::once_flag g_init_once;
bo_status g_init_status = BO_GPU_UNINITIALIZED;
bo_device_info g_info;

void do_init() {
    std::memset(&g_info, 0, sizeof(g_info));
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        std::snprintf(g_info.device_name, sizeof(g_info.device_name), "None");
        std::snprintf(g_info.status_message, sizeof(g_info.status_message),
                      "No HIP device visible (%s)", hipGetErrorString(e));
        g_init_status = BO_GPU_UNAVAILABLE;
        return;
    }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, 0);
    if (e != hipSuccess) {
        std::snprintf(g_info.status_message, sizeof(g_info.status_message),
                      "hipGetDeviceProperties failed: %s", hipGetErrorString(e));
        g_init_status = map_error(e);
        return;
    }
    hipSetDevice(0);
    size_t free_b = 0, total_b = 0;
    hipMemGetInfo(&free_b, &total_b);
    g_info.is_available = 1;
    std::snprintf(g_info.device_name, sizeof(g_info.device_name), "%s (%s)", prop.name,
                  prop.gcnArchName);
    g_info.total_memory_bytes = total_b ? total_b : prop.totalGlobalMem;
    g_info.free_memory_bytes = free_b;
    g_info.compute_capability_major = prop.major;
    g_info.compute_capability_minor = prop.minor;
    g_info.multiprocessor_count = prop.multiProcessorCount;
    std::snprintf(g_info.status_message, sizeof(g_info.status_message),
                  "%d CUs, %.1f GiB HBM", prop.multiProcessorCount,
                  static_cast<double>(g_info.total_memory_bytes) / (1024.0 * 1024.0 * 1024.0));
    g_init_status = BO_SUCCESS;
}

// ---- FP64 kernels ----------------------------------------------------------
// One thread per row, columns in stored order: the same sequence of roundings
// as CSRMatrix::multiply (src/sparse_matrix.cpp:153-163).
__global__ void k_spmv_csr_ordered(int rows, const int* __restrict__ rp, const int* __restrict__ ci,
                                   const double* __restrict__ v, const double* __restrict__ x,
                                   double* __restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    double sum = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) sum = __dadd_rn(sum, __dmul_rn(v[k], x[ci[k]]));
    y[i] = sum;
}

__global__ void k_vec_binop(int op, const double* __restrict__ a, const double* __restrict__ b,
                            double alpha, double* __restrict__ out, size_t n) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) {
        double r;
        switch (op) {
            case 0: r = a[i] + b[i]; break;
            case 1: r = a[i] - b[i]; break;
            case 2: r = alpha * a[i]; break;
            default: r = __dadd_rn(__dmul_rn(alpha, a[i]), b[i]); break;  // axpy
        }
        out[i] = r;
    }
}

// Deterministic dot: each block reduces a fixed slice in a fixed tree; the
// partials are summed in block order by a second single-block pass.
__global__ void k_dot_partial(const double* __restrict__ a, const double* __restrict__ b, size_t n,
                              double* __restrict__ partial) {
    __shared__ double s[kBlock / kWave];
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    double acc = 0.0;
    for (; i < n; i += stride) acc += a[i] * b[i];
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void k_sum_partials(const double* __restrict__ partial, int count, double* __restrict__ out,
                               int take_sqrt) {
    __shared__ double s[kBlock / kWave];
    double acc = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) acc += partial[i];
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(t) : t;
}

// Plain streaming copy, 16 B per lane per step (calibration of byte counters).
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

struct EventTimer {
    hipEvent_t a{}, b{};
    bool on;
    explicit EventTimer(bool enabled) : on(enabled) {
        if (on) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, 0);
        }
    }
    double stop() {
        if (!on) return 0.0;
        float ms = 0.f;
        hipEventRecord(b, 0);
        hipEventSynchronize(b);
        hipEventElapsedTime(&ms, a, b);
        hipEventDestroy(a);
        hipEventDestroy(b);
        on = false;
        return ms;
    }
    ~EventTimer() {
        if (on) {
            hipEventDestroy(a);
            hipEventDestroy(b);
        }
    }
};

bo_status ensure_device() {
    std::call_once(g_init_once, do_init);
    if (g_init_status != BO_SUCCESS) set_last_error(g_info.status_message);
    return g_init_status;
}

bo_status vec_op(int op, const double* a, const double* b, double alpha, double* out, size_t n,
                 bo_timing* t) {
    bo_status st = ensure_device();
    if (st != BO_SUCCESS) return st;
    if ((n > 0 && (!a || !out)) || (op <= 1 && n > 0 && !b) || (op == 3 && n > 0 && !b))
        return BO_INVALID_ARGUMENT;
    auto t0 = std::chrono::high_resolution_clock::now();
    EventTimer ev(t != nullptr);
    if (n > 0) {
        unsigned g = grid_for(n);
        if (g > 4096) g = 4096;
        hipLaunchKernelGGL(k_vec_binop, dim3(g), dim3(kBlock), 0, 0, op, a, b, alpha, out, n);
        BO_TRY(hipGetLastError());
    }
    double kms = ev.stop();
    BO_TRY(hipDeviceSynchronize());
    if (t) {
        t->kernel_time_ms = kms;
        t->total_gpu_time_ms =
            std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
    }
    return BO_SUCCESS;
}

bo_status reduce_op(const double* a, const double* b, double* result, size_t n, int take_sqrt,
                    bo_timing* t) {
    bo_status st = ensure_device();
    if (st != BO_SUCCESS) return st;
    if (!result || (n > 0 && (!a || !b))) return BO_INVALID_ARGUMENT;
    auto t0 = std::chrono::high_resolution_clock::now();
    unsigned g = grid_for(n);
    if (g > 1024) g = 1024;
    double* d_part = nullptr;
    double* d_out = nullptr;
    BO_TRY(hipMalloc(&d_part, sizeof(double) * (g + 1)));
    d_out = d_part + g;
    EventTimer ev(t != nullptr);
    hipLaunchKernelGGL(k_dot_partial, dim3(g), dim3(kBlock), 0, 0, a, b, n, d_part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, 0, d_part, static_cast<int>(g), d_out,
                       take_sqrt);
    hipError_t le = hipGetLastError();
    double kms = ev.stop();
    double h = 0.0;
    hipError_t ce = hipMemcpy(&h, d_out, sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d_part);
    BO_TRY(le);
    BO_TRY(ce);
    *result = h;
    if (t) {
        t->kernel_time_ms = kms;
        t->total_gpu_time_ms =
            std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
    }
    return BO_SUCCESS;
}

}  // namespace

bo_status ensure_device_public() { return ensure_device(); }

}  // namespace bo

using namespace bo;

extern "C" {

const char* bo_last_error(void) { return g_last_error.c_str(); }

bo_status bo_device_init(bo_device_info* info) {
    bo_status st = ensure_device();
    if (info) *info = g_info;
    return st;
}

bo_status bo_buf_alloc(size_t bytes, void** d_ptr) {
    if (!d_ptr) return BO_INVALID_ARGUMENT;
    *d_ptr = nullptr;
    bo_status st = ensure_device();
    if (st != BO_SUCCESS) return st;
    if (bytes == 0) return BO_SUCCESS;
    BO_TRY(hipMalloc(d_ptr, bytes));
    return BO_SUCCESS;
}

void bo_buf_free(void* d_ptr) {
    if (d_ptr) hipFree(d_ptr);
}

bo_status bo_buf_h2d(void* d_dst, const void* src, size_t bytes) {
    if (bytes == 0) return BO_SUCCESS;
    if (!d_dst || !src) return BO_INVALID_ARGUMENT;
    BO_TRY(hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice));
    return BO_SUCCESS;
}

bo_status bo_buf_d2h(void* dst, const void* d_src, size_t bytes) {
    if (bytes == 0) return BO_SUCCESS;
    if (!dst || !d_src) return BO_INVALID_ARGUMENT;
    BO_TRY(hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost));
    return BO_SUCCESS;
}

bo_status bo_spmv_csr_f64(size_t rows, size_t cols, size_t nnz, const int* d_row_ptr,
                          const int* d_col_idx, const double* d_val, const double* d_x,
                          double* d_y, bo_timing* timing) {
    bo_status st = ensure_device();
    if (st != BO_SUCCESS) return st;
    if (rows == 0 || cols == 0 || !d_row_ptr || !d_x || !d_y || (nnz > 0 && (!d_col_idx || !d_val)))
        return BO_INVALID_ARGUMENT;
    auto t0 = std::chrono::high_resolution_clock::now();
    EventTimer ev(timing != nullptr);
    hipLaunchKernelGGL(k_spmv_csr_ordered, dim3(grid_for(rows)), dim3(kBlock), 0, 0,
                       static_cast<int>(rows), d_row_ptr, d_col_idx, d_val, d_x, d_y);
    BO_TRY(hipGetLastError());
    double kms = ev.stop();
    auto ts = std::chrono::high_resolution_clock::now();
    BO_TRY(hipDeviceSynchronize());
    auto te = std::chrono::high_resolution_clock::now();
    if (timing) {
        timing->h2d_time_ms = 0.0;  // operands are already resident
        timing->d2h_time_ms = 0.0;
        timing->kernel_time_ms = kms;
        timing->sync_time_ms = std::chrono::duration<double, std::milli>(te - ts).count();
        timing->total_gpu_time_ms = std::chrono::duration<double, std::milli>(te - t0).count();
    }
    return BO_SUCCESS;
}

bo_status bo_vec_add_f64(const double* a, const double* b, double* out, size_t n, bo_timing* t) {
    return vec_op(0, a, b, 0.0, out, n, t);
}
bo_status bo_vec_sub_f64(const double* a, const double* b, double* out, size_t n, bo_timing* t) {
    return vec_op(1, a, b, 0.0, out, n, t);
}
bo_status bo_vec_scale_f64(const double* in, double alpha, double* out, size_t n, bo_timing* t) {
    return vec_op(2, in, nullptr, alpha, out, n, t);
}
bo_status bo_axpy_f64(double alpha, const double* x, const double* y_in, double* y_out, size_t n,
                      bo_timing* t) {
    return vec_op(3, x, y_in, alpha, y_out, n, t);
}
bo_status bo_dot_f64(const double* a, const double* b, double* result, size_t n, bo_timing* t) {
    return reduce_op(a, b, result, n, 0, t);
}
bo_status bo_norm2_f64(const double* v, double* result, size_t n, bo_timing* t) {
    return reduce_op(v, v, result, n, 1, t);
}

bo_status bo_copy_kernel_calibrate(size_t bytes, int reps, double* ms) {
    bo_status st = ensure_device();
    if (st != BO_SUCCESS) return st;
    if (bytes < 16 || reps < 1 || !ms) return BO_INVALID_ARGUMENT;
    const size_t n16 = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    BO_TRY(hipMalloc(&a, n16 * 16));
    hipError_t e = hipMalloc(&b, n16 * 16);
    if (e != hipSuccess) {
        hipFree(a);
        BO_TRY(e);
    }
    hipMemset(a, 1, n16 * 16);
    const unsigned g = 256u * 64u;  // 64 workgroups per CU on a 256-CU part
    hipLaunchKernelGGL(k_copy16, dim3(g), dim3(kBlock), 0, 0, a, b, n16);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy16, dim3(g), dim3(kBlock), 0, 0, a, b, n16);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float t = 0.f;
    hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipError_t le = hipGetLastError();
    hipFree(a);
    hipFree(b);
    BO_TRY(le);
    *ms = static_cast<double>(t) / reps;
    return BO_SUCCESS;
}

}  // extern "C"
