// gpu_backend.hpp -- device layer of the drop-in, with the public interface of
// the reference's include/bharatopt/gpu_backend.hpp:15-233.  GpuBuffer<T> now
// owns HIP device memory (data() is a device pointer) and GpuBackend runs the
// FP64 kernels of libbo_hip.so through the C ABI in include/bo_capi.h.
#ifndef BHARATOPT_GPU_BACKEND_HPP
#define BHARATOPT_GPU_BACKEND_HPP

#include <cstddef>
#include <string>
#include <utility>
#include <vector>

#include <bharatopt/config.hpp>
#include <bharatopt/sparse_matrix.hpp>

#include <bo_capi.h>

namespace bharatopt {

enum class GPUStatus { SUCCESS, GPU_UNAVAILABLE, GPU_UNINITIALIZED, OUT_OF_MEMORY, INVALID_ARGUMENT, EXECUTION_FAILED };

inline GPUStatus to_gpu_status(bo_status s) { return static_cast<GPUStatus>(static_cast<int>(s)); }

struct GPUDeviceInfo {
    bool is_available{false};
    std::string device_name{"None"};
    size_t total_memory_bytes{0};
    size_t free_memory_bytes{0};
    int compute_capability_major{0};
    int compute_capability_minor{0};
    std::string status_message;
};

struct GpuTimingResult {
    double h2d_time_ms{0.0};
    double kernel_time_ms{0.0};
    double d2h_time_ms{0.0};
    double sync_time_ms{0.0};
    double total_gpu_time_ms{0.0};
    double cpu_reference_time_ms{0.0};
};

// Move-only owner of `count` elements of device memory.
template <typename T>
class GpuBuffer {
public:
    GpuBuffer() = default;
    explicit GpuBuffer(size_t count) { allocate(count); }
    ~GpuBuffer() { deallocate(); }
    GpuBuffer(const GpuBuffer&) = delete;
    GpuBuffer& operator=(const GpuBuffer&) = delete;
    GpuBuffer(GpuBuffer&& o) noexcept : count_(o.count_), ptr_(o.ptr_) {
        o.count_ = 0;
        o.ptr_ = nullptr;
    }
    GpuBuffer& operator=(GpuBuffer&& o) noexcept {
        if (this != &o) {
            deallocate();
            count_ = o.count_;
            ptr_ = o.ptr_;
            o.count_ = 0;
            o.ptr_ = nullptr;
        }
        return *this;
    }

    GPUStatus allocate(size_t count) {
        deallocate();
        if (count == 0) return GPUStatus::SUCCESS;
        void* p = nullptr;
        bo_status s = bo_buf_alloc(count * sizeof(T), &p);
        if (s != BO_SUCCESS) return to_gpu_status(s);
        ptr_ = static_cast<T*>(p);
        count_ = count;
        return GPUStatus::SUCCESS;
    }
    void deallocate() {
        if (ptr_) bo_buf_free(ptr_);
        ptr_ = nullptr;
        count_ = 0;
    }
    GPUStatus copy_host_to_device(const T* host, size_t count) {
        if (!host && count > 0) return GPUStatus::INVALID_ARGUMENT;
        if (count > count_) {
            GPUStatus s = allocate(count);
            if (s != GPUStatus::SUCCESS) return s;
        }
        return to_gpu_status(bo_buf_h2d(ptr_, host, count * sizeof(T)));
    }
    GPUStatus copy_device_to_host(T* host, size_t count) const {
        if (!host && count > 0) return GPUStatus::INVALID_ARGUMENT;
        if (count > count_) return GPUStatus::INVALID_ARGUMENT;
        return to_gpu_status(bo_buf_d2h(host, ptr_, count * sizeof(T)));
    }
    size_t size() const { return count_; }
    size_t bytes() const { return count_ * sizeof(T); }
    T* data() { return ptr_; }              // device pointer
    const T* data() const { return ptr_; }  // device pointer
    bool is_allocated() const { return count_ > 0; }

private:
    size_t count_{0};
    T* ptr_{nullptr};
};

struct GpuMatrixCSR {
    size_t rows{0};
    size_t cols{0};
    size_t nnz{0};
    GpuBuffer<index_t> row_offsets;
    GpuBuffer<index_t> col_indices;
    GpuBuffer<real_t> values;

    GPUStatus upload_from_cpu(const CSRMatrix& a) {
        rows = a.rows();
        cols = a.cols();
        nnz = a.nnz();
        GPUStatus s1 = row_offsets.copy_host_to_device(a.row_offsets().data(), a.row_offsets().size());
        GPUStatus s2 = col_indices.copy_host_to_device(a.col_indices().data(), a.col_indices().size());
        GPUStatus s3 = values.copy_host_to_device(a.values().data(), a.values().size());
        if (s1 != GPUStatus::SUCCESS) return s1;
        if (s2 != GPUStatus::SUCCESS) return s2;
        return s3;
    }
    bool is_valid() const {
        return rows > 0 && cols > 0 && row_offsets.is_allocated() &&
               (nnz == 0 || (col_indices.is_allocated() && values.is_allocated()));
    }
};

class GpuBackend {
public:
    static GpuBackend& instance();
    GPUStatus initialize();
    bool is_available() const { return info_.is_available; }
    const GPUDeviceInfo& device_info() const { return info_; }

    GPUStatus spmv(const GpuMatrixCSR& A, const GpuBuffer<real_t>& x, GpuBuffer<real_t>& y,
                   GpuTimingResult* timing = nullptr);
    GPUStatus spmv_with_cpu_comparison(const CSRMatrix& cpu_A, const std::vector<real_t>& cpu_x,
                                       std::vector<real_t>& gpu_y, real_t& max_abs_error,
                                       GpuTimingResult* timing = nullptr);
    GPUStatus vector_add(const GpuBuffer<real_t>& a, const GpuBuffer<real_t>& b, GpuBuffer<real_t>& result, size_t n,
                         GpuTimingResult* timing = nullptr);
    GPUStatus vector_sub(const GpuBuffer<real_t>& a, const GpuBuffer<real_t>& b, GpuBuffer<real_t>& result, size_t n,
                         GpuTimingResult* timing = nullptr);
    GPUStatus vector_scale(const GpuBuffer<real_t>& input, real_t alpha, GpuBuffer<real_t>& result, size_t n,
                           GpuTimingResult* timing = nullptr);
    GPUStatus axpy(real_t alpha, const GpuBuffer<real_t>& x, const GpuBuffer<real_t>& y_in, GpuBuffer<real_t>& y_out,
                   size_t n, GpuTimingResult* timing = nullptr);
    GPUStatus dot_product(const GpuBuffer<real_t>& a, const GpuBuffer<real_t>& b, real_t& result, size_t n,
                          GpuTimingResult* timing = nullptr);
    GPUStatus norm2(const GpuBuffer<real_t>& v, real_t& result, size_t n, GpuTimingResult* timing = nullptr);

private:
    GpuBackend() = default;
    bool initialized_{false};
    GPUStatus init_status_{GPUStatus::GPU_UNINITIALIZED};
    GPUDeviceInfo info_;
};

}  // namespace bharatopt

#endif
