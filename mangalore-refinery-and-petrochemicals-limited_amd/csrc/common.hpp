// common.hpp -- shared HIP helpers for the MI355X (gfx950) PDHG engine.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/bo_capi.h"

namespace bo {

constexpr int kWave = 64;      // CDNA wavefront width
constexpr int kBlock = 256;    // 4 waves per workgroup

// hipError_t -> bo_status (the GPUStatus ordinals of gpu_backend.hpp:15-22).
inline bo_status map_error(hipError_t e) {
    switch (e) {
        case hipSuccess: return BO_SUCCESS;
        case hipErrorOutOfMemory: return BO_OUT_OF_MEMORY;
        case hipErrorNoDevice:
        case hipErrorInsufficientDriver: return BO_GPU_UNAVAILABLE;
        case hipErrorInvalidValue: return BO_INVALID_ARGUMENT;
        default: return BO_EXECUTION_FAILED;
    }
}

void set_last_error(const std::string& msg);

#define BO_TRY(expr)                                                              \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ::bo::set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
            return ::bo::map_error(_e);                                           \
        }                                                                         \
    } while (0)

inline unsigned grid_for(size_t n, int block = kBlock) {
    size_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

// A kernel launch, optionally timed by a pair of events bound to the
// dispatch itself (ev[0] when it starts, ev[1] when it ends: the kernel's own
// duration, without the launch gap an event recorded before it would add).
template <class F, class... A>
inline void launch_ev(F kernel, dim3 grid, dim3 block, unsigned lds, hipStream_t stream, const hipEvent_t* ev,
                      A... args) {
    if (ev)
        hipExtLaunchKernelGGL(kernel, grid, block, lds, stream, ev[0], ev[1], 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, stream, args...);
}

// Runs an entry point's body so that no C++ exception leaves the C ABI: a
// host allocation that fails returns BO_OUT_OF_MEMORY, any other exception
// BO_EXECUTION_FAILED, each with its message in the last-error text.
template <class F>
inline bo_status host_guard(F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        set_last_error("out of host memory");
        return BO_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(std::string("host exception: ") + e.what());
        return BO_EXECUTION_FAILED;
    } catch (...) {
        set_last_error("host exception");
        return BO_EXECUTION_FAILED;
    }
}

// Host arrays that are filled completely before they are read: the
// allocator default-initialises (no zeroing), so the threads that fill an
// array touch its pages first.  Zeroing a few hundred MB on one thread cost
// about 0.2 s of handle creation at 17M nonzeros (page faults included).
// resize(n, v) and assign(n, v) still write v.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
template <class T>
using HostVec = std::vector<T, NoInitAlloc<T>>;

// Host threads for set-up work: the hardware's, at most 16 (the host share
// of one GPU on an 8-GPU node), at most `cap`.
inline int host_thread_count(long long cap) {
    const long long hw = std::max(1u, std::thread::hardware_concurrency());
    return static_cast<int>(std::max(1LL, std::min(std::min(hw, 16LL), cap)));
}

// f(t, nt) on nt threads, t = 0 .. nt - 1 (the caller's thread runs t = 0).
// An exception thrown by any part (std::bad_alloc, say) is rethrown in the
// caller once every thread has finished.
template <class F>
inline void host_parallel(int nt, F&& f) {
    std::vector<std::exception_ptr> err(static_cast<size_t>(nt));
    auto run = [&f, &err, nt](int t) {
        try {
            f(t, nt);
        } catch (...) {
            err[static_cast<size_t>(t)] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(run, t);
    run(0);
    for (auto& th : pool) th.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// [begin, end) of part t of nt equal parts of [0, count).
inline std::pair<long long, long long> host_part(long long count, int t, int nt) {
    return {count * t / nt, count * (t + 1) / nt};
}

// An absolute time limit shared by every round of one solve (main batch,
// tail rounds, FP64 polish), so a solve's wall time stays within its
// time_limit_s however many rounds it takes.
struct Deadline {
    bool on = false;
    std::chrono::steady_clock::time_point at{};
    static Deadline after(double seconds) {
        Deadline d;
        if (seconds > 0.0) {
            d.on = true;
            d.at = std::chrono::steady_clock::now() +
                   std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                       std::chrono::duration<double>(seconds));
        }
        return d;
    }
    bool passed() const { return on && std::chrono::steady_clock::now() > at; }
};

// A bo_lp whose CSR can be read safely: sizes non-negative, every array its
// sizes require present, row_ptr[0] == 0, row_ptr non-decreasing and ending
// at nnz (nnz == 0 when there are no rows), every column index in [0, n).
// Every entry point that takes a bo_lp calls this before it reads the matrix
// or allocates anything.
inline bool valid_csr_lp(const bo_lp* lp) {
    if (!lp || lp->m < 0 || lp->n < 0 || lp->nnz < 0) return false;
    if ((lp->m > 0 && (!lp->row_ptr || !lp->b || !lp->is_eq)) || (lp->n > 0 && (!lp->c || !lp->lb || !lp->ub)) ||
        (lp->nnz > 0 && (!lp->col_idx || !lp->val)))
        return false;
    if (lp->m == 0) return lp->nnz == 0;
    if (lp->row_ptr[0] != 0 || lp->row_ptr[lp->m] != lp->nnz) return false;
    for (int i = 0; i < lp->m; ++i)
        if (lp->row_ptr[i + 1] < lp->row_ptr[i]) return false;
    for (long long k = 0; k < lp->nnz; ++k)
        if (lp->col_idx[k] < 0 || lp->col_idx[k] >= lp->n) return false;
    return true;
}

// ---- wave / block reductions (fixed tree order => deterministic) ----------
// Whole-workgroup skip: an epilogue with a member wg_dead(block) can tell a
// kernel that every segment of its workgroup belongs to finished work.
template <class E>
__device__ inline auto wg_dead(const E& e, int block, int) -> decltype(e.wg_dead(block)) {
    return e.wg_dead(block);
}
template <class E>
__device__ inline bool wg_dead(const E&, int, long) {
    return false;
}

// Split epilogues: an epilogue with load(s) / finish(s, acc, pre) has the
// data of segment s it reads besides the segment's sum loaded by the kernel
// before that sum exists, so those loads overlap the pass's own; others run
// as epi(s, acc, part) once the sum is known.
template <class E, class = void>
struct EpiHasLoad : std::false_type {};
template <class E>
struct EpiHasLoad<E, std::void_t<decltype(std::declval<const E&>().load(0))>> : std::true_type {};

// An epilogue whose two gathered vectors are interleaved (entry j of both
// in one 16-byte double2, vec2()) declares kPair = true: one gather per
// entry instead of two.
template <class E, class = void>
struct EpiPair : std::false_type {};
template <class E>
struct EpiPair<E, std::void_t<decltype(E::kPair)>> : std::bool_constant<E::kPair> {};

template <class E>
__device__ inline auto epi_preload(const E& e, int s, int) -> decltype(e.load(s)) {
    return e.load(s);
}
template <class E>
__device__ inline int epi_preload(const E&, int, long) {
    return 0;
}
template <class E, class P>
__device__ inline auto epi_finish(const E& e, int s, double acc, const P& p, int)
    -> decltype(e.finish(s, acc, p)) {
    e.finish(s, acc, p);
}
template <class E, class P>
__device__ inline void epi_finish(const E& e, int s, double acc, const P&, long) {
    double r[1] = {acc};
    e(s, r, nullptr);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// std::max with the reference's NaN behaviour: (a < b) ? b : a drops a NaN b.
__device__ inline double ref_max(double a, double b) { return (a < b) ? b : a; }

__device__ inline double wave_max_ref(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = ref_max(v, __shfl_xor(v, off, kWave));
    return v;
}

// Block-wide sum; every thread gets the result.  `scratch` holds kBlock/kWave doubles.
__device__ inline double block_sum(double v, double* scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) t += scratch[w];
    return t;
}

// Non-negative doubles (including +inf) order like their bit patterns, so a
// 64-bit unsigned atomic max is an exact, order-independent max.
__device__ inline void atomic_max_nonneg(unsigned long long* dst, double v) {
    atomicMax(dst, static_cast<unsigned long long>(__double_as_longlong(v)));
}

}  // namespace bo
