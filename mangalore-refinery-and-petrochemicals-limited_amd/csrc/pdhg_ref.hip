// pdhg_ref.hip -- the reference PDHG loop on the GPU in FP64, bit-equal to
// CPUFirstOrderSolver::solve (src/first_order_solver.cpp:103-279).
//
// Bit-equality comes from three rules: (1) every dot product is summed by one
// thread in the reference's order -- rows in column order for A x
// (src/sparse_matrix.cpp:153-163), columns in ascending row order for A'y,
// which is the order the reference's row-order scatter (:175-185) adds into
// each y[col]; (2) no fused multiply-adds (__dmul_rn/__dadd_rn, and this file
// is built with -ffp-contract=off); (3) max reductions use the reference's
// std::max NaN behaviour and an exact 64-bit atomic max, so they are
// order-independent.  This engine is the drop-in's default: it reproduces
// the reference's iterates, iteration counts and statuses.  Throughput work
// lives in pdlp.hip.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace bo {
namespace {

__device__ inline double dclamp(double v, double lo, double hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

// Steps (src/first_order_solver.cpp:103-128): row sums in column order, column
// sums in ascending row order.
__global__ void k_ref_steps(int m, int n, const int* __restrict__ rp, const double* __restrict__ rv,
                            const int* __restrict__ cp, const double* __restrict__ cv, int adaptive,
                            double primal_step, double dual_step, double* __restrict__ tau,
                            double* __restrict__ sigma) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        double s = 0.0;
        for (int k = rp[t]; k < rp[t + 1]; ++k) s = __dadd_rn(s, fabs(rv[k]));
        sigma[t] = adaptive ? ((s > 1e-12) ? (0.95 / s) : 1.0) : dual_step;
    }
    if (t < n) {
        double s = 0.0;
        for (int k = cp[t]; k < cp[t + 1]; ++k) s = __dadd_rn(s, fabs(cv[k]));
        tau[t] = adaptive ? ((s > 1e-12) ? (0.95 / s) : 1.0) : primal_step;
    }
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
    if (threadIdx.x < count) p[threadIdx.x] = 0ull;
}

// Wave-level max then one atomic per wave (values are >= 0 and never NaN).
__device__ inline void reduce_max_to(unsigned long long* dst, double v) {
    v = wave_max_ref(v);
    if ((threadIdx.x % kWave) == 0) atomic_max_nonneg(dst, v);
}

// A'y (ordered) + primal step + ergodic sum + extrapolation
// (src/first_order_solver.cpp:163-180).  x_prev equals x on entry.
__global__ void k_ref_primal(int n, const int* __restrict__ cp, const int* __restrict__ ri,
                             const double* __restrict__ cv, const double* __restrict__ y,
                             const double* __restrict__ c, const double* __restrict__ tau,
                             const double* __restrict__ lb, const double* __restrict__ ub,
                             double* __restrict__ x, double* __restrict__ x_sum,
                             double* __restrict__ x_hat, unsigned long long* __restrict__ pmax,
                             int track) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    double delta = 0.0;
    if (j < n) {
        double aty = 0.0;
        for (int k = cp[j]; k < cp[j + 1]; ++k) aty = __dadd_rn(aty, __dmul_rn(cv[k], y[ri[k]]));
        double xo = x[j];
        double grad = __dadd_rn(c[j], aty);
        double next_x = __dsub_rn(xo, __dmul_rn(tau[j], grad));
        double new_x = dclamp(next_x, lb[j], ub[j]);
        delta = fabs(__dsub_rn(new_x, xo));
        x[j] = new_x;
        x_sum[j] = __dadd_rn(x_sum[j], new_x);
        x_hat[j] = __dsub_rn(__dmul_rn(2.0, new_x), xo);
    }
    if (track) reduce_max_to(pmax, ref_max(0.0, delta));
}

// A x_hat (ordered) + dual step (src/first_order_solver.cpp:183-192).
__global__ void k_ref_dual(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ rv, const double* __restrict__ x_hat,
                           const double* __restrict__ b, const double* __restrict__ sigma,
                           const unsigned char* __restrict__ is_eq, double* __restrict__ y,
                           unsigned long long* __restrict__ dmax, int track) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double delta = 0.0;
    if (i < m) {
        double ax = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) ax = __dadd_rn(ax, __dmul_rn(rv[k], x_hat[ci[k]]));
        double yo = y[i];
        double next_y = __dadd_rn(yo, __dmul_rn(sigma[i], __dsub_rn(ax, b[i])));
        double new_y = (!is_eq[i]) ? ref_max(0.0, next_y) : next_y;
        delta = fabs(__dsub_rn(new_y, yo));
        y[i] = new_y;
    }
    if (track) reduce_max_to(dmax, ref_max(0.0, delta));
}

// x_avg = clamp(x_sum / count) and the bound violation of x (:208-218).
__global__ void k_ref_avg(int n, const double* __restrict__ x, const double* __restrict__ x_sum,
                          double count, const double* __restrict__ lb, const double* __restrict__ ub,
                          double* __restrict__ x_avg, unsigned long long* __restrict__ bmax,
                          int copy_x) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    double bv = 0.0;
    if (j < n) {
        double xj = x[j];
        if (xj < lb[j] - 1e-12) bv = ref_max(bv, lb[j] - xj);
        if (xj > ub[j] + 1e-12) bv = ref_max(bv, xj - ub[j]);
        x_avg[j] = copy_x ? xj : dclamp(x_sum[j] / count, lb[j], ub[j]);
    }
    reduce_max_to(bmax, bv);
}

__device__ inline double row_viol(double ax, double b, unsigned char eq) {
    double r = __dsub_rn(ax, b);
    return (!eq) ? ref_max(0.0, r) : fabs(r);
}

// Row violations of x and x_avg (:199-225).
__global__ void k_ref_viol(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ rv, const double* __restrict__ x,
                           const double* __restrict__ x_avg, const double* __restrict__ b,
                           const unsigned char* __restrict__ is_eq,
                           unsigned long long* __restrict__ vmax, unsigned long long* __restrict__ amax,
                           int with_x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v1 = 0.0, v2 = 0.0;
    if (i < m) {
        double s1 = 0.0, s2 = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            double a = rv[k];
            int j = ci[k];
            if (with_x) s1 = __dadd_rn(s1, __dmul_rn(a, x[j]));
            s2 = __dadd_rn(s2, __dmul_rn(a, x_avg[j]));
        }
        if (with_x) v1 = row_viol(s1, b[i], is_eq[i]);
        v2 = row_viol(s2, b[i], is_eq[i]);
    }
    if (with_x) reduce_max_to(vmax, v1);
    reduce_max_to(amax, v2);
}


// ---------------------------------------------------- one-workgroup loop
// The same loop for small LPs in one workgroup of 1024 threads, with
// workgroup barriers in place of kernel launches and host round trips:
// identical per-row / per-column arithmetic (so the same bits), maxima
// reduced exactly (order-independent), the stop test on the device.  y and
// x_hat, the vectors the passes gather, are mirrored in LDS.
constexpr int kRefSmallThreads = 1024;
constexpr int kRefSmallMaxEntries = 7680;  // n + m of LPs that run this way (their gathered vectors: 60 KB)
constexpr long long kRefSmallMaxNnz = 1LL << 16;

struct RefSmallArgs {
    int m, n, nnz;
    const int *rp, *ci, *cp, *ri;
    const double *rv, *cv, *b, *c, *lb, *ub, *tau, *sigma;
    const unsigned char* eq;
    double *x, *y, *xsum, *xhat, *xavg;
    long long max_iter, check_freq;
    double feas_tol;
    double* out;  // [0] iterations [1] loop iterations [2] status [3] constraint violation [4] bound violation [5] avg count
};

// sum_k vals[k] * vec[idx[k]] over [k0, k1) in ascending k with no fused
// multiply-adds.  The loads and products of up to eight entries are issued
// ahead; the additions stay one at a time in the reference's order, so the
// result is the same bits as the plain loop.
template <int W>
__device__ inline void dot_chunk(double& s, int k, const double* vals, const int* idx, const double* vec) {
    int i[W];
    double p[W];
#pragma unroll
    for (int u = 0; u < W; ++u) i[u] = idx[k + u];
#pragma unroll
    for (int u = 0; u < W; ++u) p[u] = __dmul_rn(vals[k + u], vec[i[u]]);
#pragma unroll
    for (int u = 0; u < W; ++u) s = __dadd_rn(s, p[u]);
}

__device__ inline double ordered_dot(int k0, int k1, const double* vals, const int* idx, const double* vec) {
    double s = 0.0;
    int k = k0;
    for (; k + 8 <= k1; k += 8) dot_chunk<8>(s, k, vals, idx, vec);
    if (k + 4 <= k1) {
        dot_chunk<4>(s, k, vals, idx, vec);
        k += 4;
    }
    for (; k < k1; ++k) s = __dadd_rn(s, __dmul_rn(vals[k], vec[idx[k]]));
    return s;
}

__device__ inline double block_max_ref(double v, double* sh) {
    v = wave_max_ref(v);
    const int lane = static_cast<int>(threadIdx.x) % kWave, wid = static_cast<int>(threadIdx.x) / kWave;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) t = ref_max(t, sh[w]);
    return t;
}

__global__ void __launch_bounds__(kRefSmallThreads) k_ref_small(RefSmallArgs a) {
    __shared__ double sh[kRefSmallThreads / kWave];
    __shared__ double ly[kRefSmallMaxEntries];  // y copy [0, m), x_hat copy [m, m + n)
    double* lxh = ly + a.m;
    const int tid = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
    for (int i = tid; i < a.m; i += nt) ly[i] = a.y[i];
    __syncthreads();
    long long iter = 0, avg_count = 0;
    int status = 1;
    double best = 0.0, bviol = 0.0;
    long long to_check = a.check_freq;
    while (iter < a.max_iter) {
        iter++;
        const bool track = (--to_check == 0 || iter == a.max_iter);  // iter % check_freq == 0 || last
        if (to_check == 0) to_check = a.check_freq;
        double dp = 0.0, dd = 0.0;
        for (int j = tid; j < a.n; j += nt) {  // A'y + primal step + ergodic sum + extrapolation
            const double aty = ordered_dot(a.cp[j], a.cp[j + 1], a.cv, a.ri, ly);
            const double xo = a.x[j];
            const double grad = __dadd_rn(a.c[j], aty);
            const double next_x = __dsub_rn(xo, __dmul_rn(a.tau[j], grad));
            const double new_x = dclamp(next_x, a.lb[j], a.ub[j]);
            dp = ref_max(dp, ref_max(0.0, fabs(__dsub_rn(new_x, xo))));
            a.x[j] = new_x;
            a.xsum[j] = __dadd_rn(a.xsum[j], new_x);
            const double xh = __dsub_rn(__dmul_rn(2.0, new_x), xo);
            a.xhat[j] = xh;
            lxh[j] = xh;
        }
        avg_count++;
        __syncthreads();
        for (int i = tid; i < a.m; i += nt) {  // A x_hat + dual step
            const double ax = ordered_dot(a.rp[i], a.rp[i + 1], a.rv, a.ci, lxh);
            const double yo = ly[i];
            const double next_y = __dadd_rn(yo, __dmul_rn(a.sigma[i], __dsub_rn(ax, a.b[i])));
            const double new_y = (!a.eq[i]) ? ref_max(0.0, next_y) : next_y;
            dd = ref_max(dd, ref_max(0.0, fabs(__dsub_rn(new_y, yo))));
            a.y[i] = new_y;
            ly[i] = new_y;
        }
        __syncthreads();
        if (!track) continue;
        const double pmax = block_max_ref(dp, sh);
        const double dmax = block_max_ref(dd, sh);
        double bv = 0.0;
        for (int j = tid; j < a.n; j += nt) {  // x_avg and the bound violation of x
            const double xj = a.x[j];
            if (xj < a.lb[j] - 1e-12) bv = ref_max(bv, a.lb[j] - xj);
            if (xj > a.ub[j] + 1e-12) bv = ref_max(bv, xj - a.ub[j]);
            a.xavg[j] = dclamp(a.xsum[j] / static_cast<double>(avg_count), a.lb[j], a.ub[j]);
        }
        __syncthreads();
        double v1 = 0.0, v2 = 0.0;
        for (int i = tid; i < a.m; i += nt) {  // row violations of x and x_avg
            double s1 = 0.0, s2 = 0.0;
            for (int k = a.rp[i]; k < a.rp[i + 1]; ++k) {
                const double w = a.rv[k];
                const int j = a.ci[k];
                s1 = __dadd_rn(s1, __dmul_rn(w, a.x[j]));
                s2 = __dadd_rn(s2, __dmul_rn(w, a.xavg[j]));
            }
            v1 = ref_max(v1, row_viol(s1, a.b[i], a.eq[i]));
            v2 = ref_max(v2, row_viol(s2, a.b[i], a.eq[i]));
        }
        bviol = block_max_ref(bv, sh);
        const double max_viol = block_max_ref(v1, sh);
        const double avg_viol = block_max_ref(v2, sh);
        best = (avg_viol < max_viol) ? avg_viol : max_viol;  // std::min
        if (isnan(best) || best > 1e15) {
            status = 2;
            break;
        }
        const bool feasible = best <= a.feas_tol && bviol <= a.feas_tol;
        const bool stationary = pmax <= 1e-7 && dmax <= 1e-7;
        if (feasible && stationary && iter >= 100) {
            status = 0;
            break;
        }
    }
    if (tid == 0) {
        a.out[0] = static_cast<double>(iter);
        a.out[1] = static_cast<double>(iter);
        a.out[2] = static_cast<double>(status);
        a.out[3] = best;
        a.out[4] = bviol;
        a.out[5] = static_cast<double>(avg_count);
    }
}

// The one-workgroup loop with the whole LP in LDS: both orientations of the
// matrix and the gathered vectors (y, x_hat; x at checks) live in LDS, and
// each thread keeps the data of the R columns and R rows it owns (j = tid +
// r * 1024) in registers, so x, x_sum and y reach memory only at checks and
// at the end.  The sparse dots load four entries ahead and then add them in
// the reference's order, one at a time: the same additions, the same bits.
constexpr int kRefLdsMaxPerThread = 2;  // 4 would spill registers

inline size_t ref_lds_bytes(long long m, long long n, long long nnz) {
    return sizeof(double) * static_cast<size_t>(m + n + 2 * nnz) + sizeof(int) * static_cast<size_t>(m + n + 2 + 2 * nnz);
}

template <int R>
__global__ void __launch_bounds__(kRefSmallThreads) k_ref_lds(RefSmallArgs a) {
    extern __shared__ double dyn[];
    __shared__ double sh[kRefSmallThreads / kWave];
    const int m = a.m, n = a.n, nnz = a.nnz;
    double* ly = dyn;                                // y [m]
    double* lxh = ly + m;                            // x_hat [n]; x during a check
    double* lrv = lxh + n;                           // CSR values [nnz]
    double* lcv = lrv + nnz;                         // CSC values [nnz]
    int* lrp = reinterpret_cast<int*>(lcv + nnz);    // [m + 1]
    int* lcp = lrp + m + 1;                          // [n + 1]
    int* lci = lcp + n + 1;                          // [nnz]
    int* lri = lci + nnz;                            // [nnz]
    const int tid = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
    for (int k = tid; k < nnz; k += nt) {
        lrv[k] = a.rv[k];
        lcv[k] = a.cv[k];
        lci[k] = a.ci[k];
        lri[k] = a.ri[k];
    }
    for (int i = tid; i <= m; i += nt) lrp[i] = a.rp[i];
    for (int j = tid; j <= n; j += nt) lcp[j] = a.cp[j];
    for (int i = tid; i < m; i += nt) ly[i] = a.y[i];
    double x[R], xs[R], cc[R], tau[R], lo[R], hi[R];
    double y[R], bb[R], sg[R];
    bool eq[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = tid + r * nt, i = tid + r * nt;
        const bool cj = j < n, ri = i < m;
        x[r] = cj ? a.x[j] : 0.0;
        xs[r] = cj ? a.xsum[j] : 0.0;
        cc[r] = cj ? a.c[j] : 0.0;
        tau[r] = cj ? a.tau[j] : 0.0;
        lo[r] = cj ? a.lb[j] : 0.0;
        hi[r] = cj ? a.ub[j] : 0.0;
        y[r] = ri ? a.y[i] : 0.0;
        bb[r] = ri ? a.b[i] : 0.0;
        sg[r] = ri ? a.sigma[i] : 0.0;
        eq[r] = ri ? (a.eq[i] != 0) : false;
    }
    __syncthreads();
    long long iter = 0, avg_count = 0;
    int status = 1;
    double best = 0.0, bviol = 0.0;
    long long to_check = a.check_freq;
    while (iter < a.max_iter) {
        iter++;
        const bool track = (--to_check == 0 || iter == a.max_iter);  // iter % check_freq == 0 || last
        if (to_check == 0) to_check = a.check_freq;
        double dp = 0.0, dd = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {  // A'y + primal step + ergodic sum + extrapolation
            const int j = tid + r * nt;
            if (j < n) {
                const double aty = ordered_dot(lcp[j], lcp[j + 1], lcv, lri, ly);
                const double xo = x[r];
                const double grad = __dadd_rn(cc[r], aty);
                const double next_x = __dsub_rn(xo, __dmul_rn(tau[r], grad));
                const double new_x = dclamp(next_x, lo[r], hi[r]);
                dp = ref_max(dp, ref_max(0.0, fabs(__dsub_rn(new_x, xo))));
                x[r] = new_x;
                xs[r] = __dadd_rn(xs[r], new_x);
                lxh[j] = __dsub_rn(__dmul_rn(2.0, new_x), xo);
            }
        }
        avg_count++;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {  // A x_hat + dual step
            const int i = tid + r * nt;
            if (i < m) {
                const double ax = ordered_dot(lrp[i], lrp[i + 1], lrv, lci, lxh);
                const double yo = y[r];
                const double next_y = __dadd_rn(yo, __dmul_rn(sg[r], __dsub_rn(ax, bb[r])));
                const double new_y = (!eq[r]) ? ref_max(0.0, next_y) : next_y;
                dd = ref_max(dd, ref_max(0.0, fabs(__dsub_rn(new_y, yo))));
                y[r] = new_y;
                ly[i] = new_y;
            }
        }
        __syncthreads();
        if (!track) continue;
        const double pmax = block_max_ref(dp, sh);
        const double dmax = block_max_ref(dd, sh);
        double bv = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {  // x (into LDS), x_avg and the bound violation of x
            const int j = tid + r * nt;
            if (j < n) {
                const double xj = x[r];
                if (xj < lo[r] - 1e-12) bv = ref_max(bv, lo[r] - xj);
                if (xj > hi[r] + 1e-12) bv = ref_max(bv, xj - hi[r]);
                lxh[j] = xj;
                a.xavg[j] = dclamp(xs[r] / static_cast<double>(avg_count), lo[r], hi[r]);
            }
        }
        __syncthreads();
        double v1 = 0.0, v2 = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {  // row violations of x and x_avg
            const int i = tid + r * nt;
            if (i < m) {
                double s1 = 0.0, s2 = 0.0;
                for (int k = lrp[i]; k < lrp[i + 1]; ++k) {
                    const double w = lrv[k];
                    const int j = lci[k];
                    s1 = __dadd_rn(s1, __dmul_rn(w, lxh[j]));
                    s2 = __dadd_rn(s2, __dmul_rn(w, a.xavg[j]));
                }
                v1 = ref_max(v1, row_viol(s1, bb[r], eq[r] ? 1 : 0));
                v2 = ref_max(v2, row_viol(s2, bb[r], eq[r] ? 1 : 0));
            }
        }
        bviol = block_max_ref(bv, sh);
        const double max_viol = block_max_ref(v1, sh);
        const double avg_viol = block_max_ref(v2, sh);
        best = (avg_viol < max_viol) ? avg_viol : max_viol;  // std::min
        if (isnan(best) || best > 1e15) {
            status = 2;
            break;
        }
        const bool feasible = best <= a.feas_tol && bviol <= a.feas_tol;
        const bool stationary = pmax <= 1e-7 && dmax <= 1e-7;
        if (feasible && stationary && iter >= 100) {
            status = 0;
            break;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = tid + r * nt, i = tid + r * nt;
        if (j < n) {
            a.x[j] = x[r];
            a.xsum[j] = xs[r];
        }
        if (i < m) a.y[i] = y[r];
    }
    if (tid == 0) {
        a.out[0] = static_cast<double>(iter);
        a.out[1] = static_cast<double>(iter);
        a.out[2] = static_cast<double>(status);
        a.out[3] = best;
        a.out[4] = bviol;
        a.out[5] = static_cast<double>(avg_count);
    }
}

struct DevBufs {
    std::vector<void*> ptrs;
    ~DevBufs() {
        for (void* p : ptrs) hipFree(p);
    }
    template <class T>
    hipError_t alloc(T** p, size_t count) {
        *p = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

inline double u64_to_double(unsigned long long u) {
    double d;
    std::memcpy(&d, &u, sizeof(d));
    return d;
}

__global__ void k_fill_zero(double* __restrict__ p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// Device arrays of one solve of the reference loop.
struct RefDev {
    int *rp, *ci, *cp, *ri;
    double *rv, *cv, *b, *c, *lb, *ub, *tau, *sigma, *x, *y, *xsum, *xhat, *xavg, *out;
    unsigned char* eq;
    unsigned long long* red;  // [0] pmax [1] dmax [2] viol [3] avg viol [4] bound viol
};

// Host side of a solve, into caller-provided arrays: the CSC (scanning rows
// in order keeps each column's rows ascending, the order the reference's
// row-order scatter adds into each A'y entry; pos holds n counters) and the
// starting point (src/first_order_solver.cpp:131-142).
void ref_host_prep(const bo_lp* lp, int* cp, int* ri, double* cv, int* pos, double* x0) {
    const int m = lp->m, n = lp->n;
    const long long nnz = lp->nnz;
    for (int j = 0; j <= n; ++j) cp[j] = 0;
    for (long long k = 0; k < nnz; ++k) cp[lp->col_idx[k] + 1]++;
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    for (int j = 0; j < n; ++j) pos[j] = cp[j];
    for (int i = 0; i < m; ++i)
        for (int k = lp->row_ptr[i]; k < lp->row_ptr[i + 1]; ++k) {
            const int q = pos[lp->col_idx[k]]++;
            ri[q] = i;
            cv[q] = lp->val[k];
        }
    for (int j = 0; j < n; ++j) {
        const double l = lp->lb[j], u = lp->ub[j];
        double target = 0.0;
        if (l > -INFINITY && u < INFINITY) target = 0.5 * (l + u);
        else if (l > -INFINITY) target = l;
        else if (u < INFINITY) target = u;
        x0[j] = (target < l) ? l : ((u < target) ? u : target);
    }
}

// Uploads the LP, its CSC and starting point into d, clears y, the ergodic
// sum and x_hat, runs the loop (one workgroup when the LP fits, else a
// kernel per pass with the checks on the host), chooses the iterate and
// downloads x and y.  Allocates nothing: every array is in d.  All work is
// on the null stream; transfer_time_ms runs from t_xfer0 (the caller's start
// of the host preparation) to the end of the uploads, kernel_time_ms is the
// host's clock from the first launch to the last download.
bo_status ref_run(const bo_lp* lp, const RefDev& d, const int* cp, const int* ri, const double* cv,
                  const double* x0, const bo_ref_options* o, double* x_out, double* y_out, bo_ref_result* res,
                  std::chrono::high_resolution_clock::time_point t_xfer0) {
    const int m = lp->m, n = lp->n;
    const long long nnz = lp->nnz;
    std::memset(res, 0, sizeof(*res));
    auto h2d = [](void* dst, const void* src, size_t bytes) {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, 0) : hipSuccess;
    };
    if (m > 0) BO_TRY(h2d(d.rp, lp->row_ptr, sizeof(int) * (m + 1)));
    BO_TRY(h2d(d.ci, lp->col_idx, sizeof(int) * nnz));
    BO_TRY(h2d(d.rv, lp->val, sizeof(double) * nnz));
    BO_TRY(h2d(d.cp, cp, sizeof(int) * (n + 1)));
    BO_TRY(h2d(d.ri, ri, sizeof(int) * nnz));
    BO_TRY(h2d(d.cv, cv, sizeof(double) * nnz));
    BO_TRY(h2d(d.b, lp->b, sizeof(double) * m));
    BO_TRY(h2d(d.eq, lp->is_eq, m));
    BO_TRY(h2d(d.c, lp->c, sizeof(double) * n));
    BO_TRY(h2d(d.lb, lp->lb, sizeof(double) * n));
    BO_TRY(h2d(d.ub, lp->ub, sizeof(double) * n));
    BO_TRY(h2d(d.x, x0, sizeof(double) * n));
    if (m > 0) hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(m)), dim3(kBlock), 0, 0, d.y, m);
    if (n > 0) {
        hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n)), dim3(kBlock), 0, 0, d.xsum, n);
        hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n)), dim3(kBlock), 0, 0, d.xhat, n);
    }
    BO_TRY(hipGetLastError());
    BO_TRY(hipStreamSynchronize(0));
    const auto t_xfer1 = std::chrono::high_resolution_clock::now();

    const unsigned gm = grid_for(m), gn = grid_for(n), gmax = grid_for(m > n ? m : n);
    hipLaunchKernelGGL(k_ref_steps, dim3(gmax), dim3(kBlock), 0, 0, m, n, d.rp, d.rv, d.cp, d.cv,
                       o->adaptive_step_size ? 1 : 0, o->primal_step_size, o->dual_step_size, d.tau, d.sigma);

    long long check_freq = o->check_frequency < 100 ? o->check_frequency : 100;
    if (check_freq < 1) check_freq = 1;  // the reference divides by zero here
    long long iter = 0, avg_count = 0;
    bool converged = false, failed = false;
    unsigned long long red[8];

    // Which loop: the whole LP in one workgroup's LDS, one workgroup over the
    // matrix in memory, or a kernel per pass with the checks on the host.
    int dev = 0, lds_max = 0;
    BO_TRY(hipGetDevice(&dev));
    BO_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t lds_bytes = ref_lds_bytes(m, n, nnz);
    const int per_thread = static_cast<int>(((m > n ? m : n) + kRefSmallThreads - 1) / kRefSmallThreads);
    const bool lds_fit = per_thread <= kRefLdsMaxPerThread && nnz < (1LL << 30) &&
                         lds_bytes + 1024 <= static_cast<size_t>(lds_max);
    const bool one_wg = lds_fit || (static_cast<long long>(n) + m <= kRefSmallMaxEntries && nnz <= kRefSmallMaxNnz);
    res->loop = lds_fit ? 2 : (one_wg ? 1 : 0);
    if (one_wg && o->max_iterations > 0) {
        RefSmallArgs a{m, n, static_cast<int>(nnz), d.rp, d.ci, d.cp, d.ri, d.rv, d.cv, d.b, d.c, d.lb, d.ub,
                       d.tau, d.sigma, d.eq, d.x, d.y, d.xsum, d.xhat, d.xavg, o->max_iterations, check_freq,
                       o->feasibility_tolerance, d.out};
        if (lds_fit) {
            const void* fn = per_thread <= 1 ? reinterpret_cast<const void*>(&k_ref_lds<1>)
                                             : reinterpret_cast<const void*>(&k_ref_lds<2>);
            BO_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_bytes)));
            if (per_thread <= 1)
                hipLaunchKernelGGL(k_ref_lds<1>, dim3(1), dim3(kRefSmallThreads), lds_bytes, 0, a);
            else
                hipLaunchKernelGGL(k_ref_lds<2>, dim3(1), dim3(kRefSmallThreads), lds_bytes, 0, a);
        } else {
            hipLaunchKernelGGL(k_ref_small, dim3(1), dim3(kRefSmallThreads), 0, 0, a);
        }
        BO_TRY(hipGetLastError());
        double out[8];
        BO_TRY(hipMemcpyAsync(out, d.out, sizeof(out), hipMemcpyDeviceToHost, 0));
        BO_TRY(hipStreamSynchronize(0));
        iter = static_cast<long long>(out[0]);
        avg_count = static_cast<long long>(out[5]);
        const int status = static_cast<int>(out[2]);
        converged = status == 0;
        failed = status == 2;
        res->constraint_violation = out[3];
        res->bound_violation = out[4];
    }
    while (!one_wg && iter < o->max_iterations) {
        iter++;
        const int track = (iter % check_freq == 0 || iter == o->max_iterations) ? 1 : 0;
        if (track) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, 0, d.red, 8);
        hipLaunchKernelGGL(k_ref_primal, dim3(gn), dim3(kBlock), 0, 0, n, d.cp, d.ri, d.cv, d.y, d.c,
                           d.tau, d.lb, d.ub, d.x, d.xsum, d.xhat, d.red + 0, track);
        avg_count++;
        hipLaunchKernelGGL(k_ref_dual, dim3(gm), dim3(kBlock), 0, 0, m, d.rp, d.ci, d.rv, d.xhat, d.b,
                           d.sigma, d.eq, d.y, d.red + 1, track);
        if (track) {
            hipLaunchKernelGGL(k_ref_avg, dim3(gn), dim3(kBlock), 0, 0, n, d.x, d.xsum,
                               static_cast<double>(avg_count), d.lb, d.ub, d.xavg, d.red + 4, 0);
            hipLaunchKernelGGL(k_ref_viol, dim3(gm), dim3(kBlock), 0, 0, m, d.rp, d.ci, d.rv, d.x, d.xavg,
                               d.b, d.eq, d.red + 2, d.red + 3, 1);
            BO_TRY(hipGetLastError());
            BO_TRY(hipMemcpyAsync(red, d.red, sizeof(red), hipMemcpyDeviceToHost, 0));
            BO_TRY(hipStreamSynchronize(0));
            const double pmax = u64_to_double(red[0]), dmax = u64_to_double(red[1]);
            const double max_viol = u64_to_double(red[2]), avg_viol = u64_to_double(red[3]);
            const double bviol = u64_to_double(red[4]);
            const double best = (avg_viol < max_viol) ? avg_viol : max_viol;  // std::min
            res->constraint_violation = best;
            res->bound_violation = bviol;
            if (std::isnan(best) || best > 1e15) {
                failed = true;
                break;
            }
            const bool feasible = best <= o->feasibility_tolerance && bviol <= o->feasibility_tolerance;
            const bool stationary = pmax <= 1e-7 && dmax <= 1e-7;
            if (feasible && stationary && iter >= 100) {
                converged = true;
                break;
            }
        }
    }
    res->loop_iterations = iter;
    if (failed) {
        res->status = 2;
        res->iterations = 0;  // the reference returns before recording the count
    } else {
        res->status = converged ? 0 : 1;
        res->iterations = iter;
        // Iterate selection (src/first_order_solver.cpp:261-279).
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, 0, d.red, 8);
        hipLaunchKernelGGL(k_ref_avg, dim3(gn), dim3(kBlock), 0, 0, n, d.x, d.xsum,
                           static_cast<double>(avg_count), d.lb, d.ub, d.xavg, d.red + 4, avg_count > 0 ? 0 : 1);
        hipLaunchKernelGGL(k_ref_viol, dim3(gm), dim3(kBlock), 0, 0, m, d.rp, d.ci, d.rv, d.x, d.xavg,
                           d.b, d.eq, d.red + 2, d.red + 3, 0);
        BO_TRY(hipGetLastError());
        BO_TRY(hipMemcpyAsync(red, d.red, sizeof(red), hipMemcpyDeviceToHost, 0));
        BO_TRY(hipStreamSynchronize(0));
        const double avg_viol = u64_to_double(red[3]);
        const bool use_avg = !converged && (avg_viol <= res->constraint_violation);
        res->used_average = use_avg ? 1 : 0;
        if (n > 0)
            BO_TRY(hipMemcpyAsync(x_out, use_avg ? d.xavg : d.x, sizeof(double) * n, hipMemcpyDeviceToHost, 0));
        if (m > 0) BO_TRY(hipMemcpyAsync(y_out, d.y, sizeof(double) * m, hipMemcpyDeviceToHost, 0));
        BO_TRY(hipStreamSynchronize(0));
    }
    BO_TRY(hipGetLastError());
    const auto t_end = std::chrono::high_resolution_clock::now();
    res->kernel_time_ms = std::chrono::duration<double, std::milli>(t_end - t_xfer1).count();
    res->transfer_time_ms = std::chrono::duration<double, std::milli>(t_xfer1 - t_xfer0).count();
    return BO_SUCCESS;
}

// ---- the allocation-free variant (bo_pdhg_ref_solve_noalloc) ---------------
// LPs that run in one workgroup: m + n <= 7,680 and nnz <= 65,536.  Every
// device array of a solve is carved from `scratch`, a fixed array in each
// GPU's memory (kNoAllocScratchBytes, held until the process ends); the
// host-side CSC and starting point go to fixed static host arrays.
constexpr int kNoAllocMaxEntries = kRefSmallMaxEntries;           // m + n
constexpr long long kNoAllocMaxNnz = kRefSmallMaxNnz;               // nonzeros
constexpr size_t kScratchAlign = 256;
constexpr size_t scratch_round(size_t b) { return (b + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }
// int arrays: rp (m+1), ci, cp (n+1), ri; double arrays: rv, cv, b, sigma,
// y (m each), c, lb, ub, tau, x, xsum, xhat, xavg (n each), out (8); eq (m);
// red (8 x u64); each rounded up to the alignment.
constexpr size_t kNoAllocScratchBytes =
    4 * (kNoAllocMaxEntries + 2 + 2 * kNoAllocMaxNnz) + 8 * (2 * kNoAllocMaxNnz + 8 * kNoAllocMaxEntries + 8) +
    kNoAllocMaxEntries + 64 + 21 * kScratchAlign;

}  // namespace

bo_status ensure_device_public();

}  // namespace bo

// One per GPU (each device holds its own copy at its own address).  Host
// code never uses this name as a device address: it looks the address up
// with hipGetSymbolAddress and passes that to the kernels.
static __device__ __align__(256) unsigned char scratch[bo::kNoAllocScratchBytes];

using namespace bo;

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_lock;               // one call at a time: the host staging arrays are shared
void* g_scratch_ptr[kMaxDevices] = {};   // scratch's address on each device, looked up once
int g_h_cp[kNoAllocMaxEntries + 1];
int g_h_pos[kNoAllocMaxEntries];
int g_h_ri[kNoAllocMaxNnz];
double g_h_cv[kNoAllocMaxNnz];
double g_h_x0[kNoAllocMaxEntries];

// scratch's device address on the current device (looked up with that device
// current, the first time it is needed there).  Called under g_noalloc_lock.
bo_status scratch_address(void** out) {
    int dev = 0;
    BO_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_last_error("bo_pdhg_ref_solve_noalloc: device index beyond the library's table");
        return BO_INVALID_ARGUMENT;
    }
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        BO_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_last_error("bo_pdhg_ref_solve_noalloc: hipGetSymbolAddress returned no address");
            return BO_EXECUTION_FAILED;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BO_SUCCESS;
}

// Carves the solve's arrays from the scratch address `base`, in a fixed order.
RefDev carve(unsigned char* base, int m, int n, long long nnz, size_t* used) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += scratch_round(bytes ? bytes : 1);
        return p;
    };
    RefDev d{};
    d.rp = reinterpret_cast<int*>(take(sizeof(int) * (m + 1)));
    d.ci = reinterpret_cast<int*>(take(sizeof(int) * nnz));
    d.cp = reinterpret_cast<int*>(take(sizeof(int) * (n + 1)));
    d.ri = reinterpret_cast<int*>(take(sizeof(int) * nnz));
    d.rv = reinterpret_cast<double*>(take(sizeof(double) * nnz));
    d.cv = reinterpret_cast<double*>(take(sizeof(double) * nnz));
    d.b = reinterpret_cast<double*>(take(sizeof(double) * m));
    d.sigma = reinterpret_cast<double*>(take(sizeof(double) * m));
    d.y = reinterpret_cast<double*>(take(sizeof(double) * m));
    d.c = reinterpret_cast<double*>(take(sizeof(double) * n));
    d.lb = reinterpret_cast<double*>(take(sizeof(double) * n));
    d.ub = reinterpret_cast<double*>(take(sizeof(double) * n));
    d.tau = reinterpret_cast<double*>(take(sizeof(double) * n));
    d.x = reinterpret_cast<double*>(take(sizeof(double) * n));
    d.xsum = reinterpret_cast<double*>(take(sizeof(double) * n));
    d.xhat = reinterpret_cast<double*>(take(sizeof(double) * n));
    d.xavg = reinterpret_cast<double*>(take(sizeof(double) * n));
    d.out = reinterpret_cast<double*>(take(sizeof(double) * 8));
    d.eq = take(m);
    d.red = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 8));
    *used = off;
    return d;
}

}  // namespace

extern "C" bo_status bo_pdhg_ref_solve(const bo_lp* lp, const bo_ref_options* o, double* x_out,
                                       double* y_out, bo_ref_result* res) {
    return host_guard([&]() -> bo_status {
        if (!o || !res || !valid_csr_lp(lp) || (lp->n > 0 && !x_out) || (lp->m > 0 && !y_out))
            return BO_INVALID_ARGUMENT;
        if (lp->nnz >= (1LL << 31)) return BO_INVALID_ARGUMENT;  // int32 offsets
        bo_status st = ensure_device_public();
        if (st != BO_SUCCESS) return st;
        const int m = lp->m, n = lp->n;
        const long long nnz = lp->nnz;
        const auto t0 = std::chrono::high_resolution_clock::now();
        std::vector<int> cp(static_cast<size_t>(n) + 1), ri(static_cast<size_t>(nnz)), pos(static_cast<size_t>(n) + 1);
        std::vector<double> cv(static_cast<size_t>(nnz)), x0(static_cast<size_t>(n) + 1);
        ref_host_prep(lp, cp.data(), ri.data(), cv.data(), pos.data(), x0.data());
        DevBufs mem;
        RefDev d{};
        BO_TRY(mem.alloc(&d.rp, m + 1));
        BO_TRY(mem.alloc(&d.ci, nnz));
        BO_TRY(mem.alloc(&d.rv, nnz));
        BO_TRY(mem.alloc(&d.cp, n + 1));
        BO_TRY(mem.alloc(&d.ri, nnz));
        BO_TRY(mem.alloc(&d.cv, nnz));
        BO_TRY(mem.alloc(&d.b, m));
        BO_TRY(mem.alloc(&d.eq, m));
        BO_TRY(mem.alloc(&d.sigma, m));
        BO_TRY(mem.alloc(&d.y, m));
        BO_TRY(mem.alloc(&d.c, n));
        BO_TRY(mem.alloc(&d.lb, n));
        BO_TRY(mem.alloc(&d.ub, n));
        BO_TRY(mem.alloc(&d.tau, n));
        BO_TRY(mem.alloc(&d.x, n));
        BO_TRY(mem.alloc(&d.xsum, n));
        BO_TRY(mem.alloc(&d.xhat, n));
        BO_TRY(mem.alloc(&d.xavg, n));
        BO_TRY(mem.alloc(&d.out, 8));
        BO_TRY(mem.alloc(&d.red, 8));
        return ref_run(lp, d, cp.data(), ri.data(), cv.data(), x0.data(), o, x_out, y_out, res, t0);
    });
}

extern "C" bo_status bo_pdhg_ref_solve_noalloc(const bo_lp* lp, const bo_ref_options* o, double* x_out,
                                               double* y_out, bo_ref_result* res) {
    return host_guard([&]() -> bo_status {
        if (!o || !res || !valid_csr_lp(lp) || (lp->n > 0 && !x_out) || (lp->m > 0 && !y_out))
            return BO_INVALID_ARGUMENT;
        if (static_cast<long long>(lp->m) + lp->n > kNoAllocMaxEntries || lp->nnz > kNoAllocMaxNnz) {
            set_last_error("bo_pdhg_ref_solve_noalloc: the LP exceeds the fixed scratch (needs m + n <= " +
                           std::to_string(kNoAllocMaxEntries) + " and nnz <= " + std::to_string(kNoAllocMaxNnz) +
                           "; got m + n = " + std::to_string(static_cast<long long>(lp->m) + lp->n) +
                           ", nnz = " + std::to_string(lp->nnz) + "); use bo_pdhg_ref_solve");
            return BO_INVALID_ARGUMENT;
        }
        bo_status st = ensure_device_public();
        if (st != BO_SUCCESS) return st;
        std::lock_guard<std::mutex> hold(g_noalloc_lock);
        void* base = nullptr;
        if ((st = scratch_address(&base)) != BO_SUCCESS) return st;
        size_t used = 0;
        const RefDev d = carve(static_cast<unsigned char*>(base), lp->m, lp->n, lp->nnz, &used);
        if (used > kNoAllocScratchBytes) {  // cannot happen within the limits above
            set_last_error("bo_pdhg_ref_solve_noalloc: scratch layout exceeds its size");
            return BO_EXECUTION_FAILED;
        }
        const auto t0 = std::chrono::high_resolution_clock::now();
        ref_host_prep(lp, g_h_cp, g_h_ri, g_h_cv, g_h_pos, g_h_x0);
        return ref_run(lp, d, g_h_cp, g_h_ri, g_h_cv, g_h_x0, o, x_out, y_out, res, t0);
    });
}
