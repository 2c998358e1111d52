// pdlp.hip -- restarted, reflected Halpern PDHG engine for one LP on one MI355X.
//
// Replaces the loop of CPUFirstOrderSolver::solve (src/first_order_solver.cpp:158-247)
// when the caller asks for a relative-KKT target.  The algorithm comes from
// the PDLP line of work (Applegate et al. 2021; Lu & Yang, cuPDLP 2023 and
// the restarted Halpern PDHG of cuPDLPx 2025), not from the reference, which
// has none of it:
//   * Ruiz equilibration (inf-norm, 10 sweeps) + Pock-Chambolle (alpha = 1),
//     computed on the device from the resident FP64 matrix on every solve;
//   * constant step eta = 0.95 / ||A_hat||_2 (device power iteration),
//     tau = eta / omega, sigma = eta * omega;
//   * one PDHG step T(x, y) = (x+, y+) in the reference sign convention
//     (min c'x, Ax <= b or = b, y_I >= 0):
//         x+ = proj_[l,u](x - tau (c + A'y))             CSC pass, fused epilogue
//         y+ = proj_Y(y + sigma (A (2 x+ - x) - b))       CSR pass, fused epilogue
//     followed by the reflected Halpern update towards the restart anchor z0:
//         z <- w (2 T(z) - z) + (1 - w) z0,   w = (k+1)/(k+2), k = steps since restart.
//     A x is carried (A z is linear in the update), so each step costs
//     exactly one CSC and one CSR pass, two launches;
//   * every `check_every` steps an FP64 check on the *unscaled* problem for
//     the current iterate and for T(z) (rel-KKT and weighted KKT), then a
//     one-workgroup decision kernel applies the KKT restart rule
//     (beta 0.2 / 0.8 / 0.36), the primal-weight update (theta 0.5) and the
//     stopping test -- no host round trip inside an interval, which is
//     captured once as a hipGraph and replayed.
// Precision: matrix values and the vectors the SpMVs gather (x+ and a copy of
// y) are FP32 (FP64 in the error-comparison mode); accumulations, the
// Halpern state (iterate, anchor, carried A x), checks and the returned
// solution are FP64.  Rows and columns are stored sorted by descending
// length so one length-binned segment launch (segments.hpp) covers a product.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "pdlp_core.hpp"
#include "segments.hpp"
#include "stream.hpp"
#include "tiles.hpp"

namespace bo {

bo_status ensure_device_public();

namespace {

// ---------------------------------------------------------------- epilogues
// out = A v, partial sum of squares of out.
template <class TX, class TO>
struct EpiStore {
    using TXv = TX;
    static constexpr int NV = 1;
    static constexpr int Q = 1;
    const TX* v;
    TO* out;
    __device__ const TX* vec(int) const { return v; }
    __device__ void operator()(int seg, const double* acc, double* part) const {
        out[seg] = static_cast<TO>(acc[0]);
        part[0] += acc[0] * acc[0];
    }
};

// norm_out[seg] = own_scale[seg] * acc   (row/column norms of the scaled matrix)
struct EpiNorm {
    using TXv = double;
    static constexpr int NV = 1;
    static constexpr int Q = 0;
    const double* other_scale;
    const double* own_scale;
    double* norm_out;
    __device__ const double* vec(int) const { return other_scale; }
    __device__ void operator()(int seg, const double* acc, double*) const {
        norm_out[seg] = own_scale[seg] * acc[0];
    }
};

// out[k] = own_scale[seg] * val[k] * other_scale[j]   (materialise scaled values)
template <class TO>
struct EpiWrite {
    using TXv = double;
    static constexpr int NV = 1;
    static constexpr int Q = 0;
    const double* own_scale;
    const double* other_scale;
    TO* out;
    __device__ const double* vec(int) const { return other_scale; }
    __device__ void write(int seg, int k, double a, int j) const {
        out[k] = static_cast<TO>(own_scale[seg] * a * other_scale[j]);
    }
    __device__ void operator()(int, const double*, double*) const {}
};

// FP64 check, row side: gathers unscaled x (current, candidate).
// partials: [0,1] ||viol||^2 (cur, cand)  [2,3] scaled  [4,5] b'y (cur, cand)
template <class T>
struct EpiCheckRows {
    using TXv = double;
    static constexpr int NV = 2;
    static constexpr int Q = 6;
    const double* xo_cur;
    const double* xo_cand;
    const double* b;
    const unsigned char* eq;
    const double* rs;
    const double* yk;
    const T* yp;
    double* yo_cur;
    double* yo_cand;
    double* ax_cand;  // scaled A x_cand, taken over when restarting to the candidate
    __device__ const double* vec(int v) const { return v == 0 ? xo_cur : xo_cand; }
    __device__ void operator()(int i, const double* acc, double* part) const {
        const double bi = b[i], ri = rs[i];
        const bool e = eq[i] != 0;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double r = acc[p] - bi;
            const double v = e ? fabs(r) : (r > 0.0 ? r : 0.0);
            part[p] += v * v;
            part[2 + p] += (ri * v) * (ri * v);
        }
        ax_cand[i] = ri * acc[1];
        double yc = ri * yk[i];
        if (!e && yc < 0.0) yc = 0.0;  // the reported dual is sign-feasible
        const double ya = ri * static_cast<double>(yp[i]);
        yo_cur[i] = yc;
        yo_cand[i] = ya;
        part[4] += bi * yc;
        part[5] += bi * ya;
    }
};

// FP64 check, column side: gathers unscaled y (current, candidate).
// partials: [0,1] ||dual res||^2  [2,3] scaled  [4,5] sum l*plam^+ + u*plam^-
struct EpiCheckCols {
    using TXv = double;
    [[maybe_unused]] static constexpr int NV = 2;  // (host code of this file no longer names them)
    [[maybe_unused]] static constexpr int Q = 6;
    const double* yo_cur;
    const double* yo_cand;
    const double* c;
    const double* lb;
    const double* ub;
    const double* cs;
    __device__ const double* vec(int v) const { return v == 0 ? yo_cur : yo_cand; }
    __device__ void operator()(int j, const double* acc, double* part) const {
        const double l = lb[j], u = ub[j], sj = cs[j];
        const bool lf = isfinite(l), uf = isfinite(u);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double lam = c[j] + acc[p];
            double pl;
            if (lf && uf) pl = lam;
            else if (lf) pl = lam > 0.0 ? lam : 0.0;
            else if (uf) pl = lam < 0.0 ? lam : 0.0;
            else pl = 0.0;
            const double d = lam - pl;
            part[p] += d * d;
            part[2 + p] += (sj * d) * (sj * d);
            part[4 + p] += pl > 0.0 ? l * pl : (pl < 0.0 ? u * pl : 0.0);
        }
    }
};

// The same two check passes with the current point and the candidate
// interleaved (entry j of both in one double2): the multi-kernel loop's
// checks gather one 16 bytes per nonzero instead of two 8 bytes, so the
// passes issue half the gather requests.  Same arithmetic, same sums.
template <class T>
struct EpiCheckRows2 {
    using TXv = double;
    static constexpr int NV = 2;
    static constexpr int Q = 6;
    static constexpr bool kPair = true;
    const double2* xo2;  // unscaled x (current, candidate)
    const double* b;
    const unsigned char* eq;
    const double* rs;
    const double* yk;
    const T* yp;
    double2* yo2;     // unscaled y (current, candidate), gathered by the column pass
    double* ax_cand;  // scaled A x_cand, taken over when restarting to the candidate
    __device__ const double2* vec2() const { return xo2; }
    __device__ void operator()(int i, const double* acc, double* part) const {
        const double bi = b[i], ri = rs[i];
        const bool e = eq[i] != 0;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double r = acc[p] - bi;
            const double v = e ? fabs(r) : (r > 0.0 ? r : 0.0);
            part[p] += v * v;
            part[2 + p] += (ri * v) * (ri * v);
        }
        ax_cand[i] = ri * acc[1];
        double yc = ri * yk[i];
        if (!e && yc < 0.0) yc = 0.0;  // the reported dual is sign-feasible
        const double ya = ri * static_cast<double>(yp[i]);
        yo2[i] = make_double2(yc, ya);
        part[4] += bi * yc;
        part[5] += bi * ya;
    }
};

struct EpiCheckCols2 {
    using TXv = double;
    static constexpr int NV = 2;
    static constexpr int Q = 6;
    static constexpr bool kPair = true;
    const double2* yo2;
    const double* c;
    const double* lb;
    const double* ub;
    const double* cs;
    __device__ const double2* vec2() const { return yo2; }
    __device__ void operator()(int j, const double* acc, double* part) const {
        EpiCheckCols{nullptr, nullptr, c, lb, ub, cs}(j, acc, part);
    }
};

// ---------------------------------------------------------------- elementwise
__global__ void k_fill_f64(double* p, double v, size_t n) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// scale[i] /= sqrt(norm[i]) where norm > 0
__global__ void k_apply_norm(double* scale, const double* norm, size_t n) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const double v = norm[i];
        if (v > 0.0 && isfinite(v)) scale[i] /= sqrt(v);
    }
}

// Scaled column data c_hat = s c, l_hat = l / s, u_hat = u / s and the
// starting point (projection of 0 or of the warm start) as iterate + anchor.
template <class T>
__global__ void k_scale_cols(int n, const double* s, const double* c, const double* lb, const double* ub,
                             const double* wx, int has_warm, T* c_hat, T* l_hat, T* u_hat, double* xk,
                             double* x0, T* xp) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double sj = s[j];
    c_hat[j] = static_cast<T>(c[j] * sj);
    const double lo = lb[j] / sj, hi = ub[j] / sj;
    l_hat[j] = static_cast<T>(lo);
    u_hat[j] = static_cast<T>(hi);
    const double xs = clampT(has_warm ? wx[j] / sj : 0.0, lo, hi);  // the FP64 state keeps a warm start exact
    xk[j] = xs;
    x0[j] = xs;
    xp[j] = static_cast<T>(xs);
}

template <class T>
__global__ void k_scale_rows(int m, const double* r, const double* b, const unsigned char* eq,
                             const double* wy, int has_warm, T* b_hat, double* yk, double* y0, T* yg, T* yp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double ri = r[i];
    b_hat[i] = static_cast<T>(b[i] * ri);
    double ys = has_warm ? wy[i] / ri : 0.0;
    if (!eq[i] && ys < 0.0) ys = 0.0;
    yk[i] = ys;
    y0[i] = ys;
    yg[i] = static_cast<T>(ys);
    yp[i] = static_cast<T>(ys);
}

__global__ void k_copy_f64(double* dst, const double* src, size_t n) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Power iteration helpers.
template <class T>
__global__ void k_pw_init(T* v, int n) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    unsigned h = static_cast<unsigned>(j) * 2654435761u + 12345u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    h ^= h >> 15;
    v[j] = static_cast<T>(0.5 + static_cast<double>(h & 0xffffu) / 65536.0);
}

__global__ void k_sum_partials_to(const double* partials, int nblocks, int Q, int q, double* out) {
    __shared__ double s[kBlock / kWave];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += partials[static_cast<size_t>(b) * Q + q];
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) *out = t;
}

template <class T>
__global__ void k_pw_normalize(T* v, int n, const double* norm2) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double nv = sqrt(*norm2);
    if (nv > 0.0) v[j] = static_cast<T>(static_cast<double>(v[j]) / nv);
}

// eta, omega, tau, sigma from the power-iteration estimate and ||c_hat||, ||b_hat||.
__global__ void k_init_state(PdlpState* st, const double* pw_u2, const double* c2, const double* b2) {
    if (threadIdx.x != 0) return;
    init_state_core(st, *pw_u2, *c2, *b2);
}

template <class T>
__global__ void k_sq_partial(const T* v, int n, double* partials) {
    __shared__ double s[kBlock / kWave];
    double acc = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double t = static_cast<double>(v[j]);
        acc += t * t;
    }
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void k_advance(PdlpState* st, int k) {
    if (threadIdx.x == 0 && !st->done) st->iter += k;
}

// Unscaled x for the check: xo = clamp(s x, l, u) for the iterate and for x+;
// partials [0,1] = c'xo.
template <class T>
__global__ void k_check_x(int n, const PdlpState* st, const double* xk, const T* xp, const double* s,
                          const double* c, const double* lb, const double* ub, double* xo_cur,
                          double* xo_cand, double* partials) {
    if (st->done) return;
    __shared__ double sh[kBlock / kWave];
    double p0 = 0.0, p1 = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double sj = s[j], l = lb[j], u = ub[j];
        double xc = sj * xk[j];
        xc = xc < l ? l : (xc > u ? u : xc);
        double xa = sj * static_cast<double>(xp[j]);
        xa = xa < l ? l : (xa > u ? u : xa);
        xo_cur[j] = xc;
        xo_cand[j] = xa;
        p0 += c[j] * xc;
        p1 += c[j] * xa;
    }
    double t0 = block_sum(p0, sh);
    double t1 = block_sum(p1, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2 + 0] = t0;
        partials[blockIdx.x * 2 + 1] = t1;
    }
}

// k_check_x with the two points interleaved (EpiCheckRows2 gathers them).
template <class T>
__global__ void k_check_x2(int n, const PdlpState* st, const double* xk, const T* xp, const double* s,
                           const double* c, const double* lb, const double* ub, double2* xo2, double* partials) {
    if (st->done) return;
    __shared__ double sh[kBlock / kWave];
    double p0 = 0.0, p1 = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double sj = s[j], l = lb[j], u = ub[j];
        double xc = sj * xk[j];
        xc = xc < l ? l : (xc > u ? u : xc);
        double xa = sj * static_cast<double>(xp[j]);
        xa = xa < l ? l : (xa > u ? u : xa);
        xo2[j] = make_double2(xc, xa);
        p0 += c[j] * xc;
        p1 += c[j] * xa;
    }
    double t0 = block_sum(p0, sh);
    double t1 = block_sum(p1, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2 + 0] = t0;
        partials[blockIdx.x * 2 + 1] = t1;
    }
}

// out[i] = component `which` (0: current point, 1: candidate) of v2[i].
__global__ void k_pick2(size_t n, const double2* v2, int which, double* out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = which ? v2[i].y : v2[i].x;
}

// N block-wide sums at once, each the same as block_sum (xor tree in the
// wavefront, then the wavefronts in order), with one pair of barriers for all.
template <int N>
__device__ inline void block_sums(double (&v)[N], double* shn) {
    const int lane = static_cast<int>(threadIdx.x) % kWave, wid = static_cast<int>(threadIdx.x) / kWave;
    const int nw = static_cast<int>(blockDim.x) / kWave;
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) shn[wid * N + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += shn[w * N + k];
        v[k] = t;
    }
}

struct CheckDims {
    int nb_x, nb_rows, nb_cols;
};

// The check's 14 sums in two levels, each in a fixed order.  k_fold_check:
// workgroup g adds its own contiguous range of each partial array (the same
// ranges every time).  k_decide: one workgroup adds the kFoldGroups results,
// then runs the stopping test, the KKT restart rule and the bookkeeping.
// (One workgroup reading every check workgroup's partials -- 45,000
// workgroups, 5 MB at the 200,000 x 500,000 Zipf LP -- took about 40 us.)
constexpr int kFoldGroups = 256;
constexpr int kDecideThreads = 256;

__global__ void __launch_bounds__(kBlock)
k_fold_check(const PdlpState* st, CheckDims d, const double* px, const double* pr, const double* pc, double* fold) {
    if (st->done) return;
    __shared__ double sh[(kBlock / kWave) * 14];
    double a[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) a[k] = 0.0;
    const int g = static_cast<int>(blockIdx.x), ng = static_cast<int>(gridDim.x);
    int b0 = 0, b1 = 0;
    auto range = [&](int nb) {
        const int per = (nb + ng - 1) / ng;
        b0 = min(nb, g * per);
        b1 = min(nb, b0 + per);
    };
    range(d.nb_x);
    for (int b = b0 + static_cast<int>(threadIdx.x); b < b1; b += blockDim.x) {
        a[0] += px[b * 2 + 0];
        a[1] += px[b * 2 + 1];
    }
    range(d.nb_rows);
    for (int b = b0 + static_cast<int>(threadIdx.x); b < b1; b += blockDim.x) {
        const double* p = pr + static_cast<size_t>(b) * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) a[2 + k] += p[k];
    }
    range(d.nb_cols);
    for (int b = b0 + static_cast<int>(threadIdx.x); b < b1; b += blockDim.x) {
        const double* p = pc + static_cast<size_t>(b) * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) a[8 + k] += p[k];
    }
    block_sums(a, sh);  // the 14 sums of block_sum, with one pair of barriers
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 14; ++k) fold[g * 14 + k] = a[k];
    }
}

__global__ void __launch_bounds__(kDecideThreads)
k_decide(PdlpState* st, const double* fold, int groups) {
    if (st->done) return;
    __shared__ double sh[(kDecideThreads / kWave) * 14];
    double a[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) a[k] = 0.0;
    for (int g = threadIdx.x; g < groups; g += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 14; ++k) a[k] += fold[g * 14 + k];
    }
    block_sums(a, sh);
    if (threadIdx.x != 0) return;
    decide_core(st, a);
}

// Restart: move to the chosen point, make it the new anchor, measure the
// move of the anchor.  partials [0] = ||dx||^2 (n side) or ||dy||^2 (m side).
template <class T>
__global__ void k_restart_x(int n, const PdlpState* st, double* xk, double* x0, const T* xp, double* partials) {
    if (st->done || !st->restart) return;
    __shared__ double sh[kBlock / kWave];
    double acc = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double xn = st->restart_to_cand ? static_cast<double>(xp[j]) : xk[j];
        const double d = xn - x0[j];
        acc += d * d;
        xk[j] = xn;
        x0[j] = xn;
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

template <class T>
__global__ void k_restart_y(int m, const PdlpState* st, double* yk, double* y0, const T* yp, T* yg, double* ax,
                            double* ax0, const double* ax_cand, double* partials) {
    if (st->done || !st->restart) return;
    __shared__ double sh[kBlock / kWave];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        double yn = yk[i];
        if (st->restart_to_cand) {
            yn = static_cast<double>(yp[i]);
            ax[i] = ax_cand[i];
        }
        const double d = yn - y0[i];
        acc += d * d;
        yk[i] = yn;
        y0[i] = yn;
        yg[i] = static_cast<T>(yn);
        ax0[i] = ax[i];
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void k_restart_finish(PdlpState* st, const double* pdx, int nbx, const double* pdy, int nby) {
    if (st->done || !st->restart) return;
    __shared__ double sh[kBlock / kWave];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nbx; k += blockDim.x) a += pdx[k];
    for (int k = threadIdx.x; k < nby; k += blockDim.x) b += pdy[k];
    const double dx2 = block_sum(a, sh);
    const double dy2 = block_sum(b, sh);
    if (threadIdx.x != 0) return;
    restart_finish_core(st, dx2, dy2);
}


// ------------------------------------------------------- one-workgroup solve
constexpr long long kSmallTuneNnz = 1LL << 18;  // LPs up to this size try the one-workgroup solve
constexpr int kSmallMaxLds = 96 * 1024;         // LDS for the gathered vectors of a one-workgroup solve

// A small LP's whole solve in one workgroup of 1024 threads: the step-0
// check, then intervals of K steps, each followed by the FP64 check, the
// decision and the restart, exactly as the multi-kernel loop runs them, with
// workgroup barriers in place of kernel launches.  The multi-kernel loop pays
// a few microseconds of launch latency per pass, which dominates LPs of a
// few thousand nonzeros.  Threads stride over rows / columns and add each
// segment in order in FP64, four entries in flight.
constexpr int kSmallThreads = 1024;

template <class T>
struct SmallArgs {
    int m, n, K, nnz;
    const int *rp, *ci, *cp, *ri;
    const T *rv, *cv;
    const double *rv64, *cv64;
    EpiPrimal<T> ep;
    EpiDual<T> ed;
    EpiCheckRows<T> er;
    EpiCheckCols ec;
    const double *cs, *c64, *lb64, *ub64;
    double *xo_cur, *xo_cand;
    double *x0, *y0, *ax0;  // anchors (written at restarts)
    PdlpState* st;
    int lds;    // 1: copies of x+ and y (the gathered vectors) live in LDS, n + m entries of T
    int power;  // k_small_lds: 1 = run the power iteration, the state's start and A x0 in the kernel
    const double* scal;  // [2] ||c_hat||^2, [3] ||b_hat||^2
};

template <class TM, class TV>
__device__ inline double seg_dot(int b, int e, const int* __restrict__ idx, const TM* __restrict__ val,
                                 const TV* __restrict__ v) {
    double a = 0.0;
    int q = b;
    for (; q + 4 <= e; q += 4) {
        const double w0 = static_cast<double>(val[q]), w1 = static_cast<double>(val[q + 1]);
        const double w2 = static_cast<double>(val[q + 2]), w3 = static_cast<double>(val[q + 3]);
        const double x0 = static_cast<double>(v[idx[q]]), x1 = static_cast<double>(v[idx[q + 1]]);
        const double x2 = static_cast<double>(v[idx[q + 2]]), x3 = static_cast<double>(v[idx[q + 3]]);
        a += w0 * x0;
        a += w1 * x1;
        a += w2 * x2;
        a += w3 * x3;
    }
    for (; q < e; ++q) a += static_cast<double>(val[q]) * static_cast<double>(v[idx[q]]);
    return a;
}

template <class T>
__device__ inline void small_check(const SmallArgs<T>& a, double* sh, T* ly) {
    __shared__ double sh14[(kSmallThreads / kWave) * 14];
    const int tid = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
    PdlpState* st = a.st;
    double p[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) p[k] = 0.0;
    for (int j = tid; j < a.n; j += nt) {  // unscaled x: current point, candidate
        const double sj = a.cs[j], l = a.lb64[j], u = a.ub64[j];
        const double xc = clampT(sj * a.ep.xk[j], l, u);
        const double xa = clampT(sj * static_cast<double>(a.ep.xp[j]), l, u);
        a.xo_cur[j] = xc;
        a.xo_cand[j] = xa;
        p[0] += a.c64[j] * xc;
        p[1] += a.c64[j] * xa;
    }
    __syncthreads();
    for (int i = tid; i < a.m; i += nt) {  // rows: A xo (FP64), violations, duals
        const int b = a.rp[i], e = a.rp[i + 1];
        const double acc[2] = {seg_dot(b, e, a.ci, a.rv64, a.xo_cur), seg_dot(b, e, a.ci, a.rv64, a.xo_cand)};
        a.er(i, acc, p + 2);
    }
    __syncthreads();
    for (int j = tid; j < a.n; j += nt) {  // columns: A'yo (FP64), dual residuals
        const int b = a.cp[j], e = a.cp[j + 1];
        const double acc[2] = {seg_dot(b, e, a.ri, a.cv64, a.er.yo_cur), seg_dot(b, e, a.ri, a.cv64, a.er.yo_cand)};
        a.ec(j, acc, p + 8);
    }
    block_sums(p, sh14);
    if (tid == 0 && !st->done) decide_core(st, p);
    __syncthreads();
    if (st->done || !st->restart) return;
    const bool cand = st->restart_to_cand != 0;
    double dx = 0.0, dy = 0.0;
    for (int j = tid; j < a.n; j += nt) {
        const double xn = cand ? static_cast<double>(a.ep.xp[j]) : a.ep.xk[j];
        const double d = xn - a.x0[j];
        dx += d * d;
        a.ep.xk[j] = xn;
        a.x0[j] = xn;
    }
    for (int i = tid; i < a.m; i += nt) {
        double yn = a.ed.yk[i];
        if (cand) {
            yn = static_cast<double>(a.ed.yp[i]);
            a.ed.ax[i] = a.er.ax_cand[i];
        }
        const double d = yn - a.y0[i];
        dy += d * d;
        a.ed.yk[i] = yn;
        a.y0[i] = yn;
        a.ed.yg[i] = static_cast<T>(yn);
        if (ly) ly[i] = static_cast<T>(yn);
        a.ax0[i] = a.ed.ax[i];
    }
    const double dx2 = block_sum(dx, sh);
    const double dy2 = block_sum(dy, sh);
    if (tid == 0) restart_finish_core(st, dx2, dy2);
    __syncthreads();
}

template <class T>
__global__ void __launch_bounds__(kSmallThreads) k_small(SmallArgs<T> a) {
    __shared__ double sh[kSmallThreads / kWave];
    extern __shared__ __attribute__((aligned(16))) unsigned char small_lds[];
    const int tid = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
    PdlpState* st = a.st;
    T* lx = a.lds ? reinterpret_cast<T*>(small_lds) : nullptr;  // x+ copy
    T* ly = a.lds ? lx + a.n : nullptr;                          // y copy
    if (a.lds) {
        for (int j = tid; j < a.n; j += nt) lx[j] = a.ep.xp[j];
        for (int i = tid; i < a.m; i += nt) ly[i] = a.ed.yg[i];
    }
    small_check(a, sh, ly);  // step 0 (its barriers order the copies above)
    const T* gy = a.lds ? ly : a.ep.yg;
    const T* gx = a.lds ? lx : a.ed.xp;
    while (!st->done) {
        const long long left = st->max_iter - st->iter;
        const int k = static_cast<int>(left < a.K ? left : a.K);
        if (k <= 0) break;
        for (int t = 0; t < k; ++t) {
            EpiPrimal<T> ep = a.ep;
            ep.t = t;
            for (int j = tid; j < a.n; j += nt) {
                const double acc = seg_dot(a.cp[j], a.cp[j + 1], a.ri, a.cv, gy);
                ep(j, &acc, nullptr);
                if (lx) lx[j] = ep.xp[j];
            }
            __syncthreads();
            EpiDual<T> ed = a.ed;
            ed.t = t;
            for (int i = tid; i < a.m; i += nt) {
                const double acc = seg_dot(a.rp[i], a.rp[i + 1], a.ci, a.rv, gx);
                ed(i, &acc, nullptr);
                if (ly) ly[i] = ed.yg[i];
            }
            __syncthreads();
        }
        if (tid == 0) st->iter += k;
        __syncthreads();
        small_check(a, sh, ly);
    }
}

// The one-workgroup solve with the whole scaled matrix resident in LDS (both
// orientations, T values and 32-bit indices) beside the gathered vectors,
// and each thread's R columns and R rows (j = tid + r * 1024) of Halpern
// state in registers.  The steps are those of EpiPrimal / EpiDual with the
// same arithmetic; the state goes back to memory for each check
// (small_check, unchanged) and is reloaded after it, since a restart moves it.
constexpr int kSmallLdsMaxPerThread = 2;

template <class T>
size_t small_resident_bytes(long long m, long long n, long long nnz) {
    return sizeof(T) * static_cast<size_t>(2 * nnz + n + m) + sizeof(int) * static_cast<size_t>(m + n + 2 + 2 * nnz);
}

// sum over k = b, b + S, b + 2S, ... < e of val[k] * v[idx[k]], four in flight
template <int S, class TM, class TV>
__device__ inline double seg_dot_strided(int b, int e, const int* __restrict__ idx, const TM* __restrict__ val,
                                         const TV* __restrict__ v) {
    double a = 0.0;
    int q = b;
    for (; q + 3 * S < e; q += 4 * S) {
        const double w0 = static_cast<double>(val[q]), w1 = static_cast<double>(val[q + S]);
        const double w2 = static_cast<double>(val[q + 2 * S]), w3 = static_cast<double>(val[q + 3 * S]);
        const double x0 = static_cast<double>(v[idx[q]]), x1 = static_cast<double>(v[idx[q + S]]);
        const double x2 = static_cast<double>(v[idx[q + 2 * S]]), x3 = static_cast<double>(v[idx[q + 3 * S]]);
        a += w0 * x0;
        a += w1 * x1;
        a += w2 * x2;
        a += w3 * x3;
    }
    for (; q < e; q += S) a += static_cast<double>(val[q]) * static_cast<double>(v[idx[q]]);
    return a;
}

// R: columns (and, with LR = 1, rows) per thread.  LR > 1 (with R = 1): LR
// consecutive lanes share a row, each summing every LR-th entry, the
// partial sums added by a fixed xor tree -- a row's dependent chain of LDS
// loads is LR times shorter, and LR times more wavefronts work on the rows.
template <class T, int R, int LR>
__global__ void __launch_bounds__(kSmallThreads) k_small_lds(SmallArgs<T> a) {
    static_assert(LR == 1 || R == 1, "lanes per row need one row per thread group");
    __shared__ double sh[kSmallThreads / kWave];
    extern __shared__ __attribute__((aligned(16))) unsigned char small_lds[];
    const int m = a.m, n = a.n, nnz = a.nnz;
    const int tid = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
    PdlpState* st = a.st;
    T* lrv = reinterpret_cast<T*>(small_lds);       // CSR values [nnz]
    T* lcv = lrv + nnz;                             // CSC values [nnz]
    T* lx = lcv + nnz;                              // x+ [n]
    T* ly = lx + n;                                 // y copy [m]
    int* lrp = reinterpret_cast<int*>(ly + m);      // [m + 1]
    int* lcp = lrp + m + 1;                         // [n + 1]
    int* lci = lcp + n + 1;                         // [nnz]
    int* lri = lci + nnz;                           // [nnz]
    for (int k = tid; k < nnz; k += nt) {
        lrv[k] = a.rv[k];
        lcv[k] = a.cv[k];
        lci[k] = a.ci[k];
        lri[k] = a.ri[k];
    }
    for (int i = tid; i <= m; i += nt) lrp[i] = a.rp[i];
    for (int j = tid; j <= n; j += nt) lcp[j] = a.cp[j];
    if (a.power) {
        // ||A_hat||_2 by 40 rounds of power iteration on the LDS matrix (the
        // launches of the multi-kernel set-up, same start vector and
        // roundings: v, u held in T in the x+ / y slots), then the state's
        // step sizes and A x0 with its anchor copy
        T* v = lx;
        T* u = ly;
        __syncthreads();
        double acc2 = 0.0;
        for (int j = tid; j < n; j += nt) {
            unsigned h = static_cast<unsigned>(j) * 2654435761u + 12345u;
            h ^= h >> 13;
            h *= 0x5bd1e995u;
            h ^= h >> 15;
            v[j] = static_cast<T>(0.5 + static_cast<double>(h & 0xffffu) / 65536.0);
            acc2 += static_cast<double>(v[j]) * static_cast<double>(v[j]);
        }
        double nv = sqrt(block_sum(acc2, sh));
        for (int j = tid; j < n; j += nt)
            if (nv > 0.0) v[j] = static_cast<T>(static_cast<double>(v[j]) / nv);
        double u2 = 0.0;
        for (int it = 0; it < 40; ++it) {
            __syncthreads();
            double pu = 0.0;
            for (int i = tid; i < m; i += nt) {
                const double acc = seg_dot(lrp[i], lrp[i + 1], lci, lrv, v);
                u[i] = static_cast<T>(acc);
                pu += acc * acc;
            }
            u2 = block_sum(pu, sh);
            double pv = 0.0;
            for (int j = tid; j < n; j += nt) {
                const double acc = seg_dot(lcp[j], lcp[j + 1], lri, lcv, u);
                v[j] = static_cast<T>(acc);
                pv += acc * acc;
            }
            nv = sqrt(block_sum(pv, sh));
            for (int j = tid; j < n; j += nt)
                if (nv > 0.0) v[j] = static_cast<T>(static_cast<double>(v[j]) / nv);
        }
        if (tid == 0) init_state_core(st, u2, a.scal[2], a.scal[3]);
        __syncthreads();
        for (int j = tid; j < n; j += nt) lx[j] = a.ep.xp[j];
        __syncthreads();
        for (int i = tid; i < m; i += nt) {
            const double acc = seg_dot(lrp[i], lrp[i + 1], lci, lrv, lx);
            a.ed.ax[i] = acc;
            a.ax0[i] = acc;
        }
    } else {
        for (int j = tid; j < n; j += nt) lx[j] = a.ep.xp[j];
    }
    for (int i = tid; i < m; i += nt) ly[i] = a.ed.yg[i];
    small_check(a, sh, ly);  // step 0 (its barriers order the copies above)
    while (!st->done) {
        const long long left = st->max_iter - st->iter;
        const int k = static_cast<int>(left < a.K ? left : a.K);
        if (k <= 0) break;
        // the data and state as the last check left them (nothing stays in
        // registers across a check)
        T cc[R], lo[R], hi[R], bb[R], xpr[R], ypr[R];
        bool eq[R];
        double xk[R], x0[R], yk[R], y0[R], ax[R], ax0[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = tid + r * nt, i = LR > 1 ? tid / LR : tid + r * nt;
            cc[r] = j < n ? a.ep.c[j] : T(0);
            lo[r] = j < n ? a.ep.lb[j] : T(0);
            hi[r] = j < n ? a.ep.ub[j] : T(0);
            xpr[r] = j < n ? a.ep.xp[j] : T(0);
            bb[r] = i < m ? a.ed.b[i] : T(0);
            eq[r] = i < m ? a.ed.eq[i] != 0 : false;
            ypr[r] = i < m ? a.ed.yp[i] : T(0);
            xk[r] = j < n ? a.ep.xk[j] : 0.0;
            x0[r] = j < n ? a.x0[j] : 0.0;
            yk[r] = i < m ? a.ed.yk[i] : 0.0;
            y0[r] = i < m ? a.y0[i] : 0.0;
            ax[r] = i < m ? a.ed.ax[i] : 0.0;
            ax0[r] = i < m ? a.ax0[i] : 0.0;
        }
        const double tau = st->tau, sigma = st->sigma;
        const long long base = st->iter - st->iter_last_restart;
        for (int t = 0; t < k; ++t) {
            const double kk = static_cast<double>(base + t);
            const double w = (kk + 1.0) / (kk + 2.0);
#pragma unroll
            for (int r = 0; r < R; ++r) {  // column j: x+ = proj(x - tau (c + A'y)), Halpern update of x
                const int j = tid + r * nt;
                if (j < n) {
                    const double acc = seg_dot(lcp[j], lcp[j + 1], lri, lcv, ly);
                    const double xo = xk[r];
                    const T xt = clampT(static_cast<T>(primal_move(xo, tau, static_cast<double>(cc[r]), acc)), lo[r], hi[r]);
                    xpr[r] = xt;
                    lx[j] = xt;
                    xk[r] = halpern_x(w, static_cast<double>(xt), xo, x0[r]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r) {  // row i: y+ = proj(y + sigma (A(2x+ - x) - b)), Halpern update of y, A x
                const int i = LR > 1 ? tid / LR : tid + r * nt;
                if (i < m) {  // (uniform over the LR lanes of a row)
                    double acc;
                    if constexpr (LR == 1) {
                        acc = seg_dot(lrp[i], lrp[i + 1], lci, lrv, lx);
                    } else {
                        acc = seg_dot_strided<LR>(lrp[i] + tid % LR, lrp[i + 1], lci, lrv, lx);
                        acc += __shfl_xor(acc, 1, kWave);
                        if constexpr (LR == 4) acc += __shfl_xor(acc, 2, kWave);
                    }
                    const double axo = ax[r], yo = yk[r];
                    double yn = dual_move(yo, sigma, acc, axo, static_cast<double>(bb[r]));
                    if (!eq[r] && yn < 0.0) yn = 0.0;
                    const T yt = static_cast<T>(yn);
                    ypr[r] = yt;
                    const double ynew = halpern_y(w, static_cast<double>(yt), yo, y0[r]);
                    yk[r] = ynew;
                    if (LR == 1 || tid % LR == 0) ly[i] = static_cast<T>(ynew);
                    ax[r] = halpern_ax(w, acc, axo, ax0[r]);
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {  // back to memory for the check
            const int j = tid + r * nt, i = LR > 1 ? tid / LR : tid + r * nt;
            if (j < n) {
                a.ep.xk[j] = xk[r];
                a.ep.xp[j] = xpr[r];
            }
            if (i < m && (LR == 1 || tid % LR == 0)) {
                a.ed.yk[i] = yk[r];
                a.ed.yp[i] = ypr[r];
                a.ed.yg[i] = static_cast<T>(yk[r]);
                a.ed.ax[i] = ax[r];
            }
        }
        if (tid == 0) st->iter += k;
        __syncthreads();
        small_check(a, sh, ly);
    }
}

// ---------------------------------------------------------------- host side
struct DevMem {
    std::vector<void*> ptrs;
    ~DevMem() {
        for (void* p : ptrs) hipFree(p);
    }
    template <class U>
    hipError_t alloc(U** p, size_t count) {
        *p = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(U) * (count ? count : 1));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// Bin table for segments sorted by length (descending, so lanes per segment
// fall from bin to bin and every lane count forms one contiguous bin).
BinTable make_bins(const std::vector<int>& sorted_len, int per_lane) {
    BinTable bt{};
    static const int kLanes[kMaxBins] = {1, 2, 4, 8, 16, 32, 64, 256};
    auto lanes_for = [&](int len) {
        if (len > 1024) return 256;
        for (int b = 0; b < 7; ++b)
            if (len <= kLanes[b] * per_lane) return kLanes[b];
        return 64;
    };
    const int nseg = static_cast<int>(sorted_len.size());
    int s = 0, blocks = 0;
    bt.nbins = 0;
    while (s < nseg) {
        const int L = lanes_for(sorted_len[s]);
        int e = s;
        while (e < nseg && lanes_for(sorted_len[e]) == L) ++e;
        const int b = bt.nbins++;
        bt.lanes[b] = L;
        bt.seg_start[b] = s;
        bt.block_start[b] = blocks;
        const int per_block = L <= kWave ? kBlock / L : 1;
        blocks += (e - s + per_block - 1) / per_block;
        s = e;
    }
    if (bt.nbins == 0) {  // no segments: one idle block keeps launches valid
        bt.nbins = 1;
        bt.lanes[0] = 1;
        bt.seg_start[0] = 0;
        bt.block_start[0] = 0;
        blocks = 1;
    }
    bt.seg_start[bt.nbins] = nseg;
    bt.block_start[bt.nbins] = blocks;
    bt.total_blocks = blocks;
    return bt;
}

// Stream-kernel blocks over segments of the given lengths: contiguous runs
// with at most `cap` nonzeros and at most 256 segments; a longer segment is a
// block of its own.  Returns the block start offsets plus the end sentinel.
std::vector<int> stream_blocks(const std::vector<int>& len, int cap) {
    std::vector<int> blk{0};
    const int nseg = static_cast<int>(len.size());
    long long cur = 0;
    int cnt = 0;
    for (int s = 0; s < nseg; ++s) {
        const int l = len[s];
        if (l > cap) {
            if (cnt > 0) blk.push_back(s);
            blk.push_back(s + 1);
            cur = 0;
            cnt = 0;
            continue;
        }
        if (cnt > 0 && (cur + l > cap || cnt == kBlock)) {
            blk.push_back(s);
            cur = 0;
            cnt = 0;
        }
        cur += l;
        ++cnt;
    }
    if (cnt > 0 || blk.size() == 1) blk.push_back(nseg);
    return blk;
}

constexpr int kStreamCaps[3] = {1024, 2048, 4096};

// How one fused iteration pass is launched: the length-binned segment kernel
// (bt, optional LDS copy of the hottest gathered entries), the stream kernel
// with LDS budget cap (cap > 0), or the tiled gather + reduction kernels
// (tiled > 0: the layout cut for that many tiles per CU).  bt also serves the
// FP64 check passes.
struct PassPlan {
    BinTable bt{};
    int kp = 4;
    int cap = 0;
    int hot = 0;
    int hot_grid = 0;
    int tiled = 0;
};

constexpr int kMaxTilesPerCu = 64;

struct EngineBase {
    virtual ~EngineBase() = default;
    virtual bo_status set_warm_start(const double* x0, const double* y0) = 0;
    virtual bo_status solve(bo_pdlp_stats* stats) = 0;
    virtual bo_status get_solution(double* x, double* y) = 0;
    virtual bo_status time_spmv(int reps, double* csr_ms, double* csc_ms) = 0;
    virtual bo_status time_kernels(int reps, double* ms4) = 0;
    virtual bo_status pass_info(int pass, bo_pdlp_pass_info* out) = 0;
};

// T: the iterate vectors the passes gather and write (FP32 or FP64); TM:
// the scaled matrix values the passes read (T, or FP32 under FP64 iterates:
// 6 bytes per stored entry instead of 10, with the iterates' precision).
template <class T, class TM = T>
struct Engine final : EngineBase {
    int m = 0, n = 0;
    long long nnz = 0;
    bo_pdlp_params prm{};
    hipStream_t stream = nullptr;
    DevMem mem;
    std::vector<int> row_of;  // permuted position -> original row
    std::vector<int> col_of;  // permuted position -> original column
    std::vector<int> row_len_sorted, col_len_sorted;
    // FP64 original (permuted) matrix in CSR and CSC; scaled working values
    int *rp = nullptr, *ci = nullptr, *cp = nullptr, *ri = nullptr;
    double *rv64 = nullptr, *cv64 = nullptr;
    TM *rv = nullptr, *cv = nullptr;
    static constexpr bool kSameTypes = std::is_same<T, TM>::value;  // the one-workgroup solve needs it
    // FP64 original vectors (permuted) and scaling
    double *b64 = nullptr, *c64 = nullptr, *lb64 = nullptr, *ub64 = nullptr;
    unsigned char* eq = nullptr;
    double *rs = nullptr, *cs = nullptr, *rnorm = nullptr, *cnorm = nullptr;
    // scaled problem data and iteration state
    T *b_hat = nullptr, *c_hat = nullptr, *l_hat = nullptr, *u_hat = nullptr;
    double *xk = nullptr, *x0 = nullptr, *yk = nullptr, *y0 = nullptr, *ax = nullptr, *ax0 = nullptr;
    T *xp = nullptr, *yp = nullptr, *yg = nullptr;
    T *pw_v = nullptr, *pw_u = nullptr;
    // check buffers
    double *xo_cur = nullptr, *xo_cand = nullptr, *yo_cur = nullptr, *yo_cand = nullptr, *ax_cand = nullptr;
    double2 *xo2 = nullptr, *yo2 = nullptr;  // the multi-kernel checks' points, interleaved
    double *wx = nullptr, *wy = nullptr;
    int has_warm = 0;
    double *part_x = nullptr, *part_r = nullptr, *part_c = nullptr, *part_dx = nullptr, *part_dy = nullptr;
    double* part_fold = nullptr;  // first level of the check's sums, kFoldGroups x 14
    double* scal = nullptr;  // [0] ||u||^2 [1] ||w||^2 [2] ||c_hat||^2 [3] ||b_hat||^2
    PdlpState* st = nullptr;
    PdlpState* st_host = nullptr;  // pinned: the state at the end of a solve
    PdlpState* st_ring[2] = {};    // pinned: the state after each of the two queued intervals
    hipEvent_t st_ready[2] = {};   // recorded after each ring copy
    PdlpState* st_init = nullptr;  // pinned, staged for the upload in prepare()
    PassPlan plan_rows{}, plan_cols{};  // CSR (dual) pass and CSC (primal) pass
    bool one_wg = false;                // the whole solve in one workgroup (k_small)
    bool small_lds_set = false;         // k_small's dynamic LDS limit raised
    bool power_pending = false;         // prepare() left the power iteration and the start to k_small_lds
    int4* sblk[2][3] = {};              // stream block descriptors per pass (rows, cols) and LDS budget
    int nsblk[2][3] = {};
    // tiled layout per pass (0: CSR pass gathering x+, 1: CSC pass gathering y)
    struct TilePass {
        int k = 0;  // tiles per CU the layout was cut for; 0 = not built
        int B = 0, W = 0, ntiles = 0, nslices = 0;
        unsigned lds_bytes = 0;  // dynamic LDS of the gather kernel
        long long pairs = 0;
        int max_pairs = 0;  // most partials a segment adds
        int reduce_u = 0;   // loads in flight per round of the reduction (0: cover max_pairs)
        size_t padded = 0;
        TileDesc* td = nullptr;
        unsigned* words = nullptr;  // slice words
        unsigned short* idx = nullptr;
        int* src = nullptr;
        TM* val = nullptr;
        int* seg_ptr = nullptr;     // segment -> partial slots
        unsigned* slot = nullptr;
        double* part = nullptr;     // one row of 64 partials per slice
    } tp[2];
    int lds_max = 0;
    int nb_x = 1, nb_y = 1;
    double b_norm = 0.0, c_norm = 0.0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_k = 0;
    int num_cus = 1;

    ~Engine() override {
        if (graph_exec) hipGraphExecDestroy(graph_exec);
        if (graph) hipGraphDestroy(graph);
        if (st_host) hipHostFree(st_host);
        for (int q = 0; q < 2; ++q) {
            if (st_ring[q]) hipHostFree(st_ring[q]);
            if (st_ready[q]) hipEventDestroy(st_ready[q]);
        }
        if (st_init) hipHostFree(st_init);
        if (stream) hipStreamDestroy(stream);
    }

    // Indices by descending length, equal lengths in index order (a stable
    // counting sort).
    static std::vector<int> order_by_length(const std::vector<int>& len) {
        int top = 0;
        for (int l : len) top = std::max(top, l);
        std::vector<int> start(static_cast<size_t>(top) + 2, 0);
        for (int l : len) start[static_cast<size_t>(top - l) + 1]++;
        for (int v = 0; v <= top; ++v) start[v + 1] += start[v];
        std::vector<int> order(len.size());
        for (size_t i = 0; i < len.size(); ++i) order[start[static_cast<size_t>(top - len[i])]++] = static_cast<int>(i);
        return order;
    }

    // verbose >= 1: the wall time of each set-up phase on stderr
    void phase(const char* what, std::chrono::steady_clock::time_point& t) const {
        const auto now = std::chrono::steady_clock::now();
        if (prm.verbose >= 1)
            std::fprintf(stderr, "bo_pdlp_create: %-34s %8.1f ms\n", what,
                         std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }

    bo_status create(const bo_lp* lp, const bo_pdlp_params* p) {
        prm = *p;
        m = lp->m;
        n = lp->n;
        nnz = lp->nnz;
        auto tph = std::chrono::steady_clock::now();
        BO_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_host), sizeof(PdlpState)));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_init), sizeof(PdlpState)));
        for (int q = 0; q < 2; ++q) {
            BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_ring[q]), sizeof(PdlpState)));
            BO_TRY(hipEventCreateWithFlags(&st_ready[q], hipEventDisableTiming));
        }
        std::memset(st_host, 0, sizeof(PdlpState));
        int dev = 0;
        BO_TRY(hipGetDevice(&dev));
        BO_TRY(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
        BO_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        // tiled layouts: for an explicit request, or for the tuner (2 tiles per CU)
        auto tiles_for = [](int code) {
            return code == 0 ? 2 : (code < 0 && code >= -kMaxTilesPerCu ? -code : 0);
        };
        const int tiles_rows = tiles_for(prm.lanes_rows), tiles_cols = tiles_for(prm.lanes_cols);
        const int nth = host_thread_count(1 + nnz / (1 << 20));
        // ---- permutations: rows and columns by descending length (stable)
        std::vector<int> rlen(m), clen(n, 0);
        for (int i = 0; i < m; ++i) rlen[i] = lp->row_ptr[i + 1] - lp->row_ptr[i];
        {  // column lengths: per-thread counts over row ranges, then added
            HostVec<int> cc(static_cast<size_t>(nth - 1) * n);  // each thread zeroes its own row
            host_parallel(nth, [&](int t, int nt) {
                int* c = t == 0 ? clen.data() : cc.data() + static_cast<size_t>(t - 1) * n;
                if (t > 0) std::fill(c, c + n, 0);
                const auto r = host_part(m, t, nt);
                for (int k = lp->row_ptr[r.first]; k < lp->row_ptr[r.second]; ++k) c[lp->col_idx[k]]++;
            });
            host_parallel(nth, [&](int t, int nt) {
                const auto r = host_part(n, t, nt);
                for (int q = 0; q + 1 < nth; ++q)
                    for (long long j = r.first; j < r.second; ++j) clen[j] += cc[static_cast<size_t>(q) * n + j];
            });
        }
        row_of = order_by_length(rlen);
        col_of = order_by_length(clen);
        std::vector<int> col_pos(n);
        for (int p2 = 0; p2 < n; ++p2) col_pos[col_of[p2]] = p2;
        row_len_sorted.resize(m);
        col_len_sorted.resize(n);
        for (int p2 = 0; p2 < m; ++p2) row_len_sorted[p2] = rlen[row_of[p2]];
        for (int p2 = 0; p2 < n; ++p2) col_len_sorted[p2] = clen[col_of[p2]];
        // permuted CSR, rows filled on nth host threads
        std::vector<int> h_rp(m + 1);
        HostVec<int> h_ci(nnz);  // these and the CSC's: every entry written on the threads
        HostVec<double> h_rv(nnz);
        h_rp[0] = 0;
        for (int p2 = 0; p2 < m; ++p2) h_rp[p2 + 1] = h_rp[p2] + row_len_sorted[p2];
        host_parallel(nth, [&](int t, int nt) {
            const auto r = host_part(m, t, nt);
            for (long long p2 = r.first; p2 < r.second; ++p2) {
                const int i = row_of[p2];
                int o = h_rp[p2];
                for (int k = lp->row_ptr[i]; k < lp->row_ptr[i + 1]; ++k, ++o) {
                    h_ci[o] = col_pos[lp->col_idx[k]];
                    h_rv[o] = lp->val[k];
                }
            }
        });
        std::vector<int> h_cp(n + 1, 0);  // the permuted CSC, below
        HostVec<int> h_ri(nnz);
        HostVec<double> h_cv(nnz);
        // The host layouts of the tiled passes are cut on threads of their
        // own while the rest of the set-up goes on; `joiner` waits for them on
        // every way out of this function (they read the host arrays above).
        HostTiles ht[2];
        bool ht_built[2] = {false, false};
        std::exception_ptr ht_err[2];  // rethrown here once the thread is joined
        std::thread tile_thread[2];
        struct Joiner {
            std::thread* t;
            ~Joiner() {
                for (int q = 0; q < 2; ++q)
                    if (t[q].joinable()) t[q].join();
            }
        } joiner{tile_thread};
        auto start_tiles = [&](int pass, int k, const int* ptr, const int* gidx) {
            if (k <= 0) return;
            tile_thread[pass] = std::thread([this, pass, k, ptr, gidx, &ht, &ht_built, &ht_err] {
                try {
                    ht_built[pass] = host_tiles(pass, k, ptr, gidx, ht[pass]);
                } catch (...) {
                    ht_err[pass] = std::current_exception();
                }
            });
        };
        start_tiles(0, tiles_rows, h_rp.data(), h_ci.data());
        // permuted CSC (transpose of the permuted CSR): thread t counts its
        // rows' entries per column, then writes them after those of threads
        // 0 .. t-1, so every column keeps its rows in ascending order
        {
            HostVec<int> cnt(static_cast<size_t>(nth) * n);  // each thread zeroes its own row
            host_parallel(nth, [&](int t, int nt) {
                int* c = cnt.data() + static_cast<size_t>(t) * n;
                std::fill(c, c + n, 0);
                const auto r = host_part(m, t, nt);
                for (int k = h_rp[r.first]; k < h_rp[r.second]; ++k) c[h_ci[k]]++;
            });
            host_parallel(nth, [&](int t, int nt) {  // column ranges: totals, then each thread's start
                const auto r = host_part(n, t, nt);
                for (long long j = r.first; j < r.second; ++j) {
                    int tot = 0;
                    for (int q = 0; q < nth; ++q) tot += cnt[static_cast<size_t>(q) * n + j];
                    h_cp[j + 1] = tot;
                }
            });
            for (int j = 0; j < n; ++j) h_cp[j + 1] += h_cp[j];
            host_parallel(nth, [&](int t, int nt) {
                const auto r = host_part(n, t, nt);
                for (long long j = r.first; j < r.second; ++j) {
                    int o = h_cp[j];
                    for (int q = 0; q < nth; ++q) {
                        int& c = cnt[static_cast<size_t>(q) * n + j];
                        const int v = c;
                        c = o;
                        o += v;
                    }
                }
            });
            host_parallel(nth, [&](int t, int nt) {
                int* pos = cnt.data() + static_cast<size_t>(t) * n;
                const auto r = host_part(m, t, nt);
                for (long long p2 = r.first; p2 < r.second; ++p2)
                    for (int k = h_rp[p2]; k < h_rp[p2 + 1]; ++k) {
                        const int q = pos[h_ci[k]]++;
                        h_ri[q] = static_cast<int>(p2);
                        h_cv[q] = h_rv[k];
                    }
            });
        }
        start_tiles(1, tiles_cols, h_cp.data(), h_ri.data());
        std::vector<double> hb(m), hc(n), hl(n), hu(n);
        std::vector<unsigned char> he(m);
        double b2 = 0.0, c2 = 0.0;
        for (int p2 = 0; p2 < m; ++p2) {
            hb[p2] = lp->b[row_of[p2]];
            he[p2] = lp->is_eq[row_of[p2]];
            b2 += hb[p2] * hb[p2];
        }
        for (int p2 = 0; p2 < n; ++p2) {
            const int j = col_of[p2];
            hc[p2] = lp->c[j];
            hl[p2] = lp->lb[j];
            hu[p2] = lp->ub[j];
            c2 += hc[p2] * hc[p2];
        }
        b_norm = std::sqrt(b2);
        c_norm = std::sqrt(c2);
        phase("renumbering, CSR and CSC (host)", tph);
        // ---- device memory
        BO_TRY(mem.alloc(&rp, m + 1));
        BO_TRY(mem.alloc(&ci, nnz));
        BO_TRY(mem.alloc(&rv64, nnz));
        BO_TRY(mem.alloc(&rv, nnz));
        BO_TRY(mem.alloc(&cp, n + 1));
        BO_TRY(mem.alloc(&ri, nnz));
        BO_TRY(mem.alloc(&cv64, nnz));
        BO_TRY(mem.alloc(&cv, nnz));
        BO_TRY(mem.alloc(&b64, m));
        BO_TRY(mem.alloc(&eq, m));
        BO_TRY(mem.alloc(&rs, m));
        BO_TRY(mem.alloc(&rnorm, m));
        BO_TRY(mem.alloc(&b_hat, m));
        BO_TRY(mem.alloc(&yk, m));
        BO_TRY(mem.alloc(&y0, m));
        BO_TRY(mem.alloc(&yp, m));
        BO_TRY(mem.alloc(&yg, m));
        BO_TRY(mem.alloc(&ax, m));
        BO_TRY(mem.alloc(&ax0, m));
        BO_TRY(mem.alloc(&ax_cand, m));
        BO_TRY(mem.alloc(&pw_u, m));
        BO_TRY(mem.alloc(&yo_cur, m));
        BO_TRY(mem.alloc(&yo_cand, m));
        BO_TRY(mem.alloc(&yo2, m));
        BO_TRY(mem.alloc(&wy, m));
        BO_TRY(mem.alloc(&c64, n));
        BO_TRY(mem.alloc(&lb64, n));
        BO_TRY(mem.alloc(&ub64, n));
        BO_TRY(mem.alloc(&cs, n));
        BO_TRY(mem.alloc(&cnorm, n));
        BO_TRY(mem.alloc(&c_hat, n));
        BO_TRY(mem.alloc(&l_hat, n));
        BO_TRY(mem.alloc(&u_hat, n));
        BO_TRY(mem.alloc(&xk, n));
        BO_TRY(mem.alloc(&x0, n));
        BO_TRY(mem.alloc(&xp, n));
        BO_TRY(mem.alloc(&pw_v, n));
        BO_TRY(mem.alloc(&xo_cur, n));
        BO_TRY(mem.alloc(&xo_cand, n));
        BO_TRY(mem.alloc(&xo2, n));
        BO_TRY(mem.alloc(&wx, n));
        BO_TRY(mem.alloc(&scal, 8));
        BO_TRY(mem.alloc(&st, 1));
        // a zero state (done = 0) until prepare() uploads the solve's: the
        // tuner's passes read it
        BO_TRY(hipMemsetAsync(st, 0, sizeof(PdlpState), stream));
        auto up = [&](void* d, const void* h, size_t bytes) -> hipError_t {
            return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream) : hipSuccess;
        };
        BO_TRY(up(rp, h_rp.data(), sizeof(int) * (m + 1)));
        BO_TRY(up(ci, h_ci.data(), sizeof(int) * nnz));
        BO_TRY(up(rv64, h_rv.data(), sizeof(double) * nnz));
        BO_TRY(up(cp, h_cp.data(), sizeof(int) * (n + 1)));
        BO_TRY(up(ri, h_ri.data(), sizeof(int) * nnz));
        BO_TRY(up(cv64, h_cv.data(), sizeof(double) * nnz));
        BO_TRY(up(b64, hb.data(), sizeof(double) * m));
        BO_TRY(up(eq, he.data(), m));
        BO_TRY(up(c64, hc.data(), sizeof(double) * n));
        BO_TRY(up(lb64, hl.data(), sizeof(double) * n));
        BO_TRY(up(ub64, hu.data(), sizeof(double) * n));
        BO_TRY(hipStreamSynchronize(stream));
        phase("allocation and upload", tph);
        // ---- launch geometry
        nb_x = static_cast<int>(std::min<unsigned>(grid_for(n), 1024));
        nb_y = static_cast<int>(std::min<unsigned>(grid_for(m), 1024));
        BO_TRY(mem.alloc(&part_x, static_cast<size_t>(nb_x) * 2));
        BO_TRY(mem.alloc(&part_dx, nb_x));
        BO_TRY(mem.alloc(&part_dy, nb_y));
        // stream-kernel blocks for every LDS budget the plans may use
        for (int pass = 0; pass < 2; ++pass) {
            const std::vector<int>& len = pass == 0 ? row_len_sorted : col_len_sorted;
            std::vector<int> first(len.size() + 1, 0);  // first nonzero of every segment
            for (size_t q = 0; q < len.size(); ++q) first[q + 1] = first[q] + len[q];
            for (int c = 0; c < 3; ++c) {
                const std::vector<int> blk = stream_blocks(len, kStreamCaps[c]);
                std::vector<int4> desc(blk.size() - 1);
                for (size_t b = 0; b + 1 < blk.size(); ++b)
                    desc[b] = make_int4(blk[b], blk[b + 1], first[blk[b]], first[blk[b + 1]]);
                nsblk[pass][c] = static_cast<int>(desc.size());
                BO_TRY(mem.alloc(&sblk[pass][c], desc.size()));
                BO_TRY(up(sblk[pass][c], desc.data(), sizeof(int4) * desc.size()));
                BO_TRY(hipStreamSynchronize(stream));  // the host staging goes out of scope
            }
        }
        BO_TRY(hipStreamSynchronize(stream));
        phase("stream-kernel blocks", tph);
        for (int pass = 0; pass < 2; ++pass) {
            if (tile_thread[pass].joinable()) tile_thread[pass].join();
            if (ht_err[pass]) std::rethrow_exception(ht_err[pass]);  // to the entry point's guard
            if (ht_built[pass]) BO_TRY(upload_tile_pass(pass, pass == 0 ? tiles_rows : tiles_cols, ht[pass]));
            ht[pass] = HostTiles{};
            phase(pass == 0 ? "tiled layout, CSR pass (wait, upload)" : "tiled layout, CSC pass (wait, upload)", tph);
        }
        // launch plans: lanes > 0 fixes the segment kernel's nonzeros-per-lane
        // target, lanes < 0 selects the stream kernel with LDS budget -lanes,
        // 0 lets the tuner choose
        const bool tune = prm.lanes_rows == 0 || prm.lanes_cols == 0 || prm.hot_columns < 0 || prm.check_every <= 0;
        const bool tune_cadence = prm.check_every <= 0;
        plan_rows = fixed_plan(0, prm.lanes_rows);
        plan_cols = fixed_plan(1, prm.lanes_cols);
        if (prm.hot_columns > 0 && plan_rows.cap == 0 && plan_rows.tiled == 0) {
            plan_rows.hot = std::min(prm.hot_columns, n);
            plan_rows.hot_grid = grid_for_hot(plan_rows.hot);
            if (plan_rows.hot_grid <= 0) plan_rows.hot = 0;  // does not fit in LDS: plain pass
        }
        // partial buffers (check passes use the plans' bin tables)
        size_t max_r = 0, max_c = 0;
        for (int kp : {1, 2, 4, 8, 16, plan_rows.kp, plan_cols.kp}) {
            max_r = std::max<size_t>(max_r, make_bins(row_len_sorted, kp).total_blocks);
            max_c = std::max<size_t>(max_c, make_bins(col_len_sorted, kp).total_blocks);
        }
        BO_TRY(mem.alloc(&part_r, max_r * 6));
        BO_TRY(mem.alloc(&part_c, max_c * 6));
        BO_TRY(mem.alloc(&part_fold, static_cast<size_t>(kFoldGroups) * 14));
        phase("bin tables", tph);
        if (tune) {
            bo_status s2 = autotune();
            if (s2 != BO_SUCCESS) return s2;
        }
        phase("schedule tuning", tph);
        prm.lanes_rows = plan_code(plan_rows);
        prm.lanes_cols = plan_code(plan_cols);
        prm.hot_columns = plan_rows.hot;
        if (!kSameTypes) {
            // FP32 values under FP64 iterates: the multi-kernel loop only
        } else if (prm.single_workgroup > 0 && prm.time_limit_s <= 0.0) {
            // (with a time limit the multi-kernel loop runs: it reads the
            // clock between check intervals, a one-workgroup solve cannot)
            one_wg = true;
        } else if (prm.single_workgroup < 0 && nnz <= kSmallTuneNnz && prm.time_limit_s <= 0.0) {
            bo_status s3 = tune_one_wg();
            if (s3 != BO_SUCCESS) return s3;
        }
        // A one-workgroup solve checks every 64 steps unless told otherwise:
        // there a check costs little, and with FP32 iterates longer
        // intervals have failed to converge on small LPs (C0-sized random
        // LPs at 128 and 256) where 64 converges and FP64 converges at any.
        if (one_wg && tune_cadence) prm.check_every = 64;
        phase("one-workgroup tuning", tph);
        return BO_SUCCESS;
    }

    // Host part of the tiled layout of one pass (no device call: it runs on a
    // host thread of its own); false when the LDS is too small or the padded
    // layout would not fit 32-bit offsets (the pass then stays unbuilt).
    bool host_tiles(int pass, int k, const int* ptr, const int* gidx, HostTiles& h) const {
        const int nseg = pass == 0 ? m : n, ng = pass == 0 ? n : m;
        if (nseg == 0) return false;
        // LDS holds the vector block: the largest block that fits
        const int wmax = std::min(65536, lds_max / static_cast<int>(sizeof(T))) / 64 * 64;
        if (wmax < 64) return false;
        return build_tiles(ptr, gidx, nseg, ng, wmax, k * num_cus, h);
    }

    // Uploads the host layout of one pass (cut for k tiles per CU).
    hipError_t upload_tile_pass(int pass, int k, const HostTiles& h) {
        TilePass& t = tp[pass];
        // buffer ranges of the kernels are 32-bit byte counts below 2^31
        if (h.pairs * 4 >= (1LL << 31) - 64 || static_cast<long long>(h.slice_word.size()) * kWave * 8 >= (1LL << 31) - 64)
            return hipSuccess;
        const size_t lds_need = static_cast<size_t>(h.W) * sizeof(T);
        hipError_t e = pass == 0 ? set_tile_lds<0>(static_cast<int>(lds_need))
                                 : set_tile_lds<1>(static_cast<int>(lds_need));
        if (e != hipSuccess) return e;
        t.lds_bytes = static_cast<unsigned>(lds_need);
        t.B = h.B;
        t.W = h.W;
        t.ntiles = static_cast<int>(h.tiles.size());
        t.nslices = static_cast<int>(h.slice_word.size());
        t.padded = h.padded;
        t.pairs = h.pairs;
        t.max_pairs = h.max_seg_pairs;
        if ((e = mem.alloc(&t.td, h.tiles.size())) != hipSuccess) return e;
        if ((e = mem.alloc(&t.words, h.slice_word.size())) != hipSuccess) return e;
        if ((e = mem.alloc(&t.idx, h.padded)) != hipSuccess) return e;
        if ((e = mem.alloc(&t.src, h.padded)) != hipSuccess) return e;
        if ((e = mem.alloc(&t.val, h.padded)) != hipSuccess) return e;
        if ((e = mem.alloc(&t.seg_ptr, h.seg_ptr.size())) != hipSuccess) return e;
        if ((e = mem.alloc(&t.slot, h.slot.size())) != hipSuccess) return e;
        if ((e = mem.alloc(&t.part, h.slice_word.size() * kWave)) != hipSuccess) return e;
        auto up = [&](void* d, const void* src, size_t bytes) {
            return bytes ? hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, stream) : hipSuccess;
        };
        if ((e = up(t.td, h.tiles.data(), sizeof(TileDesc) * h.tiles.size())) != hipSuccess) return e;
        if ((e = up(t.words, h.slice_word.data(), sizeof(unsigned) * h.slice_word.size())) != hipSuccess) return e;
        if ((e = up(t.idx, h.idx.data(), sizeof(unsigned short) * h.padded)) != hipSuccess) return e;
        if ((e = up(t.src, h.src.data(), sizeof(int) * h.padded)) != hipSuccess) return e;
        if ((e = up(t.seg_ptr, h.seg_ptr.data(), sizeof(int) * h.seg_ptr.size())) != hipSuccess) return e;
        if ((e = up(t.slot, h.slot.data(), sizeof(unsigned) * h.slot.size())) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;  // host staging goes out of scope
        t.k = k;
        return hipGetLastError();
    }

    static int cap_index(int cap) { return cap <= 1024 ? 0 : (cap <= 2048 ? 1 : 2); }

    // Schedule code: k > 0 segment kernel with k nonzeros per lane; -1..-64
    // tiled kernels (layout cut for that many tiles per CU, used when built);
    // <= -1024 stream kernel with LDS budget -code; 0 = segment kernel, k = 4
    // (the tuner's starting point).
    PassPlan fixed_plan(int pass, int code) const {
        const std::vector<int>& len = pass == 0 ? row_len_sorted : col_len_sorted;
        PassPlan p;
        if (code <= -1024) {
            p.cap = kStreamCaps[cap_index(-code)];
        } else if (code < 0) {
            if (tp[pass].k == -code) p.tiled = -code;
        } else {
            p.kp = code > 0 ? code : 4;
        }
        p.bt = make_bins(len, p.kp);
        return p;
    }

    static int plan_code(const PassPlan& p) { return p.tiled > 0 ? -p.tiled : (p.cap > 0 ? -p.cap : p.kp); }

    // ---- launch helpers
    // ev (optional): events timing the launch (launch_ev)
    template <int MODE, class TV, class Epi>
    void launch_rows(const BinTable& bt, const TV* vals, const Epi& epi, double* partials, const int* done,
                     const hipEvent_t* ev = nullptr) {
        launch_ev(k_segments<MODE, TV, typename Epi::TXv, Epi>, dim3(bt.total_blocks), dim3(kBlock), 0, stream, ev,
                  bt, static_cast<const int*>(rp), static_cast<const int*>(ci), vals, epi, partials, done);
    }
    template <int MODE, class TV, class Epi>
    void launch_cols(const BinTable& bt, const TV* vals, const Epi& epi, double* partials, const int* done,
                     const hipEvent_t* ev = nullptr) {
        launch_ev(k_segments<MODE, TV, typename Epi::TXv, Epi>, dim3(bt.total_blocks), dim3(kBlock), 0, stream, ev,
                  bt, static_cast<const int*>(cp), static_cast<const int*>(ri), vals, epi, partials, done);
    }

    const int* done_flag() const { return &st->done; }

    template <class Epi>
    void launch_stream(int pass, int cap, const int* ptr, const int* idx, const TM* vals, const Epi& e,
                       const int* done, const hipEvent_t* ev = nullptr) {
        const int c = cap_index(cap);
        const int nb = nsblk[pass][c];
        const int4* blk = sblk[pass][c];
        switch (c) {
            case 0:
                launch_ev(k_stream<1024, TM, T, Epi>, dim3(nb), dim3(kBlock), 0, stream, ev, blk, ptr, idx, vals, e,
                          done);
                break;
            case 1:
                launch_ev(k_stream<2048, TM, T, Epi>, dim3(nb), dim3(kBlock), 0, stream, ev, blk, ptr, idx, vals, e,
                          done);
                break;
            default:
                launch_ev(k_stream<4096, TM, T, Epi>, dim3(nb), dim3(kBlock), 0, stream, ev, blk, ptr, idx, vals, e,
                          done);
                break;
        }
    }

    template <int PASS>
    static hipError_t set_tile_lds(int bytes) {
        return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile<TM, T, kTileRows, PASS>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    }

    // Tiled pass: gather kernel into the partials, `mid` (optional event),
    // reduction + epilogue kernel.  Once the LP's done flag is set the
    // epilogue stores nothing (the gather runs: its partials go nowhere).
    // ev4 (optional): start / end events of the gather and of the reduction
    template <int PASS, class Epi>
    void launch_tiled(const T* gvec, const Epi& e, hipEvent_t mid, const hipEvent_t* ev4 = nullptr) {
        const TilePass& t = tp[PASS];
        const int nseg = PASS == 0 ? m : n, ng = PASS == 0 ? n : m;
        if (t.ntiles > 0)
            launch_ev(k_tile<TM, T, kTileRows, PASS>, dim3(t.ntiles), dim3(kTileThreads), t.lds_bytes, stream, ev4,
                      static_cast<const unsigned char*>(nullptr), static_cast<const TileDesc*>(t.td),
                      static_cast<const unsigned*>(t.words), static_cast<const TM*>(t.val),
                      static_cast<const unsigned short*>(t.idx), gvec, ng, t.W, t.part);
        if (mid) hipEventRecord(mid, stream);
        launch_tile_reduce(stream, t.max_pairs, nseg, t.seg_ptr, t.slot, static_cast<int>(t.pairs), t.part,
                           t.nslices * kWave, e, t.reduce_u, ev4 ? ev4 + 2 : nullptr);
    }

    // ev4 (optional): start / end events of the pass's kernels (a fused pass
    // uses the first pair only)
    void launch_primal(const PassPlan& p, const int* done, int t, hipEvent_t mid = nullptr,
                       const hipEvent_t* ev4 = nullptr) {
        EpiPrimal<T> e{yg, st, c_hat, l_hat, u_hat, xk, x0, xp, t, nullptr};
        if (p.tiled > 0) {
            launch_tiled<1>(yg, e, mid, ev4);
            return;
        }
        if (p.cap > 0) launch_stream(1, p.cap, cp, ri, cv, e, done, ev4);
        else launch_cols<kSumProd>(p.bt, cv, e, nullptr, done, ev4);
        if (mid) hipEventRecord(mid, stream);
    }
    void launch_dual(const PassPlan& p, const int* done, int t, hipEvent_t mid = nullptr,
                     const hipEvent_t* ev4 = nullptr) {
        EpiDual<T> e{xp, st, b_hat, eq, ax, ax0, yk, y0, yp, yg, t, nullptr};
        if (p.tiled > 0) {
            launch_tiled<0>(xp, e, mid, ev4);
            return;
        }
        if (p.cap > 0) {
            launch_stream(0, p.cap, rp, ci, rv, e, done, ev4);
        } else if (p.hot > 0) {
            launch_ev(k_segments_hot<kSumProd, TM, T, EpiDual<T>>, dim3(std::min(p.hot_grid, p.bt.total_blocks)),
                      dim3(kBlock), static_cast<unsigned>(p.hot * sizeof(T)), stream, ev4, p.bt,
                      static_cast<const int*>(rp), static_cast<const int*>(ci), static_cast<const TM*>(rv), e, p.hot,
                      done);
        } else {
            launch_rows<kSumProd>(p.bt, rv, e, nullptr, done, ev4);
        }
        if (mid) hipEventRecord(mid, stream);
    }

    // Resident workgroups of the persistent CSR pass with `hot` LDS entries.
    int grid_for_hot(int hot) {
        int per_cu = 0;
        hipError_t err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_segments_hot<kSumProd, TM, T, EpiDual<T>>), kBlock,
            hot * sizeof(T));
        if (err != hipSuccess || per_cu < 1) return 0;
        return per_cu * num_cus;
    }

    void launch_check() {
        hipLaunchKernelGGL(k_check_x2<T>, dim3(nb_x), dim3(kBlock), 0, stream, n, st, xk, xp, cs, c64, lb64, ub64,
                           xo2, part_x);
        EpiCheckRows2<T> er{xo2, b64, eq, rs, yk, yp, yo2, ax_cand};
        launch_rows<kSumProd>(plan_rows.bt, rv64, er, part_r, done_flag());
        EpiCheckCols2 ec{yo2, c64, lb64, ub64, cs};
        launch_cols<kSumProd>(plan_cols.bt, cv64, ec, part_c, done_flag());
        CheckDims d{nb_x, plan_rows.bt.total_blocks, plan_cols.bt.total_blocks};
        hipLaunchKernelGGL(k_fold_check, dim3(kFoldGroups), dim3(kBlock), 0, stream, st, d, part_x, part_r, part_c,
                           part_fold);
        hipLaunchKernelGGL(k_decide, dim3(1), dim3(kDecideThreads), 0, stream, st,
                           static_cast<const double*>(part_fold), kFoldGroups);
        hipLaunchKernelGGL(k_restart_x<T>, dim3(nb_x), dim3(kBlock), 0, stream, n, st, xk, x0, xp, part_dx);
        hipLaunchKernelGGL(k_restart_y<T>, dim3(nb_y), dim3(kBlock), 0, stream, m, st, yk, y0, yp, yg, ax, ax0,
                           ax_cand, part_dy);
        hipLaunchKernelGGL(k_restart_finish, dim3(1), dim3(kBlock), 0, stream, st, part_dx, nb_x, part_dy, nb_y);
    }

    // k steps followed by the check.
    void launch_interval(int k) {
        for (int t = 0; t < k; ++t) {
            launch_primal(plan_cols, done_flag(), t);
            launch_dual(plan_rows, done_flag(), t);
        }
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, stream, st, k);
        launch_check();
    }

    // Alternating launches, as inside a solve (primal, dual, primal, ...).
    // ms4 (optional): average of [CSC gather, CSC reduce, CSR gather, CSR
    // reduce]; a single-kernel pass reports its time as gather, 0 as reduce.
    bo_status time_pair(const PassPlan& pr, const PassPlan& pc, int reps, double* csr_ms, double* csc_ms,
                        double* ms4 = nullptr, double* step_ms = nullptr) {
        std::vector<hipEvent_t> ev(4 * static_cast<size_t>(reps) + 1);
        for (auto& e : ev) BO_TRY(hipEventCreate(&e));
        launch_primal(pc, nullptr, 0);
        launch_dual(pr, nullptr, 0);
        for (int r = 0; r < reps; ++r) {
            hipEventRecord(ev[4 * r], stream);
            launch_primal(pc, nullptr, r, ev[4 * r + 1]);
            hipEventRecord(ev[4 * r + 2], stream);
            launch_dual(pr, nullptr, r, ev[4 * r + 3]);
        }
        hipEventRecord(ev[4 * reps], stream);
        hipEventSynchronize(ev[4 * reps]);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int r = 0; r < reps; ++r)
            for (int q = 0; q < 4; ++q) {
                float a = 0.f;
                hipEventElapsedTime(&a, ev[4 * r + q], ev[4 * r + q + 1]);
                acc[q] += a;
            }
        for (auto& e : ev) hipEventDestroy(e);
        BO_TRY(hipGetLastError());
        if (csc_ms) *csc_ms = (acc[0] + acc[1]) / reps;
        if (csr_ms) *csr_ms = (acc[2] + acc[3]) / reps;
        if (ms4)
            for (int q = 0; q < 4; ++q) ms4[q] = acc[q] / reps;
        if (step_ms) *step_ms = (acc[0] + acc[1] + acc[2] + acc[3]) / reps;
        return BO_SUCCESS;
    }

    // The schedule of each pass -- segment kernel with a nonzeros-per-lane
    // target (for the CSR pass optionally with an LDS copy of the hottest x+
    // entries) or stream kernel with an LDS budget -- chosen by timing the
    // fused iteration kernels in alternation, the cache state of a real solve.
    bo_status autotune() {
        bo_status s0 = prepare();
        if (s0 != BO_SUCCESS) return s0;
        const int codes[] = {2, 4, 8, 16, -1024, -2048, -4096, -2};
        std::vector<PassPlan> cr, cc;
        if (prm.lanes_rows == 0) {
            for (int code : codes) cr.push_back(fixed_plan(0, code));
        } else {
            cr.push_back(plan_rows);
        }
        if (prm.lanes_cols == 0) {
            for (int code : codes) cc.push_back(fixed_plan(1, code));
        } else {
            cc.push_back(plan_cols);
        }
        // the other pass of each timed pair: the tiled one where it is built
        // (it is what a solve alternates with at the sizes that build it, and
        // the cheapest to run beside each candidate), else the segment kernel
        const PassPlan base_r = fixed_plan(0, tp[0].k > 0 ? -tp[0].k : 4);
        const PassPlan base_c = fixed_plan(1, tp[1].k > 0 ? -tp[1].k : 4);
        double best_r = 1e30, best_c = 1e30;
        PassPlan pr = cr[0], pc = cc[0];
        for (const PassPlan& cand : cr) {
            double t_ms = 0.0;
            bo_status s = time_pair(cand, base_c, 4, &t_ms, nullptr);
            if (s != BO_SUCCESS) return s;
            if (t_ms < best_r) { best_r = t_ms; pr = cand; }
        }
        for (const PassPlan& cand : cc) {
            double t_ms = 0.0;
            bo_status s = time_pair(base_r, cand, 4, nullptr, &t_ms);
            if (s != BO_SUCCESS) return s;
            if (t_ms < best_c) { best_c = t_ms; pc = cand; }
        }
        if (pr.cap > 0 || pr.tiled > 0) {
            // no hot-column variant
        } else if (prm.hot_columns > 0) {  // fixed by the caller
            pr.hot = plan_rows.hot;
            pr.hot_grid = plan_rows.hot_grid;
        } else if (prm.hot_columns < 0) {  // LDS copy of the hottest x+ entries
            for (int hot : {4096, 8192, 16384}) {
                if (hot > n) break;
                PassPlan h = pr;
                h.hot = hot;
                h.hot_grid = grid_for_hot(hot);
                if (h.hot_grid <= 0) continue;
                double t_ms = 0.0;
                bo_status s = time_pair(h, pc, 4, &t_ms, nullptr);
                if (s != BO_SUCCESS) return s;
                if (t_ms < 0.97 * best_r) { best_r = t_ms; pr = h; }
            }
        }
        plan_rows = pr;
        plan_cols = pc;
        // Loads in flight per round of each tiled reduction: the fewest
        // that cover a segment's partials wastes address-unit work when most
        // segments have far fewer partials than the longest (C1's column
        // pass: at most 5, mostly 1); timed on the real layout, kept when
        // 3% faster.
        for (int pass = 0; pass < 2; ++pass) {
            TilePass& t = tp[pass];
            if ((pass == 0 ? plan_rows.tiled : plan_cols.tiled) <= 0 || t.max_pairs <= 1) continue;
            int cover = 1;
            while (cover < t.max_pairs && cover < 16) cover *= 2;
            double best = 1e30;
            int best_u = 0;
            for (int u = cover; u >= 1; u /= 2) {
                t.reduce_u = u;
                double ms4[4];
                bo_status s = time_pair(plan_rows, plan_cols, 4, nullptr, nullptr, ms4);
                if (s != BO_SUCCESS) return s;
                const double red = pass == 0 ? ms4[3] : ms4[1];
                if (best_u == 0 || red < 0.97 * best) {
                    best = red;
                    best_u = u;
                }
            }
            t.reduce_u = best_u;
        }
        if (prm.check_every <= 0) {
            // Check cadence 64, as in the batched and family engines.  A
            // cadence chosen by cost (128 where a check costs four steps or
            // more) was tried: restarts can only happen at checks, and the
            // iteration counts it gave were no better in aggregate (seven
            // large LPs: 6.18 s against 5.87 s at 64, 5.91 s at 128).
            prm.check_every = 64;
        }
        return BO_SUCCESS;
    }

    SmallArgs<T> small_args() {
        SmallArgs<T> a{};
        a.m = m;
        a.n = n;
        a.K = prm.check_every > 0 ? prm.check_every : 64;
        a.nnz = static_cast<int>(nnz);
        a.rp = rp;
        a.ci = ci;
        a.cp = cp;
        a.ri = ri;
        a.rv = rv;
        a.cv = cv;
        a.rv64 = rv64;
        a.cv64 = cv64;
        a.ep = EpiPrimal<T>{yg, st, c_hat, l_hat, u_hat, xk, x0, xp, 0, nullptr};
        a.ed = EpiDual<T>{xp, st, b_hat, eq, ax, ax0, yk, y0, yp, yg, 0, nullptr};
        a.er = EpiCheckRows<T>{xo_cur, xo_cand, b64, eq, rs, yk, yp, yo_cur, yo_cand, ax_cand};
        a.ec = EpiCheckCols{yo_cur, yo_cand, c64, lb64, ub64, cs};
        a.cs = cs;
        a.c64 = c64;
        a.lb64 = lb64;
        a.ub64 = ub64;
        a.xo_cur = xo_cur;
        a.xo_cand = xo_cand;
        a.x0 = x0;
        a.y0 = y0;
        a.ax0 = ax0;
        a.st = st;
        return a;
    }

    // LDS for the gathered vectors when they fit beside the kernel's own
    // (the attribute is raised once, at the first use)
    unsigned small_lds_bytes() const {
        const size_t b = (static_cast<size_t>(n) + m) * sizeof(T);
        return b <= static_cast<size_t>(kSmallMaxLds) ? static_cast<unsigned>(b) : 0u;
    }

    // 0: the matrix does not fit the LDS (or too many columns / rows per
    // thread); else the columns / rows per thread of k_small_lds
    int small_resident_per_thread() const {
        const long long per = ((m > n ? m : n) + kSmallThreads - 1) / kSmallThreads;
        if (per > kSmallLdsMaxPerThread || nnz >= (1LL << 30)) return 0;
        if (small_resident_bytes<T>(m, n, nnz) + 4096 > static_cast<size_t>(lds_max)) return 0;  // + static LDS
        return per < 1 ? 1 : static_cast<int>(per);
    }

    void launch_one_wg() {
        if constexpr (kSameTypes) {  // (never chosen under FP32 values with FP64 iterates)
            if (const int per = small_resident_per_thread()) {
                SmallArgs<T> a = small_args();
                a.power = power_pending ? 1 : 0;
                a.scal = scal;
                const size_t bytes = small_resident_bytes<T>(m, n, nnz);
                // lanes per row: as many as the workgroup has threads for
                const int lr = per > 1 ? 1 : (4LL * m <= kSmallThreads ? 4 : (2LL * m <= kSmallThreads ? 2 : 1));
                const void* fn = per > 1     ? reinterpret_cast<const void*>(&k_small_lds<T, 2, 1>)
                                 : lr == 4 ? reinterpret_cast<const void*>(&k_small_lds<T, 1, 4>)
                                 : lr == 2 ? reinterpret_cast<const void*>(&k_small_lds<T, 1, 2>)
                                           : reinterpret_cast<const void*>(&k_small_lds<T, 1, 1>);
                if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)) ==
                    hipSuccess) {
                    if (per > 1)
                        hipLaunchKernelGGL((k_small_lds<T, 2, 1>), dim3(1), dim3(kSmallThreads), bytes, stream, a);
                    else if (lr == 4)
                        hipLaunchKernelGGL((k_small_lds<T, 1, 4>), dim3(1), dim3(kSmallThreads), bytes, stream, a);
                    else if (lr == 2)
                        hipLaunchKernelGGL((k_small_lds<T, 1, 2>), dim3(1), dim3(kSmallThreads), bytes, stream, a);
                    else
                        hipLaunchKernelGGL((k_small_lds<T, 1, 1>), dim3(1), dim3(kSmallThreads), bytes, stream, a);
                    power_pending = false;
                    return;
                }
            }
            if (power_pending) {  // the LDS-resident kernel could not be set up: the launches do it
                const BinTable br = make_bins(row_len_sorted, 4);
                launch_power(br);
                launch_start(br);
                power_pending = false;
            }
            SmallArgs<T> a = small_args();
            const unsigned lds = small_lds_bytes();
            a.lds = lds > 0 ? 1 : 0;
            if (lds > 0 && !small_lds_set) {
                if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small<T>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kSmallMaxLds)) ==
                    hipSuccess)
                    small_lds_set = true;
            }
            if (!small_lds_set) {
                a.lds = 0;
                hipLaunchKernelGGL(k_small<T>, dim3(1), dim3(kSmallThreads), 0, stream, a);
            } else {
                hipLaunchKernelGGL(k_small<T>, dim3(1), dim3(kSmallThreads), a.lds ? lds : 0u, stream, a);
            }
        }
    }

    // One workgroup or the multi-kernel loop: both run 4 check intervals from
    // the cold start; the one-workgroup solve is kept when it takes less than
    // 0.8 of the other (the loop's hipGraph replay saves some launch time).
    bo_status tune_one_wg() {
        const long long keep = prm.max_iterations;
        const int K = prm.check_every > 0 ? prm.check_every : 64;
        prm.max_iterations = 4LL * K;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        float t_multi = 0.0f, t_one = 0.0f;
        bo_status s = BO_SUCCESS;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess && (s = prepare()) == BO_SUCCESS) {
            e = hipEventRecord(e0, stream);
            launch_check();
            for (int r = 0; r < 4; ++r) launch_interval(K);
            if (e == hipSuccess) e = hipEventRecord(e1, stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&t_multi, e0, e1);
        }
        if (e == hipSuccess && s == BO_SUCCESS && (s = prepare(small_resident_per_thread() > 0)) == BO_SUCCESS) {
            e = hipEventRecord(e0, stream);
            launch_one_wg();
            if (e == hipSuccess) e = hipEventRecord(e1, stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&t_one, e0, e1);
            if (e == hipSuccess) e = hipGetLastError();
        }
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        prm.max_iterations = keep;
        if (s != BO_SUCCESS) return s;
        BO_TRY(e);
        one_wg = t_one < 0.8f * t_multi;
        return BO_SUCCESS;
    }

    bo_status set_warm_start(const double* wx_h, const double* wy_h) override {
        if (!wx_h && !wy_h) {
            has_warm = 0;
            return BO_SUCCESS;
        }
        std::vector<double> hx(n, 0.0), hy(m, 0.0);
        if (wx_h)
            for (int p2 = 0; p2 < n; ++p2) hx[p2] = wx_h[col_of[p2]];
        if (wy_h)
            for (int p2 = 0; p2 < m; ++p2) hy[p2] = wy_h[row_of[p2]];
        if (n) BO_TRY(hipMemcpyAsync(wx, hx.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream));
        if (m) BO_TRY(hipMemcpyAsync(wy, hy.data(), sizeof(double) * m, hipMemcpyHostToDevice, stream));
        BO_TRY(hipStreamSynchronize(stream));
        has_warm = 1;
        return BO_SUCCESS;
    }

    // Scaling, step size, starting point, A x0: everything device-side.
    bo_status prepare(bool kernel_power = false) {
        const unsigned gm = grid_for(m), gn = grid_for(n);
        hipLaunchKernelGGL(k_fill_f64, dim3(gm), dim3(kBlock), 0, stream, rs, 1.0, static_cast<size_t>(m));
        hipLaunchKernelGGL(k_fill_f64, dim3(gn), dim3(kBlock), 0, stream, cs, 1.0, static_cast<size_t>(n));
        const BinTable br = make_bins(row_len_sorted, 4), bc = make_bins(col_len_sorted, 4);
        for (int it = 0; it < prm.ruiz_iterations; ++it) {
            EpiNorm er{cs, rs, rnorm};
            launch_rows<kMaxAbs>(br, rv64, er, nullptr, nullptr);
            EpiNorm ec{rs, cs, cnorm};
            launch_cols<kMaxAbs>(bc, cv64, ec, nullptr, nullptr);
            hipLaunchKernelGGL(k_apply_norm, dim3(gm), dim3(kBlock), 0, stream, rs, rnorm, static_cast<size_t>(m));
            hipLaunchKernelGGL(k_apply_norm, dim3(gn), dim3(kBlock), 0, stream, cs, cnorm, static_cast<size_t>(n));
        }
        if (prm.pock_chambolle) {
            EpiNorm er{cs, rs, rnorm};
            launch_rows<kSumAbs>(br, rv64, er, nullptr, nullptr);
            EpiNorm ec{rs, cs, cnorm};
            launch_cols<kSumAbs>(bc, cv64, ec, nullptr, nullptr);
            hipLaunchKernelGGL(k_apply_norm, dim3(gm), dim3(kBlock), 0, stream, rs, rnorm, static_cast<size_t>(m));
            hipLaunchKernelGGL(k_apply_norm, dim3(gn), dim3(kBlock), 0, stream, cs, cnorm, static_cast<size_t>(n));
        }
        EpiWrite<TM> wr{rs, cs, rv};
        launch_rows<kWrite>(br, rv64, wr, nullptr, nullptr);
        EpiWrite<TM> wc{cs, rs, cv};
        launch_cols<kWrite>(bc, cv64, wc, nullptr, nullptr);
        for (int pass = 0; pass < 2; ++pass) {
            const TilePass& t = tp[pass];
            if (t.k > 0 && t.padded > 0)
                hipLaunchKernelGGL(k_tile_values<TM>, dim3(grid_for(t.padded)), dim3(kBlock), 0, stream, t.src, t.padded,
                                   pass == 0 ? rv : cv, t.val);
        }
        hipLaunchKernelGGL(k_scale_cols<T>, dim3(gn), dim3(kBlock), 0, stream, n, cs, c64, lb64, ub64, wx,
                           has_warm, c_hat, l_hat, u_hat, xk, x0, xp);
        hipLaunchKernelGGL(k_scale_rows<T>, dim3(gm), dim3(kBlock), 0, stream, m, rs, b64, eq, wy, has_warm,
                           b_hat, yk, y0, yg, yp);
        // power iteration for ||A_hat||_2 (the LDS-resident one-workgroup
        // solve runs it itself: kernel_power)
        if (!kernel_power) launch_power(br);
        hipLaunchKernelGGL(k_sq_partial<T>, dim3(nb_x), dim3(kBlock), 0, stream, c_hat, n, part_dx);
        hipLaunchKernelGGL(k_sum_partials_to, dim3(1), dim3(kBlock), 0, stream, part_dx, nb_x, 1, 0, scal + 2);
        hipLaunchKernelGGL(k_sq_partial<T>, dim3(nb_y), dim3(kBlock), 0, stream, b_hat, m, part_dy);
        hipLaunchKernelGGL(k_sum_partials_to, dim3(1), dim3(kBlock), 0, stream, part_dy, nb_y, 1, 0, scal + 3);
        // state
        BO_TRY(hipStreamSynchronize(stream));  // st_init may still feed a previous upload
        std::memset(st_init, 0, sizeof(PdlpState));
        st_init->b_norm = b_norm;
        st_init->c_norm = c_norm;
        st_init->eps = prm.eps_rel_kkt;
        st_init->eps_obj = prm.eps_objective;
        st_init->max_iter = prm.max_iterations;
        BO_TRY(hipMemcpyAsync(st, st_init, sizeof(PdlpState), hipMemcpyHostToDevice, stream));
        power_pending = kernel_power;
        if (!kernel_power) launch_start(br);
        BO_TRY(hipGetLastError());
        return BO_SUCCESS;
    }

    // ||A_hat||_2 by power iteration into scal[0]
    void launch_power(const BinTable& br) {
        const BinTable bc = make_bins(col_len_sorted, 4);
        hipLaunchKernelGGL(k_pw_init<T>, dim3(grid_for(n)), dim3(kBlock), 0, stream, pw_v, n);
        hipLaunchKernelGGL(k_sq_partial<T>, dim3(nb_x), dim3(kBlock), 0, stream, pw_v, n, part_dx);
        hipLaunchKernelGGL(k_sum_partials_to, dim3(1), dim3(kBlock), 0, stream, part_dx, nb_x, 1, 0, scal + 1);
        hipLaunchKernelGGL(k_pw_normalize<T>, dim3(grid_for(n)), dim3(kBlock), 0, stream, pw_v, n, scal + 1);
        for (int it = 0; it < 40; ++it) {
            EpiStore<T, T> eu{pw_v, pw_u};
            launch_rows<kSumProd>(br, rv, eu, part_r, nullptr);
            hipLaunchKernelGGL(k_sum_partials_to, dim3(1), dim3(kBlock), 0, stream, part_r, br.total_blocks, 1, 0,
                               scal + 0);
            EpiStore<T, T> ew{pw_u, pw_v};
            launch_cols<kSumProd>(bc, cv, ew, part_c, nullptr);
            hipLaunchKernelGGL(k_sum_partials_to, dim3(1), dim3(kBlock), 0, stream, part_c, bc.total_blocks, 1, 0,
                               scal + 1);
            hipLaunchKernelGGL(k_pw_normalize<T>, dim3(grid_for(n)), dim3(kBlock), 0, stream, pw_v, n, scal + 1);
        }
    }

    // step sizes and primal weight from scal, then A x0 (the starting point
    // is also the first anchor)
    void launch_start(const BinTable& br) {
        hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, stream, st, scal + 0, scal + 2, scal + 3);
        EpiStore<T, double> ea{xp, ax};
        launch_rows<kSumProd>(br, rv, ea, part_r, nullptr);
        hipLaunchKernelGGL(k_copy_f64, dim3(grid_for(m)), dim3(kBlock), 0, stream, ax0, ax, static_cast<size_t>(m));
    }

    bo_status solve(bo_pdlp_stats* out) override {
        auto t0 = std::chrono::high_resolution_clock::now();
        bo_status s0 = prepare(one_wg && small_resident_per_thread() > 0);
        if (s0 != BO_SUCCESS) return s0;
        if (!one_wg) launch_check();  // step 0: KKT of the starting point (k_small runs its own)
        BO_TRY(hipMemcpyAsync(st_host, st, sizeof(PdlpState), hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        auto t1 = std::chrono::high_resolution_clock::now();
        const int K = prm.check_every > 0 ? prm.check_every : 64;
        long long iters = 0;
        if (one_wg) {
            launch_one_wg();
            BO_TRY(hipGetLastError());
            BO_TRY(hipMemcpyAsync(st_host, st, sizeof(PdlpState), hipMemcpyDeviceToHost, stream));
            BO_TRY(hipStreamSynchronize(stream));
        }
        // Intervals are queued two deep while the last state read is far from
        // the stopping test (best rel-KKT above 10 eps): the next one is
        // launched before the host reads the state the previous one left, so
        // the device does not wait on the host between them.  Closer in they
        // run one at a time, so that no interval runs after convergence.  One
        // that does (a misjudged step) changes nothing: once the done flag is
        // set its epilogues store nothing and its check and restart kernels
        // return, so the final state is the converged one.
        auto depth_after = [&](const PdlpState& h) {
            return std::min(h.rel_kkt[0], h.rel_kkt[1]) > 10.0 * prm.eps_rel_kkt ? 2 : 1;
        };
        bool stop = !one_wg && st_host->done, timed_out = false;
        int queued = 0, oldest = 0, depth = depth_after(*st_host);
        while (!one_wg && !stop) {
            while (queued < depth) {
                const long long left = prm.max_iterations - iters;
                const int k = static_cast<int>(std::min<long long>(K, left));
                if (k <= 0) break;
                if (prm.use_graph && k == K) {
                    if (!graph_exec || graph_k != K) {
                        if (graph_exec) hipGraphExecDestroy(graph_exec);
                        if (graph) hipGraphDestroy(graph);
                        graph_exec = nullptr;
                        graph = nullptr;
                        BO_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                        launch_interval(K);
                        BO_TRY(hipStreamEndCapture(stream, &graph));
                        BO_TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
                        graph_k = K;
                    }
                    BO_TRY(hipGraphLaunch(graph_exec, stream));
                } else {
                    launch_interval(k);
                }
                BO_TRY(hipGetLastError());
                iters += k;
                const int q = (oldest + queued) % 2;
                BO_TRY(hipMemcpyAsync(st_ring[q], st, sizeof(PdlpState), hipMemcpyDeviceToHost, stream));
                BO_TRY(hipEventRecord(st_ready[q], stream));
                ++queued;
            }
            if (queued == 0) break;  // max_iterations reached
            BO_TRY(hipEventSynchronize(st_ready[oldest]));
            stop = st_ring[oldest]->done != 0;
            depth = depth_after(*st_ring[oldest]);
            oldest ^= 1;
            --queued;
            if (!stop && prm.time_limit_s > 0.0 &&
                std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count() >
                    prm.time_limit_s) {
                timed_out = true;
                stop = true;
            }
        }
        if (!one_wg) {  // the queue drained: the state as the last interval left it
            BO_TRY(hipMemcpyAsync(st_host, st, sizeof(PdlpState), hipMemcpyDeviceToHost, stream));
            BO_TRY(hipStreamSynchronize(stream));
            if (timed_out && !st_host->done) st_host->status = 3;
        }
        auto t2 = std::chrono::high_resolution_clock::now();
        if (out) {
            const PdlpState& h = *st_host;
            const int b = h.best;
            out->status = h.done ? h.status : (h.status == 3 ? 3 : 1);
            out->iterations = h.iter;
            out->restarts = h.restarts;
            out->used_average = b;
            out->rel_kkt = h.rel_kkt[b];
            out->rel_primal = h.rel_primal[b];
            out->rel_dual = h.rel_dual[b];
            out->rel_gap = h.rel_gap[b];
            out->primal_obj = h.pobj[b];
            out->dual_obj = h.dobj[b];
            out->primal_weight = h.omega;
            out->step_size = h.eta;
            out->setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            out->solve_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
            out->spmv_csr_bytes = 8.0 * nnz + 4.0 * (m + 1) + 4.0 * n + 4.0 * m;
            out->spmv_csc_bytes = 8.0 * nnz + 4.0 * (n + 1) + 4.0 * m + 4.0 * n;
            out->lanes_rows = prm.lanes_rows;
            out->lanes_cols = prm.lanes_cols;
            out->hot_columns = plan_rows.hot;
            out->check_every = prm.check_every;
            out->single_workgroup = one_wg ? (small_resident_per_thread() ? 2 : 1) : 0;
        }
        return BO_SUCCESS;
    }

    bo_status get_solution(double* hx, double* hy) override {
        const int b = st_host->best;
        std::vector<double> tx(n), ty(m);
        if (!one_wg) {  // the multi-kernel checks keep both points interleaved
            if (n) hipLaunchKernelGGL(k_pick2, dim3(grid_for(n)), dim3(kBlock), 0, stream, static_cast<size_t>(n),
                                      static_cast<const double2*>(xo2), b, b ? xo_cand : xo_cur);
            if (m) hipLaunchKernelGGL(k_pick2, dim3(grid_for(m)), dim3(kBlock), 0, stream, static_cast<size_t>(m),
                                      static_cast<const double2*>(yo2), b, b ? yo_cand : yo_cur);
            BO_TRY(hipGetLastError());
        }
        if (n) BO_TRY(hipMemcpyAsync(tx.data(), b ? xo_cand : xo_cur, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
        if (m) BO_TRY(hipMemcpyAsync(ty.data(), b ? yo_cand : yo_cur, sizeof(double) * m, hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        if (hx)
            for (int p2 = 0; p2 < n; ++p2) hx[col_of[p2]] = tx[p2];
        if (hy)
            for (int p2 = 0; p2 < m; ++p2) hy[row_of[p2]] = ty[p2];
        return BO_SUCCESS;
    }

    // Runs f with the device state's done flag cleared, so that the pass
    // epilogues (which store nothing for a finished LP) do a solve's work
    // after a solve, then puts the state back.
    template <class F>
    bo_status with_live_state(F&& f) {
        PdlpState keep, live;
        BO_TRY(hipMemcpyAsync(&keep, st, sizeof(PdlpState), hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        live = keep;
        live.done = 0;
        BO_TRY(hipMemcpyAsync(st, &live, sizeof(PdlpState), hipMemcpyHostToDevice, stream));
        const bo_status r = f();
        BO_TRY(hipMemcpyAsync(st, &keep, sizeof(PdlpState), hipMemcpyHostToDevice, stream));
        BO_TRY(hipStreamSynchronize(stream));
        return r;
    }

    // Times the two fused iteration kernels as they alternate in a solve.
    // Leaves the iterate buffers modified: call after get_solution (solve()
    // re-initialises everything).
    bo_status time_spmv(int reps, double* csr_ms, double* csc_ms) override {
        return with_live_state([&] { return time_pair(plan_rows, plan_cols, reps < 1 ? 1 : reps, csr_ms, csc_ms); });
    }

    // Each iteration kernel's own duration (events bound to its dispatch,
    // launch_ev), passes alternating as in a solve: CSC gather, CSC
    // reduction, CSR gather, CSR reduction (a fused pass: its kernel in the
    // gather entry, 0 for the reduction).
    bo_status time_kernels(int reps, double* ms4) override {
        return with_live_state([&] { return time_kernels_live(reps, ms4); });
    }
    bo_status time_kernels_live(int reps, double* ms4) {
        reps = reps < 1 ? 1 : reps;
        std::vector<hipEvent_t> ev(8 * static_cast<size_t>(reps));
        for (auto& e : ev) BO_TRY(hipEventCreate(&e));
        launch_primal(plan_cols, nullptr, 0);
        launch_dual(plan_rows, nullptr, 0);
        for (int r = 0; r < reps; ++r) {
            launch_primal(plan_cols, nullptr, r, nullptr, &ev[8 * r]);
            launch_dual(plan_rows, nullptr, r, nullptr, &ev[8 * r + 4]);
        }
        BO_TRY(hipStreamSynchronize(stream));
        BO_TRY(hipGetLastError());
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const bool two[2] = {plan_cols.tiled > 0, plan_rows.tiled > 0};
        for (int r = 0; r < reps; ++r)
            for (int q = 0; q < 4; ++q) {
                if (q % 2 == 1 && !two[q / 2]) continue;  // a fused pass: no reduction kernel
                float a = 0.f;
                BO_TRY(hipEventElapsedTime(&a, ev[8 * r + 2 * q], ev[8 * r + 2 * q + 1]));
                acc[q] += a;
            }
        for (auto& e : ev) hipEventDestroy(e);
        for (int q = 0; q < 4; ++q) ms4[q] = acc[q] / reps;
        return BO_SUCCESS;
    }

    // Schedule of a pass and the algorithmic bytes of its kernels: what each
    // launch must move at least -- matrix entries in the stored format, the
    // gathered vector once, pointers or slice descriptors, partials, and the
    // epilogue's per-segment reads and writes.
    bo_status pass_info(int pass, bo_pdlp_pass_info* out) override {
        const PassPlan& p = pass == 0 ? plan_rows : plan_cols;
        const int nseg = pass == 0 ? m : n, ng = pass == 0 ? n : m;
        const double tb = sizeof(T), tm = sizeof(TM);  // iterate and matrix value bytes
        const double epi = pass == 0 ? 49.0 + 3.0 * tb : 24.0 + 4.0 * tb;
        std::memset(out, 0, sizeof(*out));
        out->schedule = plan_code(p);
        if (p.tiled > 0) {
            const TilePass& t = tp[pass];
            out->kernels = 2;
            out->blocks = t.B;
            out->block_entries = t.W;
            out->tiles = t.ntiles;
            out->slices = t.nslices;
            out->pairs = t.pairs;
            out->padded = static_cast<long long>(t.padded);
            out->gather_bytes = (tm + 2.0) * nnz + tb * ng + 4.0 * t.nslices +
                                static_cast<double>(sizeof(TileDesc)) * t.ntiles + 8.0 * t.pairs;
            out->reduce_bytes = 4.0 * (nseg + 1) + 12.0 * t.pairs + epi * nseg;
            int cover = 1;
            while (cover < t.max_pairs && cover < 16) cover *= 2;
            out->reduce_loads = t.reduce_u > 0 ? t.reduce_u : cover;
        } else {
            out->kernels = 1;
            out->gather_bytes = (tm + 4.0) * nnz + 4.0 * (nseg + 1) + tb * ng + epi * nseg;
        }
        return BO_SUCCESS;
    }
};

}  // namespace
}  // namespace bo

struct bo_pdlp {
    std::unique_ptr<bo::EngineBase> eng;
};

using namespace bo;

extern "C" {

void bo_pdlp_default_params(bo_pdlp_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->max_iterations = 100000;
    p->eps_rel_kkt = 1e-6;
    p->eps_objective = 0.0;
    p->ruiz_iterations = 10;
    p->pock_chambolle = 1;
    p->check_every = 0;
    p->lanes_rows = 0;
    p->lanes_cols = 0;
    p->fp64_iterates = 0;
    p->use_graph = 1;
    p->time_limit_s = 0.0;
    p->verbose = 0;
    p->hot_columns = -1;
    p->stall_iterations = 16384;
    p->fp64_polish = 1;
    p->compact_tail = 1;
    p->single_workgroup = -1;
}

bo_status bo_pdlp_create(const bo_lp* lp, const bo_pdlp_params* params, bo_pdlp** out) {
    return host_guard([&]() -> bo_status {
        if (!out) return BO_INVALID_ARGUMENT;
        *out = nullptr;
        if (!valid_csr_lp(lp)) return BO_INVALID_ARGUMENT;
        if (lp->nnz >= (1LL << 31) - 4096) return BO_INVALID_ARGUMENT;  // int32 offsets with unrolled strides
        bo_status s = ensure_device_public();
        if (s != BO_SUCCESS) return s;
        bo_pdlp_params p;
        if (params) p = *params;
        else bo_pdlp_default_params(&p);
        auto h = std::make_unique<bo_pdlp>();
        bo_status st;
        if (p.fp64_iterates == 2) {  // FP64 iterates over FP32 matrix values
            auto e = std::make_unique<Engine<double, float>>();
            st = e->create(lp, &p);
            h->eng = std::move(e);
        } else if (p.fp64_iterates) {
            auto e = std::make_unique<Engine<double>>();
            st = e->create(lp, &p);
            h->eng = std::move(e);
        } else {
            auto e = std::make_unique<Engine<float>>();
            st = e->create(lp, &p);
            h->eng = std::move(e);
        }
        if (st != BO_SUCCESS) return st;
        *out = h.release();
        return BO_SUCCESS;
    });
}

bo_status bo_pdlp_set_warm_start(bo_pdlp* h, const double* x0, const double* y0) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        return h->eng->set_warm_start(x0, y0);
    });
}

bo_status bo_pdlp_solve(bo_pdlp* h, bo_pdlp_stats* stats) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        return h->eng->solve(stats);
    });
}

bo_status bo_pdlp_get_solution(bo_pdlp* h, double* x, double* y) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        return h->eng->get_solution(x, y);
    });
}

void bo_pdlp_destroy(bo_pdlp* h) { delete h; }

bo_status bo_pdlp_time_kernels(bo_pdlp* h, int reps, double* ms4) {
    return host_guard([&]() -> bo_status {
        if (!h || !h->eng || !ms4) return BO_INVALID_ARGUMENT;
        return h->eng->time_kernels(reps, ms4);
    });
}

bo_status bo_pdlp_get_pass_info(bo_pdlp* h, int pass, bo_pdlp_pass_info* out) {
    if (!h || !h->eng || !out || pass < 0 || pass > 1) return BO_INVALID_ARGUMENT;
    return h->eng->pass_info(pass, out);
}

bo_status bo_pdlp_time_spmv(bo_pdlp* h, int reps, double* csr_ms, double* csc_ms) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        return h->eng->time_spmv(reps, csr_ms, csc_ms);
    });
}

}  // extern "C"
