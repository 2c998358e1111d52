// pdlp.hip -- restarted PDHG (PDLP-style) engine for one LP on one MI355X.
//
// Replaces the loop of CPUFirstOrderSolver::solve (src/first_order_solver.cpp:158-247)
// when the caller asks for a rel-KKT target.  Algorithm (from the PDLP papers,
// Applegate et al. 2021/2023, not from the reference, which has none of it):
//   * Ruiz equilibration (inf-norm, 10 sweeps) + Pock-Chambolle (alpha = 1),
//     all on the device from the resident FP64 matrix;
//   * constant step eta = 0.95 / ||A_hat||_2 (power iteration on the device),
//     tau = eta / omega, sigma = eta * omega;
//   * iteration (reference sign convention: min c'x, Ax <= b / = b, y_I >= 0):
//         x+ = proj_[l,u](x - tau (c + A'y))            CSC pass, fused epilogue
//         y+ = proj_Y(y + sigma (2 A x+ - A x - b))     CSR pass, fused epilogue
//     two launches per iteration; A x is carried so the extrapolation needs no
//     third product;
//   * every `check_every` iterations an FP64 check on the *unscaled* problem
//     for the current and the average iterate (rel-KKT, weighted KKT), then a
//     one-workgroup decision kernel applies the KKT restart rule
//     (beta 0.2 / 0.8 / 0.36), the primal-weight update (theta 0.5) and the
//     stopping test -- no host round trip inside an interval, which is
//     captured once as a hipGraph and replayed.
// Iterates and matrix values are FP32 (or FP64 for the error-comparison run);
// all accumulations, checks, averages and the returned solution are FP64.
// Rows and columns are stored sorted by length so the length-binned segment
// kernels (segments.hpp) cover a whole product in one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "segments.hpp"

namespace bo {

bo_status ensure_device_public();

namespace {

// ---------------------------------------------------------------- state
struct PdlpState {
    double eta, omega, tau, sigma;
    double b_norm, c_norm, eps;
    double kkt_last_restart, kkt_prev_candidate;
    double rel_kkt[2], rel_primal[2], rel_dual[2], rel_gap[2], pobj[2], dobj[2], kkt_w[2];
    double sigma_max, pw_u2, pw_w2;
    long long iter, iter_last_restart, max_iter;
    int done, status, restart, restart_to_avg, best, restarts, checks, pad;
};

__device__ inline double avg_count(const PdlpState* st) {
    return static_cast<double>(st->iter - st->iter_last_restart);
}

template <class T>
__device__ inline T clampT(T v, T lo, T hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// ---------------------------------------------------------------- epilogues
// y = A v (optionally scaled by seg_scale), partial sum of squares of y.
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

// Primal step on column j (CSC pass gathers y).
template <class T>
struct EpiPrimal {
    using TXv = T;
    static constexpr int NV = 1;
    static constexpr int Q = 0;
    const T* y;
    const PdlpState* st;
    const T* c;
    const T* lb;
    const T* ub;
    T* x;
    double* xsum;
    __device__ const T* vec(int) const { return y; }
    __device__ void operator()(int j, const double* acc, double*) const {
        const double tau = st->tau;
        const double xo = static_cast<double>(x[j]);
        const double xn = xo - tau * (static_cast<double>(c[j]) + acc[0]);
        const T xt = clampT(static_cast<T>(xn), lb[j], ub[j]);
        x[j] = xt;
        xsum[j] += static_cast<double>(xt);
    }
};

// Dual step on row i (CSR pass gathers the new x); carries A x.
template <class T>
struct EpiDual {
    using TXv = T;
    static constexpr int NV = 1;
    static constexpr int Q = 0;
    const T* x;
    const PdlpState* st;
    const T* b;
    const unsigned char* eq;
    T* ax;
    T* y;
    double* ysum;
    __device__ const T* vec(int) const { return x; }
    __device__ void operator()(int i, const double* acc, double*) const {
        const double sig = st->sigma;
        const double axo = static_cast<double>(ax[i]);
        double yn = static_cast<double>(y[i]) + sig * (2.0 * acc[0] - axo - static_cast<double>(b[i]));
        if (!eq[i] && yn < 0.0) yn = 0.0;
        const T yt = static_cast<T>(yn);
        y[i] = yt;
        ax[i] = static_cast<T>(acc[0]);
        ysum[i] += static_cast<double>(yt);
    }
};

// FP64 check, row side: gathers unscaled x (current, average).
// partials: [0,1] ||viol||^2 (cur, avg)  [2,3] scaled  [4,5] b'y (cur, avg)
template <class T>
struct EpiCheckRows {
    using TXv = double;
    static constexpr int NV = 2;
    static constexpr int Q = 6;
    const double* xo_cur;
    const double* xo_avg;
    const PdlpState* st;
    const double* b;
    const unsigned char* eq;
    const double* rs;
    const T* y;
    const double* ysum;
    double* yo_cur;
    double* yo_avg;
    T* ax_avg;  // scaled A x_avg, used when restarting to the average
    __device__ const double* vec(int v) const { return v == 0 ? xo_cur : xo_avg; }
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
        ax_avg[i] = static_cast<T>(ri * acc[1]);
        const double cnt = avg_count(st);
        const double yc = ri * static_cast<double>(y[i]);
        const double ya = cnt > 0.0 ? ri * (ysum[i] / cnt) : yc;
        yo_cur[i] = yc;
        yo_avg[i] = ya;
        part[4] += bi * yc;
        part[5] += bi * ya;
    }
};

// FP64 check, column side: gathers unscaled y (current, average).
// partials: [0,1] ||dual res||^2  [2,3] scaled  [4,5] sum l*plam^+ + u*plam^-
struct EpiCheckCols {
    using TXv = double;
    static constexpr int NV = 2;
    static constexpr int Q = 6;
    const double* yo_cur;
    const double* yo_avg;
    const double* c;
    const double* lb;
    const double* ub;
    const double* cs;
    __device__ const double* vec(int v) const { return v == 0 ? yo_cur : yo_avg; }
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

// Scaled vectors: c_hat = s c, l_hat = l / s, u_hat = u / s  (columns);
// b_hat = r b (rows).  Also the initial point.
template <class T>
__global__ void k_scale_cols(int n, const double* s, const double* c, const double* lb, const double* ub,
                             const double* wx, int has_warm, T* c_hat, T* l_hat, T* u_hat, T* x,
                             T* x_last, double* xsum) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double sj = s[j];
    c_hat[j] = static_cast<T>(c[j] * sj);
    const T lo = static_cast<T>(lb[j] / sj), hi = static_cast<T>(ub[j] / sj);
    l_hat[j] = lo;
    u_hat[j] = hi;
    const T x0 = clampT(static_cast<T>(has_warm ? wx[j] / sj : 0.0), lo, hi);
    x[j] = x0;
    x_last[j] = x0;
    xsum[j] = 0.0;
}

template <class T>
__global__ void k_scale_rows(int m, const double* r, const double* b, const unsigned char* eq,
                             const double* wy, int has_warm, T* b_hat, T* y, T* y_last, double* ysum) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double ri = r[i];
    b_hat[i] = static_cast<T>(b[i] * ri);
    double y0 = has_warm ? wy[i] / ri : 0.0;
    if (!eq[i] && y0 < 0.0) y0 = 0.0;
    y[i] = static_cast<T>(y0);
    y_last[i] = static_cast<T>(y0);
    ysum[i] = 0.0;
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
    const double smax = sqrt(*pw_u2);
    st->sigma_max = smax;
    st->eta = smax > 0.0 ? 0.95 / smax : 1.0;
    const double cn = sqrt(*c2), bn = sqrt(*b2);
    st->omega = (cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0;
    st->tau = st->eta / st->omega;
    st->sigma = st->eta * st->omega;
    st->iter = 0;
    st->iter_last_restart = 0;
    st->done = 0;
    st->status = 1;
    st->restart = 0;
    st->restart_to_avg = 0;
    st->best = 0;
    st->restarts = 0;
    st->checks = 0;
    st->kkt_last_restart = 0.0;
    st->kkt_prev_candidate = INFINITY;
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

// Unscaled x for the check: xo = clamp(s x_hat, l, u) (current and average);
// partials [0,1] = c'xo.
template <class T>
__global__ void k_check_x(int n, const PdlpState* st, const T* x, const double* xsum, const double* s,
                          const double* c, const double* lb, const double* ub, double* xo_cur,
                          double* xo_avg, double* partials) {
    if (st->done) return;
    __shared__ double sh[kBlock / kWave];
    const double cnt = avg_count(st);
    double p0 = 0.0, p1 = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double sj = s[j], l = lb[j], u = ub[j];
        double xc = sj * static_cast<double>(x[j]);
        xc = xc < l ? l : (xc > u ? u : xc);
        double xa = cnt > 0.0 ? sj * (xsum[j] / cnt) : xc;
        xa = xa < l ? l : (xa > u ? u : xa);
        xo_cur[j] = xc;
        xo_avg[j] = xa;
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

struct CheckDims {
    int nb_x, nb_rows, nb_cols;
};

// One workgroup: reduce the check partials in a fixed order, then the
// stopping test, the KKT restart rule and the bookkeeping.
__global__ void k_decide(PdlpState* st, CheckDims d, const double* px, const double* pr, const double* pc) {
    if (st->done) return;
    __shared__ double sh[kBlock / kWave];
    double q[14];
    for (int k = 0; k < 2; ++k) {
        double a = 0.0;
        for (int b = threadIdx.x; b < d.nb_x; b += blockDim.x) a += px[b * 2 + k];
        q[k] = block_sum(a, sh);
    }
    for (int k = 0; k < 6; ++k) {
        double a = 0.0;
        for (int b = threadIdx.x; b < d.nb_rows; b += blockDim.x) a += pr[static_cast<size_t>(b) * 6 + k];
        q[2 + k] = block_sum(a, sh);
    }
    for (int k = 0; k < 6; ++k) {
        double a = 0.0;
        for (int b = threadIdx.x; b < d.nb_cols; b += blockDim.x) a += pc[static_cast<size_t>(b) * 6 + k];
        q[8 + k] = block_sum(a, sh);
    }
    if (threadIdx.x != 0) return;
    // q: [0,1] c'x  [2,3] ||pviol||^2 [4,5] scaled [6,7] b'y  [8,9] ||dres||^2 [10,11] scaled [12,13] dterm
    const double w = st->omega;
    bool bad = false;
    for (int p = 0; p < 2; ++p) {
        const double pobj = q[0 + p];
        const double dobj = -q[6 + p] + q[12 + p];
        const double rp = sqrt(q[2 + p]), rd = sqrt(q[8 + p]);
        st->pobj[p] = pobj;
        st->dobj[p] = dobj;
        st->rel_primal[p] = rp / (1.0 + st->b_norm);
        st->rel_dual[p] = rd / (1.0 + st->c_norm);
        st->rel_gap[p] = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
        double k = st->rel_primal[p];
        if (st->rel_dual[p] > k) k = st->rel_dual[p];
        if (st->rel_gap[p] > k) k = st->rel_gap[p];
        st->rel_kkt[p] = k;
        const double g = pobj - dobj;
        st->kkt_w[p] = sqrt(w * q[4 + p] + q[10 + p] / w + g * g);
        if (!isfinite(k)) bad = true;
    }
    st->checks += 1;
    const int best = st->rel_kkt[1] < st->rel_kkt[0] ? 1 : 0;
    st->best = best;
    if (bad) {
        st->done = 1;
        st->status = 2;
        return;
    }
    if (st->rel_kkt[best] <= st->eps) {
        st->done = 1;
        st->status = 0;
        return;
    }
    if (st->iter >= st->max_iter) {
        st->done = 1;
        st->status = 1;
        return;
    }
    if (st->iter == 0) {  // initial point
        st->kkt_last_restart = st->kkt_w[0];
        st->kkt_prev_candidate = st->kkt_w[0];
        return;
    }
    const int cand = st->kkt_w[1] < st->kkt_w[0] ? 1 : 0;
    const double kc = st->kkt_w[cand];
    const long long since = st->iter - st->iter_last_restart;
    const bool restart = (kc <= 0.2 * st->kkt_last_restart) ||
                         (kc <= 0.8 * st->kkt_last_restart && kc > st->kkt_prev_candidate) ||
                         (static_cast<double>(since) >= 0.36 * static_cast<double>(st->iter));
    st->kkt_prev_candidate = kc;
    if (restart) {
        st->restart = 1;
        st->restart_to_avg = cand;
        st->restarts += 1;
        st->kkt_last_restart = kc;
    }
}

// Restart: move to the candidate, measure the move since the last restart,
// reset the averages.  partials [0] = ||dx||^2 (n side) or ||dy||^2 (m side).
template <class T>
__global__ void k_restart_x(int n, const PdlpState* st, T* x, T* x_last, double* xsum, double* partials) {
    if (st->done || !st->restart) return;
    __shared__ double sh[kBlock / kWave];
    const double cnt = avg_count(st);
    double acc = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        T xn = x[j];
        if (st->restart_to_avg && cnt > 0.0) xn = static_cast<T>(xsum[j] / cnt);
        x[j] = xn;
        const double d = static_cast<double>(xn) - static_cast<double>(x_last[j]);
        acc += d * d;
        x_last[j] = xn;
        xsum[j] = 0.0;
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

template <class T>
__global__ void k_restart_y(int m, const PdlpState* st, T* y, T* y_last, double* ysum, T* ax,
                            const T* ax_avg, double* partials) {
    if (st->done || !st->restart) return;
    __shared__ double sh[kBlock / kWave];
    const double cnt = avg_count(st);
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        T yn = y[i];
        if (st->restart_to_avg && cnt > 0.0) {
            yn = static_cast<T>(ysum[i] / cnt);
            ax[i] = ax_avg[i];
        }
        y[i] = yn;
        const double d = static_cast<double>(yn) - static_cast<double>(y_last[i]);
        acc += d * d;
        y_last[i] = yn;
        ysum[i] = 0.0;
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
    const double dx = sqrt(dx2), dy = sqrt(dy2);
    if (dx > 1e-10 && dy > 1e-10) {
        st->omega = exp(0.5 * log(dy / dx) + 0.5 * log(st->omega));
        st->tau = st->eta / st->omega;
        st->sigma = st->eta * st->omega;
    }
    st->iter_last_restart = st->iter;
    st->restart = 0;
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

// Bin table for segments sorted by ascending length.
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

struct EngineBase {
    virtual ~EngineBase() = default;
    virtual bo_status set_warm_start(const double* x0, const double* y0) = 0;
    virtual bo_status solve(bo_pdlp_stats* stats) = 0;
    virtual bo_status get_solution(double* x, double* y) = 0;
    virtual bo_status time_spmv(int reps, double* csr_ms, double* csc_ms) = 0;
};

template <class T>
struct Engine final : EngineBase {
    int m = 0, n = 0;
    long long nnz = 0;
    bo_pdlp_params prm{};
    hipStream_t stream = nullptr;
    DevMem mem;
    std::vector<int> row_of;  // permuted position -> original row
    std::vector<int> col_of;  // permuted position -> original column
    std::vector<int> row_len_sorted, col_len_sorted;
    // device: FP64 original (permuted) matrix in CSR and CSC
    int *rp = nullptr, *ci = nullptr, *cp = nullptr, *ri = nullptr;
    double *rv64 = nullptr, *cv64 = nullptr;
    T *rv = nullptr, *cv = nullptr;  // scaled working values
    // FP64 original vectors (permuted)
    double *b64 = nullptr, *c64 = nullptr, *lb64 = nullptr, *ub64 = nullptr;
    unsigned char* eq = nullptr;
    double *rs = nullptr, *cs = nullptr, *rnorm = nullptr, *cnorm = nullptr;
    // working vectors
    T *b_hat = nullptr, *c_hat = nullptr, *l_hat = nullptr, *u_hat = nullptr;
    T *x = nullptr, *y = nullptr, *ax = nullptr, *ax_avg = nullptr, *x_last = nullptr, *y_last = nullptr;
    T *pw_v = nullptr, *pw_u = nullptr;
    double *xsum = nullptr, *ysum = nullptr, *xo_cur = nullptr, *xo_avg = nullptr, *yo_cur = nullptr,
           *yo_avg = nullptr;
    double *wx = nullptr, *wy = nullptr;
    int has_warm = 0;
    double *part_x = nullptr, *part_r = nullptr, *part_c = nullptr, *part_dx = nullptr, *part_dy = nullptr;
    double* scal = nullptr;  // small scalars: [0] ||u||^2 [1] ||w||^2 [2] ||c_hat||^2 [3] ||b_hat||^2
    PdlpState* st = nullptr;
    PdlpState* st_host = nullptr;  // pinned, read back after each interval
    PdlpState* st_init = nullptr;  // pinned, staged for the upload in prepare()
    BinTable bt_rows{}, bt_cols{};
    int nb_x = 1, nb_y = 1;
    double b_norm = 0.0, c_norm = 0.0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_k = 0;

    ~Engine() override {
        if (graph_exec) hipGraphExecDestroy(graph_exec);
        if (graph) hipGraphDestroy(graph);
        if (st_host) hipHostFree(st_host);
        if (st_init) hipHostFree(st_init);
        if (stream) hipStreamDestroy(stream);
    }

    bo_status create(const bo_lp* lp, const bo_pdlp_params* p) {
        prm = *p;
        m = lp->m;
        n = lp->n;
        nnz = lp->nnz;
        BO_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_host), sizeof(PdlpState)));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_init), sizeof(PdlpState)));
        // ---- permutations: rows and columns by ascending length (stable)
        std::vector<int> rlen(m), clen(n, 0);
        for (int i = 0; i < m; ++i) rlen[i] = lp->row_ptr[i + 1] - lp->row_ptr[i];
        for (long long k = 0; k < nnz; ++k) clen[lp->col_idx[k]]++;
        row_of.resize(m);
        col_of.resize(n);
        std::iota(row_of.begin(), row_of.end(), 0);
        std::iota(col_of.begin(), col_of.end(), 0);
        std::stable_sort(row_of.begin(), row_of.end(), [&](int a, int b) { return rlen[a] < rlen[b]; });
        std::stable_sort(col_of.begin(), col_of.end(), [&](int a, int b) { return clen[a] < clen[b]; });
        std::vector<int> col_pos(n);
        for (int p2 = 0; p2 < n; ++p2) col_pos[col_of[p2]] = p2;
        row_len_sorted.resize(m);
        col_len_sorted.resize(n);
        for (int p2 = 0; p2 < m; ++p2) row_len_sorted[p2] = rlen[row_of[p2]];
        for (int p2 = 0; p2 < n; ++p2) col_len_sorted[p2] = clen[col_of[p2]];
        // permuted CSR
        std::vector<int> h_rp(m + 1), h_ci(nnz);
        std::vector<double> h_rv(nnz);
        h_rp[0] = 0;
        for (int p2 = 0; p2 < m; ++p2) {
            const int i = row_of[p2];
            int o = h_rp[p2];
            for (int k = lp->row_ptr[i]; k < lp->row_ptr[i + 1]; ++k, ++o) {
                h_ci[o] = col_pos[lp->col_idx[k]];
                h_rv[o] = lp->val[k];
            }
            h_rp[p2 + 1] = o;
        }
        // permuted CSC (transpose of the permuted CSR)
        std::vector<int> h_cp(n + 1, 0), h_ri(nnz);
        std::vector<double> h_cv(nnz);
        for (long long k = 0; k < nnz; ++k) h_cp[h_ci[k] + 1]++;
        for (int j = 0; j < n; ++j) h_cp[j + 1] += h_cp[j];
        {
            std::vector<int> pos(h_cp.begin(), h_cp.end() - 1);
            for (int p2 = 0; p2 < m; ++p2)
                for (int k = h_rp[p2]; k < h_rp[p2 + 1]; ++k) {
                    const int q = pos[h_ci[k]]++;
                    h_ri[q] = p2;
                    h_cv[q] = h_rv[k];
                }
        }
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
        BO_TRY(mem.alloc(&y, m));
        BO_TRY(mem.alloc(&ax, m));
        BO_TRY(mem.alloc(&ax_avg, m));
        BO_TRY(mem.alloc(&y_last, m));
        BO_TRY(mem.alloc(&pw_u, m));
        BO_TRY(mem.alloc(&ysum, m));
        BO_TRY(mem.alloc(&yo_cur, m));
        BO_TRY(mem.alloc(&yo_avg, m));
        BO_TRY(mem.alloc(&wy, m));
        BO_TRY(mem.alloc(&c64, n));
        BO_TRY(mem.alloc(&lb64, n));
        BO_TRY(mem.alloc(&ub64, n));
        BO_TRY(mem.alloc(&cs, n));
        BO_TRY(mem.alloc(&cnorm, n));
        BO_TRY(mem.alloc(&c_hat, n));
        BO_TRY(mem.alloc(&l_hat, n));
        BO_TRY(mem.alloc(&u_hat, n));
        BO_TRY(mem.alloc(&x, n));
        BO_TRY(mem.alloc(&x_last, n));
        BO_TRY(mem.alloc(&pw_v, n));
        BO_TRY(mem.alloc(&xsum, n));
        BO_TRY(mem.alloc(&xo_cur, n));
        BO_TRY(mem.alloc(&xo_avg, n));
        BO_TRY(mem.alloc(&wx, n));
        BO_TRY(mem.alloc(&scal, 8));
        BO_TRY(mem.alloc(&st, 1));
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
        // ---- launch geometry
        nb_x = static_cast<int>(std::min<unsigned>(grid_for(n), 1024));
        nb_y = static_cast<int>(std::min<unsigned>(grid_for(m), 1024));
        if (prm.lanes_rows < 0) prm.lanes_rows = 0;
        if (prm.lanes_cols < 0) prm.lanes_cols = 0;
        const int pr_rows = prm.lanes_rows > 0 ? prm.lanes_rows : 4;
        const int pr_cols = prm.lanes_cols > 0 ? prm.lanes_cols : 4;
        bt_rows = make_bins(row_len_sorted, pr_rows);
        bt_cols = make_bins(col_len_sorted, pr_cols);
        BO_TRY(mem.alloc(&part_x, static_cast<size_t>(nb_x) * 2));
        BO_TRY(mem.alloc(&part_dx, nb_x));
        BO_TRY(mem.alloc(&part_dy, nb_y));
        // partial buffers sized for the largest bin tables the tuner may try
        size_t max_r = 0, max_c = 0;
        for (int kp : {1, 2, 4, 8, 16, pr_rows, pr_cols}) {
            max_r = std::max<size_t>(max_r, make_bins(row_len_sorted, kp).total_blocks);
            max_c = std::max<size_t>(max_c, make_bins(col_len_sorted, kp).total_blocks);
        }
        BO_TRY(mem.alloc(&part_r, max_r * 6));
        BO_TRY(mem.alloc(&part_c, max_c * 6));
        if (prm.lanes_rows <= 0 || prm.lanes_cols <= 0) {
            bo_status s2 = autotune();
            if (s2 != BO_SUCCESS) return s2;
        }
        return BO_SUCCESS;
    }

    // ---- launch helpers
    template <int MODE, class TV, class Epi>
    void launch_rows(const BinTable& bt, const TV* vals, const Epi& epi, double* partials, const int* done) {
        hipLaunchKernelGGL((k_segments<MODE, TV, typename Epi::TXv, Epi>), dim3(bt.total_blocks), dim3(kBlock), 0,
                           stream, bt, rp, ci, vals, epi, partials, done);
    }
    template <int MODE, class TV, class Epi>
    void launch_cols(const BinTable& bt, const TV* vals, const Epi& epi, double* partials, const int* done) {
        hipLaunchKernelGGL((k_segments<MODE, TV, typename Epi::TXv, Epi>), dim3(bt.total_blocks), dim3(kBlock), 0,
                           stream, bt, cp, ri, vals, epi, partials, done);
    }

    const int* done_flag() const { return &st->done; }

    void launch_primal(const BinTable& bt, const int* done) {
        EpiPrimal<T> e{y, st, c_hat, l_hat, u_hat, x, xsum};
        launch_cols<kSumProd>(bt, cv, e, nullptr, done);
    }
    void launch_dual(const BinTable& bt, const int* done) {
        EpiDual<T> e{x, st, b_hat, eq, ax, y, ysum};
        launch_rows<kSumProd>(bt, rv, e, nullptr, done);
    }

    void launch_check() {
        hipLaunchKernelGGL(k_check_x<T>, dim3(nb_x), dim3(kBlock), 0, stream, n, st, x, xsum, cs, c64, lb64,
                           ub64, xo_cur, xo_avg, part_x);
        EpiCheckRows<T> er{xo_cur, xo_avg, st, b64, eq, rs, y, ysum, yo_cur, yo_avg, ax_avg};
        launch_rows<kSumProd>(bt_rows, rv64, er, part_r, done_flag());
        EpiCheckCols ec{yo_cur, yo_avg, c64, lb64, ub64, cs};
        launch_cols<kSumProd>(bt_cols, cv64, ec, part_c, done_flag());
        CheckDims d{nb_x, bt_rows.total_blocks, bt_cols.total_blocks};
        hipLaunchKernelGGL(k_decide, dim3(1), dim3(kBlock), 0, stream, st, d, part_x, part_r, part_c);
        hipLaunchKernelGGL(k_restart_x<T>, dim3(nb_x), dim3(kBlock), 0, stream, n, st, x, x_last, xsum, part_dx);
        hipLaunchKernelGGL(k_restart_y<T>, dim3(nb_y), dim3(kBlock), 0, stream, m, st, y, y_last, ysum, ax,
                           ax_avg, part_dy);
        hipLaunchKernelGGL(k_restart_finish, dim3(1), dim3(kBlock), 0, stream, st, part_dx, nb_x, part_dy, nb_y);
    }

    // K iterations followed by the check.
    void launch_interval(int k) {
        for (int it = 0; it < k; ++it) {
            launch_primal(bt_cols, done_flag());
            launch_dual(bt_rows, done_flag());
        }
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, stream, st, k);
        launch_check();
    }

    bo_status autotune() {
        // time the two iteration kernels at candidate nonzeros-per-lane targets
        const int cands[] = {1, 2, 4, 8, 16};
        bo_status s0 = prepare();
        if (s0 != BO_SUCCESS) return s0;
        hipEvent_t e0, e1;
        BO_TRY(hipEventCreate(&e0));
        BO_TRY(hipEventCreate(&e1));
        auto time_it = [&](auto&& fn) {
            fn();
            hipEventRecord(e0, stream);
            for (int r = 0; r < 5; ++r) fn();
            hipEventRecord(e1, stream);
            hipEventSynchronize(e1);
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            return ms;
        };
        float best_r = 1e30f, best_c = 1e30f;
        int kr = 4, kc = 4;
        for (int kp : cands) {
            if (prm.lanes_rows <= 0) {
                BinTable bt = make_bins(row_len_sorted, kp);
                float t = time_it([&] { launch_dual(bt, nullptr); });
                if (t < best_r) { best_r = t; kr = kp; }
            }
            if (prm.lanes_cols <= 0) {
                BinTable bt = make_bins(col_len_sorted, kp);
                float t = time_it([&] { launch_primal(bt, nullptr); });
                if (t < best_c) { best_c = t; kc = kp; }
            }
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        BO_TRY(hipGetLastError());
        if (prm.lanes_rows <= 0) { bt_rows = make_bins(row_len_sorted, kr); prm.lanes_rows = kr; }
        if (prm.lanes_cols <= 0) { bt_cols = make_bins(col_len_sorted, kc); prm.lanes_cols = kc; }
        return BO_SUCCESS;
    }

    bo_status set_warm_start(const double* x0, const double* y0) override {
        if (!x0 && !y0) {
            has_warm = 0;
            return BO_SUCCESS;
        }
        std::vector<double> hx(n, 0.0), hy(m, 0.0);
        if (x0)
            for (int p2 = 0; p2 < n; ++p2) hx[p2] = x0[col_of[p2]];
        if (y0)
            for (int p2 = 0; p2 < m; ++p2) hy[p2] = y0[row_of[p2]];
        if (n) BO_TRY(hipMemcpyAsync(wx, hx.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream));
        if (m) BO_TRY(hipMemcpyAsync(wy, hy.data(), sizeof(double) * m, hipMemcpyHostToDevice, stream));
        BO_TRY(hipStreamSynchronize(stream));
        has_warm = 1;
        return BO_SUCCESS;
    }

    // Scaling, step size, initial point, initial Ax: everything device-side.
    bo_status prepare() {
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
        EpiWrite<T> wr{rs, cs, rv};
        launch_rows<kWrite>(br, rv64, wr, nullptr, nullptr);
        EpiWrite<T> wc{cs, rs, cv};
        launch_cols<kWrite>(bc, cv64, wc, nullptr, nullptr);
        hipLaunchKernelGGL(k_scale_cols<T>, dim3(gn), dim3(kBlock), 0, stream, n, cs, c64, lb64, ub64, wx,
                           has_warm, c_hat, l_hat, u_hat, x, x_last, xsum);
        hipLaunchKernelGGL(k_scale_rows<T>, dim3(gm), dim3(kBlock), 0, stream, m, rs, b64, eq, wy, has_warm,
                           b_hat, y, y_last, ysum);
        // power iteration for ||A_hat||_2
        hipLaunchKernelGGL(k_pw_init<T>, dim3(gn), dim3(kBlock), 0, stream, pw_v, n);
        hipLaunchKernelGGL(k_sq_partial<T>, dim3(nb_x), dim3(kBlock), 0, stream, pw_v, n, part_dx);
        hipLaunchKernelGGL(k_sum_partials_to, dim3(1), dim3(kBlock), 0, stream, part_dx, nb_x, 1, 0, scal + 1);
        hipLaunchKernelGGL(k_pw_normalize<T>, dim3(gn), dim3(kBlock), 0, stream, pw_v, n, scal + 1);
        for (int it = 0; it < 40; ++it) {
            EpiStore<T, T> eu{pw_v, pw_u};
            launch_rows<kSumProd>(br, rv, eu, part_r, nullptr);
            hipLaunchKernelGGL(k_sum_partials_to, dim3(1), dim3(kBlock), 0, stream, part_r, br.total_blocks, 1, 0,
                               scal + 0);
            EpiStore<T, T> ew{pw_u, pw_v};
            launch_cols<kSumProd>(bc, cv, ew, part_c, nullptr);
            hipLaunchKernelGGL(k_sum_partials_to, dim3(1), dim3(kBlock), 0, stream, part_c, bc.total_blocks, 1, 0,
                               scal + 1);
            hipLaunchKernelGGL(k_pw_normalize<T>, dim3(gn), dim3(kBlock), 0, stream, pw_v, n, scal + 1);
        }
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
        st_init->max_iter = prm.max_iterations;
        BO_TRY(hipMemcpyAsync(st, st_init, sizeof(PdlpState), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, stream, st, scal + 0, scal + 2, scal + 3);
        // initial A x
        EpiStore<T, T> ea{x, ax};
        launch_rows<kSumProd>(br, rv, ea, part_r, nullptr);
        BO_TRY(hipGetLastError());
        return BO_SUCCESS;
    }

    bo_status solve(bo_pdlp_stats* out) override {
        auto t0 = std::chrono::high_resolution_clock::now();
        bo_status s0 = prepare();
        if (s0 != BO_SUCCESS) return s0;
        launch_check();  // iteration 0: KKT of the starting point
        BO_TRY(hipMemcpyAsync(st_host, st, sizeof(PdlpState), hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        auto t1 = std::chrono::high_resolution_clock::now();
        const int K = prm.check_every > 0 ? prm.check_every : 64;
        long long iters = 0;
        while (!st_host->done) {
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
            BO_TRY(hipMemcpyAsync(st_host, st, sizeof(PdlpState), hipMemcpyDeviceToHost, stream));
            BO_TRY(hipStreamSynchronize(stream));
            if (prm.time_limit_s > 0.0) {
                double el = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
                if (el > prm.time_limit_s && !st_host->done) {
                    st_host->status = 3;
                    break;
                }
            }
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
        }
        return BO_SUCCESS;
    }

    bo_status get_solution(double* hx, double* hy) override {
        const int b = st_host->best;
        std::vector<double> tx(n), ty(m);
        if (n) BO_TRY(hipMemcpyAsync(tx.data(), b ? xo_avg : xo_cur, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
        if (m) BO_TRY(hipMemcpyAsync(ty.data(), b ? yo_avg : yo_cur, sizeof(double) * m, hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        if (hx)
            for (int p2 = 0; p2 < n; ++p2) hx[col_of[p2]] = tx[p2];
        if (hy)
            for (int p2 = 0; p2 < m; ++p2) hy[row_of[p2]] = ty[p2];
        return BO_SUCCESS;
    }

    // Times the two fused iteration kernels (CSR pass + dual step, CSC pass +
    // primal step) on the handle's stream.  Leaves the iterate buffers
    // modified: call after get_solution (solve() re-initialises everything).
    bo_status time_spmv(int reps, double* csr_ms, double* csc_ms) override {
        if (reps < 1) reps = 1;
        hipEvent_t e0, e1;
        BO_TRY(hipEventCreate(&e0));
        BO_TRY(hipEventCreate(&e1));
        float ms = 0.f;
        launch_dual(bt_rows, nullptr);
        hipEventRecord(e0, stream);
        for (int r = 0; r < reps; ++r) launch_dual(bt_rows, nullptr);
        hipEventRecord(e1, stream);
        hipEventSynchronize(e1);
        hipEventElapsedTime(&ms, e0, e1);
        if (csr_ms) *csr_ms = ms / reps;
        launch_primal(bt_cols, nullptr);
        hipEventRecord(e0, stream);
        for (int r = 0; r < reps; ++r) launch_primal(bt_cols, nullptr);
        hipEventRecord(e1, stream);
        hipEventSynchronize(e1);
        hipEventElapsedTime(&ms, e0, e1);
        if (csc_ms) *csc_ms = ms / reps;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        BO_TRY(hipGetLastError());
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
    p->ruiz_iterations = 10;
    p->pock_chambolle = 1;
    p->check_every = 64;
    p->lanes_rows = 0;
    p->lanes_cols = 0;
    p->fp64_iterates = 0;
    p->use_graph = 1;
    p->time_limit_s = 0.0;
    p->verbose = 0;
}

bo_status bo_pdlp_create(const bo_lp* lp, const bo_pdlp_params* params, bo_pdlp** out) {
    if (!out) return BO_INVALID_ARGUMENT;
    *out = nullptr;
    bo_status s = ensure_device_public();
    if (s != BO_SUCCESS) return s;
    if (!lp || lp->m < 0 || lp->n < 0 || lp->nnz < 0) return BO_INVALID_ARGUMENT;
    if ((lp->m > 0 && (!lp->row_ptr || !lp->b || !lp->is_eq)) || (lp->n > 0 && (!lp->c || !lp->lb || !lp->ub)) ||
        (lp->nnz > 0 && (!lp->col_idx || !lp->val)))
        return BO_INVALID_ARGUMENT;
    if (lp->nnz >= (1LL << 31)) return BO_INVALID_ARGUMENT;
    if (lp->m > 0 && (lp->row_ptr[0] != 0 || lp->row_ptr[lp->m] != lp->nnz)) return BO_INVALID_ARGUMENT;
    for (int i = 0; i < lp->m; ++i)
        if (lp->row_ptr[i + 1] < lp->row_ptr[i]) return BO_INVALID_ARGUMENT;
    for (long long k = 0; k < lp->nnz; ++k)
        if (lp->col_idx[k] < 0 || lp->col_idx[k] >= lp->n) return BO_INVALID_ARGUMENT;
    bo_pdlp_params p;
    if (params) p = *params;
    else bo_pdlp_default_params(&p);
    auto h = std::make_unique<bo_pdlp>();
    bo_status st;
    if (p.fp64_iterates) {
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
}

bo_status bo_pdlp_set_warm_start(bo_pdlp* h, const double* x0, const double* y0) {
    if (!h) return BO_INVALID_ARGUMENT;
    return h->eng->set_warm_start(x0, y0);
}

bo_status bo_pdlp_solve(bo_pdlp* h, bo_pdlp_stats* stats) {
    if (!h) return BO_INVALID_ARGUMENT;
    return h->eng->solve(stats);
}

bo_status bo_pdlp_get_solution(bo_pdlp* h, double* x, double* y) {
    if (!h) return BO_INVALID_ARGUMENT;
    return h->eng->get_solution(x, y);
}

void bo_pdlp_destroy(bo_pdlp* h) { delete h; }

bo_status bo_pdlp_time_spmv(bo_pdlp* h, int reps, double* csr_ms, double* csc_ms) {
    if (!h) return BO_INVALID_ARGUMENT;
    return h->eng->time_spmv(reps, csr_ms, csc_ms);
}

}  // extern "C"
