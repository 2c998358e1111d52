// family.hip -- PDLP over a parametric family of LPs: one constraint matrix,
// one set of bounds and row senses, and up to 64 LPs that differ in their
// costs c_k and right-hand sides b_k, solved together.
//
// Layout: every per-LP vector is stored LP-minor -- entry (j, k) at
// j * 64 + k -- so one wavefront owns one row (or column) and its 64 lanes
// are the 64 LPs.  A product pass reads each matrix entry once, with
// wave-uniform (scalar) loads, and applies it to all LPs; the entries of the
// other side's vector that it gathers for that index form one coalesced
// 256-byte line (FP32), and every epilogue access (FP64 iterates, anchors,
// A x) is a contiguous 512-byte line.  The per-LP state machine (steps,
// restarts, checks, stall test) is the single-LP engine's (pdlp_core.hpp),
// one state per lane.
//
// The scaling (Ruiz, then Pock-Chambolle) and the power-iteration estimate of
// ||A_hat||_2 depend on the matrix alone: they are computed once, at create,
// on the host.  Each LP's primal weight starts from its own ||c_hat|| / ||b_hat||.
// The scaled matrix is kept in FP64: shared by all LPs, its 12 bytes per
// entry are a few percent of a pass, and its products carry no rounding of
// A (the iterates and gathered vectors are FP32).
//
// No counterpart exists in the reference, which solves one LP per call
// (src/first_order_solver.cpp:15-291).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"
#include "pdlp_core.hpp"

namespace bo {

bo_status ensure_device_public();

namespace {

constexpr int kL = 64;          // LP slots: one per lane
constexpr int kSegWaves = 4;    // segments (rows / columns) per 256-thread workgroup
constexpr int kSumBlocks = 512; // workgroups of the elementwise check kernels
constexpr int kPowerIters = 40;

// Per-lane step data for the current check interval: tau, sigma, the
// Halpern step count since the last restart, and whether the LP still runs.
struct LaneCoef {
    double tau[kL];
    double sigma[kL];
    double kbase[kL];
    int live[kL];
};

__device__ inline double halpern_at(const LaneCoef* cf, int lane, int t) {
    const double k = cf->kbase[lane] + static_cast<double>(t);
    return (k + 1.0) / (k + 2.0);
}

// ----------------------------------------------------------------- products
constexpr int kXcds = 8;    // L2 domains of the chip; workgroup b runs on XCD b % 8
constexpr int kUnroll = 8;  // matrix entries in flight per wavefront

// Launch grid of a product pass over nseg segments: a multiple of kXcds.
inline unsigned spmm_grid(int nseg) {
    const int blocks = (nseg + kSegWaves - 1) / kSegWaves;
    return static_cast<unsigned>((blocks + kXcds - 1) / kXcds * kXcds);
}

// One wavefront per segment s (row of the CSR or column of the CSC):
// acc[v] = sum_q val[q] * vec_v[idx[q] * 64 + lane] in FP64, then epi(s, lane, acc).
// Workgroups are renumbered so that each XCD takes one contiguous eighth of
// the segments: neighbouring rows (or columns) gather neighbouring entries,
// which then stay in that XCD's L2.  Lanes of finished LPs load nothing.
template <class TM, class Epi>
__global__ void __launch_bounds__(256)
kf_spmm(int nseg, const int* __restrict__ ptr, const int* __restrict__ idx, const TM* __restrict__ val, Epi epi) {
    const int per = static_cast<int>(gridDim.x) / kXcds;
    const int lb = (static_cast<int>(blockIdx.x) % kXcds) * per + static_cast<int>(blockIdx.x) / kXcds;
    const int s = __builtin_amdgcn_readfirstlane(lb * kSegWaves + static_cast<int>(threadIdx.x) / kWave);
    if (s >= nseg) return;
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    if (!epi.live(lane)) return;
    const int b = __builtin_amdgcn_readfirstlane(ptr[s]);
    const int e = __builtin_amdgcn_readfirstlane(ptr[s + 1]);
    using TV = typename Epi::TV;
    const TV* v0 = epi.vec(0) + lane;
    const TV* v1 = epi.vec(Epi::NV - 1) + lane;
    // the epilogue's own operands are fetched first, so their latency hides
    // behind the gathers
    const typename Epi::Pre pre = epi.load(s, lane);
    double a0 = 0.0, a1 = 0.0;
    // kUnroll entries per round trip; entries past the segment (the arrays
    // carry kUnroll entries of padding) get weight 0.  The sum runs in
    // entry order.
    for (int q = b; q < e; q += kUnroll) {
        int j[kUnroll];
        double w[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            j[u] = idx[q + u];
            const double a = static_cast<double>(val[q + u]);
            w[u] = q + u < e ? a : 0.0;
        }
        double x[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) x[u] = static_cast<double>(v0[static_cast<size_t>(j[u]) * kL]);
        if (Epi::NV == 2) {
            double z[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) z[u] = static_cast<double>(v1[static_cast<size_t>(j[u]) * kL]);
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) a1 += w[u] * z[u];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) a0 += w[u] * x[u];
    }
    const double acc[2] = {a0, a1};
    epi(s, lane, acc, pre);
}

// Column j of LP `lane`: primal step from A'y, Halpern update of x.
template <class T>
struct EpiFPrimal {
    using TV = T;
    static constexpr int NV = 1;
    const T* yg;
    const LaneCoef* cf;
    const T* c_hat;
    const T* l_hat;
    const T* u_hat;
    double* xk;
    const T* x0;       // Halpern anchor (working precision: it enters with weight 1 / (k + 2))
    T* xp;
    int t;
    struct Pre {
        double xo, x0, c;
        T lo, hi;
    };
    __device__ const T* vec(int) const { return yg; }
    __device__ bool live(int lane) const { return cf->live[lane] != 0; }
    __device__ Pre load(int j, int lane) const {
        const size_t e = static_cast<size_t>(j) * kL + lane;
        return Pre{xk[e], static_cast<double>(x0[e]), static_cast<double>(c_hat[e]), l_hat[j], u_hat[j]};
    }
    __device__ void operator()(int j, int lane, const double* acc, const Pre& p) const {
        const size_t e = static_cast<size_t>(j) * kL + lane;
        const T xt = clampT(static_cast<T>(primal_move(p.xo, cf->tau[lane], p.c, acc[0])), p.lo, p.hi);
        xp[e] = xt;
        xk[e] = halpern_x(halpern_at(cf, lane, t), static_cast<double>(xt), p.xo, p.x0);
    }
};

// Row i of LP `lane`: dual step from A x+, Halpern update of y and of A x.
template <class T>
struct EpiFDual {
    using TV = T;
    static constexpr int NV = 1;
    const T* xp;
    const LaneCoef* cf;
    const T* b_hat;
    const unsigned char* eq;
    double* ax;
    const T* ax0;      // anchors in working precision
    double* yk;
    const T* y0;
    T* yp;
    T* yg;
    int t;
    struct Pre {
        double axo, yo, b, y0, ax0;
        bool eq;
    };
    __device__ const T* vec(int) const { return xp; }
    __device__ bool live(int lane) const { return cf->live[lane] != 0; }
    __device__ Pre load(int i, int lane) const {
        const size_t e = static_cast<size_t>(i) * kL + lane;
        return Pre{ax[e], yk[e], static_cast<double>(b_hat[e]), static_cast<double>(y0[e]), static_cast<double>(ax0[e]),
                   eq[i] != 0};
    }
    __device__ void operator()(int i, int lane, const double* acc, const Pre& p) const {
        const size_t e = static_cast<size_t>(i) * kL + lane;
        double yn = dual_move(p.yo, cf->sigma[lane], acc[0], p.axo, p.b);
        if (!p.eq && yn < 0.0) yn = 0.0;
        const T yt = static_cast<T>(yn);
        yp[e] = yt;
        const double w = halpern_at(cf, lane, t);
        const double ynew = halpern_y(w, static_cast<double>(yt), p.yo, p.y0);
        yk[e] = ynew;
        yg[e] = static_cast<T>(ynew);
        ax[e] = halpern_ax(w, acc[0], p.axo, p.ax0);
    }
};

// out0 (and out1) = the product(s), for every lane (used at set-up).
template <class TV_>
struct EpiFStore {
    using TV = TV_;
    static constexpr int NV = 1;
    const TV* v;
    double* out0;
    TV* out1;  // may be null
    struct Pre {};
    __device__ const TV* vec(int) const { return v; }
    __device__ bool live(int) const { return true; }
    __device__ Pre load(int, int) const { return Pre{}; }
    __device__ void operator()(int s, int lane, const double* acc, const Pre&) const {
        const size_t e = static_cast<size_t>(s) * kL + lane;
        out0[e] = acc[0];
        if (out1) out1[e] = static_cast<TV>(acc[0]);
    }
};

// FP64 check products of two vectors (current point, candidate), live LPs.
struct EpiFCheck2 {
    using TV = double;
    static constexpr int NV = 2;
    const double* v0;
    const double* v1;
    const LaneCoef* cf;
    double* out0;
    double* out1;
    struct Pre {};
    __device__ const double* vec(int k) const { return k == 0 ? v0 : v1; }
    __device__ bool live(int lane) const { return cf->live[lane] != 0; }
    __device__ Pre load(int, int) const { return Pre{}; }
    __device__ void operator()(int s, int lane, const double* acc, const Pre&) const {
        const size_t e = static_cast<size_t>(s) * kL + lane;
        out0[e] = acc[0];
        out1[e] = acc[1];
    }
};

// -------------------------------------------------------- per-lane reductions
// Each workgroup (4 waves) strides over segments; lane sums of NQ quantities
// are reduced across its waves and stored as part[(block * NQ + q) * 64 + lane].
template <int NQ>
__device__ inline void store_lane_sums(const double* q, double* part) {
    __shared__ double sh[NQ][kSegWaves][kL];
    const int lane = static_cast<int>(threadIdx.x) % kWave, wave = static_cast<int>(threadIdx.x) / kWave;
#pragma unroll
    for (int k = 0; k < NQ; ++k) sh[k][wave][lane] = q[k];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kSegWaves; ++w) s += sh[k][w][lane];
            part[(static_cast<size_t>(blockIdx.x) * NQ + k) * kL + lane] = s;
        }
    }
}

// sum over the workgroups' partials, in workgroup order: out[q] for this lane
template <int NQ>
__device__ inline void sum_lane_parts(const double* part, int nblocks, int lane, double* out) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) out[k] = 0.0;
    for (int bk = 0; bk < nblocks; ++bk)
#pragma unroll
        for (int k = 0; k < NQ; ++k) out[k] += part[(static_cast<size_t>(bk) * NQ + k) * kL + lane];
}

// Quantity q of this lane summed over the workgroups' partials, with 8
// interleaved accumulators combined in a fixed order (one wavefront per
// quantity keeps the loads of a check's sums in flight together).
__device__ inline double sum_lane_q(const double* part, int nblocks, int nq, int q, int lane) {
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int b = 0;
    for (; b + 8 <= nblocks; b += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += part[(static_cast<size_t>(b + u) * nq + q) * kL + lane];
    for (; b < nblocks; ++b) a[0] += part[(static_cast<size_t>(b) * nq + q) * kL + lane];
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

__device__ inline int first_seg() { return static_cast<int>(blockIdx.x) * kSegWaves + static_cast<int>(threadIdx.x) / kWave; }
__device__ inline int seg_stride() { return static_cast<int>(gridDim.x) * kSegWaves; }

// ------------------------------------------------------------------- set-up
// Scaled costs / right-hand sides, starting points (warm start or zero), and
// lane sums ||c_hat||^2 (cols) or ||b_hat||^2 (rows).
template <class T>
__global__ void __launch_bounds__(256)
kf_scale_cols(int n, const double* cs, const double* c, const T* l_hat, const T* u_hat, const double* wx,
              int has_warm, T* c_hat, double* xk, T* x0, T* xp, double* part) {
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    double q[1] = {0.0};
    for (int j = first_seg(); j < n; j += seg_stride()) {
        const size_t e = static_cast<size_t>(j) * kL + lane;
        const double sj = cs[j];
        const T ch = static_cast<T>(c[e] * sj);
        c_hat[e] = ch;
        q[0] += static_cast<double>(ch) * static_cast<double>(ch);
        const double xs = clampT(has_warm ? wx[e] / sj : 0.0, static_cast<double>(l_hat[j]), static_cast<double>(u_hat[j]));
        xk[e] = xs;
        x0[e] = static_cast<T>(xs);
        xp[e] = static_cast<T>(xs);
    }
    store_lane_sums<1>(q, part);
}

template <class T>
__global__ void __launch_bounds__(256)
kf_scale_rows(int m, const double* rs, const double* b, const unsigned char* eq, const double* wy, int has_warm,
              T* b_hat, double* yk, T* y0, T* yg, T* yp, double* part) {
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    double q[1] = {0.0};
    for (int i = first_seg(); i < m; i += seg_stride()) {
        const size_t e = static_cast<size_t>(i) * kL + lane;
        const double ri = rs[i];
        const T bh = static_cast<T>(b[e] * ri);
        b_hat[e] = bh;
        q[0] += static_cast<double>(bh) * static_cast<double>(bh);
        double ys = has_warm ? wy[e] / ri : 0.0;
        if (!eq[i] && ys < 0.0) ys = 0.0;
        yk[e] = ys;
        y0[e] = static_cast<T>(ys);
        yg[e] = static_cast<T>(ys);
        yp[e] = static_cast<T>(ys);
    }
    store_lane_sums<1>(q, part);
}

// One thread per lane: the initial state of every LP (lanes >= count: done).
__global__ void kf_init_state(int count, PdlpState* st, const PdlpState* init, double pw_u2, const double* part_c,
                              int nc, const double* part_b, int nb) {
    const int lane = static_cast<int>(threadIdx.x);
    if (lane >= kL) return;
    st[lane] = init[lane];
    double c2, b2;
    sum_lane_parts<1>(part_c, nc, lane, &c2);
    sum_lane_parts<1>(part_b, nb, lane, &b2);
    init_state_core(st + lane, pw_u2, c2, b2);
    if (lane >= count) {
        st[lane].done = 1;
        st[lane].status = 0;
    }
}

// Step data of the next interval and the number of LPs still running.
__global__ void kf_coef(const PdlpState* st, LaneCoef* cf, int* running) {
    const int lane = static_cast<int>(threadIdx.x);
    if (lane >= kL) return;
    const PdlpState& s = st[lane];
    cf->tau[lane] = s.tau;
    cf->sigma[lane] = s.sigma;
    cf->kbase[lane] = static_cast<double>(s.iter - s.iter_last_restart);
    cf->live[lane] = s.done ? 0 : 1;
    const unsigned long long alive = __builtin_amdgcn_ballot_w64(!s.done);
    if (lane == 0) *running = __popcll(alive);
}

__global__ void kf_advance(PdlpState* st, int steps) {
    const int lane = static_cast<int>(threadIdx.x);
    if (lane < kL && !st[lane].done) st[lane].iter += steps;
}

// ------------------------------------------------------------------ checks
// xo = clamp(s x, l, u) for the iterate and for x+ (unscaled); sums c'xo.
template <class T>
__global__ void __launch_bounds__(256)
kf_check_x(int n, const LaneCoef* cf, const double* xk, const T* xp, const double* cs, const double* c,
           const double* lb, const double* ub, double* xo_cur, double* xo_cand, double* part) {
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    double q[2] = {0.0, 0.0};
    if (cf->live[lane])
        for (int j = first_seg(); j < n; j += seg_stride()) {
            const size_t e = static_cast<size_t>(j) * kL + lane;
            const double sj = cs[j], l = lb[j], u = ub[j];
            const double xc = clampT(sj * xk[e], l, u);
            const double xa = clampT(sj * static_cast<double>(xp[e]), l, u);
            xo_cur[e] = xc;
            xo_cand[e] = xa;
            q[0] += c[e] * xc;
            q[1] += c[e] * xa;
        }
    store_lane_sums<2>(q, part);
}

// Row side from a0 = A xo_cur, a1 = A xo_cand (unscaled FP64 matrix):
// [0,1] ||viol||^2  [2,3] scaled  [4,5] b'y; also the unscaled duals and
// A x+ scaled (for a restart to the candidate).
template <class T>
__global__ void __launch_bounds__(256)
kf_check_rows(int m, const LaneCoef* cf, const double* ar0, const double* ar1, const double* b, const unsigned char* eq,
              const double* rs, const double* yk, const T* yp, double* yo_cur, double* yo_cand, double* ax_cand,
              double* part) {
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (cf->live[lane])
        for (int i = first_seg(); i < m; i += seg_stride()) {
            const size_t e = static_cast<size_t>(i) * kL + lane;
            const double bi = b[e], ri = rs[i];
            const bool isq = eq[i] != 0;
            const double acc[2] = {ar0[e], ar1[e]};
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const double r = acc[p] - bi;
                const double v = isq ? fabs(r) : (r > 0.0 ? r : 0.0);
                q[p] += v * v;
                q[2 + p] += (ri * v) * (ri * v);
            }
            ax_cand[e] = ri * acc[1];
            double yc = ri * yk[e];
            if (!isq && yc < 0.0) yc = 0.0;
            const double ya = ri * static_cast<double>(yp[e]);
            yo_cur[e] = yc;
            yo_cand[e] = ya;
            q[4] += bi * yc;
            q[5] += bi * ya;
        }
    store_lane_sums<6>(q, part);
}

// Column side from a0 = A'yo_cur, a1 = A'yo_cand:
// [0,1] ||dual residual||^2  [2,3] scaled  [4,5] sum l plam+ + u plam-
__global__ void __launch_bounds__(256)
kf_check_cols(int n, const LaneCoef* cf, const double* ac0, const double* ac1, const double* c, const double* lb,
              const double* ub, const double* cs, double* part) {
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (cf->live[lane])
        for (int j = first_seg(); j < n; j += seg_stride()) {
            const size_t e = static_cast<size_t>(j) * kL + lane;
            const double l = lb[j], u = ub[j], sj = cs[j];
            const bool lf = isfinite(l), uf = isfinite(u);
            const double acc[2] = {ac0[e], ac1[e]};
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const double lam = c[e] + acc[p];
                double pl;
                if (lf && uf) pl = lam;
                else if (lf) pl = lam > 0.0 ? lam : 0.0;
                else if (uf) pl = lam < 0.0 ? lam : 0.0;
                else pl = 0.0;
                const double d = lam - pl;
                q[p] += d * d;
                q[2 + p] += (sj * d) * (sj * d);
                q[4 + p] += pl > 0.0 ? l * pl : (pl < 0.0 ? u * pl : 0.0);
            }
        }
    store_lane_sums<6>(q, part);
}

// The check's lane sums in two levels, each in a fixed order (one
// workgroup walking every workgroup's partials took about 0.5 ms per check).
// kf_fold_check: workgroup g, wavefront q adds quantity q of the blocks in
// g's contiguous range, for all lanes, into fold[g][q][lane].  kf_decide:
// wavefront q adds the kFoldGroupsF results of quantity q; then one thread
// per lane runs the stopping test, stall test and restart rule of the
// single-LP engine.
constexpr int kFoldGroupsF = 16;

__device__ inline double sum_lane_range(const double* part, int nblocks, int nq, int q, int lane) {
    const int g = static_cast<int>(blockIdx.x), ng = static_cast<int>(gridDim.x);
    const int per = (nblocks + ng - 1) / ng, b0 = min(nblocks, g * per), b1 = min(nblocks, b0 + per);
    return sum_lane_q(part + static_cast<size_t>(b0) * nq * kL, b1 - b0, nq, q, lane);
}

__global__ void __launch_bounds__(1024)
kf_fold_check(const double* px, const double* pr, const double* pc, int ncol_blocks, int nrow_blocks, double* fold) {
    const int lane = static_cast<int>(threadIdx.x) % kWave, w = static_cast<int>(threadIdx.x) / kWave;
    double v = 0.0;
    if (w < 2) v = sum_lane_range(px, ncol_blocks, 2, w, lane);
    else if (w < 8) v = sum_lane_range(pr, nrow_blocks, 6, w - 2, lane);
    else if (w < 14) v = sum_lane_range(pc, ncol_blocks, 6, w - 8, lane);
    if (w < 14) fold[(static_cast<size_t>(blockIdx.x) * 14 + w) * kL + lane] = v;
}

__global__ void __launch_bounds__(1024)
kf_decide(PdlpState* st, const double* fold, int groups) {
    __shared__ double sh[14][kL];
    const int lane = static_cast<int>(threadIdx.x) % kWave, w = static_cast<int>(threadIdx.x) / kWave;
    if (w < 14) sh[w][lane] = sum_lane_q(fold, groups, 14, w, lane);
    __syncthreads();
    if (w != 0 || st[lane].done) return;
    double a[14];
#pragma unroll
    for (int q = 0; q < 14; ++q) a[q] = sh[q][lane];
    decide_core(st + lane, a);
}

// Restarts of the LPs that decided to: move to the chosen point, make it the
// anchor; lane sums of ||dx||^2 and ||dy||^2 of the anchor move.
template <class T>
__global__ void __launch_bounds__(256)
kf_restart_x(int n, const PdlpState* st, double* xk, T* x0, const T* xp, double* part) {
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    const PdlpState& s = st[lane];
    double q[1] = {0.0};
    if (!s.done && s.restart)
        for (int j = first_seg(); j < n; j += seg_stride()) {
            const size_t e = static_cast<size_t>(j) * kL + lane;
            const double xn = s.restart_to_cand ? static_cast<double>(xp[e]) : xk[e];
            const double d = xn - static_cast<double>(x0[e]);
            q[0] += d * d;
            xk[e] = xn;
            x0[e] = static_cast<T>(xn);
        }
    store_lane_sums<1>(q, part);
}

template <class T>
__global__ void __launch_bounds__(256)
kf_restart_y(int m, const PdlpState* st, double* yk, T* y0, const T* yp, T* yg, double* ax, T* ax0,
             const double* ax_cand, double* part) {
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    const PdlpState& s = st[lane];
    double q[1] = {0.0};
    if (!s.done && s.restart)
        for (int i = first_seg(); i < m; i += seg_stride()) {
            const size_t e = static_cast<size_t>(i) * kL + lane;
            double yn = yk[e];
            if (s.restart_to_cand) {
                yn = static_cast<double>(yp[e]);
                ax[e] = ax_cand[e];
            }
            const double d = yn - static_cast<double>(y0[e]);
            q[0] += d * d;
            yk[e] = yn;
            y0[e] = static_cast<T>(yn);
            yg[e] = static_cast<T>(yn);
            ax0[e] = static_cast<T>(ax[e]);
        }
    store_lane_sums<1>(q, part);
}

// ||dx||^2 and ||dy||^2 of every lane in the same two levels: kf_fold_restart
// (wavefront 0: dx, 1: dy) into fold[g][2][lane], then kf_restart_finish
// adds the groups and runs the primal-weight update.
__global__ void kf_fold_restart(const double* pdx, int ncol_blocks, const double* pdy, int nrow_blocks,
                                double* fold) {
    const int lane = static_cast<int>(threadIdx.x) % kWave, w = static_cast<int>(threadIdx.x) / kWave;
    const double v = w == 0 ? sum_lane_range(pdx, ncol_blocks, 1, 0, lane) : sum_lane_range(pdy, nrow_blocks, 1, 0, lane);
    fold[(static_cast<size_t>(blockIdx.x) * 2 + w) * kL + lane] = v;
}

__global__ void kf_restart_finish(PdlpState* st, const double* fold, int groups) {
    __shared__ double sh[2][kL];
    const int lane = static_cast<int>(threadIdx.x) % kWave, w = static_cast<int>(threadIdx.x) / kWave;
    sh[w][lane] = sum_lane_q(fold, groups, 2, w, lane);
    __syncthreads();
    if (w != 0 || st[lane].done || !st[lane].restart) return;
    restart_finish_core(st + lane, sh[0][lane], sh[1][lane]);
}

// ------------------------------------------------------------- data movement
// LP-major rows (count x len, row-major) <-> LP-minor (len x 64); lanes >=
// count are filled with `fill`.
__global__ void kf_to_minor(int count, int len, const double* src, double* dst, double fill) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<size_t>(len) * kL) return;
    const int j = static_cast<int>(e / kL), k = static_cast<int>(e % kL);
    dst[e] = k < count ? src[static_cast<size_t>(k) * len + j] : fill;
}

// The returned point of every LP (iterate or candidate, by its `best`), LP-major.
__global__ void kf_from_minor(int count, int len, const PdlpState* st, const double* cur, const double* cand,
                              double* dst) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<size_t>(len) * count) return;
    const int k = static_cast<int>(e / len), j = static_cast<int>(e % len);
    const size_t s = static_cast<size_t>(j) * kL + k;
    dst[e] = st[k].best ? cand[s] : cur[s];
}

struct DevMemF {
    std::vector<void*> ptrs;
    ~DevMemF() {
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

// ------------------------------------------------------------- host scaling
// Ruiz equilibration (inf-norms, `ruiz` rounds) then Pock-Chambolle (l1
// norms, alpha = 1): in each round both norms use the previous round's
// scales, and every scale is divided by the square root of its norm.
void host_scaling(int m, int n, const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<double>& v,
                  int ruiz, bool pock_chambolle, std::vector<double>& rs, std::vector<double>& cs) {
    rs.assign(m, 1.0);
    cs.assign(n, 1.0);
    std::vector<double> rn(m), cn(n);
    auto round = [&](bool l1) {
        std::fill(rn.begin(), rn.end(), 0.0);
        std::fill(cn.begin(), cn.end(), 0.0);
        for (int i = 0; i < m; ++i)
            for (int q = rp[i]; q < rp[i + 1]; ++q) {
                const double a = std::fabs(v[q] * cs[ci[q]]) * rs[i];
                const double b = std::fabs(v[q] * rs[i]) * cs[ci[q]];
                if (l1) {
                    rn[i] += a;
                    cn[ci[q]] += b;
                } else {
                    rn[i] = std::max(rn[i], a);
                    cn[ci[q]] = std::max(cn[ci[q]], b);
                }
            }
        for (int i = 0; i < m; ++i)
            if (rn[i] > 0.0 && std::isfinite(rn[i])) rs[i] /= std::sqrt(rn[i]);
        for (int j = 0; j < n; ++j)
            if (cn[j] > 0.0 && std::isfinite(cn[j])) cs[j] /= std::sqrt(cn[j]);
    };
    for (int it = 0; it < ruiz; ++it) round(false);
    if (pock_chambolle) round(true);
}

// ||A_hat v||^2 after kPowerIters rounds of v <- A'A v / ||.|| from a hashed
// positive start.
double host_power(int m, int n, const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<double>& a) {
    std::vector<double> v(n), u(m);
    for (int j = 0; j < n; ++j) {
        unsigned h = static_cast<unsigned>(j) * 2654435761u + 12345u;
        h ^= h >> 13;
        h *= 0x5bd1e995u;
        h ^= h >> 15;
        v[j] = 0.5 + static_cast<double>(h & 0xffffu) / 65536.0;
    }
    auto normalize = [&]() {
        double s = 0.0;
        for (double t : v) s += t * t;
        s = std::sqrt(s);
        if (s > 0.0)
            for (double& t : v) t /= s;
    };
    auto mul = [&]() {
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int q = rp[i]; q < rp[i + 1]; ++q) s += a[q] * v[ci[q]];
            u[i] = s;
        }
    };
    normalize();
    for (int it = 0; it < kPowerIters; ++it) {
        mul();
        std::fill(v.begin(), v.end(), 0.0);
        for (int i = 0; i < m; ++i)
            for (int q = rp[i]; q < rp[i + 1]; ++q) v[ci[q]] += a[q] * u[i];
        normalize();
    }
    mul();
    double s = 0.0;
    for (double t : u) s += t * t;
    return s;
}

}  // namespace

struct FamilyBase {
    virtual ~FamilyBase() = default;
    virtual bo_status set_lps(int count, const double* c, const double* b) = 0;
    virtual bo_status set_warm_start(const double* x, const double* y) = 0;
    virtual bo_status set_primal_weights(const double* w) = 0;
    virtual void set_iteration_limits(const long long* lim) = 0;
    virtual void set_deadline(const Deadline& d) = 0;
    virtual bo_status solve(bo_pdlp_stats* per_lp, bo_pdlp_batch_stats* total) = 0;
    virtual bo_status get_solutions(double* x, double* y) = 0;
    virtual bo_status time_passes(int reps, double* ms2) = 0;
    virtual int count() const = 0;
};

namespace {

template <class T>
struct Family final : FamilyBase {
    int K = 0, cap = 0, m = 0, n = 0;
    long long nnz = 0;
    bo_pdlp_params prm{};
    hipStream_t stream = nullptr;
    DevMemF mem;
    int *rp = nullptr, *ci = nullptr, *cp = nullptr, *ri = nullptr;
    double *rv = nullptr, *cv = nullptr;  // scaled values, FP64 (the matrix is shared: 4 more bytes per entry)
    double *rv64 = nullptr, *cv64 = nullptr;
    double *rs = nullptr, *cs = nullptr, *lb64 = nullptr, *ub64 = nullptr;
    T *l_hat = nullptr, *u_hat = nullptr;
    unsigned char* eq = nullptr;
    double pw_u2 = 0.0;
    // LP-minor per-LP data
    double *c64 = nullptr, *b64 = nullptr;
    T *c_hat = nullptr, *b_hat = nullptr;
    double *xk = nullptr, *yk = nullptr, *ax = nullptr;
    T *x0 = nullptr, *y0 = nullptr, *ax0 = nullptr;  // Halpern anchors
    T *xp = nullptr, *yp = nullptr, *yg = nullptr;
    double *xo_cur = nullptr, *xo_cand = nullptr, *yo_cur = nullptr, *yo_cand = nullptr;
    double *ar0 = nullptr, *ar1 = nullptr, *ac0 = nullptr, *ac1 = nullptr, *ax_cand = nullptr;
    double *wx = nullptr, *wy = nullptr, *stage = nullptr;
    int has_warm = 0;
    double *part_a = nullptr, *part_b = nullptr, *part_c = nullptr;
    double* part_fold = nullptr;  // first level of the check's and restart's lane sums
    LaneCoef* cf = nullptr;
    PdlpState* st = nullptr;
    PdlpState* st_host = nullptr;
    PdlpState* st_init = nullptr;
    int* running = nullptr;
    int* running_host = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_k = 0;
    Deadline deadline;  // unset: time_limit_s counts from the start of solve()

    ~Family() override {
        if (graph_exec) hipGraphExecDestroy(graph_exec);
        if (graph) hipGraphDestroy(graph);
        if (st_host) hipHostFree(st_host);
        if (st_init) hipHostFree(st_init);
        if (running_host) hipHostFree(running_host);
        if (stream) hipStreamDestroy(stream);
    }

    int count() const override { return K; }
    unsigned sum_grid(int nseg) const {
        return static_cast<unsigned>(std::max(1, std::min<int>(kSumBlocks, (nseg + kSegWaves - 1) / kSegWaves)));
    }

    hipError_t up(void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream) : hipSuccess;
    }

    bo_status create(const bo_lp& lp, int capacity, const bo_pdlp_params& p) {
        prm = p;
        m = lp.m;
        n = lp.n;
        nnz = lp.nnz;
        K = capacity;
        cap = capacity;
        BO_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_host), sizeof(PdlpState) * kL));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_init), sizeof(PdlpState) * kL));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&running_host), sizeof(int)));
        std::memset(st_host, 0, sizeof(PdlpState) * kL);
        std::memset(st_init, 0, sizeof(PdlpState) * kL);
        for (int k = 0; k < kL; ++k) {
            st_init[k].eps = prm.eps_rel_kkt;
            st_init[k].eps_obj = prm.eps_objective;
            st_init[k].max_iter = prm.max_iterations;
            st_init[k].stall_window = sizeof(T) == sizeof(float) ? prm.stall_iterations : 0;
        }
        // host matrix: CSR as given, CSC by counting; scaling and ||A_hat||
        std::vector<int> h_rp(lp.row_ptr, lp.row_ptr + m + 1), h_ci(lp.col_idx, lp.col_idx + nnz);
        std::vector<double> h_rv(lp.val, lp.val + nnz);
        std::vector<int> h_cp(n + 1, 0), h_ri(nnz);
        std::vector<double> h_cv(nnz);
        for (long long q = 0; q < nnz; ++q) h_cp[h_ci[q] + 1]++;
        for (int j = 0; j < n; ++j) h_cp[j + 1] += h_cp[j];
        {
            std::vector<int> pos(h_cp.begin(), h_cp.end() - 1);
            for (int i = 0; i < m; ++i)
                for (int q = h_rp[i]; q < h_rp[i + 1]; ++q) {
                    const int d = pos[h_ci[q]]++;
                    h_ri[d] = i;
                    h_cv[d] = h_rv[q];
                }
        }
        std::vector<double> hrs, hcs;
        host_scaling(m, n, h_rp, h_ci, h_rv, prm.ruiz_iterations, prm.pock_chambolle != 0, hrs, hcs);
        std::vector<double> h_rsc(nnz), h_csc(nnz);  // scaled values
        for (int i = 0; i < m; ++i)
            for (int q = h_rp[i]; q < h_rp[i + 1]; ++q) h_rsc[q] = hrs[i] * h_rv[q] * hcs[h_ci[q]];
        for (int j = 0; j < n; ++j)
            for (int q = h_cp[j]; q < h_cp[j + 1]; ++q) h_csc[q] = hcs[j] * h_cv[q] * hrs[h_ri[q]];
        pw_u2 = host_power(m, n, h_rp, h_ci, h_rsc);
        std::vector<T> hl(n), hu(n);
        for (int j = 0; j < n; ++j) {
            hl[j] = static_cast<T>(lp.lb[j] / hcs[j]);
            hu[j] = static_cast<T>(lp.ub[j] / hcs[j]);
        }
        // device
        const size_t NL = static_cast<size_t>(n) * kL, ML = static_cast<size_t>(m) * kL;
        BO_TRY(mem.alloc(&rp, m + 1));
        BO_TRY(mem.alloc(&ci, nnz + kUnroll));
        BO_TRY(mem.alloc(&rv, nnz + kUnroll));
        BO_TRY(mem.alloc(&rv64, nnz + kUnroll));
        BO_TRY(mem.alloc(&cp, n + 1));
        BO_TRY(mem.alloc(&ri, nnz + kUnroll));
        BO_TRY(mem.alloc(&cv, nnz + kUnroll));
        BO_TRY(mem.alloc(&cv64, nnz + kUnroll));
        // the padding: index 0, value 0 (read by the last round trip of the last segment)
        BO_TRY(hipMemsetAsync(ci + nnz, 0, sizeof(int) * kUnroll, stream));
        BO_TRY(hipMemsetAsync(ri + nnz, 0, sizeof(int) * kUnroll, stream));
        BO_TRY(hipMemsetAsync(rv + nnz, 0, sizeof(double) * kUnroll, stream));
        BO_TRY(hipMemsetAsync(cv + nnz, 0, sizeof(double) * kUnroll, stream));
        BO_TRY(hipMemsetAsync(rv64 + nnz, 0, sizeof(double) * kUnroll, stream));
        BO_TRY(hipMemsetAsync(cv64 + nnz, 0, sizeof(double) * kUnroll, stream));
        BO_TRY(mem.alloc(&rs, m));
        BO_TRY(mem.alloc(&cs, n));
        BO_TRY(mem.alloc(&lb64, n));
        BO_TRY(mem.alloc(&ub64, n));
        BO_TRY(mem.alloc(&l_hat, n));
        BO_TRY(mem.alloc(&u_hat, n));
        BO_TRY(mem.alloc(&eq, m));
        for (double** v : {&c64, &xk, &xo_cur, &xo_cand, &ac0, &ac1, &wx}) BO_TRY(mem.alloc(v, NL));
        for (double** v : {&b64, &yk, &ax, &yo_cur, &yo_cand, &ar0, &ar1, &ax_cand, &wy}) BO_TRY(mem.alloc(v, ML));
        BO_TRY(mem.alloc(&x0, NL));
        BO_TRY(mem.alloc(&y0, ML));
        BO_TRY(mem.alloc(&ax0, ML));
        BO_TRY(mem.alloc(&c_hat, NL));
        BO_TRY(mem.alloc(&xp, NL));
        BO_TRY(mem.alloc(&b_hat, ML));
        BO_TRY(mem.alloc(&yp, ML));
        BO_TRY(mem.alloc(&yg, ML));
        BO_TRY(mem.alloc(&stage, std::max(NL, ML)));
        BO_TRY(mem.alloc(&part_a, static_cast<size_t>(kSumBlocks) * 6 * kL));
        BO_TRY(mem.alloc(&part_fold, static_cast<size_t>(kFoldGroupsF) * 14 * kL));
        BO_TRY(mem.alloc(&part_b, static_cast<size_t>(kSumBlocks) * 6 * kL));
        BO_TRY(mem.alloc(&part_c, static_cast<size_t>(kSumBlocks) * 6 * kL));
        BO_TRY(mem.alloc(&cf, 1));
        BO_TRY(mem.alloc(&st, kL));
        BO_TRY(mem.alloc(&running, 1));
        BO_TRY(up(rp, h_rp.data(), sizeof(int) * (m + 1)));
        BO_TRY(up(ci, h_ci.data(), sizeof(int) * nnz));
        BO_TRY(up(rv, h_rsc.data(), sizeof(double) * nnz));
        BO_TRY(up(rv64, h_rv.data(), sizeof(double) * nnz));
        BO_TRY(up(cp, h_cp.data(), sizeof(int) * (n + 1)));
        BO_TRY(up(ri, h_ri.data(), sizeof(int) * nnz));
        BO_TRY(up(cv, h_csc.data(), sizeof(double) * nnz));
        BO_TRY(up(cv64, h_cv.data(), sizeof(double) * nnz));
        BO_TRY(up(rs, hrs.data(), sizeof(double) * m));
        BO_TRY(up(cs, hcs.data(), sizeof(double) * n));
        BO_TRY(up(lb64, lp.lb, sizeof(double) * n));
        BO_TRY(up(ub64, lp.ub, sizeof(double) * n));
        BO_TRY(up(l_hat, hl.data(), sizeof(T) * n));
        BO_TRY(up(u_hat, hu.data(), sizeof(T) * n));
        BO_TRY(up(eq, lp.is_eq, m));
        BO_TRY(hipStreamSynchronize(stream));
        // every LP starts as the base LP
        std::vector<double> cb(static_cast<size_t>(K) * n), bb(static_cast<size_t>(K) * m);
        for (int k = 0; k < K; ++k) {
            std::copy(lp.c, lp.c + n, cb.begin() + static_cast<size_t>(k) * n);
            std::copy(lp.b, lp.b + m, bb.begin() + static_cast<size_t>(k) * m);
        }
        return set_lps(K, cb.data(), bb.data());
    }

    // LP-major host rows -> LP-minor device array through the staging buffer
    bo_status to_minor(int count, int len, const double* host, double* dev, double fill) {
        if (len == 0) return BO_SUCCESS;
        BO_TRY(hipMemcpyAsync(stage, host, sizeof(double) * count * len, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(kf_to_minor, dim3(grid_for(static_cast<size_t>(len) * kL)), dim3(kBlock), 0, stream, count,
                           len, stage, dev, fill);
        BO_TRY(hipGetLastError());
        BO_TRY(hipStreamSynchronize(stream));
        return BO_SUCCESS;
    }

    bo_status set_lps(int count, const double* c, const double* b) override {
        if (count <= 0 || count > cap) return BO_INVALID_ARGUMENT;
        if ((c == nullptr) != (b == nullptr) || (count != K && !c)) return BO_INVALID_ARGUMENT;
        if (c) {
            for (size_t e = 0; e < static_cast<size_t>(count) * n; ++e)
                if (!std::isfinite(c[e])) return BO_INVALID_ARGUMENT;
            for (size_t e = 0; e < static_cast<size_t>(count) * m; ++e)
                if (!std::isfinite(b[e])) return BO_INVALID_ARGUMENT;
        }
        BO_TRY(hipStreamSynchronize(stream));
        K = count;
        has_warm = 0;
        for (int k = 0; k < kL; ++k) {
            st_init[k].omega_init = 0.0;
            st_init[k].max_iter = prm.max_iterations;
        }
        if (!c) return BO_SUCCESS;
        bo_status s = to_minor(count, n, c, c64, 0.0);
        if (s != BO_SUCCESS) return s;
        if ((s = to_minor(count, m, b, b64, 0.0)) != BO_SUCCESS) return s;
        for (int k = 0; k < count; ++k) {
            double c2 = 0.0, b2 = 0.0;
            for (int j = 0; j < n; ++j) c2 += c[static_cast<size_t>(k) * n + j] * c[static_cast<size_t>(k) * n + j];
            for (int i = 0; i < m; ++i) b2 += b[static_cast<size_t>(k) * m + i] * b[static_cast<size_t>(k) * m + i];
            st_init[k].c_norm = std::sqrt(c2);
            st_init[k].b_norm = std::sqrt(b2);
        }
        return BO_SUCCESS;
    }

    bo_status set_warm_start(const double* x, const double* y) override {
        if (!x && !y) {
            has_warm = 0;
            return BO_SUCCESS;
        }
        if (!x || !y) return BO_INVALID_ARGUMENT;
        bo_status s = to_minor(K, n, x, wx, 0.0);
        if (s != BO_SUCCESS) return s;
        if ((s = to_minor(K, m, y, wy, 0.0)) != BO_SUCCESS) return s;
        has_warm = 1;
        return BO_SUCCESS;
    }

    bo_status set_primal_weights(const double* w) override {
        if (w)
            for (int k = 0; k < K; ++k)
                if (!(w[k] >= 0.0) || !std::isfinite(w[k])) return BO_INVALID_ARGUMENT;
        BO_TRY(hipStreamSynchronize(stream));
        for (int k = 0; k < kL; ++k) st_init[k].omega_init = (w && k < K) ? w[k] : 0.0;
        return BO_SUCCESS;
    }

    // The absolute time limit of the solve this engine is a round of; LPs
    // still running when it passes stop with status 3 (TIME_LIMIT).
    void set_deadline(const Deadline& d) override { deadline = d; }

    void set_iteration_limits(const long long* lim) override {
        for (int k = 0; k < K; ++k) st_init[k].max_iter = std::min(lim[k], prm.max_iterations);
    }

    void launch_primal(int t) {
        EpiFPrimal<T> e{yg, cf, c_hat, l_hat, u_hat, xk, x0, xp, t};
        hipLaunchKernelGGL((kf_spmm<double, EpiFPrimal<T>>), dim3(spmm_grid(n)), dim3(256), 0, stream, n, cp, ri, cv, e);
    }
    void launch_dual(int t) {
        EpiFDual<T> e{xp, cf, b_hat, eq, ax, ax0, yk, y0, yp, yg, t};
        hipLaunchKernelGGL((kf_spmm<double, EpiFDual<T>>), dim3(spmm_grid(m)), dim3(256), 0, stream, m, rp, ci, rv, e);
    }

    void launch_check() {
        const unsigned gn = sum_grid(n), gm = sum_grid(m);
        hipLaunchKernelGGL(kf_check_x<T>, dim3(gn), dim3(256), 0, stream, n, cf, xk, xp, cs, c64, lb64, ub64, xo_cur,
                           xo_cand, part_a);
        hipLaunchKernelGGL((kf_spmm<double, EpiFCheck2>), dim3(spmm_grid(m)), dim3(256), 0, stream, m, rp, ci, rv64,
                           EpiFCheck2{xo_cur, xo_cand, cf, ar0, ar1});
        hipLaunchKernelGGL(kf_check_rows<T>, dim3(gm), dim3(256), 0, stream, m, cf, ar0, ar1, b64, eq, rs, yk, yp,
                           yo_cur, yo_cand, ax_cand, part_b);
        hipLaunchKernelGGL((kf_spmm<double, EpiFCheck2>), dim3(spmm_grid(n)), dim3(256), 0, stream, n, cp, ri, cv64,
                           EpiFCheck2{yo_cur, yo_cand, cf, ac0, ac1});
        hipLaunchKernelGGL(kf_check_cols, dim3(gn), dim3(256), 0, stream, n, cf, ac0, ac1, c64, lb64, ub64, cs,
                           part_c);
        hipLaunchKernelGGL(kf_fold_check, dim3(kFoldGroupsF), dim3(1024), 0, stream,
                           static_cast<const double*>(part_a), static_cast<const double*>(part_b),
                           static_cast<const double*>(part_c), static_cast<int>(gn), static_cast<int>(gm), part_fold);
        hipLaunchKernelGGL(kf_decide, dim3(1), dim3(1024), 0, stream, st, static_cast<const double*>(part_fold),
                           kFoldGroupsF);
        hipLaunchKernelGGL(kf_restart_x<T>, dim3(gn), dim3(256), 0, stream, n, st, xk, x0, xp, part_a);
        hipLaunchKernelGGL(kf_restart_y<T>, dim3(gm), dim3(256), 0, stream, m, st, yk, y0, yp, yg, ax, ax0, ax_cand,
                           part_b);
        hipLaunchKernelGGL(kf_fold_restart, dim3(kFoldGroupsF), dim3(2 * kL), 0, stream,
                           static_cast<const double*>(part_a), static_cast<int>(gn), static_cast<const double*>(part_b),
                           static_cast<int>(gm), part_fold);
        hipLaunchKernelGGL(kf_restart_finish, dim3(1), dim3(2 * kL), 0, stream, st,
                           static_cast<const double*>(part_fold), kFoldGroupsF);
        hipLaunchKernelGGL(kf_coef, dim3(1), dim3(kL), 0, stream, st, cf, running);
    }

    void launch_interval(int k) {
        for (int t = 0; t < k; ++t) {
            launch_primal(t);
            launch_dual(t);
        }
        hipLaunchKernelGGL(kf_advance, dim3(1), dim3(kL), 0, stream, st, k);
        launch_check();
    }

    bo_status prepare() {
        const unsigned gn = sum_grid(n), gm = sum_grid(m);
        hipLaunchKernelGGL(kf_scale_cols<T>, dim3(gn), dim3(256), 0, stream, n, cs, c64, l_hat, u_hat, wx, has_warm,
                           c_hat, xk, x0, xp, part_a);
        hipLaunchKernelGGL(kf_scale_rows<T>, dim3(gm), dim3(256), 0, stream, m, rs, b64, eq, wy, has_warm, b_hat, yk,
                           y0, yg, yp, part_b);
        BO_TRY(hipMemcpyAsync(st, st_init, sizeof(PdlpState) * kL, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(kf_init_state, dim3(1), dim3(kL), 0, stream, K, st, st, pw_u2, part_a,
                           static_cast<int>(gn), part_b, static_cast<int>(gm));
        hipLaunchKernelGGL(kf_coef, dim3(1), dim3(kL), 0, stream, st, cf, running);
        // A x0 (the starting point is also the first anchor)
        hipLaunchKernelGGL((kf_spmm<double, EpiFStore<T>>), dim3(spmm_grid(m)), dim3(256), 0, stream, m, rp, ci, rv,
                           EpiFStore<T>{xp, ax, ax0});
        BO_TRY(hipGetLastError());
        return BO_SUCCESS;
    }

    bo_status solve(bo_pdlp_stats* per_lp, bo_pdlp_batch_stats* total) override {
        auto t0 = std::chrono::high_resolution_clock::now();
        const Deadline dl = deadline.on ? deadline : Deadline::after(prm.time_limit_s);
        bool timed_out = false;
        bo_status s0 = prepare();
        if (s0 != BO_SUCCESS) return s0;
        launch_check();  // step 0: KKT of the starting points
        BO_TRY(hipMemcpyAsync(running_host, running, sizeof(int), hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        const int Kc = prm.check_every > 0 ? prm.check_every : 64;
        long long iters = 0;
        while (*running_host > 0) {
            if (dl.passed()) {
                timed_out = true;
                break;
            }
            const long long left = prm.max_iterations - iters;
            const int k = static_cast<int>(std::min<long long>(Kc, left));
            if (k <= 0) break;
            if (prm.use_graph && k == Kc) {
                if (!graph_exec || graph_k != Kc) {
                    if (graph_exec) hipGraphExecDestroy(graph_exec);
                    if (graph) hipGraphDestroy(graph);
                    graph_exec = nullptr;
                    graph = nullptr;
                    BO_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                    launch_interval(Kc);
                    BO_TRY(hipStreamEndCapture(stream, &graph));
                    BO_TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
                    graph_k = Kc;
                }
                BO_TRY(hipGraphLaunch(graph_exec, stream));
            } else {
                launch_interval(k);
            }
            BO_TRY(hipGetLastError());
            iters += k;
            BO_TRY(hipMemcpyAsync(running_host, running, sizeof(int), hipMemcpyDeviceToHost, stream));
            BO_TRY(hipStreamSynchronize(stream));
        }
        BO_TRY(hipMemcpyAsync(st_host, st, sizeof(PdlpState) * kL, hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
        int solved = 0;
        long long max_it = 0, sum_it = 0;
        for (int k = 0; k < K; ++k) {
            const PdlpState& h = st_host[k];
            const int b = h.best;
            const int status = h.done ? h.status : (timed_out ? 3 : 1);
            solved += status == 0;
            max_it = std::max(max_it, h.iter);
            sum_it += h.iter;
            if (per_lp) {
                bo_pdlp_stats& o = per_lp[k];
                std::memset(&o, 0, sizeof(o));
                o.status = status;
                o.iterations = h.iter;
                o.restarts = h.restarts;
                o.used_average = b;
                o.rel_kkt = h.rel_kkt[b];
                o.rel_primal = h.rel_primal[b];
                o.rel_dual = h.rel_dual[b];
                o.rel_gap = h.rel_gap[b];
                o.primal_obj = h.pobj[b];
                o.dual_obj = h.dobj[b];
                o.primal_weight = h.omega;
                o.step_size = h.eta;
                o.solve_ms = ms;
            }
        }
        if (total) {
            std::memset(total, 0, sizeof(*total));
            total->count = K;
            total->solved = solved;
            total->max_iterations = max_it;
            total->sum_iterations = sum_it;
            total->loop_steps = iters;
            total->solve_ms = ms;
            total->tiled = 0;
        }
        return BO_SUCCESS;
    }

    bo_status get_solutions(double* x, double* y) override {
        if (x && n) {
            hipLaunchKernelGGL(kf_from_minor, dim3(grid_for(static_cast<size_t>(n) * K)), dim3(kBlock), 0, stream, K, n,
                               st, xo_cur, xo_cand, stage);
            BO_TRY(hipMemcpyAsync(x, stage, sizeof(double) * K * n, hipMemcpyDeviceToHost, stream));
            BO_TRY(hipStreamSynchronize(stream));
        }
        if (y && m) {
            hipLaunchKernelGGL(kf_from_minor, dim3(grid_for(static_cast<size_t>(m) * K)), dim3(kBlock), 0, stream, K, m,
                               st, yo_cur, yo_cand, stage);
            BO_TRY(hipMemcpyAsync(y, stage, sizeof(double) * K * m, hipMemcpyDeviceToHost, stream));
            BO_TRY(hipStreamSynchronize(stream));
        }
        BO_TRY(hipGetLastError());
        return BO_SUCCESS;
    }

    // Mean time of one primal (CSC) and one dual (CSR) product pass over the
    // state of the last solve (HIP events on the engine's stream); restores
    // nothing, so call it after the solutions have been read.
    bo_status time_passes(int reps, double* ms2) override {
        hipEvent_t e0, e1, e2;
        BO_TRY(hipEventCreate(&e0));
        BO_TRY(hipEventCreate(&e1));
        BO_TRY(hipEventCreate(&e2));
        hipLaunchKernelGGL(kf_coef, dim3(1), dim3(kL), 0, stream, st, cf, running);
        // every lane live for the timing (the passes' full work)
        std::vector<PdlpState> keep(kL);
        BO_TRY(hipMemcpyAsync(keep.data(), st, sizeof(PdlpState) * kL, hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        std::vector<PdlpState> live(keep);
        for (int k = 0; k < K; ++k) live[k].done = 0;
        BO_TRY(hipMemcpyAsync(st, live.data(), sizeof(PdlpState) * kL, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(kf_coef, dim3(1), dim3(kL), 0, stream, st, cf, running);
        float t1 = 0.0f, t2 = 0.0f;
        double a = 0.0, b = 0.0;
        for (int r = 0; r < reps; ++r) {
            BO_TRY(hipEventRecord(e0, stream));
            launch_primal(0);
            BO_TRY(hipEventRecord(e1, stream));
            launch_dual(0);
            BO_TRY(hipEventRecord(e2, stream));
            BO_TRY(hipEventSynchronize(e2));
            BO_TRY(hipEventElapsedTime(&t1, e0, e1));
            BO_TRY(hipEventElapsedTime(&t2, e1, e2));
            a += t1;
            b += t2;
        }
        BO_TRY(hipMemcpyAsync(st, keep.data(), sizeof(PdlpState) * kL, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(kf_coef, dim3(1), dim3(kL), 0, stream, st, cf, running);
        BO_TRY(hipStreamSynchronize(stream));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        hipEventDestroy(e2);
        ms2[0] = a / reps;
        ms2[1] = b / reps;
        return BO_SUCCESS;
    }
};

}  // namespace
}  // namespace bo

// ---------------------------------------------------------------- C ABI
// The handle keeps a copy of the base LP and of the LPs' costs and
// right-hand sides, for the FP64 polish of LPs the FP32 engine stops as
// STALLED: one FP64 engine over the same matrix, built on first use.
struct bo_pdlp_family {
    std::unique_ptr<bo::FamilyBase> eng;
    std::unique_ptr<bo::FamilyBase> pol;
    bo_pdlp_params prm{};
    int m = 0, n = 0;
    std::vector<int> rp, ci;
    std::vector<double> val, lb, ub, c, b;  // c: count x n, b: count x m
    std::vector<unsigned char> eq;
    std::vector<int> pol_of;  // LP -> its lane in pol, or -1 (last solve)
    bool polishing() const { return !prm.fp64_iterates && prm.fp64_polish && prm.stall_iterations > 0; }
    bo_lp base() const {
        bo_lp v{};
        v.m = m;
        v.n = n;
        v.nnz = static_cast<long long>(val.size());
        v.row_ptr = rp.data();
        v.col_idx = ci.data();
        v.val = val.data();
        v.b = b.data();
        v.c = c.data();
        v.lb = lb.data();
        v.ub = ub.data();
        v.is_eq = eq.data();
        return v;
    }
};

using namespace bo;

namespace {
template <class T>
bo_status make_family(const bo_lp& lp, int capacity, const bo_pdlp_params& p, std::unique_ptr<FamilyBase>& out) {
    auto e = std::make_unique<Family<T>>();
    bo_status s = e->create(lp, capacity, p);
    if (s != BO_SUCCESS) return s;
    out = std::move(e);
    return BO_SUCCESS;
}
}  // namespace

extern "C" {

bo_status bo_pdlp_family_create(const bo_lp* lp, int capacity, const bo_pdlp_params* params, bo_pdlp_family** out) {
    return host_guard([&]() -> bo_status {
        if (!out) return BO_INVALID_ARGUMENT;
        *out = nullptr;
        if (!valid_csr_lp(lp) || capacity <= 0 || capacity > kL || lp->m <= 0 || lp->n <= 0 || lp->nnz <= 0 ||
            lp->nnz >= (1LL << 31) - 1)
            return BO_INVALID_ARGUMENT;
        for (long long q = 0; q < lp->nnz; ++q)
            if (!std::isfinite(lp->val[q])) return BO_INVALID_ARGUMENT;
        for (int j = 0; j < lp->n; ++j)
            if (!(lp->lb[j] <= lp->ub[j]) || !std::isfinite(lp->c[j])) return BO_INVALID_ARGUMENT;
        if (static_cast<long long>(std::max(lp->m, lp->n)) * kL >= (1LL << 31)) return BO_INVALID_ARGUMENT;
        bo_status s = ensure_device_public();
        if (s != BO_SUCCESS) return s;
        bo_pdlp_params p;
        if (params) p = *params;
        else bo_pdlp_default_params(&p);
        auto h = std::make_unique<bo_pdlp_family>();
        s = p.fp64_iterates ? make_family<double>(*lp, capacity, p, h->eng) : make_family<float>(*lp, capacity, p, h->eng);
        if (s != BO_SUCCESS) return s;
        h->prm = p;
        h->m = lp->m;
        h->n = lp->n;
        h->rp.assign(lp->row_ptr, lp->row_ptr + lp->m + 1);
        h->ci.assign(lp->col_idx, lp->col_idx + lp->nnz);
        h->val.assign(lp->val, lp->val + lp->nnz);
        h->lb.assign(lp->lb, lp->lb + lp->n);
        h->ub.assign(lp->ub, lp->ub + lp->n);
        h->eq.assign(lp->is_eq, lp->is_eq + lp->m);
        h->c.clear();
        h->b.clear();
        for (int k = 0; k < capacity; ++k) {
            h->c.insert(h->c.end(), lp->c, lp->c + lp->n);
            h->b.insert(h->b.end(), lp->b, lp->b + lp->m);
        }
        h->pol_of.assign(capacity, -1);
        *out = h.release();
        return BO_SUCCESS;
    });
}

bo_status bo_pdlp_family_set_lps(bo_pdlp_family* h, int count, const double* c, const double* b) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        bo_status s = h->eng->set_lps(count, c, b);
        if (s != BO_SUCCESS) return s;
        if (c) {
            h->c.assign(c, c + static_cast<size_t>(count) * h->n);
            h->b.assign(b, b + static_cast<size_t>(count) * h->m);
        }
        h->pol_of.assign(count, -1);
        return BO_SUCCESS;
    });
}

bo_status bo_pdlp_family_set_warm_start(bo_pdlp_family* h, const double* x, const double* y) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        return h->eng->set_warm_start(x, y);
    });
}

bo_status bo_pdlp_family_set_primal_weights(bo_pdlp_family* h, const double* w) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        return h->eng->set_primal_weights(w);
    });
}

bo_status bo_pdlp_family_solve(bo_pdlp_family* h, bo_pdlp_stats* per_lp, bo_pdlp_batch_stats* total) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        auto t0 = std::chrono::high_resolution_clock::now();
        const int K = h->eng->count();
        std::vector<bo_pdlp_stats> per(K);
        bo_pdlp_batch_stats tot{};
        h->pol_of.assign(K, -1);
        const Deadline dl = Deadline::after(h->prm.time_limit_s);
        h->eng->set_deadline(dl);
        bo_status s = h->eng->solve(per.data(), &tot);
        if (s != BO_SUCCESS) return s;
        std::vector<int> stalled;
        if (h->polishing())
            for (int k = 0; k < K; ++k)
                if (per[k].status == 4 && per[k].iterations < h->prm.max_iterations) stalled.push_back(k);
        if (!stalled.empty() && dl.passed()) {  // no time left for the polish: stopped by the limit
            for (int k : stalled) per[k].status = 3;
            stalled.clear();
        }
        if (!stalled.empty()) {
            auto tp0 = std::chrono::high_resolution_clock::now();
            const int P = static_cast<int>(stalled.size());
            if (!h->pol) {
                bo_pdlp_params p = h->prm;
                p.fp64_iterates = 1;
                std::vector<double> c0(h->c.begin(), h->c.begin() + h->n), b0(h->b.begin(), h->b.begin() + h->m);
                bo_lp v = h->base();
                v.c = c0.data();
                v.b = b0.data();
                if ((s = make_family<double>(v, kL, p, h->pol)) != BO_SUCCESS) return s;
            }
            std::vector<double> X(static_cast<size_t>(K) * h->n), Y(static_cast<size_t>(K) * h->m);
            if ((s = h->eng->get_solutions(X.data(), Y.data())) != BO_SUCCESS) return s;
            std::vector<double> pc, pb, px, py, w(P);
            std::vector<long long> lim(P);
            for (int j = 0; j < P; ++j) {
                const int k = stalled[j];
                const size_t on = static_cast<size_t>(k) * h->n, om = static_cast<size_t>(k) * h->m;
                pc.insert(pc.end(), h->c.begin() + on, h->c.begin() + on + h->n);
                pb.insert(pb.end(), h->b.begin() + om, h->b.begin() + om + h->m);
                px.insert(px.end(), X.begin() + on, X.begin() + on + h->n);
                py.insert(py.end(), Y.begin() + om, Y.begin() + om + h->m);
                w[j] = per[k].primal_weight;
                lim[j] = h->prm.max_iterations - per[k].iterations;
            }
            FamilyBase* f = h->pol.get();
            if ((s = f->set_lps(P, pc.data(), pb.data())) != BO_SUCCESS) return s;
            f->set_iteration_limits(lim.data());
            f->set_deadline(dl);
            if ((s = f->set_primal_weights(w.data())) != BO_SUCCESS) return s;
            if ((s = f->set_warm_start(px.data(), py.data())) != BO_SUCCESS) return s;
            std::vector<bo_pdlp_stats> per2(P);
            bo_pdlp_batch_stats tot2{};
            if ((s = f->solve(per2.data(), &tot2)) != BO_SUCCESS) return s;
            for (int j = 0; j < P; ++j) {
                const int k = stalled[j];
                const long long fp32_its = per[k].iterations;
                per[k] = per2[j];
                per[k].iterations += fp32_its;
                per[k].fp64_polish_iterations = per2[j].iterations;
                h->pol_of[k] = j;
            }
            tot.polished = P;
            tot.tail_lps = P;
            tot.tail_ms =
                std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - tp0).count();
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
        tot.solved = 0;
        tot.max_iterations = 0;
        tot.sum_iterations = 0;
        for (int k = 0; k < K; ++k) {
            per[k].solve_ms = ms;
            tot.solved += per[k].status == 0;
            tot.max_iterations = std::max(tot.max_iterations, per[k].iterations);
            tot.sum_iterations += per[k].iterations;
        }
        tot.solve_ms = ms;
        if (per_lp) std::copy(per.begin(), per.end(), per_lp);
        if (total) *total = tot;
        return BO_SUCCESS;
    });
}

bo_status bo_pdlp_family_get_solutions(bo_pdlp_family* h, double* x, double* y) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        const int K = h->eng->count();
        bo_status s = h->eng->get_solutions(x, y);
        if (s != BO_SUCCESS) return s;
        bool any = false;
        for (int k = 0; k < K; ++k) any = any || h->pol_of[k] >= 0;
        if (!any || !h->pol) return BO_SUCCESS;
        const int P = h->pol->count();
        std::vector<double> X(static_cast<size_t>(P) * h->n), Y(static_cast<size_t>(P) * h->m);
        if ((s = h->pol->get_solutions(x ? X.data() : nullptr, y ? Y.data() : nullptr)) != BO_SUCCESS) return s;
        for (int k = 0; k < K; ++k) {
            const int j = h->pol_of[k];
            if (j < 0) continue;
            if (x) std::copy(X.begin() + static_cast<size_t>(j) * h->n, X.begin() + static_cast<size_t>(j + 1) * h->n,
                             x + static_cast<size_t>(k) * h->n);
            if (y) std::copy(Y.begin() + static_cast<size_t>(j) * h->m, Y.begin() + static_cast<size_t>(j + 1) * h->m,
                             y + static_cast<size_t>(k) * h->m);
        }
        return BO_SUCCESS;
    });
}

bo_status bo_pdlp_family_time_passes(bo_pdlp_family* h, int reps, double* ms2) {
    return host_guard([&]() -> bo_status {
        if (!h || reps <= 0 || !ms2) return BO_INVALID_ARGUMENT;
        return h->eng->time_passes(reps, ms2);
    });
}

void bo_pdlp_family_destroy(bo_pdlp_family* h) { delete h; }

}  // extern "C"
