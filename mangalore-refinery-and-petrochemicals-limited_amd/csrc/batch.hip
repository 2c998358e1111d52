// batch.hip -- many LPs of uneven size solved together by the restarted,
// reflected Halpern PDHG of pdlp.hip.
//
// The LPs are laid out as one block-diagonal LP: rows of LP k follow those of
// LP k-1, and so do its columns.  Every launch covers all LPs at once.
//  * The two iteration passes run on the tiled layout (tiles.hpp).  A tile
//    that holds rows or columns of finished LPs only is skipped.
//  * The epilogues read the state of the row's or column's own LP (inst maps).
//  * Scaling is row- and column-local, so it is the same as scaling each LP
//    on its own.
//  * Power iteration, norms and the FP64 checks reduce over "chunks": up to
//    256 consecutive rows or columns of one LP.  Each LP sums its chunk
//    partials in a fixed order, in one wavefront, and then runs the same
//    stopping test, restart rule and primal-weight update as the single-LP
//    engine (pdlp_core.hpp).
// One interval of steps, its check and its restarts are captured once as a
// hipGraph and replayed until every LP has stopped.  The replacement for the
// reference's one-model-at-a-time FirstOrderLPSolver::solve
// (src/first_order_solver.cpp:311-319) when many independent LPs are solved.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <vector>

#include "common.hpp"
#include "pdlp_core.hpp"
#include "segments.hpp"
#include "tiles.hpp"

namespace bo {

bo_status ensure_device_public();

namespace {

// f(k) for k in [0, count) on up to 16 host threads (LPs handed out one at a
// time, so large and small LPs balance).
template <class F>
void parallel_for_lps(int count, F&& f) {
    const int hw = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    const int nt = std::min(std::min(hw, 16), count);
    if (nt <= 1) {
        for (int k = 0; k < count; ++k) f(k);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    pool.reserve(static_cast<size_t>(nt));
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&] {
            for (int k = next++; k < count; k = next++) f(k);
        });
    for (auto& th : pool) th.join();
}

// Up to kBlock consecutive rows or columns of one LP.
struct Chunk {
    int inst;
    int begin;
    int end;
};

// ---------------------------------------------------------------- epilogues
// norm_out[seg] = own_scale[seg] * acc   (row/column norms of the scaled matrix)
struct EpiNormB {
    using TXv = double;
    static constexpr int NV = 1;
    static constexpr int Q = 0;
    const double* other_scale;
    const double* own_scale;
    double* norm_out;
    __device__ const double* vec(int) const { return other_scale; }
    __device__ void operator()(int seg, const double* acc, double*) const { norm_out[seg] = own_scale[seg] * acc[0]; }
};

// out[k] = own_scale[seg] * val[k] * other_scale[j]
template <class TO>
struct EpiWriteB {
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

// out = A v
template <class TX, class TO>
struct EpiStoreB {
    using TXv = TX;
    static constexpr int NV = 1;
    static constexpr int Q = 0;
    const TX* v;
    TO* out;
    __device__ const TX* vec(int) const { return v; }
    __device__ void operator()(int seg, const double* acc, double*) const { out[seg] = static_cast<TO>(acc[0]); }
};

// o0 = A v0, o1 = A v1 (the FP64 check products)
struct EpiStore2 {
    using TXv = double;
    static constexpr int NV = 2;
    static constexpr int Q = 0;
    const double* v0;
    const double* v1;
    double* o0;
    double* o1;
    const unsigned char* wg_live;  // 0 for a workgroup whose LPs all stopped
    __device__ const double* vec(int v) const { return v ? v1 : v0; }
    __device__ bool wg_dead(int b) const { return !wg_live[b]; }
    __device__ void operator()(int seg, const double* acc, double*) const {
        o0[seg] = acc[0];
        o1[seg] = acc[1];
    }
};

// ---------------------------------------------------------------- elementwise
__global__ void kb_fill_f64(double* p, double v, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void kb_apply_norm(double* scale, const double* norm, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const double v = norm[i];
        if (v > 0.0 && isfinite(v)) scale[i] /= sqrt(v);
    }
}

template <class T>
__global__ void kb_scale_cols(int n, const double* s, const double* c, const double* lb, const double* ub,
                              const double* wx, int has_warm, T* c_hat, T* l_hat, T* u_hat, double* xk, double* x0,
                              T* xp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double sj = s[j];
    c_hat[j] = static_cast<T>(c[j] * sj);
    const double lo = lb[j] / sj, hi = ub[j] / sj;
    l_hat[j] = static_cast<T>(lo);
    u_hat[j] = static_cast<T>(hi);
    const double xs = clampT(has_warm ? wx[j] / sj : 0.0, lo, hi);
    xk[j] = xs;
    x0[j] = xs;
    xp[j] = static_cast<T>(xs);
}

template <class T>
__global__ void kb_scale_rows(int m, const double* r, const double* b, const unsigned char* eq, const double* wy,
                              int has_warm, T* b_hat, double* yk, double* y0, T* yg, T* yp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
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

__global__ void kb_copy_f64(double* dst, const double* src, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <class T>
__global__ void kb_pw_init(T* v, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    unsigned h = static_cast<unsigned>(j) * 2654435761u + 12345u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    h ^= h >> 15;
    v[j] = static_cast<T>(0.5 + static_cast<double>(h & 0xffffu) / 65536.0);
}

// part[c] = sum of v^2 over chunk c
template <class T>
__global__ void __launch_bounds__(kBlock) kb_sq_chunks(const Chunk* chunks, const T* v, double* part) {
    __shared__ double sh[kBlock / kWave];
    const Chunk ch = chunks[blockIdx.x];
    const int i = ch.begin + static_cast<int>(threadIdx.x);
    double t = 0.0;
    if (i < ch.end) {
        const double a = static_cast<double>(v[i]);
        t = a * a;
    }
    const double s = block_sum(t, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[k] = sum of the Q-wide chunk partials of LP k (slot q), chunks in order;
// one wavefront per LP.
__global__ void kb_inst_sum(int K, const int* chunk_ptr, const double* part, int Q, int q, double* out) {
    const int k = blockIdx.x * (blockDim.x / kWave) + static_cast<int>(threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    if (k >= K) return;
    double a = 0.0;
    for (int c = chunk_ptr[k] + lane; c < chunk_ptr[k + 1]; c += kWave) a += part[static_cast<size_t>(c) * Q + q];
    a = wave_sum(a);
    if (lane == 0) out[k] = a;
}

template <class T>
__global__ void kb_pw_normalize(T* v, int n, const int* inst, const double* norm2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double nv = sqrt(norm2[inst[j]]);
    if (nv > 0.0) v[j] = static_cast<T>(static_cast<double>(v[j]) / nv);
}

__global__ void kb_init_state(int K, PdlpState* st, const PdlpState* init, const double* pw_u2, const double* c2,
                              const double* b2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    st[k] = init[k];
    init_state_core(st + k, pw_u2[k], c2[k], b2[k]);
}

__global__ void kb_advance(int K, PdlpState* st, int steps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K && !st[k].done) st[k].iter += steps;
}

// ---------------------------------------------------------------- FP64 check
// Unscaled x: xo = clamp(s x, l, u) for the iterate and for x+; partials
// [2c, 2c+1] = c'xo over chunk c (0 for a finished LP).
template <class T>
__global__ void __launch_bounds__(kBlock)
kb_check_x(const Chunk* chunks, const PdlpState* st, const double* xk, const T* xp, const double* s, const double* c,
           const double* lb, const double* ub, double* xo_cur, double* xo_cand, double* part) {
    __shared__ double sh[kBlock / kWave];
    const Chunk ch = chunks[blockIdx.x];
    const bool live = !st[ch.inst].done;
    const int j = ch.begin + static_cast<int>(threadIdx.x);
    double p0 = 0.0, p1 = 0.0;
    if (live && j < ch.end) {
        const double sj = s[j], l = lb[j], u = ub[j];
        const double xc = clampT(sj * xk[j], l, u);
        const double xa = clampT(sj * static_cast<double>(xp[j]), l, u);
        xo_cur[j] = xc;
        xo_cand[j] = xa;
        p0 = c[j] * xc;
        p1 = c[j] * xa;
    }
    const double t0 = block_sum(p0, sh);
    const double t1 = block_sum(p1, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x + 0] = t0;
        part[2 * blockIdx.x + 1] = t1;
    }
}

// Row side, from the products a0 = A xo_cur, a1 = A xo_cand (unscaled FP64
// matrix).  partials per chunk: [0,1] ||viol||^2  [2,3] scaled  [4,5] b'y.
template <class T>
__global__ void __launch_bounds__(kBlock)
kb_check_rows(const Chunk* chunks, const PdlpState* st, const double* ar0, const double* ar1, const double* b,
              const unsigned char* eq, const double* rs, const double* yk, const T* yp, double* yo_cur,
              double* yo_cand, double* ax_cand, double* part) {
    __shared__ double sh[kBlock / kWave];
    const Chunk ch = chunks[blockIdx.x];
    const bool live = !st[ch.inst].done;
    const int i = ch.begin + static_cast<int>(threadIdx.x);
    double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (live && i < ch.end) {
        const double bi = b[i], ri = rs[i];
        const bool e = eq[i] != 0;
        const double acc[2] = {ar0[i], ar1[i]};
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double r = acc[p] - bi;
            const double v = e ? fabs(r) : (r > 0.0 ? r : 0.0);
            q[p] = v * v;
            q[2 + p] = (ri * v) * (ri * v);
        }
        ax_cand[i] = ri * acc[1];
        double yc = ri * yk[i];
        if (!e && yc < 0.0) yc = 0.0;
        const double ya = ri * static_cast<double>(yp[i]);
        yo_cur[i] = yc;
        yo_cand[i] = ya;
        q[4] = bi * yc;
        q[5] = bi * ya;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double t = block_sum(q[k], sh);
        if (threadIdx.x == 0) part[6 * blockIdx.x + k] = t;
    }
}

// Column side, from the products a0 = A'yo_cur, a1 = A'yo_cand.
// partials per chunk: [0,1] ||dual res||^2  [2,3] scaled  [4,5] sum l plam+ + u plam-
__global__ void __launch_bounds__(kBlock)
kb_check_cols(const Chunk* chunks, const PdlpState* st, const double* ac0, const double* ac1, const double* c,
              const double* lb, const double* ub, const double* cs, double* part) {
    __shared__ double sh[kBlock / kWave];
    const Chunk ch = chunks[blockIdx.x];
    const bool live = !st[ch.inst].done;
    const int j = ch.begin + static_cast<int>(threadIdx.x);
    double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (live && j < ch.end) {
        const double l = lb[j], u = ub[j], sj = cs[j];
        const bool lf = isfinite(l), uf = isfinite(u);
        const double acc[2] = {ac0[j], ac1[j]};
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double lam = c[j] + acc[p];
            double pl;
            if (lf && uf) pl = lam;
            else if (lf) pl = lam > 0.0 ? lam : 0.0;
            else if (uf) pl = lam < 0.0 ? lam : 0.0;
            else pl = 0.0;
            const double d = lam - pl;
            q[p] = d * d;
            q[2 + p] = (sj * d) * (sj * d);
            q[4 + p] = pl > 0.0 ? l * pl : (pl < 0.0 ? u * pl : 0.0);
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double t = block_sum(q[k], sh);
        if (threadIdx.x == 0) part[6 * blockIdx.x + k] = t;
    }
}

// One wavefront per LP: its check sums (chunks in order, fixed lane tree),
// then the stopping test and restart rule of the single-LP engine.
__global__ void kb_decide(int K, PdlpState* st, const int* col_ptr, const double* px, const double* pc,
                          const int* row_ptr, const double* pr) {
    const int k = blockIdx.x * (blockDim.x / kWave) + static_cast<int>(threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    if (k >= K || st[k].done) return;
    double a[14];
#pragma unroll
    for (int q = 0; q < 14; ++q) a[q] = 0.0;
    for (int c = col_ptr[k] + lane; c < col_ptr[k + 1]; c += kWave) {
        a[0] += px[2 * c + 0];
        a[1] += px[2 * c + 1];
#pragma unroll
        for (int q = 0; q < 6; ++q) a[8 + q] += pc[6 * static_cast<size_t>(c) + q];
    }
    for (int c = row_ptr[k] + lane; c < row_ptr[k + 1]; c += kWave) {
#pragma unroll
        for (int q = 0; q < 6; ++q) a[2 + q] += pr[6 * static_cast<size_t>(c) + q];
    }
#pragma unroll
    for (int q = 0; q < 14; ++q) a[q] = wave_sum(a[q]);
    if (lane == 0) decide_core(st + k, a);
}

// Restart of the LPs that decided to: move to the chosen point, make it the
// anchor; per-chunk ||dx||^2 / ||dy||^2 of the anchor move.
template <class T>
__global__ void __launch_bounds__(kBlock)
kb_restart_x(const Chunk* chunks, const PdlpState* st, double* xk, double* x0, const T* xp, double* part) {
    __shared__ double sh[kBlock / kWave];
    const Chunk ch = chunks[blockIdx.x];
    const PdlpState* s = st + ch.inst;
    const bool go = !s->done && s->restart;
    const int j = ch.begin + static_cast<int>(threadIdx.x);
    double d2 = 0.0;
    if (go && j < ch.end) {
        const double xn = s->restart_to_cand ? static_cast<double>(xp[j]) : xk[j];
        const double d = xn - x0[j];
        d2 = d * d;
        xk[j] = xn;
        x0[j] = xn;
    }
    const double t = block_sum(d2, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
kb_restart_y(const Chunk* chunks, const PdlpState* st, double* yk, double* y0, const T* yp, T* yg, double* ax,
             double* ax0, const double* ax_cand, double* part) {
    __shared__ double sh[kBlock / kWave];
    const Chunk ch = chunks[blockIdx.x];
    const PdlpState* s = st + ch.inst;
    const bool go = !s->done && s->restart;
    const int i = ch.begin + static_cast<int>(threadIdx.x);
    double d2 = 0.0;
    if (go && i < ch.end) {
        double yn = yk[i];
        if (s->restart_to_cand) {
            yn = static_cast<double>(yp[i]);
            ax[i] = ax_cand[i];
        }
        const double d = yn - y0[i];
        d2 = d * d;
        yk[i] = yn;
        y0[i] = yn;
        yg[i] = static_cast<T>(yn);
        ax0[i] = ax[i];
    }
    const double t = block_sum(d2, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void kb_restart_finish(int K, PdlpState* st, const int* col_ptr, const double* pdx, const int* row_ptr,
                                  const double* pdy) {
    const int k = blockIdx.x * (blockDim.x / kWave) + static_cast<int>(threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    if (k >= K || st[k].done || !st[k].restart) return;
    double a = 0.0, b = 0.0;
    for (int c = col_ptr[k] + lane; c < col_ptr[k + 1]; c += kWave) a += pdx[c];
    for (int c = row_ptr[k] + lane; c < row_ptr[k + 1]; c += kWave) b += pdy[c];
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) restart_finish_core(st + k, a, b);
}

// Tile flags for the next interval (1 while any of its LPs runs) and the
// number of LPs still running.
__global__ void kb_tile_active(int ntiles, const TileDesc* tiles, const int* inst, const PdlpState* st,
                               unsigned char* active) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int k0 = inst[tiles[t].seg_lo], k1 = inst[tiles[t].seg_hi];
    unsigned char a = 0;
    for (int k = k0; k <= k1; ++k) a |= st[k].done ? 0 : 1;
    active[t] = a;
}

// Workgroup flags: workgroup b of a launch that gives `per` consecutive
// segments to each workgroup is live while any LP among its segments runs
// (segments are in LP order, so its LPs are inst[first] .. inst[last]).
__global__ void kb_wg_live(int nwg, int per, int nseg, const int* inst, const PdlpState* st, unsigned char* flags) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nwg) return;
    const int s0 = b * per;
    const int s1 = min(nseg, s0 + per) - 1;
    unsigned char a = 0;
    if (s0 <= s1)
        for (int k = inst[s0]; k <= inst[s1] && !a; ++k) a = st[k].done ? 0 : 1;
    flags[b] = a;
}

__global__ void kb_count_running(int K, const PdlpState* st, int* running) {
    __shared__ double sh[kBlock / kWave];
    double c = 0.0;
    for (int k = threadIdx.x; k < K; k += blockDim.x) c += st[k].done ? 0.0 : 1.0;
    const double t = block_sum(c, sh);
    if (threadIdx.x == 0) *running = static_cast<int>(t);
}

// ---------------------------------------------------------------- host side
struct DevMemB {
    std::vector<void*> ptrs;
    ~DevMemB() {
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

// One bin holding every segment with L lanes each: segment order is free.
BinTable one_bin(int nseg, int L) {
    BinTable bt{};
    bt.nbins = 1;
    bt.lanes[0] = L;
    bt.seg_start[0] = 0;
    bt.seg_start[1] = nseg;
    bt.block_start[0] = 0;
    const int per = L <= kWave ? kBlock / L : 1;
    bt.block_start[1] = std::max(1, (nseg + per - 1) / per);
    bt.total_blocks = bt.block_start[1];
    return bt;
}

std::vector<Chunk> make_chunks(const std::vector<int>& off, std::vector<int>& ptr) {
    std::vector<Chunk> ch;
    const int K = static_cast<int>(off.size()) - 1;
    ptr.assign(K + 1, 0);
    for (int k = 0; k < K; ++k) {
        ptr[k] = static_cast<int>(ch.size());
        for (int b = off[k]; b < off[k + 1]; b += kBlock) ch.push_back(Chunk{k, b, std::min(off[k + 1], b + kBlock)});
    }
    ptr[K] = static_cast<int>(ch.size());
    if (ch.empty()) ch.push_back(Chunk{0, 0, 0});  // keeps launches valid
    return ch;
}

struct BatchBase {
    virtual ~BatchBase() = default;
    virtual bo_status set_warm_start(const double* x, const double* y) = 0;
    virtual bo_status set_primal_weights(const double* w) = 0;
    virtual bo_status set_objective_rhs(const double* c, const double* b) = 0;
    virtual void set_iteration_limits(const long long* lim) = 0;
    virtual void set_stop_running(int r) = 0;
    virtual void set_deadline(const Deadline& d) = 0;
    virtual bo_status time_kernels(int reps, double* ms4, double* bytes4) = 0;
    virtual bo_status solve(bo_pdlp_stats* per_lp, bo_pdlp_batch_stats* total) = 0;
    virtual bo_status get_solution(int k, double* x, double* y) = 0;
};

template <class T>
struct Batch final : BatchBase {
    int K = 0, m = 0, n = 0;
    long long nnz = 0;
    bo_pdlp_params prm{};
    hipStream_t stream = nullptr;
    DevMemB mem;
    std::vector<int> roff, coff;  // row / column offset of every LP (K + 1)
    // FP64 matrix (CSR and CSC), scaled working values
    int *rp = nullptr, *ci = nullptr, *cp = nullptr, *ri = nullptr;
    double *rv64 = nullptr, *cv64 = nullptr;
    T *rv = nullptr, *cv = nullptr;
    double *b64 = nullptr, *c64 = nullptr, *lb64 = nullptr, *ub64 = nullptr;
    unsigned char* eq = nullptr;
    int *inst_r = nullptr, *inst_c = nullptr;
    double *rs = nullptr, *cs = nullptr, *rnorm = nullptr, *cnorm = nullptr;
    T *b_hat = nullptr, *c_hat = nullptr, *l_hat = nullptr, *u_hat = nullptr;
    double *xk = nullptr, *x0 = nullptr, *yk = nullptr, *y0 = nullptr, *ax = nullptr, *ax0 = nullptr;
    T *xp = nullptr, *yp = nullptr, *yg = nullptr, *pw_v = nullptr, *pw_u = nullptr;
    double *xo_cur = nullptr, *xo_cand = nullptr, *yo_cur = nullptr, *yo_cand = nullptr, *ax_cand = nullptr;
    double *wx = nullptr, *wy = nullptr;
    double *ar0 = nullptr, *ar1 = nullptr, *ac0 = nullptr, *ac1 = nullptr;  // check products
    int has_warm = 0;
    // chunks
    Chunk *rch = nullptr, *cch = nullptr;
    int nrch = 0, ncch = 0;
    int *rch_ptr = nullptr, *cch_ptr = nullptr;
    double *part_x = nullptr, *part_r = nullptr, *part_c = nullptr, *part_dx = nullptr, *part_dy = nullptr;
    double *part_sq = nullptr, *inst_a = nullptr, *inst_b = nullptr, *inst_c2 = nullptr;
    PdlpState* st = nullptr;
    PdlpState* st_host = nullptr;
    PdlpState* st_init = nullptr;
    int* running = nullptr;
    int* running_host = nullptr;
    BinTable bt_r{}, bt_c{};
    // live flags of the workgroups of the reduce (256 segments each) and
    // check-product (bt_r / bt_c) launches
    unsigned char *live_red_r = nullptr, *live_red_c = nullptr, *live_chk_r = nullptr, *live_chk_c = nullptr;
    // tiled passes
    struct TilePass {
        bool built = false;
        int B = 0, W = 0, ntiles = 0, nslices = 0;
        long long pairs = 0;
        int max_pairs = 0;  // most partials a segment adds
        size_t padded = 0;
        TileDesc* td = nullptr;
        unsigned* words = nullptr;
        unsigned short* idx = nullptr;
        int* src = nullptr;
        T* val = nullptr;
        int* seg_ptr = nullptr;
        unsigned* slot = nullptr;
        double* part = nullptr;
        unsigned char* active = nullptr;
    } tp[2];
    int num_cus = 1, lds_max = 0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_k = 0;
    int stop_running = 0;
    Deadline deadline;  // unset: time_limit_s counts from the start of solve()

    ~Batch() override {
        if (graph_exec) hipGraphExecDestroy(graph_exec);
        if (graph) hipGraphDestroy(graph);
        if (st_host) hipHostFree(st_host);
        if (st_init) hipHostFree(st_init);
        if (running_host) hipHostFree(running_host);
        if (stream) hipStreamDestroy(stream);
    }

    hipError_t up(void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream) : hipSuccess;
    }

    hipError_t build_tiles_pass(int pass, const int* ptr, const int* gidx) {
        TilePass& t = tp[pass];
        const int nseg = pass == 0 ? m : n, ng = pass == 0 ? n : m;
        if (nseg == 0 || ng == 0) return hipSuccess;
        const int wmax = std::min(65536, lds_max / static_cast<int>(sizeof(T))) / 64 * 64;
        if (wmax < 64) return hipSuccess;
        HostTiles h;
        if (!build_tiles(ptr, gidx, nseg, ng, wmax, 2 * num_cus, h)) return hipSuccess;
        if (h.pairs * 4 >= (1LL << 31) - 64 || static_cast<long long>(h.slice_word.size()) * kWave * 8 >= (1LL << 31) - 64)
            return hipSuccess;
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(
                                               pass == 0 ? &k_tile<T, T, kTileRows, 2> : &k_tile<T, T, kTileRows, 3>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           static_cast<int>(h.W * sizeof(T)));
        if (e != hipSuccess) return e;
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
        if ((e = mem.alloc(&t.active, h.tiles.size())) != hipSuccess) return e;
        if ((e = up(t.td, h.tiles.data(), sizeof(TileDesc) * h.tiles.size())) != hipSuccess) return e;
        if ((e = up(t.words, h.slice_word.data(), sizeof(unsigned) * h.slice_word.size())) != hipSuccess) return e;
        if ((e = up(t.idx, h.idx.data(), sizeof(unsigned short) * h.padded)) != hipSuccess) return e;
        if ((e = up(t.src, h.src.data(), sizeof(int) * h.padded)) != hipSuccess) return e;
        if ((e = up(t.seg_ptr, h.seg_ptr.data(), sizeof(int) * h.seg_ptr.size())) != hipSuccess) return e;
        if ((e = up(t.slot, h.slot.data(), sizeof(unsigned) * h.slot.size())) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        t.built = true;
        return hipSuccess;
    }

    bo_status create(const bo_lp* lps, int count, const bo_pdlp_params* p) {
        prm = *p;
        K = count;
        roff.assign(K + 1, 0);
        coff.assign(K + 1, 0);
        for (int k = 0; k < K; ++k) {
            roff[k + 1] = roff[k] + lps[k].m;
            coff[k + 1] = coff[k] + lps[k].n;
            nnz += lps[k].nnz;
        }
        m = roff[K];
        n = coff[K];
        BO_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_host), sizeof(PdlpState) * std::max(K, 1)));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&st_init), sizeof(PdlpState) * std::max(K, 1)));
        BO_TRY(hipHostMalloc(reinterpret_cast<void**>(&running_host), sizeof(int)));
        std::memset(st_host, 0, sizeof(PdlpState) * std::max(K, 1));
        // block-diagonal CSR (rows in LP order, columns shifted) and its CSC.
        // The blocks are independent -- LP k's entries fill [zoff[k],
        // zoff[k+1]) of both orientations -- so host threads build them LP by
        // LP; the arrays are the same as a serial pass over the whole matrix.
        std::vector<long long> zoff(K + 1, 0);
        for (int k = 0; k < K; ++k) zoff[k + 1] = zoff[k] + lps[k].nnz;
        std::vector<int> h_rp(m + 1, 0), h_ir(m), h_ic(n), h_cp(n + 1, 0);
        std::vector<double> hb(m), hc(n), hl(n), hu(n);
        HostVec<int> h_ci(nnz), h_ri(nnz);  // every entry written by its LP's thread
        HostVec<double> h_rv(nnz), h_cv(nnz);
        std::vector<unsigned char> he(m);
        std::vector<double> bn(K), cn(K);
        parallel_for_lps(K, [&](int k) {
            const bo_lp& lp = lps[k];
            double b2 = 0.0, c2 = 0.0;
            long long o = zoff[k];
            for (int i = 0; i < lp.m; ++i) {
                const int gi = roff[k] + i;
                for (int q = lp.row_ptr[i]; q < lp.row_ptr[i + 1]; ++q, ++o) {
                    h_ci[o] = coff[k] + lp.col_idx[q];
                    h_rv[o] = lp.val[q];
                }
                h_rp[gi + 1] = static_cast<int>(o);
                hb[gi] = lp.b[i];
                he[gi] = lp.is_eq[i];
                h_ir[gi] = k;
                b2 += lp.b[i] * lp.b[i];
            }
            for (int j = 0; j < lp.n; ++j) {
                const int gj = coff[k] + j;
                hc[gj] = lp.c[j];
                hl[gj] = lp.lb[j];
                hu[gj] = lp.ub[j];
                h_ic[gj] = k;
                c2 += lp.c[j] * lp.c[j];
            }
            bn[k] = std::sqrt(b2);
            cn[k] = std::sqrt(c2);
            // CSC of this block: columns coff[k].., rows ascending
            std::vector<int> pos(static_cast<size_t>(lp.n) + 1, 0);
            for (long long q = zoff[k]; q < zoff[k + 1]; ++q) pos[h_ci[q] - coff[k] + 1]++;
            for (int j = 0; j < lp.n; ++j) pos[j + 1] += pos[j];
            for (int j = 0; j < lp.n; ++j) h_cp[coff[k] + j + 1] = static_cast<int>(zoff[k] + pos[j + 1]);
            for (int i = 0; i < lp.m; ++i) {
                const int gi = roff[k] + i;
                // the first row's start is the previous LP's end, which another
                // thread writes: take it from zoff instead
                const long long q0 = i == 0 ? zoff[k] : h_rp[gi];
                for (long long q = q0; q < h_rp[gi + 1]; ++q) {
                    const long long d = zoff[k] + pos[h_ci[q] - coff[k]]++;
                    h_ri[d] = gi;
                    h_cv[d] = h_rv[q];
                }
            }
        });
        for (int k = 0; k < K; ++k) {
            PdlpState& s = st_init[k];
            std::memset(&s, 0, sizeof(s));
            s.b_norm = bn[k];
            s.c_norm = cn[k];
            s.eps = prm.eps_rel_kkt;
            s.eps_obj = prm.eps_objective;
            s.max_iter = prm.max_iterations;
            s.stall_window = sizeof(T) == sizeof(float) ? prm.stall_iterations : 0;
        }
        // device memory
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
        BO_TRY(mem.alloc(&inst_r, m));
        BO_TRY(mem.alloc(&rs, m));
        BO_TRY(mem.alloc(&rnorm, m));
        BO_TRY(mem.alloc(&b_hat, m));
        for (double** v : {&yk, &y0, &ax, &ax0, &ax_cand, &yo_cur, &yo_cand, &wy, &ar0, &ar1}) BO_TRY(mem.alloc(v, m));
        for (T** v : {&yp, &yg, &pw_u}) BO_TRY(mem.alloc(v, m));
        BO_TRY(mem.alloc(&c64, n));
        BO_TRY(mem.alloc(&lb64, n));
        BO_TRY(mem.alloc(&ub64, n));
        BO_TRY(mem.alloc(&inst_c, n));
        BO_TRY(mem.alloc(&cs, n));
        BO_TRY(mem.alloc(&cnorm, n));
        for (T** v : {&c_hat, &l_hat, &u_hat, &xp, &pw_v}) BO_TRY(mem.alloc(v, n));
        for (double** v : {&xk, &x0, &xo_cur, &xo_cand, &wx, &ac0, &ac1}) BO_TRY(mem.alloc(v, n));
        BO_TRY(mem.alloc(&st, std::max(K, 1)));
        // zero states (done = 0) until prepare() uploads the solve's: the
        // reduction tuner's passes read them
        BO_TRY(hipMemsetAsync(st, 0, sizeof(PdlpState) * std::max(K, 1), stream));
        BO_TRY(mem.alloc(&running, 1));
        BO_TRY(up(rp, h_rp.data(), sizeof(int) * (m + 1)));
        BO_TRY(up(ci, h_ci.data(), sizeof(int) * nnz));
        BO_TRY(up(rv64, h_rv.data(), sizeof(double) * nnz));
        BO_TRY(up(cp, h_cp.data(), sizeof(int) * (n + 1)));
        BO_TRY(up(ri, h_ri.data(), sizeof(int) * nnz));
        BO_TRY(up(cv64, h_cv.data(), sizeof(double) * nnz));
        BO_TRY(up(b64, hb.data(), sizeof(double) * m));
        BO_TRY(up(eq, he.data(), m));
        BO_TRY(up(inst_r, h_ir.data(), sizeof(int) * m));
        BO_TRY(up(c64, hc.data(), sizeof(double) * n));
        BO_TRY(up(lb64, hl.data(), sizeof(double) * n));
        BO_TRY(up(ub64, hu.data(), sizeof(double) * n));
        BO_TRY(up(inst_c, h_ic.data(), sizeof(int) * n));
        // chunks
        std::vector<int> hrp, hcp;
        const std::vector<Chunk> rc = make_chunks(roff, hrp), cc = make_chunks(coff, hcp);
        nrch = static_cast<int>(rc.size());
        ncch = static_cast<int>(cc.size());
        BO_TRY(mem.alloc(&rch, rc.size()));
        BO_TRY(mem.alloc(&cch, cc.size()));
        BO_TRY(mem.alloc(&rch_ptr, K + 1));
        BO_TRY(mem.alloc(&cch_ptr, K + 1));
        BO_TRY(up(rch, rc.data(), sizeof(Chunk) * rc.size()));
        BO_TRY(up(cch, cc.data(), sizeof(Chunk) * cc.size()));
        BO_TRY(up(rch_ptr, hrp.data(), sizeof(int) * (K + 1)));
        BO_TRY(up(cch_ptr, hcp.data(), sizeof(int) * (K + 1)));
        BO_TRY(mem.alloc(&part_x, 2 * static_cast<size_t>(ncch)));
        BO_TRY(mem.alloc(&part_c, 6 * static_cast<size_t>(ncch)));
        BO_TRY(mem.alloc(&part_r, 6 * static_cast<size_t>(nrch)));
        BO_TRY(mem.alloc(&part_dx, ncch));
        BO_TRY(mem.alloc(&part_dy, nrch));
        BO_TRY(mem.alloc(&part_sq, std::max(nrch, ncch)));
        BO_TRY(mem.alloc(&inst_a, std::max(K, 1)));
        BO_TRY(mem.alloc(&inst_b, std::max(K, 1)));
        BO_TRY(mem.alloc(&inst_c2, std::max(K, 1)));
        BO_TRY(hipStreamSynchronize(stream));
        bt_r = one_bin(m, 4);
        bt_c = one_bin(n, 4);
        BO_TRY(mem.alloc(&live_red_r, grid_for(m)));
        BO_TRY(mem.alloc(&live_red_c, grid_for(n)));
        BO_TRY(mem.alloc(&live_chk_r, bt_r.total_blocks));
        BO_TRY(mem.alloc(&live_chk_c, bt_c.total_blocks));
        int dev = 0;
        BO_TRY(hipGetDevice(&dev));
        BO_TRY(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
        BO_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        // a positive schedule code keeps that pass on the segment kernel (no tile layout to build)
        if (prm.lanes_rows <= 0) BO_TRY(build_tiles_pass(0, h_rp.data(), h_ci.data()));
        if (prm.lanes_cols <= 0) BO_TRY(build_tiles_pass(1, h_cp.data(), h_ri.data()));
        return BO_SUCCESS;
    }

    template <int MODE, class TV, class Epi>
    void rows(const TV* vals, const Epi& e) {
        hipLaunchKernelGGL((k_segments<MODE, TV, typename Epi::TXv, Epi>), dim3(bt_r.total_blocks), dim3(kBlock), 0,
                           stream, bt_r, rp, ci, vals, e, nullptr, nullptr);
    }
    template <int MODE, class TV, class Epi>
    void cols(const TV* vals, const Epi& e) {
        hipLaunchKernelGGL((k_segments<MODE, TV, typename Epi::TXv, Epi>), dim3(bt_c.total_blocks), dim3(kBlock), 0,
                           stream, bt_c, cp, ri, vals, e, nullptr, nullptr);
    }
    // per-LP sum of squares of v (rows or columns side) into out[K]
    template <class TV>
    void inst_sq(bool row_side, const TV* v, double* out) {
        hipLaunchKernelGGL(kb_sq_chunks<TV>, dim3(row_side ? nrch : ncch), dim3(kBlock), 0, stream,
                           row_side ? rch : cch, v, part_sq);
        hipLaunchKernelGGL(kb_inst_sum, dim3(grid_for(static_cast<size_t>(std::max(K, 1)) * kWave)), dim3(kBlock), 0,
                           stream, K, row_side ? rch_ptr : cch_ptr, part_sq, 1, 0, out);
    }

    bo_status prepare() {
        const unsigned gm = grid_for(m), gn = grid_for(n);
        hipLaunchKernelGGL(kb_fill_f64, dim3(gm), dim3(kBlock), 0, stream, rs, 1.0, static_cast<size_t>(m));
        hipLaunchKernelGGL(kb_fill_f64, dim3(gn), dim3(kBlock), 0, stream, cs, 1.0, static_cast<size_t>(n));
        for (int it = 0; it < prm.ruiz_iterations; ++it) {
            rows<kMaxAbs>(rv64, EpiNormB{cs, rs, rnorm});
            cols<kMaxAbs>(cv64, EpiNormB{rs, cs, cnorm});
            hipLaunchKernelGGL(kb_apply_norm, dim3(gm), dim3(kBlock), 0, stream, rs, rnorm, static_cast<size_t>(m));
            hipLaunchKernelGGL(kb_apply_norm, dim3(gn), dim3(kBlock), 0, stream, cs, cnorm, static_cast<size_t>(n));
        }
        if (prm.pock_chambolle) {
            rows<kSumAbs>(rv64, EpiNormB{cs, rs, rnorm});
            cols<kSumAbs>(cv64, EpiNormB{rs, cs, cnorm});
            hipLaunchKernelGGL(kb_apply_norm, dim3(gm), dim3(kBlock), 0, stream, rs, rnorm, static_cast<size_t>(m));
            hipLaunchKernelGGL(kb_apply_norm, dim3(gn), dim3(kBlock), 0, stream, cs, cnorm, static_cast<size_t>(n));
        }
        rows<kWrite>(rv64, EpiWriteB<T>{rs, cs, rv});
        cols<kWrite>(cv64, EpiWriteB<T>{cs, rs, cv});
        for (int pass = 0; pass < 2; ++pass) {
            const TilePass& t = tp[pass];
            if (t.built && t.padded > 0)
                hipLaunchKernelGGL(k_tile_values<T>, dim3(grid_for(t.padded)), dim3(kBlock), 0, stream, t.src,
                                   t.padded, pass == 0 ? rv : cv, t.val);
        }
        hipLaunchKernelGGL(kb_scale_cols<T>, dim3(gn), dim3(kBlock), 0, stream, n, cs, c64, lb64, ub64, wx, has_warm,
                           c_hat, l_hat, u_hat, xk, x0, xp);
        hipLaunchKernelGGL(kb_scale_rows<T>, dim3(gm), dim3(kBlock), 0, stream, m, rs, b64, eq, wy, has_warm, b_hat,
                           yk, y0, yg, yp);
        // power iteration per LP
        hipLaunchKernelGGL(kb_pw_init<T>, dim3(gn), dim3(kBlock), 0, stream, pw_v, n);
        inst_sq(false, pw_v, inst_b);
        hipLaunchKernelGGL(kb_pw_normalize<T>, dim3(gn), dim3(kBlock), 0, stream, pw_v, n, inst_c, inst_b);
        for (int it = 0; it < 40; ++it) {
            rows<kSumProd>(rv, EpiStoreB<T, T>{pw_v, pw_u});
            cols<kSumProd>(cv, EpiStoreB<T, T>{pw_u, pw_v});
            inst_sq(false, pw_v, inst_b);
            hipLaunchKernelGGL(kb_pw_normalize<T>, dim3(gn), dim3(kBlock), 0, stream, pw_v, n, inst_c, inst_b);
        }
        rows<kSumProd>(rv, EpiStoreB<T, T>{pw_v, pw_u});
        inst_sq(true, pw_u, inst_a);    // ||A_hat v||^2 with ||v|| = 1
        inst_sq(false, c_hat, inst_c2);  // ||c_hat||^2
        inst_sq(true, b_hat, inst_b);   // ||b_hat||^2
        BO_TRY(hipMemcpyAsync(st, st_init, sizeof(PdlpState) * K, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(kb_init_state, dim3(grid_for(K)), dim3(kBlock), 0, stream, K, st, st, inst_a, inst_c2,
                           inst_b);
        // A x0 (the starting point is also the first anchor)
        rows<kSumProd>(rv, EpiStoreB<T, double>{xp, ax});
        hipLaunchKernelGGL(kb_copy_f64, dim3(gm), dim3(kBlock), 0, stream, ax0, ax, static_cast<size_t>(m));
        for (int pass = 0; pass < 2; ++pass)
            if (tp[pass].built)
                hipLaunchKernelGGL(kb_tile_active, dim3(grid_for(tp[pass].ntiles)), dim3(kBlock), 0, stream,
                                   tp[pass].ntiles, tp[pass].td, pass == 0 ? inst_r : inst_c, st, tp[pass].active);
        launch_live_flags();
        BO_TRY(hipGetLastError());
        return BO_SUCCESS;
    }

    void launch_primal(int t) {
        const TilePass& tl = tp[1];
        EpiPrimal<T> e{yg, st, c_hat, l_hat, u_hat, xk, x0, xp, t, inst_c, tl.built ? live_red_c : live_chk_c};
        if (tl.built) {
            hipLaunchKernelGGL((k_tile<T, T, kTileRows, 3>), dim3(tl.ntiles), dim3(kTileThreads),
                               static_cast<unsigned>(tl.W * sizeof(T)), stream,
                               static_cast<const unsigned char*>(tl.active), tl.td, tl.words, tl.val, tl.idx, yg, m,
                               tl.W, tl.part);
            launch_reduce(tl, n, e);
        } else {
            cols<kSumProd>(cv, e);
        }
    }
    void launch_dual(int t) {
        const TilePass& tl = tp[0];
        EpiDual<T> e{xp, st, b_hat, eq, ax, ax0, yk, y0, yp, yg, t, inst_r, tl.built ? live_red_r : live_chk_r};
        if (tl.built) {
            hipLaunchKernelGGL((k_tile<T, T, kTileRows, 2>), dim3(tl.ntiles), dim3(kTileThreads),
                               static_cast<unsigned>(tl.W * sizeof(T)), stream,
                               static_cast<const unsigned char*>(tl.active), tl.td, tl.words, tl.val, tl.idx, xp, n,
                               tl.W, tl.part);
            launch_reduce(tl, m, e);
        } else {
            rows<kSumProd>(rv, e);
        }
    }
    void launch_live_flags() {
        const int pr = kBlock, pc = kBlock, qr = kBlock / 4, qc = kBlock / 4;  // one_bin(., 4): 64 segments
        hipLaunchKernelGGL(kb_wg_live, dim3(grid_for(grid_for(m))), dim3(kBlock), 0, stream,
                           static_cast<int>(grid_for(m)), pr, m, inst_r, st, live_red_r);
        hipLaunchKernelGGL(kb_wg_live, dim3(grid_for(grid_for(n))), dim3(kBlock), 0, stream,
                           static_cast<int>(grid_for(n)), pc, n, inst_c, st, live_red_c);
        hipLaunchKernelGGL(kb_wg_live, dim3(grid_for(bt_r.total_blocks)), dim3(kBlock), 0, stream,
                           bt_r.total_blocks, qr, m, inst_r, st, live_chk_r);
        hipLaunchKernelGGL(kb_wg_live, dim3(grid_for(bt_c.total_blocks)), dim3(kBlock), 0, stream,
                           bt_c.total_blocks, qc, n, inst_c, st, live_chk_c);
    }

    template <class Epi>
    void launch_reduce(const TilePass& tl, int nseg, const Epi& e, const hipEvent_t* ev = nullptr) {
        launch_tile_reduce(stream, tl.max_pairs, nseg, tl.seg_ptr, tl.slot, static_cast<int>(tl.pairs), tl.part,
                           tl.nslices * kWave, e, 0, ev);
    }

    void launch_check() {
        const unsigned gw = grid_for(static_cast<size_t>(std::max(K, 1)) * kWave);
        hipLaunchKernelGGL(kb_check_x<T>, dim3(ncch), dim3(kBlock), 0, stream, cch, st, xk, xp, cs, c64, lb64, ub64,
                           xo_cur, xo_cand, part_x);
        rows<kSumProd>(rv64, EpiStore2{xo_cur, xo_cand, ar0, ar1, live_chk_r});
        hipLaunchKernelGGL(kb_check_rows<T>, dim3(nrch), dim3(kBlock), 0, stream, rch, st, ar0, ar1, b64, eq, rs, yk,
                           yp, yo_cur, yo_cand, ax_cand, part_r);
        cols<kSumProd>(cv64, EpiStore2{yo_cur, yo_cand, ac0, ac1, live_chk_c});
        hipLaunchKernelGGL(kb_check_cols, dim3(ncch), dim3(kBlock), 0, stream, cch, st, ac0, ac1, c64, lb64, ub64, cs,
                           part_c);
        hipLaunchKernelGGL(kb_decide, dim3(gw), dim3(kBlock), 0, stream, K, st, cch_ptr, part_x, part_c, rch_ptr,
                           part_r);
        hipLaunchKernelGGL(kb_restart_x<T>, dim3(ncch), dim3(kBlock), 0, stream, cch, st, xk, x0, xp, part_dx);
        hipLaunchKernelGGL(kb_restart_y<T>, dim3(nrch), dim3(kBlock), 0, stream, rch, st, yk, y0, yp, yg, ax, ax0,
                           ax_cand, part_dy);
        hipLaunchKernelGGL(kb_restart_finish, dim3(gw), dim3(kBlock), 0, stream, K, st, cch_ptr, part_dx, rch_ptr,
                           part_dy);
        for (int pass = 0; pass < 2; ++pass)
            if (tp[pass].built)
                hipLaunchKernelGGL(kb_tile_active, dim3(grid_for(tp[pass].ntiles)), dim3(kBlock), 0, stream,
                                   tp[pass].ntiles, tp[pass].td, pass == 0 ? inst_r : inst_c, st, tp[pass].active);
        launch_live_flags();
        hipLaunchKernelGGL(kb_count_running, dim3(1), dim3(kBlock), 0, stream, K, st, running);
    }

    void launch_interval(int k) {
        for (int t = 0; t < k; ++t) {
            launch_primal(t);
            launch_dual(t);
        }
        hipLaunchKernelGGL(kb_advance, dim3(grid_for(K)), dim3(kBlock), 0, stream, K, st, k);
        launch_check();
    }

    bo_status set_warm_start(const double* x, const double* y) override {
        if (!x && !y) {
            has_warm = 0;
            return BO_SUCCESS;
        }
        if (!x || !y) return BO_INVALID_ARGUMENT;
        BO_TRY(hipMemcpyAsync(wx, x, sizeof(double) * n, hipMemcpyHostToDevice, stream));
        BO_TRY(hipMemcpyAsync(wy, y, sizeof(double) * m, hipMemcpyHostToDevice, stream));
        BO_TRY(hipStreamSynchronize(stream));
        has_warm = 1;
        return BO_SUCCESS;
    }

    bo_status set_primal_weights(const double* w) override {
        if (w)
            for (int k = 0; k < K; ++k)
                if (!(w[k] >= 0.0) || !std::isfinite(w[k])) return BO_INVALID_ARGUMENT;
        BO_TRY(hipStreamSynchronize(stream));  // st_init may still feed an upload
        for (int k = 0; k < K; ++k) st_init[k].omega_init = w ? w[k] : 0.0;
        return BO_SUCCESS;
    }

    // Mean time of each iteration kernel with every LP running, on the state
    // the last solve left (restored afterwards): [CSC gather, CSC reduce +
    // primal epilogue, CSR gather, CSR reduce + dual epilogue], HIP events on
    // the engine's stream; bytes4: their algorithmic bytes (matrix entries in
    // the tiled format, the gathered vector once, slice words, tile table,
    // partials; the reduce's pointers, slots, partials, row / column -> LP
    // map and epilogue reads and writes).
    bo_status time_kernels(int reps, double* ms4, double* bytes4) override {
        if (!tp[0].built || !tp[1].built || reps <= 0) return BO_INVALID_ARGUMENT;
        std::vector<PdlpState> keep(K), live;
        BO_TRY(hipMemcpyAsync(keep.data(), st, sizeof(PdlpState) * K, hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        live = keep;
        for (PdlpState& q : live) q.done = 0;
        auto flags = [&]() {
            for (int pass = 0; pass < 2; ++pass)
                hipLaunchKernelGGL(kb_tile_active, dim3(grid_for(tp[pass].ntiles)), dim3(kBlock), 0, stream,
                                   tp[pass].ntiles, tp[pass].td, pass == 0 ? inst_r : inst_c, st, tp[pass].active);
            launch_live_flags();
        };
        BO_TRY(hipMemcpyAsync(st, live.data(), sizeof(PdlpState) * K, hipMemcpyHostToDevice, stream));
        flags();
        // each kernel's own duration: events bound to its dispatch (launch_ev)
        hipEvent_t ev[8];
        for (hipEvent_t& e : ev) BO_TRY(hipEventCreate(&e));
        double acc[4] = {0, 0, 0, 0};
        for (int r = 0; r < reps; ++r) {
            const TilePass& tc = tp[1];
            const TilePass& tr = tp[0];
            launch_ev(k_tile<T, T, kTileRows, 3>, dim3(tc.ntiles), dim3(kTileThreads),
                      static_cast<unsigned>(tc.W * sizeof(T)), stream, &ev[0],
                      static_cast<const unsigned char*>(tc.active), static_cast<const TileDesc*>(tc.td),
                      static_cast<const unsigned*>(tc.words), static_cast<const T*>(tc.val),
                      static_cast<const unsigned short*>(tc.idx), static_cast<const T*>(yg), m, tc.W, tc.part);
            launch_reduce(tc, n, EpiPrimal<T>{yg, st, c_hat, l_hat, u_hat, xk, x0, xp, 0, inst_c, live_red_c}, &ev[2]);
            launch_ev(k_tile<T, T, kTileRows, 2>, dim3(tr.ntiles), dim3(kTileThreads),
                      static_cast<unsigned>(tr.W * sizeof(T)), stream, &ev[4],
                      static_cast<const unsigned char*>(tr.active), static_cast<const TileDesc*>(tr.td),
                      static_cast<const unsigned*>(tr.words), static_cast<const T*>(tr.val),
                      static_cast<const unsigned short*>(tr.idx), static_cast<const T*>(xp), n, tr.W, tr.part);
            launch_reduce(tr, m, EpiDual<T>{xp, st, b_hat, eq, ax, ax0, yk, y0, yp, yg, 0, inst_r, live_red_r},
                          &ev[6]);
            BO_TRY(hipStreamSynchronize(stream));
            for (int q = 0; q < 4; ++q) {
                float t = 0.0f;
                BO_TRY(hipEventElapsedTime(&t, ev[2 * q], ev[2 * q + 1]));
                acc[q] += t;
            }
        }
        for (hipEvent_t& e : ev) hipEventDestroy(e);
        BO_TRY(hipMemcpyAsync(st, keep.data(), sizeof(PdlpState) * K, hipMemcpyHostToDevice, stream));
        flags();
        BO_TRY(hipStreamSynchronize(stream));
        const double tb = sizeof(T);
        for (int q = 0; q < 4; ++q) ms4[q] = acc[q] / reps;
        for (int pass = 0; pass < 2; ++pass) {
            const TilePass& t = tp[pass];
            const int nseg = pass == 0 ? m : n, ng = pass == 0 ? n : m;
            const double epi = pass == 0 ? 49.0 + 3.0 * tb : 24.0 + 4.0 * tb;
            const double g = (tb + 2.0) * static_cast<double>(nnz) + tb * ng + 4.0 * t.nslices +
                             static_cast<double>(sizeof(TileDesc)) * t.ntiles + 8.0 * static_cast<double>(t.pairs);
            const double rd = 4.0 * (nseg + 1) + 12.0 * static_cast<double>(t.pairs) + (epi + 4.0) * nseg;
            bytes4[pass == 1 ? 0 : 2] = g;
            bytes4[pass == 1 ? 1 : 3] = rd;
        }
        return BO_SUCCESS;
    }

    // The solve loop returns once at most this many LPs still run (they keep
    // status 1 with fewer iterations than their budget).
    void set_stop_running(int r) override { stop_running = r; }

    // The absolute time limit of the solve this batch is a round of; LPs
    // still running when it passes stop with status 3 (TIME_LIMIT).
    void set_deadline(const Deadline& d) override { deadline = d; }

    // Iteration budget of every LP (at most max_iterations of the parameters).
    void set_iteration_limits(const long long* lim) override {
        for (int k = 0; k < K; ++k) st_init[k].max_iter = std::min(lim[k], prm.max_iterations);
    }

    // New costs and / or right-hand sides for every LP (original space, batch
    // order); the matrices, bounds, row senses and the layout stay.
    bo_status set_objective_rhs(const double* c, const double* b) override {
        if (c)
            for (int j = 0; j < n; ++j)
                if (!std::isfinite(c[j])) return BO_INVALID_ARGUMENT;
        if (b)
            for (int i = 0; i < m; ++i)
                if (!std::isfinite(b[i])) return BO_INVALID_ARGUMENT;
        BO_TRY(hipStreamSynchronize(stream));
        if (c) BO_TRY(up(c64, c, sizeof(double) * n));
        if (b) BO_TRY(up(b64, b, sizeof(double) * m));
        for (int k = 0; k < K; ++k) {
            if (c) {
                double c2 = 0.0;
                for (int j = coff[k]; j < coff[k + 1]; ++j) c2 += c[j] * c[j];
                st_init[k].c_norm = std::sqrt(c2);
            }
            if (b) {
                double b2 = 0.0;
                for (int i = roff[k]; i < roff[k + 1]; ++i) b2 += b[i] * b[i];
                st_init[k].b_norm = std::sqrt(b2);
            }
        }
        BO_TRY(hipStreamSynchronize(stream));
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
        while (*running_host > stop_running) {
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
            if (prm.verbose >= 2)
                std::fprintf(stderr, "batch: step %lld running %d at %.3f ms\n", iters, *running_host,
                             std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0)
                                 .count());
        }
        BO_TRY(hipMemcpyAsync(st_host, st, sizeof(PdlpState) * K, hipMemcpyDeviceToHost, stream));
        BO_TRY(hipStreamSynchronize(stream));
        auto t2 = std::chrono::high_resolution_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
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
            total->tiled = tp[0].built && tp[1].built;
        }
        return BO_SUCCESS;
    }

    bo_status get_solution(int k, double* x, double* y) override {
        if (k < 0 || k >= K) return BO_INVALID_ARGUMENT;
        const int b = st_host[k].best;
        const int nk = coff[k + 1] - coff[k], mk = roff[k + 1] - roff[k];
        if (x && nk)
            BO_TRY(hipMemcpyAsync(x, (b ? xo_cand : xo_cur) + coff[k], sizeof(double) * nk, hipMemcpyDeviceToHost,
                                  stream));
        if (y && mk)
            BO_TRY(hipMemcpyAsync(y, (b ? yo_cand : yo_cur) + roff[k], sizeof(double) * mk, hipMemcpyDeviceToHost,
                                  stream));
        BO_TRY(hipStreamSynchronize(stream));
        return BO_SUCCESS;
    }
};

}  // namespace
}  // namespace bo

namespace bo {
namespace {

// Owned copy of one LP, kept for the FP64 polish of LPs the FP32 engine
// stops as STALLED (the caller's arrays need not outlive create).
struct HostLp {
    int m = 0, n = 0;
    std::vector<int> rp, ci;
    std::vector<double> val, b, c, lb, ub;
    std::vector<unsigned char> eq;
    explicit HostLp(const bo_lp& lp)
        : m(lp.m), n(lp.n), rp(lp.row_ptr, lp.row_ptr + (lp.m > 0 ? lp.m + 1 : 0)),
          ci(lp.col_idx, lp.col_idx + lp.nnz), val(lp.val, lp.val + lp.nnz), b(lp.b, lp.b + lp.m),
          c(lp.c, lp.c + lp.n), lb(lp.lb, lp.lb + lp.n), ub(lp.ub, lp.ub + lp.n), eq(lp.is_eq, lp.is_eq + lp.m) {}
    bo_lp view() const {
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

}  // namespace
}  // namespace bo

struct bo_pdlp_batch {
    std::unique_ptr<bo::BatchBase> eng;
    bo_pdlp_params prm{};
    int count = 0;
    std::vector<bo::HostLp> lps;             // kept for the tail rounds
    std::vector<std::vector<double>> fx, fy;  // solutions of LPs finished in a tail round
    std::vector<char> in_tail;               // 1: the LP's last solution is in fx / fy
    bool polishing() const { return !prm.fp64_iterates && prm.fp64_polish && prm.stall_iterations > 0; }
    bool compacting() const { return prm.compact_tail != 0; }
    bool keep_lps() const { return polishing() || compacting(); }
};

using namespace bo;

extern "C" {

bo_status bo_pdlp_batch_create(const bo_lp* lps, int count, const bo_pdlp_params* params, bo_pdlp_batch** out) {
    return host_guard([&]() -> bo_status {
        if (!out) return BO_INVALID_ARGUMENT;
        *out = nullptr;
        if (!lps || count <= 0) return BO_INVALID_ARGUMENT;
        long long m = 0, n = 0, nnz = 0;
        for (int k = 0; k < count; ++k) {
            if (!valid_csr_lp(&lps[k])) return BO_INVALID_ARGUMENT;
            m += lps[k].m;
            n += lps[k].n;
            nnz += lps[k].nnz;
        }
        if (m >= (1LL << 31) - 1 || n >= (1LL << 31) - 1 || nnz >= (1LL << 31) - 4096) return BO_INVALID_ARGUMENT;
        bo_status s = ensure_device_public();
        if (s != BO_SUCCESS) return s;
        bo_pdlp_params p;
        if (params) p = *params;
        else bo_pdlp_default_params(&p);
        auto h = std::make_unique<bo_pdlp_batch>();
        bo_status st;
        if (p.fp64_iterates) {
            auto e = std::make_unique<Batch<double>>();
            st = e->create(lps, count, &p);
            h->eng = std::move(e);
        } else {
            auto e = std::make_unique<Batch<float>>();
            st = e->create(lps, count, &p);
            h->eng = std::move(e);
        }
        if (st != BO_SUCCESS) return st;
        h->prm = p;
        h->count = count;
        h->in_tail.assign(count, 0);
        h->fx.resize(count);
        h->fy.resize(count);
        if (h->keep_lps()) {
            h->lps.reserve(count);
            for (int k = 0; k < count; ++k) h->lps.emplace_back(lps[k]);
        }
        *out = h.release();
        return BO_SUCCESS;
    });
}

bo_status bo_pdlp_batch_set_warm_start(bo_pdlp_batch* h, const double* x, const double* y) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        return h->eng->set_warm_start(x, y);
    });
}

bo_status bo_pdlp_batch_set_primal_weights(bo_pdlp_batch* h, const double* w) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        return h->eng->set_primal_weights(w);
    });
}

bo_status bo_pdlp_batch_set_objective_rhs(bo_pdlp_batch* h, const double* c, const double* b) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        bo_status s = h->eng->set_objective_rhs(c, b);
        if (s != BO_SUCCESS) return s;
        size_t oc = 0, ob = 0;
        for (HostLp& lp : h->lps) {
            if (c) std::copy(c + oc, c + oc + lp.n, lp.c.begin());
            if (b) std::copy(b + ob, b + ob + lp.m, lp.b.begin());
            oc += lp.n;
            ob += lp.m;
        }
        return BO_SUCCESS;
    });
}

}  // extern "C"

namespace {

// An LP handed from one round to the next: its starting point, primal
// weight and the iterations it has used.
struct TailLp {
    int k;
    long long used, fp64_used;
    double w;
    std::vector<double> x, y;
    bo_pdlp_stats last;  // its stats at the end of its last round
};

// A round returns once at most 1/kTailShare of its LPs still run; the rest
// go on in a compact batch.  Swept on the 512-LP ragged batch (MI355X):
// 1/2 49.9, 1/3 61.3, 1/4 62.6, 1/5 62.5, 1/6 61.9, 1/8 60.0 LPs/s -- a
// smaller share keeps more LPs on a layout whose tiles mostly idle, a larger
// one moves too many onto the segment kernels.
constexpr int kTailShare = 4;

// Cut by a round's stop rule with iterations left: status 1 only (an LP the
// time limit stopped has status 3 and is final).
bool tail_from_round(const bo_pdlp_stats& s, long long lim) { return s.status == 1 && s.iterations < lim; }

// LPs still queued for a round when the time limit has passed: final, with
// status 3 (TIME_LIMIT), at the point their last round left.
void finish_timed_out(bo_pdlp_batch* h, std::vector<TailLp>& queued, std::vector<bo_pdlp_stats>& per) {
    for (TailLp& t : queued) {
        per[t.k] = t.last;
        per[t.k].status = 3;
        per[t.k].iterations = t.used;
        per[t.k].fp64_polish_iterations = t.fp64_used;
        h->fx[t.k] = std::move(t.x);
        h->fy[t.k] = std::move(t.y);
        h->in_tail[t.k] = 1;
    }
    queued.clear();
}

// One compact round over `in` (segment kernels, no tile layout): LPs that
// finish are final (stats and solution kept on the handle), LPs cut by the
// round's stop rule go to `again`, LPs that stall (FP32 rounds) go to `to64`.
template <class T>
bo_status tail_round(bo_pdlp_batch* h, const Deadline& dl, std::vector<TailLp>& in, std::vector<TailLp>& again,
                     std::vector<TailLp>& to64, std::vector<bo_pdlp_stats>& per) {
    const bool is64 = sizeof(T) == sizeof(double);
    const int P = static_cast<int>(in.size());
    auto rc0 = std::chrono::high_resolution_clock::now();
    std::vector<bo_lp> views(P);
    std::vector<long long> lim(P);
    std::vector<double> w(P), wx, wy;
    long long most = 0;
    for (int j = 0; j < P; ++j) {
        views[j] = h->lps[in[j].k].view();
        lim[j] = h->prm.max_iterations - in[j].used;
        most = std::max(most, lim[j]);
        w[j] = in[j].w;
        wx.insert(wx.end(), in[j].x.begin(), in[j].x.end());
        wy.insert(wy.end(), in[j].y.begin(), in[j].y.end());
    }
    bo_pdlp_params p = h->prm;
    p.fp64_iterates = is64 ? 1 : 0;
    p.max_iterations = std::max(1LL, most);
    p.lanes_rows = p.lanes_cols = 4;  // segment kernels: no host tile layout for a few LPs
    Batch<T> e;
    bo_status s = e.create(views.data(), P, &p);
    if (s != BO_SUCCESS) return s;
    e.set_iteration_limits(lim.data());
    e.set_stop_running(h->compacting() && P >= 16 ? P / kTailShare : 0);
    e.set_deadline(dl);
    if ((s = e.set_primal_weights(w.data())) != BO_SUCCESS) return s;
    if ((s = e.set_warm_start(wx.data(), wy.data())) != BO_SUCCESS) return s;
    std::vector<bo_pdlp_stats> per2(P);
    bo_pdlp_batch_stats tot2{};
    auto r0 = std::chrono::high_resolution_clock::now();
    if ((s = e.solve(per2.data(), &tot2)) != BO_SUCCESS) return s;
    if (h->prm.verbose >= 2)
        std::fprintf(stderr, "batch tail round: %s, %d LPs, %lld steps, solved %d, %.1f ms (set-up %.1f ms)\n",
                     is64 ? "fp64" : "fp32", P, tot2.loop_steps, tot2.solved,
                     std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - r0).count(),
                     std::chrono::duration<double, std::milli>(r0 - rc0).count());
    for (int j = 0; j < P; ++j) {
        const int k = in[j].k;
        std::vector<double> x(h->lps[k].n), y(h->lps[k].m);
        if ((s = e.get_solution(j, x.data(), y.data())) != BO_SUCCESS) return s;
        TailLp t{k, in[j].used + per2[j].iterations, in[j].fp64_used + (is64 ? per2[j].iterations : 0),
                 per2[j].primal_weight, std::move(x), std::move(y), per2[j]};
        if (tail_from_round(per2[j], lim[j])) {
            again.push_back(std::move(t));
        } else if (per2[j].status == 4 && !is64 && h->polishing()) {
            to64.push_back(std::move(t));
        } else {
            per[k] = per2[j];
            per[k].iterations = t.used;
            per[k].fp64_polish_iterations = t.fp64_used;
            h->fx[k] = std::move(t.x);
            h->fy[k] = std::move(t.y);
            h->in_tail[k] = 1;
        }
    }
    return BO_SUCCESS;
}

}  // namespace

extern "C" {

// The main batch, then the tail: with compact_tail the main batch returns
// once at most 1/kTailShare of its LPs still run, and those continue in
// compact batches (segment kernels, no tile layout) that again hand over
// their last share, until every LP has stopped; with fp64_polish an LP whose FP32
// iterates stall continues in an FP64 batch.  Every LP carries over its
// point, its primal weight and the iterations it has used.
bo_status bo_pdlp_batch_solve(bo_pdlp_batch* h, bo_pdlp_stats* per_lp, bo_pdlp_batch_stats* total) {
    return host_guard([&]() -> bo_status {
        if (!h) return BO_INVALID_ARGUMENT;
        auto t0 = std::chrono::high_resolution_clock::now();
        const int K = h->count;
        std::vector<bo_pdlp_stats> per(K);
        bo_pdlp_batch_stats tot{};
        std::fill(h->in_tail.begin(), h->in_tail.end(), 0);
        const Deadline dl = Deadline::after(h->prm.time_limit_s);
        h->eng->set_deadline(dl);
        h->eng->set_stop_running(h->compacting() && K >= 16 ? K / kTailShare : 0);
        bo_status s = h->eng->solve(per.data(), &tot);
        if (s != BO_SUCCESS) return s;
        std::vector<TailLp> same, p64;
        const bool main64 = h->prm.fp64_iterates != 0;
        for (int k = 0; k < K; ++k) {
            const bool cut = h->compacting() && tail_from_round(per[k], h->prm.max_iterations);
            const bool stalled = per[k].status == 4 && h->polishing();
            if (!cut && !stalled) continue;
            TailLp t{k, per[k].iterations, main64 ? per[k].iterations : 0, per[k].primal_weight, {}, {}, per[k]};
            t.x.resize(h->lps[k].n);
            t.y.resize(h->lps[k].m);
            if ((s = h->eng->get_solution(k, t.x.data(), t.y.data())) != BO_SUCCESS) return s;
            (stalled ? p64 : same).push_back(std::move(t));
        }
        if (!same.empty() || !p64.empty()) {
            auto tp0 = std::chrono::high_resolution_clock::now();
            int tail_lps = static_cast<int>(same.size() + p64.size());
            std::vector<TailLp> unused;
            while (!same.empty() || !p64.empty()) {
                if (dl.passed()) {
                    finish_timed_out(h, same, per);
                    finish_timed_out(h, p64, per);
                    break;
                }
                if (!same.empty()) {
                    std::vector<TailLp> in;
                    in.swap(same);
                    s = main64 ? tail_round<double>(h, dl, in, same, unused, per)
                               : tail_round<float>(h, dl, in, same, p64, per);
                    if (s != BO_SUCCESS) return s;
                }
                if (!p64.empty() && !dl.passed()) {
                    std::vector<TailLp> in;
                    in.swap(p64);
                    if ((s = tail_round<double>(h, dl, in, p64, unused, per)) != BO_SUCCESS) return s;
                }
            }
            tot.tail_lps = tail_lps;
            tot.tail_ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - tp0).count();
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
        tot.solved = 0;
        tot.max_iterations = 0;
        tot.sum_iterations = 0;
        tot.polished = 0;
        for (int k = 0; k < K; ++k) {
            per[k].solve_ms = ms;
            tot.solved += per[k].status == 0;
            tot.max_iterations = std::max(tot.max_iterations, per[k].iterations);
            tot.sum_iterations += per[k].iterations;
            tot.polished += (!main64 && per[k].fp64_polish_iterations > 0) ? 1 : 0;
        }
        tot.solve_ms = ms;
        if (per_lp) std::copy(per.begin(), per.end(), per_lp);
        if (total) *total = tot;
        return BO_SUCCESS;
    });
}

bo_status bo_pdlp_batch_time_kernels(bo_pdlp_batch* h, int reps, double* ms4, double* bytes4) {
    return host_guard([&]() -> bo_status {
        if (!h || !ms4 || !bytes4) return BO_INVALID_ARGUMENT;
        return h->eng->time_kernels(reps, ms4, bytes4);
    });
}

bo_status bo_pdlp_batch_get_solution(bo_pdlp_batch* h, int k, double* x, double* y) {
    return host_guard([&]() -> bo_status {
        if (!h || k < 0 || k >= h->count) return BO_INVALID_ARGUMENT;
        if (h->in_tail[k]) {
            if (x) std::copy(h->fx[k].begin(), h->fx[k].end(), x);
            if (y) std::copy(h->fy[k].begin(), h->fy[k].end(), y);
            return BO_SUCCESS;
        }
        return h->eng->get_solution(k, x, y);
    });
}

void bo_pdlp_batch_destroy(bo_pdlp_batch* h) { delete h; }

}  // extern "C"
