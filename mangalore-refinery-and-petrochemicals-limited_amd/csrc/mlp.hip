// mlp.hip -- the warm-start network: a BF16 multilayer perceptron (FP32
// accumulation, FP32 master weights, Adam) trained and run by hand-written
// kernels on the matrix cores.
//
// Layers: n_in -> width -> ... -> width -> n_out, ReLU between layers, no
// activation on the output; loss = mean squared error.  Activations and
// weights are BF16 in the products; every product accumulates in FP32.
//
// The product kernel (k_gemm) computes C[M x N] = op(A)[M x K] op(B)[K x N]
// with gfx950's v_mfma_f32_32x32x16_bf16: a 256-thread workgroup owns a 64 x 64
// tile of C, each wavefront a 32 x 32 quarter; K advances 64 at a time
// through LDS, 16 per instruction.  Operand layout of the instruction (lane
// l, 0..63): A row l % 32, k = 8 (l / 32) + 0..7; B column l % 32, same k;
// result register r of lane l is C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  op(A) / op(B) read the stored matrix directly or
// transposed (strides), so forward (X W), input gradients (dZ W') and weight
// gradients (H' dZ) use the one kernel.  The epilogue adds a bias, applies
// ReLU and writes FP32 and / or BF16.
//
// No counterpart exists in the reference (which has no learned component).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "common.hpp"

namespace bo {

bo_status ensure_device_public();

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bfloat16, round to nearest even (NaN kept quiet)
__host__ __device__ inline unsigned short f2bf(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return static_cast<unsigned short>((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}

__host__ __device__ inline float bf2f(unsigned short h) {
    const unsigned u = static_cast<unsigned>(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

constexpr int kTm = 64, kTn = 64, kTk = 64;
constexpr int kLd = kTk + 8;      // LDS row pitch in BF16 elements (144 B: 16-byte aligned rows)
constexpr int kMaxSplit = 64;     // K splits of one product at most
constexpr int kFillTarget = 1024; // workgroups a split product aims for (4 per CU)

enum : int { kEpiNone = 0, kEpiBias = 1, kEpiRelu = 2 };

typedef unsigned short us8 __attribute__((ext_vector_type(8)));

// Operand layouts: the stored matrix is contiguous along k (0) or along its
// other dimension (1).  A 64 x 64 tile is 512 runs of 8 consecutive
// elements along the contiguous direction; thread t stages runs t and
// t + 256: a 16-byte load when all 8 are in range and `vec` (unit stride,
// 8-element aligned rows and base), else 8 guarded loads.  Loading (into
// registers) and storing (into LDS) are separate so the next tile's loads are
// in flight while the current tile is multiplied.
//
// LDS tiles hold both operands as [row of C or column of C][k], k
// contiguous, so each lane reads its 8 k-values with one 16-byte LDS load.
template <int L>
__device__ inline void tile_load(us8* v, const unsigned short* __restrict__ G, long long s_r, long long s_k, int r0,
                                 int R, int k0, int kend, bool vec) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int c = static_cast<int>(threadIdx.x) + 256 * u;
        if (L == 0) {  // k contiguous: run c -> row c / 8, k 8 (c % 8) .. + 7
            const int gr = r0 + (c >> 3), gk = k0 + (c & 7) * 8;
            if (vec && gr < R && gk + 7 < kend) {
                v[u] = *reinterpret_cast<const us8*>(G + gr * s_r + gk);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    v[u][e] = (gr < R && gk + e < kend) ? G[gr * s_r + static_cast<long long>(gk + e) * s_k]
                                                        : static_cast<unsigned short>(0);
            }
        } else {  // row-direction contiguous: run c -> k c / 8, rows 8 (c % 8) .. + 7
            const int gk = k0 + (c >> 3), gr = r0 + (c & 7) * 8;
            if (vec && gk < kend && gr + 7 < R) {
                v[u] = *reinterpret_cast<const us8*>(G + static_cast<long long>(gk) * s_k + gr);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    v[u][e] = (gk < kend && gr + e < R) ? G[static_cast<long long>(gr + e) * s_r + gk * s_k]
                                                        : static_cast<unsigned short>(0);
            }
        }
    }
}

template <int L>
__device__ inline void tile_store(unsigned short (*T)[kLd], const us8* v) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int c = static_cast<int>(threadIdx.x) + 256 * u;
        if (L == 0) {
            *reinterpret_cast<us8*>(&T[c >> 3][(c & 7) * 8]) = v[u];
        } else {
            const int kk = c >> 3, r = (c & 7) * 8;
#pragma unroll
            for (int e = 0; e < 8; ++e) T[r + e][kk] = v[u][e];
        }
    }
}

// C = op(A) op(B) (+ bias, ReLU): A(m,k) = A[m*sam + k*sak], B(k,n) =
// B[k*sbk + n*sbn]; a 256-thread workgroup owns a 64 x 64 tile of C, each
// wavefront a 32 x 32 quarter, k in steps of 64.  Split K (gridDim.z > 1):
// workgroup z sums k in [z*kc, (z+1)*kc) and stores its raw FP32 tile to
// ws[z][M][N]; the epilogue then runs in k_splitk_sum.
template <int EPI, int LA, int LB>
__global__ void __launch_bounds__(256)
k_gemm(int M, int N, int K, int kc, const unsigned short* __restrict__ A, long long sam, long long sak, bool va,
       const unsigned short* __restrict__ B, long long sbk, long long sbn, bool vb, const float* __restrict__ bias,
       float* __restrict__ C, unsigned short* __restrict__ Cb, long long ldc, float* __restrict__ ws) {
    __shared__ __attribute__((aligned(16))) unsigned short As[kTm][kLd];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[kTn][kLd];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int bm = blockIdx.y * kTm, bn = blockIdx.x * kTn;
    const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
    const int kb = blockIdx.z * kc;
    const int ke = min(K, kb + kc);
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    us8 ra[2], rb[2];
    if (kb < ke) {
        tile_load<LA>(ra, A, sam, sak, bm, M, kb, ke, va);  // A tile rows are m (stride sam)
        tile_load<LB>(rb, B, sbn, sbk, bn, N, kb, ke, vb);  // B tile rows are n (stride sbn)
    }
    for (int k0 = kb; k0 < ke; k0 += kTk) {
        tile_store<LA>(As, ra);
        tile_store<LB>(Bs, rb);
        __syncthreads();
        if (k0 + kTk < ke) {  // the next tile's loads overlap this tile's products
            tile_load<LA>(ra, A, sam, sak, bm, M, k0 + kTk, ke, va);
            tile_load<LB>(rb, B, sbn, sbk, bn, N, k0 + kTk, ke, vb);
        }
#pragma unroll
        for (int kk = 0; kk < kTk; kk += 16) {
            const int kq = kk + 8 * (lane / 32);
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + lane % 32][kq]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + lane % 32][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const bool split = gridDim.z > 1;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = bm + wm + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
        const int col = bn + wn + lane % 32;
        if (row < M && col < N) {
            float v = acc[r];
            if (split) {
                ws[(static_cast<size_t>(blockIdx.z) * M + row) * N + col] = v;
                continue;
            }
            if (EPI & kEpiBias) v += bias[col];
            if (EPI & kEpiRelu) v = v > 0.0f ? v : 0.0f;
            if (C) C[row * ldc + col] = v;
            if (Cb) Cb[row * ldc + col] = f2bf(v);
        }
    }
}

// Sum of the S split partials in order (deterministic), then the epilogue.
template <int EPI>
__global__ void k_splitk_sum(int M, int N, int S, const float* __restrict__ ws, const float* __restrict__ bias,
                             float* __restrict__ C, unsigned short* __restrict__ Cb, long long ldc) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t MN = static_cast<size_t>(M) * N;
    if (e >= MN) return;
    const int row = static_cast<int>(e / N), col = static_cast<int>(e % N);
    float v = 0.0f;
    for (int z = 0; z < S; ++z) v += ws[z * MN + e];
    if (EPI & kEpiBias) v += bias[col];
    if (EPI & kEpiRelu) v = v > 0.0f ? v : 0.0f;
    if (C) C[row * ldc + col] = v;
    if (Cb) Cb[row * ldc + col] = f2bf(v);
}

// K split of an M x N x K product: enough workgroups to fill the chip, at
// least 256 k per split, kc a multiple of the 32-wide k step.
struct SplitK {
    int S, kc;
};
inline SplitK split_for(int M, int N, int K) {
    const long long tiles = static_cast<long long>((M + kTm - 1) / kTm) * ((N + kTn - 1) / kTn);
    int S = 1;
    if (tiles < kFillTarget / 2 && K >= 512) {
        S = static_cast<int>(std::min<long long>((kFillTarget + tiles - 1) / tiles, K / 256));
        S = std::max(1, std::min(S, kMaxSplit));
    }
    int kc = (K + S - 1) / S;
    kc = (kc + kTk - 1) / kTk * kTk;
    S = (K + kc - 1) / kc;
    return SplitK{S, kc};
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Launch of one product on `stream`; ws holds split partials (ws_floats
// floats at least S * M * N, checked).
template <int EPI>
hipError_t launch_gemm(hipStream_t stream, int M, int N, int K, const unsigned short* A, long long sam, long long sak,
                       const unsigned short* B, long long sbk, long long sbn, const float* bias, float* C,
                       unsigned short* Cb, long long ldc, float* ws, size_t ws_floats) {
    const SplitK sk = split_for(M, N, K);
    if (sk.S > 1 && static_cast<size_t>(sk.S) * M * N > ws_floats) return hipErrorInvalidValue;
    dim3 grid((N + kTn - 1) / kTn, (M + kTm - 1) / kTm, sk.S);
    const int la = sak == 1 ? 0 : 1, lb = sbk == 1 ? 0 : 1;
    const bool va = aligned16(A) && (la == 0 ? (sak == 1 && sam % 8 == 0) : (sam == 1 && sak % 8 == 0));
    const bool vb = aligned16(B) && (lb == 0 ? (sbk == 1 && sbn % 8 == 0) : (sbn == 1 && sbk % 8 == 0));
#define BO_GEMM_LAUNCH(LA_, LB_)                                                                                  \
    hipLaunchKernelGGL((k_gemm<EPI, LA_, LB_>), grid, dim3(256), 0, stream, M, N, K, sk.kc, A, sam, sak, va, B, sbk, \
                       sbn, vb, bias, C, Cb, ldc, ws)
    if (la == 0 && lb == 0)
        BO_GEMM_LAUNCH(0, 0);
    else if (la == 0)
        BO_GEMM_LAUNCH(0, 1);
    else if (lb == 0)
        BO_GEMM_LAUNCH(1, 0);
    else
        BO_GEMM_LAUNCH(1, 1);
#undef BO_GEMM_LAUNCH
    if (sk.S > 1) {
        const size_t MN = static_cast<size_t>(M) * N;
        hipLaunchKernelGGL(k_splitk_sum<EPI>, dim3(grid_for(MN)), dim3(kBlock), 0, stream, M, N, sk.S, ws, bias, C,
                           Cb, ldc);
    }
    return hipGetLastError();
}

__global__ void k_to_bf16(const float* __restrict__ x, unsigned short* __restrict__ y, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// dY = 2 (Y - T) / count (FP32 and BF16); loss partial sums per block
__global__ void __launch_bounds__(kBlock)
k_mse(const float* __restrict__ y, const float* __restrict__ t, size_t n, float scale, unsigned short* __restrict__ dyb,
      double* __restrict__ part) {
    __shared__ double sh[kBlock / kWave];
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < n) {
        const float d = y[i] - t[i];
        dyb[i] = f2bf(2.0f * d * scale);
        l = static_cast<double>(d) * d;
    }
    const double s = block_sum(l, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_sum(const double* __restrict__ part, int n, double scale, double* __restrict__ out) {
    __shared__ double sh[kBlock / kWave];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a += part[i];
    const double s = block_sum(a, sh);
    if (threadIdx.x == 0) *out = s * scale;
}

// dZ = dH * (H > 0), as BF16 for the next products
__global__ void k_relu_bwd(const float* __restrict__ dh, const unsigned short* __restrict__ h,
                           unsigned short* __restrict__ dz, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dz[i] = bf2f(h[i]) > 0.0f ? f2bf(dh[i]) : static_cast<unsigned short>(0);
}

// db[n] = sum over the batch of dZ[b][n] (BF16 values, FP32 sum in batch order)
__global__ void k_colsum(const unsigned short* __restrict__ dz, int B, int N, float* __restrict__ db) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int b = 0; b < B; ++b) s += bf2f(dz[static_cast<size_t>(b) * N + n]);
    db[n] = s;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on FP32 master weights, BF16 copy out
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, unsigned short* __restrict__ wb, size_t n, float lr, float bc1,
                       float bc2) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

struct DevMemM {
    std::vector<void*> ptrs;
    ~DevMemM() {
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

}  // namespace
}  // namespace bo

using namespace bo;

struct bo_mlp {
    int layers = 0, n_in = 0, width = 0, n_out = 0, max_batch = 0;
    std::vector<int> din, dout;  // per layer
    std::vector<size_t> woff, boff;  // offsets of W_l (din x dout) and b_l in the flat parameter array
    size_t nparam = 0;
    long long steps = 0;
    hipStream_t stream = nullptr;
    DevMemM mem;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 master, Adam moments, gradients
    unsigned short* wb = nullptr;                                    // BF16 copy of w (weights only used)
    // activations (max_batch rows): x, hidden BF16 per layer, output FP32
    unsigned short* xb = nullptr;
    std::vector<unsigned short*> hb;  // layers - 1 hidden activations (BF16, post-ReLU)
    float *y = nullptr, *t = nullptr, *xf = nullptr;
    float* dh = nullptr;                // FP32 input gradient of the current layer (max_batch x max dim)
    unsigned short *dzb = nullptr, *dzb2 = nullptr;  // BF16 gradients (ping-pong)
    double *part = nullptr, *loss = nullptr;
    float* ws = nullptr;  // split-K partials
    size_t ws_floats = 0;
    hipError_t err = hipSuccess;  // first launch error of the current call

    ~bo_mlp() {
        if (stream) hipStreamDestroy(stream);
    }

    template <int EPI>
    void gemm(int M, int N, int K, const unsigned short* A, long long sam, long long sak, const unsigned short* B,
              long long sbk, long long sbn, const float* bias, float* C, unsigned short* Cb, long long ldc) {
        const hipError_t e = launch_gemm<EPI>(stream, M, N, K, A, sam, sak, B, sbk, sbn, bias, C, Cb, ldc, ws, ws_floats);
        if (err == hipSuccess) err = e;
    }

    // split-K workspace for every product of a batch of up to max_batch rows
    size_t ws_needed() const {
        size_t need = 0;
        auto take = [&](int M, int N, int K) {
            const SplitK sk = split_for(M, N, K);
            if (sk.S > 1) need = std::max(need, static_cast<size_t>(sk.S) * M * N);
        };
        for (int B = 1; B <= max_batch; ++B)
            for (int l = 0; l < layers; ++l) {
                take(B, dout[l], din[l]);   // forward
                take(din[l], dout[l], B);   // weight gradient
                take(B, din[l], dout[l]);   // input gradient
            }
        return need;
    }

    bo_status create(int nin, int wdt, int nout, int nl, int maxb, unsigned long long seed) {
        layers = nl;
        n_in = nin;
        width = wdt;
        n_out = nout;
        max_batch = maxb;
        for (int l = 0; l < layers; ++l) {
            din.push_back(l == 0 ? n_in : width);
            dout.push_back(l == layers - 1 ? n_out : width);
        }
        for (int l = 0; l < layers; ++l) {
            woff.push_back(nparam);
            nparam += static_cast<size_t>(din[l]) * dout[l];
            boff.push_back(nparam);
            nparam += dout[l];
        }
        BO_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        BO_TRY(mem.alloc(&w, nparam));
        BO_TRY(mem.alloc(&m, nparam));
        BO_TRY(mem.alloc(&v, nparam));
        BO_TRY(mem.alloc(&g, nparam));
        BO_TRY(mem.alloc(&wb, nparam));
        const size_t maxd = static_cast<size_t>(std::max(std::max(n_in, n_out), width));
        BO_TRY(mem.alloc(&xf, static_cast<size_t>(maxb) * n_in));
        BO_TRY(mem.alloc(&xb, static_cast<size_t>(maxb) * n_in));
        hb.assign(layers - 1, nullptr);
        for (int l = 0; l + 1 < layers; ++l) BO_TRY(mem.alloc(&hb[l], static_cast<size_t>(maxb) * width));
        BO_TRY(mem.alloc(&y, static_cast<size_t>(maxb) * n_out));
        BO_TRY(mem.alloc(&t, static_cast<size_t>(maxb) * n_out));
        BO_TRY(mem.alloc(&dh, static_cast<size_t>(maxb) * maxd));
        BO_TRY(mem.alloc(&dzb, static_cast<size_t>(maxb) * maxd));
        BO_TRY(mem.alloc(&dzb2, static_cast<size_t>(maxb) * maxd));
        BO_TRY(mem.alloc(&part, grid_for(static_cast<size_t>(maxb) * n_out)));
        BO_TRY(mem.alloc(&loss, 1));
        ws_floats = ws_needed();
        if (ws_floats) BO_TRY(mem.alloc(&ws, ws_floats));
        BO_TRY(hipMemsetAsync(xb, 0, sizeof(unsigned short) * maxb * n_in, stream));
        // He-uniform initialisation of weights, zero biases (host, seeded)
        std::vector<float> hw(nparam, 0.0f);
        std::mt19937_64 rng(seed);
        for (int l = 0; l < layers; ++l) {
            const float a = std::sqrt(6.0f / static_cast<float>(din[l]));
            std::uniform_real_distribution<float> U(-a, a);
            for (size_t i = 0; i < static_cast<size_t>(din[l]) * dout[l]; ++i) hw[woff[l] + i] = U(rng);
        }
        return set_params(hw.data());
    }

    bo_status set_params(const float* hw) {
        BO_TRY(hipMemcpyAsync(w, hw, sizeof(float) * nparam, hipMemcpyHostToDevice, stream));
        BO_TRY(hipMemsetAsync(m, 0, sizeof(float) * nparam, stream));
        BO_TRY(hipMemsetAsync(v, 0, sizeof(float) * nparam, stream));
        hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(nparam)), dim3(kBlock), 0, stream, w, wb, nparam);
        BO_TRY(hipStreamSynchronize(stream));
        steps = 0;
        return BO_SUCCESS;
    }

    void forward(int B) {
        const unsigned short* in = xb;
        for (int l = 0; l < layers; ++l) {
            const int K = din[l], N = dout[l];
            const unsigned short* W = wb + woff[l];
            const float* bias = w + boff[l];  // FP32 master bias
            if (l + 1 < layers)
                gemm<kEpiBias | kEpiRelu>(B, N, K, in, K, 1, W, N, 1, bias, nullptr, hb[l], N);
            else
                gemm<kEpiBias>(B, N, K, in, K, 1, W, N, 1, bias, y, nullptr, N);
            if (l + 1 < layers) in = hb[l];
        }
    }

    // gradients of the mean squared error into g; loss into *loss
    void backward(int B) {
        const size_t ny = static_cast<size_t>(B) * n_out;
        const unsigned nb = grid_for(ny);
        hipLaunchKernelGGL(k_mse, dim3(nb), dim3(kBlock), 0, stream, y, t, ny, 1.0f / static_cast<float>(ny), dzb,
                           part);
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(kBlock), 0, stream, part, static_cast<int>(nb),
                           1.0 / static_cast<double>(ny), loss);
        unsigned short* dz = dzb;       // gradient w.r.t. the pre-activation of layer l (BF16, B x dout)
        unsigned short* other = dzb2;
        for (int l = layers - 1; l >= 0; --l) {
            const int K = din[l], N = dout[l];
            const unsigned short* Hin = l == 0 ? xb : hb[l - 1];
            // dW = Hin' dZ   (K x N), db = colsum dZ
            gemm<kEpiNone>(K, N, B, Hin, 1, K, dz, N, 1, nullptr, g + woff[l], nullptr, N);
            hipLaunchKernelGGL(k_colsum, dim3(grid_for(N)), dim3(kBlock), 0, stream, dz, B, N, g + boff[l]);
            if (l == 0) break;
            // dH = dZ W'   (B x K), then through the ReLU of layer l-1
            gemm<kEpiNone>(B, K, N, dz, N, 1, wb + woff[l], 1, N, nullptr, dh, nullptr, K);
            hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(static_cast<size_t>(B) * K)), dim3(kBlock), 0, stream, dh,
                               hb[l - 1], other, static_cast<size_t>(B) * K);
            std::swap(dz, other);
        }
    }

    hipError_t take_error() {
        const hipError_t e = err != hipSuccess ? err : hipGetLastError();
        err = hipSuccess;
        return e;
    }

    void adam(float lr) {
        ++steps;
        const float bc1 = 1.0f - std::pow(0.9f, static_cast<float>(steps));
        const float bc2 = 1.0f - std::pow(0.999f, static_cast<float>(steps));
        hipLaunchKernelGGL(k_adam, dim3(grid_for(nparam)), dim3(kBlock), 0, stream, w, m, v, g, wb, nparam, lr, bc1,
                           bc2);
    }

    bo_status upload_x(const float* x, int B) {
        BO_TRY(hipMemcpyAsync(xf, x, sizeof(float) * B * n_in, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(static_cast<size_t>(B) * n_in)), dim3(kBlock), 0, stream, xf, xb,
                           static_cast<size_t>(B) * n_in);
        return BO_SUCCESS;
    }
};

extern "C" {

bo_status bo_mlp_create(int n_in, int width, int n_out, int layers, int max_batch, unsigned long long seed,
                        bo_mlp** out) {
    return host_guard([&]() -> bo_status {
        if (!out) return BO_INVALID_ARGUMENT;
        *out = nullptr;
        bo_status s = ensure_device_public();
        if (s != BO_SUCCESS) return s;
        if (n_in <= 0 || width <= 0 || n_out <= 0 || layers < 2 || max_batch <= 0) return BO_INVALID_ARGUMENT;
        if (static_cast<long long>(max_batch) * std::max(std::max(n_in, n_out), width) >= (1LL << 31))
            return BO_INVALID_ARGUMENT;
        auto h = std::make_unique<bo_mlp>();
        s = h->create(n_in, width, n_out, layers, max_batch, seed);
        if (s != BO_SUCCESS) return s;
        *out = h.release();
        return BO_SUCCESS;
    });
}

long long bo_mlp_num_params(const bo_mlp* h) { return h ? static_cast<long long>(h->nparam) : 0; }

bo_status bo_mlp_get_params(bo_mlp* h, float* params) {
    return host_guard([&]() -> bo_status {
        if (!h || !params) return BO_INVALID_ARGUMENT;
        BO_TRY(hipMemcpyAsync(params, h->w, sizeof(float) * h->nparam, hipMemcpyDeviceToHost, h->stream));
        BO_TRY(hipStreamSynchronize(h->stream));
        return BO_SUCCESS;
    });
}

bo_status bo_mlp_set_params(bo_mlp* h, const float* params) {
    return host_guard([&]() -> bo_status {
        if (!h || !params) return BO_INVALID_ARGUMENT;
        return h->set_params(params);
    });
}

bo_status bo_mlp_forward(bo_mlp* h, const float* x, int batch, float* y) {
    return host_guard([&]() -> bo_status {
        if (!h || !x || !y || batch <= 0 || batch > h->max_batch) return BO_INVALID_ARGUMENT;
        bo_status s = h->upload_x(x, batch);
        if (s != BO_SUCCESS) return s;
        h->err = hipSuccess;
        h->forward(batch);
        BO_TRY(h->take_error());
        BO_TRY(hipMemcpyAsync(y, h->y, sizeof(float) * batch * h->n_out, hipMemcpyDeviceToHost, h->stream));
        BO_TRY(hipStreamSynchronize(h->stream));
        BO_TRY(hipGetLastError());
        return BO_SUCCESS;
    });
}

bo_status bo_mlp_gradients(bo_mlp* h, const float* x, const float* target, int batch, double* loss, float* grads) {
    return host_guard([&]() -> bo_status {
        if (!h || !x || !target || batch <= 0 || batch > h->max_batch) return BO_INVALID_ARGUMENT;
        bo_status s = h->upload_x(x, batch);
        if (s != BO_SUCCESS) return s;
        BO_TRY(hipMemcpyAsync(h->t, target, sizeof(float) * batch * h->n_out, hipMemcpyHostToDevice, h->stream));
        h->err = hipSuccess;
        h->forward(batch);
        h->backward(batch);
        BO_TRY(h->take_error());
        if (loss) BO_TRY(hipMemcpyAsync(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        if (grads) BO_TRY(hipMemcpyAsync(grads, h->g, sizeof(float) * h->nparam, hipMemcpyDeviceToHost, h->stream));
        BO_TRY(hipStreamSynchronize(h->stream));
        BO_TRY(hipGetLastError());
        return BO_SUCCESS;
    });
}

bo_status bo_mlp_train_step(bo_mlp* h, const float* x, const float* target, int batch, float lr, double* loss) {
    return host_guard([&]() -> bo_status {
        if (!h || !x || !target || batch <= 0 || batch > h->max_batch || !(lr > 0.0f)) return BO_INVALID_ARGUMENT;
        bo_status s = h->upload_x(x, batch);
        if (s != BO_SUCCESS) return s;
        BO_TRY(hipMemcpyAsync(h->t, target, sizeof(float) * batch * h->n_out, hipMemcpyHostToDevice, h->stream));
        h->err = hipSuccess;
        h->forward(batch);
        h->backward(batch);
        BO_TRY(h->take_error());
        h->adam(lr);
        if (loss) BO_TRY(hipMemcpyAsync(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        BO_TRY(hipStreamSynchronize(h->stream));
        BO_TRY(hipGetLastError());
        return BO_SUCCESS;
    });
}

bo_status bo_mlp_time(bo_mlp* h, int what, int batch, int reps, double* ms) {
    return host_guard([&]() -> bo_status {
        if (!h || !ms || (what != 0 && what != 1) || batch <= 0 || batch > h->max_batch || reps <= 0)
            return BO_INVALID_ARGUMENT;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        BO_TRY(hipEventCreate(&e0));
        hipError_t e = hipEventCreate(&e1);
        if (e == hipSuccess) {
            h->err = hipSuccess;
            for (int r = 0; r <= reps; ++r) {  // one untimed pass first
                if (r == 1) e = hipEventRecord(e0, h->stream);
                h->forward(batch);
                if (what == 1) h->backward(batch);
            }
            if (e == hipSuccess) e = h->take_error();
            if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            *ms = static_cast<double>(t) / reps;
        }
        hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        BO_TRY(e);
        return BO_SUCCESS;
    });
}

bo_status bo_gemm_bf16(int M, int N, int K, const unsigned short* A, long long sam, long long sak,
                       const unsigned short* B, long long sbk, long long sbn, float* C) {
    // host arrays; test and timing entry point of the product kernel
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C || sam < 0 || sak < 0 || sbk < 0 || sbn < 0)
        return BO_INVALID_ARGUMENT;
    bo_status s = ensure_device_public();
    if (s != BO_SUCCESS) return s;
    const long long na = (M - 1) * sam + (K - 1) * sak + 1, nb = (K - 1) * sbk + (N - 1) * sbn + 1;
    const SplitK sk = split_for(M, N, K);
    const size_t nws = sk.S > 1 ? static_cast<size_t>(sk.S) * M * N : 0;
    DevMemM mem;
    unsigned short *dA = nullptr, *dB = nullptr;
    float *dC = nullptr, *dW = nullptr;
    BO_TRY(mem.alloc(&dA, na));
    BO_TRY(mem.alloc(&dB, nb));
    BO_TRY(mem.alloc(&dC, static_cast<size_t>(M) * N));
    if (nws) BO_TRY(mem.alloc(&dW, nws));
    BO_TRY(hipMemcpy(dA, A, sizeof(unsigned short) * na, hipMemcpyHostToDevice));
    BO_TRY(hipMemcpy(dB, B, sizeof(unsigned short) * nb, hipMemcpyHostToDevice));
    BO_TRY((launch_gemm<kEpiNone>(nullptr, M, N, K, dA, sam, sak, dB, sbk, sbn, nullptr, dC, nullptr,
                                  static_cast<long long>(N), dW, nws)));
    BO_TRY(hipMemcpy(C, dC, sizeof(float) * M * N, hipMemcpyDeviceToHost));
    return BO_SUCCESS;
}

void bo_mlp_destroy(bo_mlp* h) { delete h; }

}  // extern "C"
