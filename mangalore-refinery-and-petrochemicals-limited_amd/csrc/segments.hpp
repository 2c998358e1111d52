// segments.hpp -- length-binned segment kernels: one launch covers every row
// (or column) of a CSR (or CSC) matrix whose segments were sorted by length.
//
// Segments are grouped into bins; a bin with L lanes per segment puts 256/L
// segments in a 256-thread workgroup (L <= 64, one wavefront holds 64/L
// segments) or spreads one segment over the whole workgroup (L == 256).
// Lanes of a segment read its nonzeros with stride L (coalesced), accumulate
// in double, and reduce with a fixed xor tree, so results are deterministic.
// The bin a workgroup serves is read from a small table passed by value.
// Segments are sorted by descending length, so the heaviest workgroups are
// dispatched first (longest-processing-time-first), and consecutive
// workgroup ids -- which the hardware deals round-robin to the 8 XCDs -- carry
// similar work, which keeps the XCDs balanced.
#pragma once

#include "common.hpp"

namespace bo {

constexpr int kMaxBins = 8;
constexpr int kNumXcd = 8;

struct BinTable {
    int nbins;
    int total_blocks;
    int block_start[kMaxBins + 1];
    int seg_start[kMaxBins + 1];
    int lanes[kMaxBins];
};

// Accumulation modes of a segment pass.
enum SegMode : int {
    kSumProd = 0,  // sum_k val[k] * vec[idx[k]]           (SpMV)
    kMaxAbs = 1,   // max_k |val[k] * vec[idx[k]]|        (Ruiz inf-norms)
    kSumAbs = 2,   // sum_k |val[k] * vec[idx[k]]|        (Pock-Chambolle l1-norms)
    kWrite = 3,    // epi.write(seg, k, val[k], idx[k]) per nonzero, no reduction
};

template <int MODE>
__device__ inline double seg_combine(double a, double b) {
    if constexpr (MODE == kMaxAbs) return a > b ? a : b;
    else return a + b;
}

template <int MODE, class TV, class TX>
__device__ inline double seg_term(TV v, TX x) {
    double p = static_cast<double>(v) * static_cast<double>(x);
    if constexpr (MODE == kSumProd) return p;
    else return fabs(p);
}

// Lanes of a segment walk its nonzeros with stride L, unroll<L>() nonzeros
// per trip: all index/value loads of a trip are issued before any gather, so
// a lane keeps that many independent loads in flight.  Narrow groups unroll
// deeper so one trip covers about 128 nonzeros of a segment.
template <int L>
constexpr int unroll() {
    return L >= 32 ? 4 : (L == 16 ? 8 : 16);
}

// Optional LDS copy of the first `count` entries of the gathered vector (the
// hottest ones: segments of the other dimension are stored longest first).
// The pointer is typed in the LDS address space, so a gather reads either
// LDS or memory with its own instruction (with a generic pointer the
// compiler merged the two into one flat load through a pointer select).
template <class TX>
struct HotCache {
    using LdsPtr = const __attribute__((address_space(3))) TX*;
    LdsPtr lds;
    int count;
};

template <int MODE, int L, int NV, bool HOT, class TV, class Epi, class TX>
__device__ inline void seg_accumulate(int k0, int k1, int lane, const int* __restrict__ idx,
                                      const TV* __restrict__ val, const Epi& epi, double* acc,
                                      HotCache<TX> hot) {
    constexpr int U = unroll<L>();
    for (int k = k0 + lane; k < k1; k += U * L) {
        if constexpr (HOT) {
            TV a[U];
            int j[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = k + u * L;
                a[u] = kk < k1 ? val[kk] : TV(0);
                j[u] = kk < k1 ? idx[kk] : -1;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (j[u] >= 0) {
                    const TX xv = j[u] < hot.count ? hot.lds[j[u]] : epi.vec(0)[j[u]];
                    acc[0] = seg_combine<MODE>(acc[0], seg_term<MODE>(a[u], xv));
                }
            }
        } else {
            // Groups of G entries per lane, with no load under a branch:
            // positions past the segment load entry k1 - 1 (k < k1 here) and
            // gather entry 0, and their terms are not added, so a group's
            // loads are all in flight together (under a per-entry branch each
            // waited for itself).  A group that no lane of the wavefront
            // reaches ends the trip, so short segments issue no wasted loads
            // beyond their last group.  Terms are added in entry order.
            constexpr int G = U < 4 ? U : 4;
#pragma unroll
            for (int u0 = 0; u0 < U; u0 += G) {
                if (!__any(k + u0 * L < k1)) break;  // uniform; later groups start further on
                TV a[G];
                int j[G];
                TX xv[NV][G];
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const int kk = k + (u0 + g) * L;
                    const int kc = kk < k1 ? kk : k1 - 1;
                    const TV av = val[kc];
                    const int jv = idx[kc];
                    a[g] = kk < k1 ? av : TV(0);
                    j[g] = kk < k1 ? jv : -1;
                }
                if constexpr (EpiPair<Epi>::value) {
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const double2 xp2 = epi.vec2()[j[g] >= 0 ? j[g] : 0];
                        xv[0][g] = xp2.x;
                        xv[1][g] = xp2.y;
                    }
                } else {
#pragma unroll
                    for (int g = 0; g < G; ++g)
#pragma unroll
                        for (int v = 0; v < NV; ++v) xv[v][g] = epi.vec(v)[j[g] >= 0 ? j[g] : 0];
                }
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int v = 0; v < NV; ++v)
                        acc[v] = j[g] >= 0 ? seg_combine<MODE>(acc[v], seg_term<MODE>(a[g], xv[v][g])) : acc[v];
            }
        }
    }
}

template <int MODE, int L>
__device__ inline double group_reduce(double v) {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v = seg_combine<MODE>(v, __shfl_xor(v, off, kWave));
    return v;
}

// Epilogue concept:
//   static constexpr int NV;        number of gathered vectors (1 or 2)
//   static constexpr int Q;         number of per-block partial sums (0..8)
//   const TX* vec(int v) const;     gathered vector v
//   __device__ void operator()(int seg, const double* acc, double* part) const;
// A split epilogue (load / finish, NV == 1, Q == 0, sums of products) has
// segment seg's own data loaded by the group's first lane together with the
// segment's bounds, before the entries; done (optional) is read at the same
// time.  Returns false when done is set (nothing has been written).
template <int MODE, int L, class TV, class TX, class Epi, bool HOT = false>
__device__ inline bool seg_process(const BinTable& bt, int bin, int vblock, const int* __restrict__ ptr,
                                   const int* __restrict__ idx, const TV* __restrict__ val,
                                   const Epi& epi, double* part, const int* __restrict__ done,
                                   HotCache<TX> hot = {nullptr, 0}) {
    constexpr int NV = Epi::NV;
    constexpr bool kSplit = EpiHasLoad<Epi>::value && NV == 1 && Epi::Q == 0 && MODE == kSumProd;
    const int seg_end = bt.seg_start[bin + 1];
    const int dn = done ? *done : 0;
    if constexpr (L <= kWave) {
        constexpr int per_block = kBlock / L;
        const int seg = bt.seg_start[bin] + (vblock - bt.block_start[bin]) * per_block +
                        static_cast<int>(threadIdx.x) / L;
        const int lane = static_cast<int>(threadIdx.x) % L;
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
        const bool live = seg < seg_end;
        int k0 = 0, k1 = 0;
        decltype(epi_preload(epi, 0, 0)) pre{};
        if (live) {
            k0 = ptr[seg];
            k1 = ptr[seg + 1];
            if constexpr (kSplit) {
                if (lane == 0) pre = epi_preload(epi, seg, 0);
            }
        }
        if (dn) return false;
        if (live) {
            if constexpr (MODE == kWrite) {
                for (int k = k0 + lane; k < k1; k += L) epi.write(seg, k, val[k], idx[k]);
            } else {
                seg_accumulate<MODE, L, NV, HOT>(k0, k1, lane, idx, val, epi, acc, hot);
            }
        }
        if constexpr (MODE != kWrite) {
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = group_reduce<MODE, L>(acc[v]);
            if (live && lane == 0) {
                if constexpr (kSplit) epi_finish(epi, seg, acc[0], pre, 0);
                else epi(seg, acc, part);
            }
        }
    } else {
        // one segment per workgroup
        __shared__ double red[NV][kBlock / kWave];
        const int seg = bt.seg_start[bin] + (vblock - bt.block_start[bin]);
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
        const int k0 = ptr[seg], k1 = ptr[seg + 1];
        decltype(epi_preload(epi, 0, 0)) pre{};
        if constexpr (kSplit) {
            if (threadIdx.x == 0) pre = epi_preload(epi, seg, 0);
        }
        if (dn) return false;
        if constexpr (MODE == kWrite) {
            for (int k = k0 + static_cast<int>(threadIdx.x); k < k1; k += kBlock) epi.write(seg, k, val[k], idx[k]);
            return true;
        } else {
            seg_accumulate<MODE, kBlock, NV, HOT>(k0, k1, static_cast<int>(threadIdx.x), idx, val, epi, acc, hot);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = group_reduce<MODE, kWave>(acc[v]);
        const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
        if (lane == 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) red[v][wid] = acc[v];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double t = red[v][0];
                for (int w = 1; w < kBlock / kWave; ++w) t = seg_combine<MODE>(t, red[v][w]);
                acc[v] = t;
            }
            (void)seg_end;
            if constexpr (kSplit) epi_finish(epi, seg, acc[0], pre, 0);
            else epi(seg, acc, part);
        }
    }
    return true;
}

template <int MODE, class TV, class TX, class Epi>
__global__ void __launch_bounds__(kBlock)
k_segments(BinTable bt, const int* __restrict__ ptr, const int* __restrict__ idx,
           const TV* __restrict__ val, Epi epi, double* __restrict__ partials,
           const int* __restrict__ done) {
    if (wg_dead(epi, static_cast<int>(blockIdx.x), 0)) return;
    constexpr int Q = Epi::Q;
    const int vblock = static_cast<int>(blockIdx.x);
    int bin = 0;
    while (bin + 1 < bt.nbins && vblock >= bt.block_start[bin + 1]) ++bin;
    double part[Q > 0 ? Q : 1];
#pragma unroll
    for (int q = 0; q < (Q > 0 ? Q : 1); ++q) part[q] = 0.0;
    bool go = true;
    switch (bt.lanes[bin]) {
        case 1: go = seg_process<MODE, 1, TV, TX>(bt, bin, vblock, ptr, idx, val, epi, part, done); break;
        case 2: go = seg_process<MODE, 2, TV, TX>(bt, bin, vblock, ptr, idx, val, epi, part, done); break;
        case 4: go = seg_process<MODE, 4, TV, TX>(bt, bin, vblock, ptr, idx, val, epi, part, done); break;
        case 8: go = seg_process<MODE, 8, TV, TX>(bt, bin, vblock, ptr, idx, val, epi, part, done); break;
        case 16: go = seg_process<MODE, 16, TV, TX>(bt, bin, vblock, ptr, idx, val, epi, part, done); break;
        case 32: go = seg_process<MODE, 32, TV, TX>(bt, bin, vblock, ptr, idx, val, epi, part, done); break;
        case 64: go = seg_process<MODE, 64, TV, TX>(bt, bin, vblock, ptr, idx, val, epi, part, done); break;
        default: go = seg_process<MODE, 256, TV, TX>(bt, bin, vblock, ptr, idx, val, epi, part, done); break;
    }
    if constexpr (Q > 0) {
        if (!go) return;  // uniform: done is one flag
        __shared__ double s[kBlock / kWave];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double t = block_sum(part[q], s);
            if (threadIdx.x == 0) partials[static_cast<size_t>(blockIdx.x) * Q + q] = t;
        }
    }
    (void)go;
}

// Persistent variant with an LDS copy of the hottest `hot_count` entries of
// the (single) gathered vector: a grid of resident workgroups walks the
// segment groups g = blockIdx.x, blockIdx.x + gridDim.x, ... in bin order,
// after each workgroup has filled its LDS copy once.  Dynamic LDS holds
// hot_count elements of TX.  No per-block partials.
template <int MODE, class TV, class TX, class Epi>
__global__ void __launch_bounds__(kBlock)
k_segments_hot(BinTable bt, const int* __restrict__ ptr, const int* __restrict__ idx,
               const TV* __restrict__ val, Epi epi, int hot_count, const int* __restrict__ done) {
    static_assert(Epi::Q == 0 && Epi::NV == 1, "hot-cache passes gather one vector and keep no partials");
    if (done && *done) return;
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    TX* lds = reinterpret_cast<TX*>(dyn_lds);
    const TX* src = epi.vec(0);
    for (int i = threadIdx.x; i < hot_count; i += kBlock) lds[i] = src[i];
    __syncthreads();
    const HotCache<TX> hot{(typename HotCache<TX>::LdsPtr)lds, hot_count};
    double part[1] = {0.0};
    int bin = 0;
    for (int g = blockIdx.x; g < bt.total_blocks; g += gridDim.x) {
        while (bin + 1 < bt.nbins && g >= bt.block_start[bin + 1]) ++bin;
        switch (bt.lanes[bin]) {
            case 1: seg_process<MODE, 1, TV, TX, Epi, true>(bt, bin, g, ptr, idx, val, epi, part, nullptr, hot); break;
            case 2: seg_process<MODE, 2, TV, TX, Epi, true>(bt, bin, g, ptr, idx, val, epi, part, nullptr, hot); break;
            case 4: seg_process<MODE, 4, TV, TX, Epi, true>(bt, bin, g, ptr, idx, val, epi, part, nullptr, hot); break;
            case 8: seg_process<MODE, 8, TV, TX, Epi, true>(bt, bin, g, ptr, idx, val, epi, part, nullptr, hot); break;
            case 16: seg_process<MODE, 16, TV, TX, Epi, true>(bt, bin, g, ptr, idx, val, epi, part, nullptr, hot); break;
            case 32: seg_process<MODE, 32, TV, TX, Epi, true>(bt, bin, g, ptr, idx, val, epi, part, nullptr, hot); break;
            case 64: seg_process<MODE, 64, TV, TX, Epi, true>(bt, bin, g, ptr, idx, val, epi, part, nullptr, hot); break;
            default:
                seg_process<MODE, 256, TV, TX, Epi, true>(bt, bin, g, ptr, idx, val, epi, part, nullptr, hot);
                __syncthreads();  // the one-segment path reuses its LDS reduction slots on its next trip
                break;
        }
    }
}

}  // namespace bo
