// stream.hpp -- "stream" segment kernel: every workgroup owns a contiguous
// range of segments whose nonzeros fit a fixed LDS budget (CSR-stream, after
// Greathouse & Daga, SC'14).  Phase 1 streams the range's nonzeros with all
// 256 lanes busy and consecutive lanes on consecutive nonzeros, gathers the
// vector and stores the FP64 products in LDS; phase 2 reduces each segment
// from LDS with a power-of-two group of lanes and runs the epilogue.  A
// segment longer than the budget gets a workgroup of its own and is reduced
// in registers.  Block boundaries are built on the host (stream_blocks).
#pragma once

#include "segments.hpp"

namespace bo {

constexpr int kStreamMaxCap = 4096;  // largest LDS budget (nonzeros) of a stream block

// lanes per segment in the reduce phase of a block holding nseg segments
__device__ inline int stream_group(int nseg) {
    int L = 64;
    while (L > 1 && L * nseg > kBlock) L >>= 1;
    return L;
}

template <int L>
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block b is blk[b] = {first segment, end segment, first nonzero, end
// nonzero}.  Before anything waits, a workgroup has in flight its
// descriptor and the done flag, then -- once the descriptor is in -- the
// matrix entries, the bounds of the segment it reduces in phase 2 and that
// segment's epilogue data (split epilogues), so a pass costs three dependent
// memory round trips (descriptor, entries, gathers) before its stores.
template <int CAP, class TV, class TX, class Epi>
__global__ void __launch_bounds__(kBlock)
k_stream(const int4* __restrict__ blk, const int* __restrict__ ptr, const int* __restrict__ idx,
         const TV* __restrict__ val, Epi epi, const int* __restrict__ done) {
    static_assert(Epi::NV == 1 && Epi::Q == 0, "stream passes gather one vector and keep no partials");
    __shared__ double prod[CAP];
    __shared__ double red[kBlock / kWave];
    const int4 d = blk[blockIdx.x];
    const int dn = done ? *done : 0;
    const int s0 = d.x, s1 = d.y, k0 = d.z, k1 = d.w;
    const int tid = threadIdx.x;
    const TX* xv = epi.vec(0);
    // branch-free loads and gathers for FP32 gathered vectors only (measured
    // faster there, in the staircase LP's row pass); the FP64 instantiations
    // keep the per-entry branches
    constexpr bool kFlat = sizeof(TX) == 4;
    if (k1 - k0 > CAP) {  // one long segment: register accumulation, block reduction
        decltype(epi_preload(epi, 0, 0)) pre{};
        if (tid == 0) pre = epi_preload(epi, s0, 0);
        if (dn) return;
        double acc = 0.0;
        for (int k = k0 + tid; k < k1; k += 4 * kBlock) {
            TV a[4];
            int j[4];
            if constexpr (kFlat) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {  // clamped loads, selected afterwards (k < k1 here)
                    const int kk = k + u * kBlock;
                    const int kc = kk < k1 ? kk : k1 - 1;
                    const TV av = val[kc];
                    const int jv = idx[kc];
                    a[u] = kk < k1 ? av : TV(0);
                    j[u] = kk < k1 ? jv : -1;
                }
                TX g[4];  // unconditional gathers (entry 0 past the end): all 4 in flight at once
#pragma unroll
                for (int u = 0; u < 4; ++u) g[u] = xv[j[u] >= 0 ? j[u] : 0];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    acc = j[u] >= 0 ? acc + static_cast<double>(a[u]) * static_cast<double>(g[u]) : acc;
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int kk = k + u * kBlock;
                    a[u] = kk < k1 ? val[kk] : TV(0);
                    j[u] = kk < k1 ? idx[kk] : -1;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (j[u] >= 0) acc += static_cast<double>(a[u]) * static_cast<double>(xv[j[u]]);
            }
        }
        acc = group_sum<kWave>(acc);
        if ((tid % kWave) == 0) red[tid / kWave] = acc;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < kBlock / kWave; ++w) t += red[w];
            epi_finish(epi, s0, t, pre, 0);
        }
        return;
    }
    // phase-2 set-up, in flight with the entries: the segment this thread
    // group reduces, its bounds, and (lane 0) its epilogue data
    const int nseg = s1 - s0;
    const int L = stream_group(nseg);
    const int seg = s0 + tid / L;
    const int lane = tid % L;
    int a0 = 0, a1 = 0;
    decltype(epi_preload(epi, 0, 0)) pre{};
    if (seg < s1) {
        a0 = ptr[seg] - k0;
        a1 = ptr[seg + 1] - k0;
        if (lane == 0) pre = epi_preload(epi, seg, 0);
    }
    // phase 1: products of the block's nonzeros into LDS
    constexpr int U = CAP / kBlock;
    if (kFlat && k1 > k0) {  // uniform; clamped loads (k1 - 1 is the block's), selected afterwards
        TV a[U];
        int j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k0 + tid + u * kBlock;
            const int kc = kk < k1 ? kk : k1 - 1;
            const TV av = val[kc];
            const int jv = idx[kc];
            a[u] = kk < k1 ? av : TV(0);
            j[u] = kk < k1 ? jv : -1;
        }
        if (dn) return;  // uniform; nothing has been written
        // unconditional gathers (entry 0 past the block's nonzeros) so that
        // all U are in flight at once; products past k1 are never read
        TX g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) g[u] = xv[j[u] >= 0 ? j[u] : 0];
#pragma unroll
        for (int u = 0; u < U; ++u) prod[tid + u * kBlock] = static_cast<double>(a[u]) * static_cast<double>(g[u]);
    } else if (!kFlat) {
        if (dn) return;
        TV a[U];
        int j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k0 + tid + u * kBlock;
            a[u] = kk < k1 ? val[kk] : TV(0);
            j[u] = kk < k1 ? idx[kk] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j[u] >= 0) prod[tid + u * kBlock] = static_cast<double>(a[u]) * static_cast<double>(xv[j[u]]);
    } else if (dn) {
        return;
    }
    __syncthreads();
    // phase 2: segment sums from LDS, L lanes per segment
    double acc = 0.0;
    for (int q = a0 + lane; q < a1; q += L) acc += prod[q];
    switch (L) {  // uniform across the block
        case 64: acc = group_sum<64>(acc); break;
        case 32: acc = group_sum<32>(acc); break;
        case 16: acc = group_sum<16>(acc); break;
        case 8: acc = group_sum<8>(acc); break;
        case 4: acc = group_sum<4>(acc); break;
        case 2: acc = group_sum<2>(acc); break;
        default: break;
    }
    if (seg < s1 && lane == 0) epi_finish(epi, seg, acc, pre, 0);
}

}  // namespace bo
