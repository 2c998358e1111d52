// tiles.hpp -- two-level ("tiled") schedule of a fused iteration pass:
// the gathered vector is cut into blocks of W consecutive entries that fit in
// LDS, and every workgroup (a tile) owns one block and a contiguous range of
// segments.  It copies its block into LDS with coalesced loads once, then
// walks the nonzeros of its segments that fall into the block: they are
// stored as SELL-64 slices (64 segments per slice, sorted by length inside
// the tile, entry j of lane l at slice_offset + 64 j + l, padded to the
// slice's longest segment) with 16-bit block-local indices, so every global
// load is a fully coalesced stream (4 + 2 bytes per nonzero for FP32) and
// every gather is an LDS read.  A (block, segment) pair with at least
// kWideMin nonzeros gets a "wide" slice of its own instead: its entries are
// contiguous, the 64 lanes of a wave stride over them and a fixed xor tree
// adds the lanes, so one long segment never pads 63 short ones.  Each tile
// writes one FP64 partial per (block, segment) pair it holds; a reduction
// kernel adds the B partials of a segment in block order and runs the pass's
// epilogue.  Pairs without nonzeros are never written: their slots stay at
// the zero they were initialised with.  Results are deterministic (fixed
// slice walk, fixed block order).
#pragma once

#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace bo {

constexpr int kTileThreads = 1024;                 // one workgroup per CU: the block fills most of the LDS
constexpr int kTileWaves = kTileThreads / kWave;
template <int CH>
constexpr int tile_slices_per_wave() { return CH > 8 ? 1 : 2; }  // SELL slices a wave loads at once
constexpr unsigned short kNoSeg = 0xFFFF;          // perm entry of an unused lane
constexpr int kTileMaxSpan = 65535;                // local segment ids stay below this (16-bit, kNoSeg excluded)
constexpr int kWideMin = 128;                      // pair length from which a pair gets a wide slice
constexpr unsigned kWideFlag = 0x80000000u;        // slice length word: wide slice

struct TileDesc {
    int block;   // gathered block: entries [block * W, block * W + W)
    int seg0;    // first segment of the tile's range
    int slice0;  // first SELL slice (the wide slices come first)
    int nslice;  // slices of the tile
    int nwide;   // wide slices among them
    int pad[3];
};

// Host-side layout of one pass.
struct HostTiles {
    int B = 0;  // blocks
    int W = 0;  // entries per block
    std::vector<TileDesc> tiles;
    std::vector<uint2> slices;          // x: first padded entry, y: length (| kWideFlag: wide slice)
    std::vector<unsigned short> perm;   // 64 per slice: segment - seg0, or kNoSeg
    std::vector<unsigned short> idx;    // padded entries: block-local gathered index
    std::vector<int> src;               // padded entries: source nonzero, -1 for padding
    size_t padded = 0;
};

// Builds the tiled layout of a segment matrix (ptr over nseg segments, gidx
// the gathered index of every nonzero, ng gathered entries).  wmax: largest
// block (a multiple of 64); target_tiles: about how many tiles to cut.
inline bool build_tiles(const int* ptr, const int* gidx, int nseg, int ng, int wmax, int target_tiles, HostTiles& h) {
    h = HostTiles{};
    if (nseg <= 0 || ng <= 0) return true;
    const int B = std::max(1, (ng + wmax - 1) / wmax);
    int W = (ng + B - 1) / B;
    W = (W + 63) / 64 * 64;
    h.B = B;
    h.W = W;
    const long long nnz = ptr[nseg];
    // nonzeros grouped by block, segment-ascending inside a block; runs of one
    // segment inside a block are its (block, segment) pairs
    std::vector<long long> bstart(B + 1, 0);
    for (long long k = 0; k < nnz; ++k) bstart[gidx[k] / W + 1]++;
    for (int c = 0; c < B; ++c) bstart[c + 1] += bstart[c];
    std::vector<int> ent(static_cast<size_t>(nnz));
    struct Pair {
        int s;
        int len;
        long long start;  // into ent
    };
    std::vector<std::vector<Pair>> pairs(B);
    {
        std::vector<long long> pos(bstart.begin(), bstart.end() - 1);
        std::vector<int> last(B, -1);
        for (int s = 0; s < nseg; ++s) {
            for (int k = ptr[s]; k < ptr[s + 1]; ++k) {
                const int c = gidx[k] / W;
                if (last[c] != s) {
                    pairs[c].push_back(Pair{s, 0, pos[c]});
                    last[c] = s;
                }
                pairs[c].back().len++;
                ent[pos[c]++] = k;
            }
        }
    }
    // tiles: consecutive pairs of one block, equal work (nonzeros plus a
    // per-pair cost for the partial and its index).  Blocks get tile counts
    // in proportion to their work (largest remainders), so the layout has
    // target_tiles tiles -- a whole number of rounds of resident workgroups --
    // unless a block needs more to keep local segment ids in 16 bits.
    std::vector<long long> bwork(B, 0);
    long long total = 0;
    for (int c = 0; c < B; ++c) {
        for (const Pair& p : pairs[c]) bwork[c] += p.len + 2;
        total += bwork[c];
    }
    // at least ~4096 units of work per tile: small problems get few tiles
    const int T = static_cast<int>(std::max<long long>(1, std::min<long long>(target_tiles, total / 4096)));
    std::vector<int> ntile(B, 0);
    {
        std::vector<std::pair<double, int>> frac;
        int used = 0;
        for (int c = 0; c < B; ++c) {
            if (bwork[c] == 0) continue;
            const double quota = total > 0 ? static_cast<double>(T) * bwork[c] / total : 0.0;
            ntile[c] = std::max(1, static_cast<int>(quota));
            used += ntile[c];
            frac.emplace_back(quota - ntile[c], c);
        }
        std::stable_sort(frac.begin(), frac.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        for (size_t q = 0; used < T && q < frac.size(); ++q, ++used) ntile[frac[q].second]++;
    }
    struct Span {
        int c;
        size_t p0, p1;
        long long work;
    };
    std::vector<Span> spans;
    for (int c = 0; c < B; ++c) {
        if (pairs[c].empty()) continue;
        size_t p0 = 0;
        long long w = 0, cum = 0;
        int cut = 1;  // next equal-work boundary: cut * bwork / ntile
        for (size_t q = 0; q < pairs[c].size(); ++q) {
            const bool full = cum * ntile[c] >= cut * bwork[c];
            if (q > p0 && (full || pairs[c][q].s - pairs[c][p0].s >= kTileMaxSpan)) {
                spans.push_back(Span{c, p0, q, w});
                p0 = q;
                w = 0;
                while (cum * ntile[c] >= cut * bwork[c]) ++cut;
            }
            w += pairs[c][q].len + 2;
            cum += pairs[c][q].len + 2;
        }
        spans.push_back(Span{c, p0, pairs[c].size(), w});
    }
    // heaviest tiles first (longest-processing-time-first dispatch)
    std::stable_sort(spans.begin(), spans.end(), [](const Span& a, const Span& b) { return a.work > b.work; });
    std::vector<size_t> order;
    for (const Span& sp : spans) {
        const std::vector<Pair>& P = pairs[sp.c];
        order.resize(sp.p1 - sp.p0);
        std::iota(order.begin(), order.end(), sp.p0);
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return P[a].len > P[b].len; });
        TileDesc td{sp.c, P[sp.p0].s, static_cast<int>(h.slices.size()), 0, 0, {0, 0, 0}};
        size_t nwide = 0;
        while (nwide < order.size() && P[order[nwide]].len >= kWideMin) ++nwide;
        for (size_t q = 0; q < nwide; ++q) {  // wide slices: one pair, contiguous entries
            const Pair& p = P[order[q]];
            const size_t off = h.padded;
            if (off + static_cast<size_t>(p.len) > 0xFFFFFFFFull || static_cast<unsigned>(p.len) >= kWideFlag) return false;
            h.slices.push_back(make_uint2(static_cast<unsigned>(off), static_cast<unsigned>(p.len) | kWideFlag));
            h.padded += static_cast<size_t>(p.len);
            h.idx.resize(h.padded, 0);
            h.src.resize(h.padded, -1);
            for (int j = 0; j < p.len; ++j) {
                const int k = ent[p.start + j];
                h.idx[off + j] = static_cast<unsigned short>(gidx[k] - sp.c * W);
                h.src[off + j] = k;
            }
            h.perm.push_back(static_cast<unsigned short>(p.s - td.seg0));
            for (int l = 1; l < kWave; ++l) h.perm.push_back(kNoSeg);
            td.nslice++;
            td.nwide++;
        }
        for (size_t q0 = nwide; q0 < order.size(); q0 += kWave) {
            const size_t q1 = std::min(order.size(), q0 + kWave);
            const int L = P[order[q0]].len;
            const size_t off = h.padded;
            if (off + static_cast<size_t>(L) * kWave > 0xFFFFFFFFull) return false;
            h.slices.push_back(make_uint2(static_cast<unsigned>(off), static_cast<unsigned>(L)));
            h.padded += static_cast<size_t>(L) * kWave;
            h.idx.resize(h.padded, 0);
            h.src.resize(h.padded, -1);
            for (size_t q = q0; q < q0 + kWave; ++q) {
                if (q >= q1) {
                    h.perm.push_back(kNoSeg);
                    continue;
                }
                const Pair& p = P[order[q]];
                h.perm.push_back(static_cast<unsigned short>(p.s - td.seg0));
                const size_t lane = q - q0;
                for (int j = 0; j < p.len; ++j) {
                    const int k = ent[p.start + j];
                    const size_t e = off + static_cast<size_t>(j) * kWave + lane;
                    h.idx[e] = static_cast<unsigned short>(gidx[k] - sp.c * W);
                    h.src[e] = k;
                }
            }
            td.nslice++;
        }
        h.tiles.push_back(td);
    }
    return true;
}

// Tile values from the (scaled) segment-matrix values; padding gets 0.
template <class T>
__global__ void k_tile_values(const int* __restrict__ src, size_t n, const T* __restrict__ vals, T* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const int k = src[i];
        out[i] = k >= 0 ? vals[k] : T(0);
    }
}

// Raw buffer resource over `bytes` bytes at `base` (gfx9 dword 3: 32-bit
// data format); loads at or past `bytes` return 0 and access no memory.
__device__ inline __amdgpu_buffer_rsrc_t slice_rsrc(const void* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, bytes, 0x00020000);
}

template <class T>
__device__ inline T buf_val(__amdgpu_buffer_rsrc_t r, int byte_off) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
    }
}

__device__ inline unsigned short buf_idx(__amdgpu_buffer_rsrc_t r, int byte_off) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, byte_off, 0, 0);
}

// NS slices in flight for one wave: buffer resources limited to each
// slice's entries, the lane's perm entries and the first CH rows of 64
// entries (a wide slice is the same layout: entry j at offset j).
template <class T, int CH, int NS>
struct SliceBatch {
    __amdgpu_buffer_rsrc_t rv[NS], ri[NS];
    int rows[NS];
    int wide[NS];
    unsigned short r[NS];
    T a[NS][CH];
    unsigned short g[NS][CH];
};

// Issues every load of the batch starting at slice s (slices s, s + 16, ...).
// Slices at or past `end` become empty ones: their descriptor and perm reads
// go to slice end - 1 (the tile has at least one slice) and their buffer
// range is 0 bytes, so their loads read 0 and touch no memory.
template <class T, int CH, int NS>
__device__ inline void batch_load(SliceBatch<T, CH, NS>& b, int s, int end, const uint2* __restrict__ slices,
                                  const unsigned short* __restrict__ perm, const T* __restrict__ val,
                                  const unsigned short* __restrict__ idx, int lane) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const int sq = s + q * kTileWaves;
        const bool live = sq < end;
        const int sr = live ? sq : end - 1;
        const uint2 d = slices[sr];
        const unsigned x = __builtin_amdgcn_readfirstlane(d.x);
        const unsigned y = __builtin_amdgcn_readfirstlane(d.y);
        const int w = (y & kWideFlag) != 0;
        const int len = live ? (w ? static_cast<int>(y & ~kWideFlag) : static_cast<int>(y) * kWave) : 0;
        b.rows[q] = (len + kWave - 1) / kWave;
        b.wide[q] = w;
        b.rv[q] = slice_rsrc(val + x, len * static_cast<int>(sizeof(T)));
        b.ri[q] = slice_rsrc(idx + x, len * 2);
        const unsigned short pr = perm[static_cast<size_t>(sr) * kWave + lane];
        b.r[q] = live ? pr : kNoSeg;
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            b.a[q][u] = buf_val<T>(b.rv[q], (u * kWave + lane) * static_cast<int>(sizeof(T)));
            b.g[q][u] = buf_idx(b.ri[q], (u * kWave + lane) * 2);
        }
    }
}

// Gathers from the LDS block, finishes slices longer than CH rows, and
// stores one partial per segment (a wide slice adds its 64 lanes first).
template <class T, int CH, int NS>
__device__ inline void batch_finish(const SliceBatch<T, CH, NS>& b, const T* lds, double* __restrict__ out,
                                    int lane) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < CH; ++u) acc += static_cast<double>(b.a[q][u]) * static_cast<double>(lds[b.g[q][u]]);
        for (int j0 = CH; j0 < b.rows[q]; j0 += CH) {
            T c[CH];
            unsigned short h[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                c[u] = buf_val<T>(b.rv[q], ((j0 + u) * kWave + lane) * static_cast<int>(sizeof(T)));
                h[u] = buf_idx(b.ri[q], ((j0 + u) * kWave + lane) * 2);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) acc += static_cast<double>(c[u]) * static_cast<double>(lds[h[u]]);
        }
        if (b.wide[q]) acc = wave_sum(acc);
        if (b.r[q] != kNoSeg) out[b.r[q]] = acc;
    }
}

// One tile per workgroup.  Dynamic LDS: W entries of T.  CH: rows of 64
// entries of a slice loaded up front; NS: slices a wave loads at once.
// PASS only names the instantiation (0: CSR pass, 1: CSC pass), so that
// profiles tell the two passes apart.
template <class T, int CH, int NS, int PASS>
__global__ void __launch_bounds__(kTileThreads)
k_tile(const TileDesc* __restrict__ tiles, const uint2* __restrict__ slices, const unsigned short* __restrict__ perm,
       const T* __restrict__ val, const unsigned short* __restrict__ idx, const T* __restrict__ gvec, int ng, int W,
       int nseg, double* __restrict__ part, const int* __restrict__ done) {
    if (done && *done) return;
    extern __shared__ __align__(16) unsigned char tile_lds[];
    T* lds = reinterpret_cast<T*>(tile_lds);
    const TileDesc td = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    {
        const int g0 = td.block * W;
        const int gn = min(W, ng - g0);
        constexpr int V = 16 / sizeof(T);  // entries per 16-byte load; g0 is a multiple of 64
        using Vec = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
        const Vec* src = reinterpret_cast<const Vec*>(gvec + g0);
        Vec* dst = reinterpret_cast<Vec*>(lds);
        const int nv = gn / V;
        for (int i = tid; i < nv; i += kTileThreads) dst[i] = src[i];
        for (int i = nv * V + tid; i < gn; i += kTileThreads) lds[i] = gvec[g0 + i];
    }
    __syncthreads();
    // wave-uniform values live in scalar registers (slice bounds, buffer
    // resources), so nothing below branches per lane
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave), lane = tid % kWave;
    double* out = part + static_cast<size_t>(td.block) * nseg + td.seg0;
    const int end = td.slice0 + td.nslice;
    // Software pipeline over batches of NS slices (wave w takes slices w,
    // w + 16, ...): the loads of the next batch are issued before the
    // current batch is reduced and its partials stored, so each wave always
    // has a batch in flight and no store sits between a load and its use.
    constexpr int step = NS * kTileWaves;
    int s = td.slice0 + wave;
    if (s >= end) return;
    SliceBatch<T, CH, NS> A, Bt;
    batch_load(A, s, end, slices, perm, val, idx, lane);
    for (;;) {
        batch_load(Bt, s + step, end, slices, perm, val, idx, lane);
        batch_finish(A, lds, out, lane);
        s += step;
        if (s >= end) break;
        batch_load(A, s + step, end, slices, perm, val, idx, lane);
        batch_finish(Bt, lds, out, lane);
        s += step;
        if (s >= end) break;
    }
}

// Segment s: the B partials in block order, then the pass epilogue.
template <class Epi>
__global__ void __launch_bounds__(kBlock)
k_tile_reduce(int nseg, int B, const double* __restrict__ part, Epi epi, const int* __restrict__ done) {
    if (done && *done) return;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= nseg) return;
    // runs of 8 partial loads in flight; added in block order
    double acc = 0.0;
    int c = 0;
    for (; c + 8 <= B; c += 8) {
        double q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = part[static_cast<size_t>(c + u) * nseg + s];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += q[u];
    }
    for (; c < B; ++c) acc += part[static_cast<size_t>(c) * nseg + s];
    double r[1] = {acc};
    epi(s, r, nullptr);
}

}  // namespace bo
