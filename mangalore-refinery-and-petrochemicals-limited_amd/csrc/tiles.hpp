// tiles.hpp -- two-level ("tiled") schedule of a fused iteration pass.
//
// The gathered vector is cut into B blocks of W consecutive entries that fit
// in LDS.  A tile is one 1024-thread workgroup: it owns one block and a
// contiguous range of segments (rows or columns), copies its block into LDS
// once with 16-byte loads, then walks the nonzeros of its segments that fall
// into the block -- its (block, segment) pairs -- gathering from LDS.
//
// Layout of a tile's pairs (built on the host):
//   * pairs sorted by length; 64 of them form a SELL slice (entry j of lane l
//     at slice start + 64 j + l, padded to the slice's longest pair).  A long
//     pair (128 or more nonzeros) gets L lanes instead, L = 8 .. 64 so that
//     it spans 16 to 31 rows: 64 / L such pairs share a slice, entry j of
//     member p at row j / L, lane p L + j % L, and at the slice's end each
//     group of L lanes adds its values with lane moves (DPP) into its last
//     lane;
//   * the slices are dealt to the 16 wavefronts longest-first (LPT) and each
//     wavefront's slices are stored back to back, so a wavefront streams ONE
//     contiguous run of 64-entry rows: FP32 values and 16-bit block-local
//     indices, fully coalesced, read with raw buffer loads bounded to the run
//     (rows past the end read 0 without touching memory) and double-buffered
//     in registers so the next CH rows are always in flight;
//   * at the end of each slice the wavefront stores its 64 partial sums
//     (FP64) as one coalesced row of the partial array, in slice order.
// A reduction kernel adds, for every segment, its partials in block order
// (through a segment -> partial-slot index built on the host) and runs the
// pass epilogue.  Results are deterministic: fixed walk, fixed block order.
#pragma once

#include <algorithm>
#include <atomic>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace bo {

constexpr int kTileThreads = 1024;           // one workgroup per CU: the block fills most of the LDS
constexpr int kTileWaves = kTileThreads / kWave;
constexpr int kWideMin = 128;                // pair length from which a pair gets L > 1 lanes
constexpr int kLaneRows = 16;                // such a pair spans kLaneRows .. 2 kLaneRows - 1 rows
constexpr unsigned kLgShift = 28;            // slice word: rows | log2(lanes per pair) << kLgShift
constexpr unsigned kRowsMask = (1u << kLgShift) - 1u;
constexpr int kFillTrips = 10;               // 16-byte loads per thread covering a 160 KB block
constexpr int kTileRows = 8;                 // rows of a wavefront's run in flight per buffer

struct WaveDesc {
    unsigned off;  // first entry of the wavefront's rows (64 entries per row)
    int nrows;     // rows of the run
    int slice0;    // first slice (global index; the partial row of slice q is part[64 q ...])
    int nslice;    // slices of the run
};

struct TileDesc {
    int block;   // gathered block: entries [block * W, block * W + W)
    int nslice;
    int seg_lo;  // first and last segment the tile holds pairs of
    int seg_hi;
    WaveDesc wave[kTileWaves];
};

// Host-side layout of one pass.
struct HostTiles {
    int B = 0;  // blocks
    int W = 0;  // entries per block
    HostVec<TileDesc> tiles;
    HostVec<unsigned> slice_word;       // per slice: rows | log2(lanes per pair) << kLgShift
    HostVec<unsigned short> idx;        // padded entries: block-local gathered index
    HostVec<int> src;                   // padded entries: source nonzero, -1 for padding
    std::vector<int> seg_ptr;           // per segment: range of its partial slots in `slot`
    HostVec<unsigned> slot;             // partial slots (64 q + lane), block order within a segment
    size_t padded = 0;
    long long pairs = 0;
    int max_wave_slices = 0;
    int max_seg_pairs = 0;              // most pairs any segment has (partials to add)
};

// Builds the tiled layout of a segment matrix (ptr over nseg segments, gidx
// the gathered index of every nonzero, ng gathered entries).  wmax: largest
// block (a multiple of 64); target_tiles: about how many tiles to cut;
// wide_min: pair length from which a pair gets L > 1 lanes.
// Returns false when the padded layout does not fit 32-bit offsets.
// The work runs on up to 16 host threads (segment ranges for the pairs,
// tiles for the slices); the layout is the same for any thread count.
inline bool build_tiles(const int* ptr, const int* gidx, int nseg, int ng, int wmax, int target_tiles, HostTiles& h,
                        int wide_min = kWideMin) {
    h = HostTiles{};
    if (nseg <= 0 || ng <= 0) {
        h.seg_ptr.assign(static_cast<size_t>(std::max(nseg, 0)) + 1, 0);
        return true;
    }
    const int B = std::max(1, (ng + wmax - 1) / wmax);
    int W = (ng + B - 1) / B;
    W = (W + 63) / 64 * 64;
    h.B = B;
    h.W = W;
    const long long nnz = ptr[nseg];
    const int nth = host_thread_count(1 + nnz / (1 << 19));
    // Nonzeros grouped by block, segment-ascending inside a block; runs of one
    // segment inside a block are its (block, segment) pairs.  Thread t takes
    // a range of segments: it counts its nonzeros and pairs per block, and
    // writes them after those of threads 0 .. t-1 in every block.
    struct Pair {
        int s;
        int len;
        long long start;  // into ent
    };
    // segment ranges of equal nonzero counts (segments may be sorted by length)
    std::vector<int> sbound(static_cast<size_t>(nth) + 1, nseg);
    sbound[0] = 0;
    for (int t = 1; t < nth; ++t)
        sbound[t] = static_cast<int>(std::lower_bound(ptr, ptr + nseg, static_cast<int>(nnz * t / nth)) - ptr);
    auto seg_range = [&](int t) { return std::pair<long long, long long>(sbound[t], sbound[t + 1]); };
    std::vector<long long> zc(static_cast<size_t>(nth) * B, 0), pc(static_cast<size_t>(nth) * B, 0);
    host_parallel(nth, [&](int t, int) {
        const auto r = seg_range(t);
        long long* z = zc.data() + static_cast<size_t>(t) * B;
        long long* q = pc.data() + static_cast<size_t>(t) * B;
        std::vector<int> last(B, -1);
        for (long long s = r.first; s < r.second; ++s)
            for (int k = ptr[s]; k < ptr[s + 1]; ++k) {
                const int c = gidx[k] / W;
                z[c]++;
                if (last[c] != s) {
                    q[c]++;
                    last[c] = static_cast<int>(s);
                }
            }
    });
    std::vector<long long> bstart(B + 1, 0);
    std::vector<std::vector<Pair>> pairs(B);
    for (int c = 0; c < B; ++c) {  // per (thread, block): first entry and first pair
        long long zs = bstart[c], ps = 0;
        for (int t = 0; t < nth; ++t) {
            const long long z = zc[static_cast<size_t>(t) * B + c], q = pc[static_cast<size_t>(t) * B + c];
            zc[static_cast<size_t>(t) * B + c] = zs;
            pc[static_cast<size_t>(t) * B + c] = ps;
            zs += z;
            ps += q;
        }
        bstart[c + 1] = zs;
        pairs[c].resize(static_cast<size_t>(ps));
    }
    HostVec<int> ent(static_cast<size_t>(nnz));  // every entry written below
    host_parallel(nth, [&](int t, int) {
        const auto r = seg_range(t);
        long long* zpos = zc.data() + static_cast<size_t>(t) * B;
        long long* ppos = pc.data() + static_cast<size_t>(t) * B;
        std::vector<int> last(B, -1);
        for (long long s = r.first; s < r.second; ++s)
            for (int k = ptr[s]; k < ptr[s + 1]; ++k) {
                const int c = gidx[k] / W;
                if (last[c] != s) {
                    pairs[c][ppos[c]++] = Pair{static_cast<int>(s), 0, zpos[c]};
                    last[c] = static_cast<int>(s);
                }
                pairs[c][ppos[c] - 1].len++;
                ent[zpos[c]++] = k;
            }
    });
    // Tiles: consecutive pairs of one block with equal work (nonzeros plus a
    // per-pair cost for its partial).  Blocks get tile counts in proportion
    // to their work (largest remainders), so the layout has target_tiles
    // tiles -- a whole number of rounds of resident workgroups.
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
            if (q > p0 && cum * ntile[c] >= cut * bwork[c]) {
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

    // lanes of a pair of `len` nonzeros: 1 below wide_min, else the power of
    // two (8 .. 64) that makes it span kLaneRows .. 2 kLaneRows - 1 rows
    auto lanes_lg = [wide_min](int len) {
        if (len < wide_min) return 0;
        int lg = 3;
        while (lg < 6 && len >= (2 << lg) * kLaneRows) ++lg;
        return lg;
    };
    // Each tile on its own (tiles on the threads), with offsets local to the
    // tile: its descriptor, slice words, padded entries and pairs.
    struct Local {
        TileDesc td{};
        std::vector<unsigned> words;
        std::vector<unsigned short> idx;
        std::vector<int> src;
        std::vector<int> pseg;       // pairs: segment,
        std::vector<unsigned> pslot;  // and slot (64 local slice + lane)
        int max_wave_slices = 0;
        bool fits = true;
    };
    const int nspan = static_cast<int>(spans.size());
    std::vector<Local> loc(static_cast<size_t>(nspan));
    std::atomic<int> next{0};
    host_parallel(host_thread_count(nspan), [&](int, int) {
        struct Slice {
            std::vector<size_t> members;  // pair indices (into pairs[c])
            int rows;
            int lg;  // log2 of the lanes each member gets
        };
        std::vector<size_t> order;
        for (int si = next++; si < nspan; si = next++) {
            const Span& sp = spans[static_cast<size_t>(si)];
            Local& L = loc[static_cast<size_t>(si)];
            const std::vector<Pair>& P = pairs[sp.c];
            order.resize(sp.p1 - sp.p0);
            std::iota(order.begin(), order.end(), sp.p0);
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return P[a].len > P[b].len; });
            // slices, longest pairs first: pairs of one lane count together, 64 / L
            // per slice (L lanes each), padded to the slice's longest member
            std::vector<Slice> sl;
            for (size_t q = 0; q < order.size();) {
                const int lg = lanes_lg(P[order[q]].len);
                const int lanes = 1 << lg;
                Slice s{{}, (P[order[q]].len + lanes - 1) / lanes, lg};
                size_t r = q;
                for (; r < order.size() && r < q + static_cast<size_t>(kWave / lanes) && lanes_lg(P[order[r]].len) == lg;
                     ++r)
                    s.members.push_back(order[r]);
                sl.push_back(std::move(s));
                q = r;
            }
            // deal the slices to the wavefronts, longest first, to the least loaded
            std::vector<int> owner(sl.size());
            std::vector<long long> load(kTileWaves, 0);
            std::vector<size_t> by_rows(sl.size());
            std::iota(by_rows.begin(), by_rows.end(), 0);
            std::stable_sort(by_rows.begin(), by_rows.end(),
                             [&](size_t a, size_t b) { return sl[a].rows > sl[b].rows; });
            for (size_t i : by_rows) {
                const int w = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
                owner[i] = w;
                load[w] += sl[i].rows + 1;
            }
            size_t padded = 0;
            for (const Slice& s : sl) padded += static_cast<size_t>(s.rows) * kWave;
            L.idx.assign(padded, 0);
            L.src.assign(padded, -1);
            L.td.block = sp.c;
            L.td.seg_lo = P[sp.p0].s;
            L.td.seg_hi = P[sp.p1 - 1].s;
            size_t off = 0;
            for (int w = 0; w < kTileWaves; ++w) {
                WaveDesc& wd = L.td.wave[w];
                wd.off = static_cast<unsigned>(off);  // local; made global below
                wd.slice0 = static_cast<int>(L.words.size());
                for (size_t i : by_rows) {
                    if (owner[i] != w) continue;
                    const Slice& s = sl[i];
                    const unsigned lslice = static_cast<unsigned>(L.words.size());
                    if (static_cast<unsigned>(s.rows) > kRowsMask) L.fits = false;
                    L.words.push_back(static_cast<unsigned>(s.rows) | (static_cast<unsigned>(s.lg) << kLgShift));
                    const int lanes = 1 << s.lg;
                    for (size_t mi = 0; mi < s.members.size(); ++mi) {
                        const Pair& p = P[s.members[mi]];
                        const size_t lane0 = mi * static_cast<size_t>(lanes);
                        for (int j = 0; j < p.len; ++j) {
                            const int k = ent[p.start + j];
                            const size_t e = off + static_cast<size_t>(j / lanes) * kWave + lane0 + (j % lanes);
                            L.idx[e] = static_cast<unsigned short>(gidx[k] - sp.c * W);
                            L.src[e] = k;
                        }
                        L.pseg.push_back(p.s);
                        // the member's sum ends in the last lane of its group
                        L.pslot.push_back(lslice * kWave + static_cast<unsigned>(lane0 + lanes - 1));
                    }
                    off += static_cast<size_t>(s.rows) * kWave;
                    wd.nrows += s.rows;
                    wd.nslice++;
                }
                L.td.nslice += wd.nslice;
                L.max_wave_slices = std::max(L.max_wave_slices, wd.nslice);
            }
        }
    });
    // Global offsets in tile order: padded entries, slices, pairs.
    std::vector<size_t> poff(static_cast<size_t>(nspan) + 1, 0), soff(static_cast<size_t>(nspan) + 1, 0),
        qoff(static_cast<size_t>(nspan) + 1, 0);
    for (int si = 0; si < nspan; ++si) {
        const Local& L = loc[static_cast<size_t>(si)];
        if (!L.fits) return false;
        poff[si + 1] = poff[si] + L.idx.size();
        soff[si + 1] = soff[si] + L.words.size();
        qoff[si + 1] = qoff[si] + L.pseg.size();
        h.max_wave_slices = std::max(h.max_wave_slices, L.max_wave_slices);
    }
    if (poff[nspan] > 0xFFFFFFFFull || soff[nspan] > (1u << 26)) return false;
    h.padded = poff[nspan];
    h.pairs = static_cast<long long>(qoff[nspan]);
    h.tiles.resize(static_cast<size_t>(nspan));
    h.slice_word.resize(soff[nspan]);
    h.idx.resize(h.padded);
    h.src.resize(h.padded);
    HostVec<int> pair_seg(qoff[nspan]);  // these and h's arrays: every entry written below
    HostVec<unsigned> pair_slot(qoff[nspan]);
    host_parallel(host_thread_count(nspan), [&](int t, int nt) {
        const auto r = host_part(nspan, t, nt);
        for (long long si = r.first; si < r.second; ++si) {
            const Local& L = loc[static_cast<size_t>(si)];
            TileDesc td = L.td;
            for (int w = 0; w < kTileWaves; ++w) {
                td.wave[w].off += static_cast<unsigned>(poff[si]);
                td.wave[w].slice0 += static_cast<int>(soff[si]);
            }
            h.tiles[static_cast<size_t>(si)] = td;
            std::copy(L.words.begin(), L.words.end(), h.slice_word.begin() + static_cast<long long>(soff[si]));
            std::copy(L.idx.begin(), L.idx.end(), h.idx.begin() + static_cast<long long>(poff[si]));
            std::copy(L.src.begin(), L.src.end(), h.src.begin() + static_cast<long long>(poff[si]));
            for (size_t q = 0; q < L.pseg.size(); ++q) {
                pair_seg[qoff[si] + q] = L.pseg[q];
                pair_slot[qoff[si] + q] = L.pslot[q] + static_cast<unsigned>(soff[si]) * kWave;
            }
        }
    });
    loc.clear();
    // segment -> partial slots, in block order: a counting sort by segment
    // over the pairs taken block by block (a segment has at most one pair
    // per block)
    h.seg_ptr.assign(static_cast<size_t>(nseg) + 1, 0);
    for (int s : pair_seg) h.seg_ptr[s + 1]++;
    for (int s = 0; s < nseg; ++s) h.seg_ptr[s + 1] += h.seg_ptr[s];
    h.slot.resize(pair_seg.size());
    {
        std::vector<int> pos(h.seg_ptr.begin(), h.seg_ptr.end() - 1);
        for (int c = 0; c < B; ++c)
            for (int si = 0; si < nspan; ++si) {
                if (spans[static_cast<size_t>(si)].c != c) continue;
                for (size_t q = qoff[si]; q < qoff[si + 1]; ++q) h.slot[pos[pair_seg[q]]++] = pair_slot[q];
            }
    }
    for (int s = 0; s < nseg; ++s) h.max_seg_pairs = std::max(h.max_seg_pairs, h.seg_ptr[s + 1] - h.seg_ptr[s]);
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

// A double moved between lanes by a DPP control (both 32-bit halves); lanes
// whose source is outside the pattern, or whose row is masked off, get 0.
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over each group of 2^lg consecutive lanes, in the group's last lane
// (the other lanes hold partial sums): pairs and quads by quad permutes, 8
// and 16 lanes by row shifts of 4 and 8, 32 lanes by broadcasting row 0 into
// row 1 and row 2 into row 3, all 64 by broadcasting lane 31 into rows 2-3.
// lg is uniform across the wavefront.  Lane moves run on the vector ALU,
// with no LDS round trip between the steps.
__device__ inline double lane_group_sum(double v, int lg) {
    if (lg >= 1) v += dpp_move<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    if (lg >= 2) v += dpp_move<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    if (lg >= 3) v += dpp_move<0x114, 0xF>(v);  // row_shr:4
    if (lg >= 4) v += dpp_move<0x118, 0xF>(v);  // row_shr:8
    if (lg >= 5) v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    if (lg >= 6) v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// CH rows of a wavefront's run, starting at row r (rows past the run read 0).
template <class T, int CH>
struct RowChunk {
    T a[CH];
    unsigned short g[CH];
};

template <class T, int CH>
__device__ inline void chunk_load(RowChunk<T, CH>& c, __amdgpu_buffer_rsrc_t rv, __amdgpu_buffer_rsrc_t ri, int r,
                                  int lane) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        c.a[u] = buf_val<T>(rv, ((r + u) * kWave + lane) * static_cast<int>(sizeof(T)));
        c.g[u] = buf_idx(ri, ((r + u) * kWave + lane) * 2);
    }
}

// Walk state of a wavefront: the running sum of the current slice, the row
// after its last row, and the slice words of the run (lane q holds the word
// of slice q of the current window of 64 slices).
struct RunState {
    double acc;
    int sl;     // current slice within the run
    int bound;  // row after the current slice's last row
    int lg;     // log2 of the lanes per member of the current slice
    unsigned words;
};

__device__ inline unsigned run_word(const unsigned* __restrict__ slice_word, int first, int nslice, int base,
                                    int lane) {
    const int q = base + lane;
    return q < nslice ? slice_word[first + q] : 0u;
}

template <class T, class TX, int CH, bool EARLY>
__device__ inline void chunk_finish(const RowChunk<T, CH>& c, int r, const TX* lds, RunState& st,
                                    const WaveDesc& wd, const unsigned* __restrict__ slice_word,
                                    double* __restrict__ part, int lane) {
    // EARLY: the chunk's LDS gathers are all issued before the first is
    // used, one LDS latency per chunk instead of one per row.  It pays when
    // slices span several chunks: on the 200,000 x 500,000 Zipf LP the row
    // pass gains 1 us, and the column pass, once long columns took 8-64
    // lanes (slices of 11 rows on average instead of 5.6), 1.6 us (27.9
    // against 29.5 us); with the short slices it had cost the column pass
    // 1.3 us.
    TX g[CH];
    if constexpr (EARLY) {
#pragma unroll
        for (int u = 0; u < CH; ++u) g[u] = lds[c.g[u]];
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const TX gv = EARLY ? g[u] : lds[c.g[u]];
        st.acc += static_cast<double>(c.a[u]) * static_cast<double>(gv);
        if (r + u + 1 == st.bound) {  // last row of the current slice (uniform)
            double* row = part + static_cast<size_t>(wd.slice0 + st.sl) * kWave;
            // one coalesced row per slice; a member's sum is in its group's last lane
            row[lane] = st.lg ? lane_group_sum(st.acc, st.lg) : st.acc;
            st.acc = 0.0;
            if (++st.sl < wd.nslice) {
                // The refill of the word window waits for its own load (and
                // so for every matrix load in flight) on its own rare path:
                // read at a join of both paths, the common path would wait
                // too, at every slice end.
                unsigned w;
                if ((st.sl & (kWave - 1)) == 0) {
                    st.words = run_word(slice_word, wd.slice0, wd.nslice, st.sl, lane);
                    w = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(st.words)));
                } else {
                    w = static_cast<unsigned>(
                        __builtin_amdgcn_readlane(static_cast<int>(st.words), st.sl & (kWave - 1)));
                }
                st.bound += static_cast<int>(w & kRowsMask);
                st.lg = static_cast<int>(w >> kLgShift);
            }
        }
    }
}

// One tile per workgroup.  T: matrix values; TX: the gathered vector, W
// entries of it in dynamic LDS (FP32 values may meet an FP64 vector).  CH:
// rows of the wavefront's run kept in flight per buffer.  PASS names the
// instantiation (0 / 1: CSR / CSC pass of the single-LP engine, 2 / 3: of
// the batched engine), so that profiles tell the passes apart; every pass
// issues a chunk's LDS gathers early (chunk_finish).  active (optional):
// per-tile flag; a tile whose flag is 0 does nothing.  It comes first, among
// the arguments the dispatch preloads into SGPRs, so testing it costs no
// argument load.  The kernel reads no done flag: its partials only reach
// the state through the reduction, whose epilogue stores nothing once the
// LP is done.  (A flag test here costs a scalar round trip: the compiler
// moves the flag's load down to the test.)
template <class T, class TX, int CH, int PASS>
__global__ void __launch_bounds__(kTileThreads)
k_tile(const unsigned char* __restrict__ active, const TileDesc* __restrict__ tiles,
       const unsigned* __restrict__ slice_word, const T* __restrict__ val, const unsigned short* __restrict__ idx,
       const TX* __restrict__ gvec, int ng, int W, double* __restrict__ part) {
    if (active && !active[blockIdx.x]) return;  // every LP of the tile has finished (batched engine)
    extern __shared__ __align__(16) unsigned char tile_lds[];
    TX* lds = reinterpret_cast<TX*>(tile_lds);
    const int tid = threadIdx.x;
    // wave-uniform values live in scalar registers, so nothing below
    // branches per lane
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave), lane = tid % kWave;
    const int block = tiles[blockIdx.x].block;
    const WaveDesc wd = tiles[blockIdx.x].wave[wave];
    // Start-up: the block's 16-byte loads (into registers), the first row
    // chunk and the run's slice words are all issued before anything waits,
    // so the LDS fill overlaps the first matrix fetch.
    const int g0 = block * W;
    const int gn = min(W, ng - g0);
    constexpr int V = 16 / sizeof(TX);
    const int nv = gn / V;  // whole 16-byte groups of the block (g0 is a multiple of 64)
    const __amdgpu_buffer_rsrc_t rx = slice_rsrc(gvec + g0, nv * 16);
    uint4 f[kFillTrips];
#pragma unroll
    for (int u = 0; u < kFillTrips; ++u)
        f[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rx, (u * kTileThreads + tid) * 16, 0, 0));
    const int nrows = wd.nrows;
    const __amdgpu_buffer_rsrc_t rv = slice_rsrc(val + wd.off, nrows * kWave * static_cast<int>(sizeof(T)));
    const __amdgpu_buffer_rsrc_t ri = slice_rsrc(idx + wd.off, nrows * kWave * 2);
    RowChunk<T, CH> A, Bc;
    chunk_load(A, rv, ri, 0, lane);
    RunState st{0.0, 0, 0, 0, run_word(slice_word, wd.slice0, wd.nslice, 0, lane)};
    uint4* dst = reinterpret_cast<uint4*>(tile_lds);
#pragma unroll
    for (int u = 0; u < kFillTrips; ++u) {
        const int i = u * kTileThreads + tid;
        if (i < nv) dst[i] = f[u];
    }
    for (int i = nv * V + tid; i < gn; i += kTileThreads) lds[i] = gvec[g0 + i];  // tail of the last block
    __syncthreads();
    if (nrows == 0) return;
    {
        const unsigned w0 = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(st.words), 0));
        st.bound = static_cast<int>(w0 & kRowsMask);
        st.lg = static_cast<int>(w0 >> kLgShift);
    }
    // Double-buffered walk: the next CH rows are loaded before the current
    // ones are gathered and their finished slices stored, so no store sits
    // between a load and its use.  A counted loop with one exit: with an
    // exit between the halves, the compiler's wait analysis merged that
    // path's unconsumed loads into the loop head and waited there for every
    // load in flight.  Rows past the run load value 0 and index 0 without
    // touching memory; they add 0 to a sum that is no longer stored (no
    // slice ends past the run's last row), and every chunk's registers are
    // consumed on every path, so no load is left pending at the loop head.
    int r = 0;
    for (int pairs = (nrows + 2 * CH - 1) / (2 * CH); pairs > 0; --pairs) {
        chunk_load(Bc, rv, ri, r + CH, lane);
        chunk_finish<T, TX, CH, true>(A, r, lds, st, wd, slice_word, part, lane);
        chunk_load(A, rv, ri, r + 2 * CH, lane);
        chunk_finish<T, TX, CH, true>(Bc, r + CH, lds, st, wd, slice_word, part, lane);
        r += 2 * CH;
    }
}

// Segment s: its partials in block order, then the pass epilogue.  The slot
// and partial loads of up to U pairs are issued at once: raw buffer loads at
// an out-of-range offset read 0, so no load sits under a per-lane branch.
// An epilogue may skip a segment before anything is loaded (Epi::skip).
// The segment's bounds are loaded before its epilogue data (vector-load
// waits are in order: the loop bounds do not wait for the epilogue data).
// The kernel reads no done flag of its own: the pass epilogues load their
// LP's with the epilogue data and store nothing once it is set, so a
// reduction that runs after convergence changes nothing, at no round trip.
template <int U, class Epi>
__global__ void __launch_bounds__(kBlock)
k_tile_reduce(int nseg, const int* __restrict__ seg_ptr, const unsigned* __restrict__ slot, int nslot,
              const double* __restrict__ part, int npart, Epi epi) {
    if (wg_dead(epi, static_cast<int>(blockIdx.x), 0)) return;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= nseg || epi.skip(s)) return;
    constexpr unsigned kOut = 0x7FFFFFF0u;  // beyond every range below (ranges are < 2^31 bytes)
    const __amdgpu_buffer_rsrc_t rs = slice_rsrc(slot, nslot * 4);
    const __amdgpu_buffer_rsrc_t rp = slice_rsrc(part, npart * 8);
    const int q0 = seg_ptr[s], q1 = seg_ptr[s + 1];
    const auto pre = epi_preload(epi, s, 0);
    double acc = 0.0;
    for (int q = q0; q < q1; q += U) {
        unsigned p[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            p[u] = __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(q + u < q1 ? (q + u) * 4u : kOut), 0, 0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            v[u] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                                  rp, static_cast<int>(q + u < q1 ? p[u] * 8u : kOut), 0, 0));
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u];
    }
    epi_finish(epi, s, acc, pre, 0);
}

// Launch of the reduction with U loads in flight per round: `u` when given
// (a tuned value: a segment's partials are still added in order, so any U
// gives the same sums), else the smallest U covering the most pairs a
// segment has.  ev (optional): events timing the launch (launch_ev).
template <class Epi>
inline void launch_tile_reduce(hipStream_t stream, int max_pairs, int nseg, const int* seg_ptr, const unsigned* slot,
                               int nslot, const double* part, int npart, const Epi& e, int u = 0,
                               const hipEvent_t* ev = nullptr) {
    const dim3 grid(grid_for(static_cast<size_t>(nseg))), block(kBlock);
    if (u > 0) max_pairs = u;
    if (max_pairs <= 1)
        launch_ev(k_tile_reduce<1, Epi>, grid, block, 0, stream, ev, nseg, seg_ptr, slot, nslot, part, npart, e);
    else if (max_pairs <= 2)
        launch_ev(k_tile_reduce<2, Epi>, grid, block, 0, stream, ev, nseg, seg_ptr, slot, nslot, part, npart, e);
    else if (max_pairs <= 4)
        launch_ev(k_tile_reduce<4, Epi>, grid, block, 0, stream, ev, nseg, seg_ptr, slot, nslot, part, npart, e);
    else if (max_pairs <= 8)
        launch_ev(k_tile_reduce<8, Epi>, grid, block, 0, stream, ev, nseg, seg_ptr, slot, nslot, part, npart, e);
    else
        launch_ev(k_tile_reduce<16, Epi>, grid, block, 0, stream, ev, nseg, seg_ptr, slot, nslot, part, npart, e);
}

}  // namespace bo
