// Host checks of the tiled layout (csrc/tiles.hpp, build_tiles): on seeded
// random segment matrices of several shapes, the layout must
//   * hold every nonzero exactly once, at its block-local gathered index,
//     with padding entries marked -1;
//   * give each wavefront a contiguous run of whole slices whose word
//     (rows, lanes per member) matches the entries stored;
//   * map every segment's (block, segment) pairs to partial slots in block
//     order, one slot per pair;
// and a CPU walk of the layout in the kernels' order (lane sums, lane-group
// sums of long pairs, the slot reduction) must reproduce each segment's dot
// product.  The layout is cut on host threads; the same matrix built twice
// must give the same bytes.  Exit status 0 when every case passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../mangalore-refinery-and-petrochemicals-limited_amd/csrc/tiles.hpp"

using namespace bo;

namespace {

struct Case {
    const char* name;
    int nseg, ng, wmax, tiles;
    int long_every;  // every long_every-th segment gets 300..2300 nonzeros
    int max_short;
};

int fails = 0;

void check(bool ok, const char* what, const char* name) {
    if (!ok) {
        std::printf("FAIL %s: %s\n", name, what);
        ++fails;
    }
}

bool same_layout(const HostTiles& a, const HostTiles& b) {
    return a.B == b.B && a.W == b.W && a.padded == b.padded && a.pairs == b.pairs &&
           a.tiles.size() == b.tiles.size() &&
           std::memcmp(a.tiles.data(), b.tiles.data(), sizeof(TileDesc) * a.tiles.size()) == 0 &&
           a.slice_word == b.slice_word && a.idx == b.idx && a.src == b.src && a.seg_ptr == b.seg_ptr &&
           a.slot == b.slot && a.max_seg_pairs == b.max_seg_pairs && a.max_wave_slices == b.max_wave_slices;
}

void run(const Case& cs, unsigned seed) {
    std::mt19937 rng(seed);
    std::vector<int> ptr(static_cast<size_t>(cs.nseg) + 1, 0), gidx;
    for (int s = 0; s < cs.nseg; ++s) {
        const int len = (cs.long_every > 0 && s % cs.long_every == 0) ? 300 + static_cast<int>(rng() % 2000)
                                                                      : 1 + static_cast<int>(rng() % cs.max_short);
        std::vector<int> cols;
        for (int k = 0; k < len; ++k) cols.push_back(static_cast<int>(rng() % cs.ng));
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        gidx.insert(gidx.end(), cols.begin(), cols.end());
        ptr[s + 1] = static_cast<int>(gidx.size());
    }
    const long long nnz = ptr[cs.nseg];
    std::vector<double> val(static_cast<size_t>(nnz)), x(static_cast<size_t>(cs.ng));
    for (auto& v : val) v = static_cast<double>(static_cast<int>(rng() % 2001) - 1000) / 64.0;
    for (auto& v : x) v = static_cast<double>(static_cast<int>(rng() % 2001) - 1000) / 128.0;

    HostTiles h, h2;
    check(build_tiles(ptr.data(), gidx.data(), cs.nseg, cs.ng, cs.wmax, cs.tiles, h), "layout fits", cs.name);
    check(build_tiles(ptr.data(), gidx.data(), cs.nseg, cs.ng, cs.wmax, cs.tiles, h2), "layout fits", cs.name);
    check(same_layout(h, h2), "two builds give the same layout", cs.name);
    check(h.W % 64 == 0 && h.W <= cs.wmax && static_cast<long long>(h.B) * h.W >= cs.ng, "block size", cs.name);

    // every nonzero exactly once, at its block-local index
    std::vector<int> seen(static_cast<size_t>(nnz), 0);
    check(h.idx.size() == h.padded && h.src.size() == h.padded, "entry arrays", cs.name);
    // walk: per tile, per wavefront, the slices of its run in order
    std::vector<double> part(h.slice_word.size() * kWave, 0.0);
    size_t rows_total = 0;
    for (const TileDesc& td : h.tiles) {
        for (int w = 0; w < kTileWaves; ++w) {
            const WaveDesc& wd = td.wave[w];
            size_t row = 0;
            for (int q = 0; q < wd.nslice; ++q) {
                const unsigned word = h.slice_word[static_cast<size_t>(wd.slice0 + q)];
                const int rows = static_cast<int>(word & kRowsMask), lg = static_cast<int>(word >> kLgShift);
                double acc[kWave] = {};
                for (int r = 0; r < rows; ++r)
                    for (int l = 0; l < kWave; ++l) {
                        const size_t e = wd.off + (row + r) * kWave + l;
                        const int k = h.src[e];
                        if (k < 0) continue;
                        seen[static_cast<size_t>(k)]++;
                        if (gidx[k] != td.block * h.W + h.idx[e]) check(false, "block-local index", cs.name);
                        if (gidx[k] / h.W != td.block) check(false, "entry in its tile's block", cs.name);
                        acc[l] += val[k] * x[static_cast<size_t>(gidx[k])];
                    }
                row += static_cast<size_t>(rows);
                const int L = 1 << lg;
                for (int g = 0; g < kWave; g += L) {  // a member's sum ends in its group's last lane
                    double s = 0.0;
                    for (int l = g; l < g + L; ++l) s += acc[l];
                    acc[g + L - 1] = s;
                }
                for (int l = 0; l < kWave; ++l) part[static_cast<size_t>(wd.slice0 + q) * kWave + l] = acc[l];
            }
            if (row != static_cast<size_t>(wd.nrows)) check(false, "run rows match its slices", cs.name);
            rows_total += row;
        }
    }
    check(rows_total * kWave == h.padded, "padded entries are whole rows", cs.name);
    bool once = true;
    for (int v : seen) once = once && v == 1;
    check(once, "every nonzero stored exactly once", cs.name);
    // slots: one per pair, segments' pairs in block order, sums reproduce A x
    check(h.seg_ptr.size() == static_cast<size_t>(cs.nseg) + 1 && h.seg_ptr[cs.nseg] == h.pairs &&
              h.slot.size() == static_cast<size_t>(h.pairs),
          "segment pointers", cs.name);
    std::vector<int> slot_used(part.size(), 0);
    double worst = 0.0;
    int most = 0;
    for (int s = 0; s < cs.nseg; ++s) {
        double ref = 0.0;
        std::vector<int> blocks;
        for (int k = ptr[s]; k < ptr[s + 1]; ++k) {
            ref += val[k] * x[static_cast<size_t>(gidx[k])];
            const int b = gidx[k] / h.W;
            if (blocks.empty() || blocks.back() != b) blocks.push_back(b);
        }
        const int q0 = h.seg_ptr[s], q1 = h.seg_ptr[s + 1];
        most = std::max(most, q1 - q0);
        if (q1 - q0 != static_cast<int>(blocks.size())) check(false, "one pair per (block, segment)", cs.name);
        double sum = 0.0;
        for (int q = q0; q < q1; ++q) {
            const unsigned sl = h.slot[static_cast<size_t>(q)];
            if (sl >= part.size()) {
                check(false, "slot in range", cs.name);
                continue;
            }
            slot_used[sl]++;
            sum += part[sl];
        }
        worst = std::max(worst, std::fabs(sum - ref) / (1.0 + std::fabs(ref)));
    }
    check(most == h.max_seg_pairs, "most pairs of a segment", cs.name);
    bool slots_once = true;
    for (int v : slot_used) slots_once = slots_once && v <= 1;
    check(slots_once, "no slot shared by two pairs", cs.name);
    check(worst < 1e-12, "walk reproduces every segment's dot product", cs.name);
    std::printf("%-28s B %3d W %6d tiles %4zu slices %6zu pairs %8lld padded %9zu  max rel err %.1e\n", cs.name, h.B,
                h.W, h.tiles.size(), h.slice_word.size(), h.pairs, h.padded, worst);
}

}  // namespace

int main() {
    const Case cases[] = {
        {"short rows, many blocks", 20000, 50000, 4096, 64, 0, 30},
        {"long rows (8-64 lanes)", 3000, 200000, 40960, 512, 7, 40},
        {"mixed, few segments", 500, 7000, 2048, 100, 13, 10},
        {"single block", 8000, 3000, 40960, 256, 97, 25},
        {"one nonzero per segment", 60000, 90000, 8192, 128, 0, 1},
    };
    unsigned seed = 11;
    for (const Case& c : cases) run(c, seed++);
    if (fails) {
        std::printf("%d check(s) failed\n", fails);
        return 1;
    }
    std::printf("all tiled-layout checks passed\n");
    return 0;
}
