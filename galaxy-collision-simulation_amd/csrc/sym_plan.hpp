// sym_plan.hpp — the work list of the symmetric mixed direct kernel
// (direct_sym.hip), plain host C++ so it can be checked without a device.
//
// The work of independent runs of 256-particle tiles is the upper triangle
// of tile pairs (I, J), J >= I, within each run.  Each row I of a run is cut
// into the diagonal item (I, I) and items of `len` consecutive J > I tiles
// (the last item of a row may be shorter); by default len = max(1, T / 16384)
// for a run with T tile pairs, about 16,384 items per run of 181 or more
// tiles.  An item's row partial
// goes to slot islot, numbered in (run, row, J) order, so row T's slots are
// [row_start[T], row_start[T + 1]) in J order.  The launch order is longest
// item first (ties in slot order, diagonal items after single tiles), one
// block per item, so the hardware dispatcher balances the tail with the
// shortest items.  The force kernel writes one FP32 row partial per item
// (the sums of the item's targets over all of its source tiles) at islot,
// and one column partial per tile pair (I, J), I < J, of a run at
// pbase + J (J - 1) / 2 + I (tiles counted from the run's first); the
// adjoint keeps its (hi, lo) row partials at the item slots too.
#pragma once

#include <algorithm>
#include <cstddef>
#include <string>
#include <vector>

namespace gc {

constexpr int kSymItemsPerRun = 16384;

// Row tile I against source tiles [jb, je); row-partial slot islot; pair
// partials of I's run start at pbase, the run's first tile is t0.
struct SymItem {
    int I, jb, je, islot;
    long long pbase;
    int t0, pad;
};

struct SymHostPlan {
    std::vector<SymItem> items;         // launch order
    std::vector<int> row_start;         // ntiles + 1 slot offsets
    std::vector<int> tile_t0;           // ntiles
    std::vector<int> tile_nt;           // ntiles: tile count of each tile's run
    std::vector<long long> tile_pbase;  // ntiles
    long long pairs = 0;                // tile pairs I < J over all runs
    int ntiles = 0;
};

// J tiles per item for a run of nt tiles: `len` if positive, else the default
inline int sym_item_tiles(int nt, int len) {
    if (len > 0) return len;
    const long long tri = static_cast<long long>(nt) * (nt + 1) / 2;
    return static_cast<int>(std::max<long long>(1, tri / kSymItemsPerRun));
}

// false (and *err) for no tiles or a run of more than max_tiles tiles
inline bool build_sym_plan(const int* run_t0, const int* run_nt, int nruns, int len,
                           int max_tiles, SymHostPlan& out, std::string* err) {
    out = SymHostPlan{};
    int ntiles = 0;
    for (int r = 0; r < nruns; ++r) {
        if (run_nt[r] < 0 || run_nt[r] > max_tiles) {
            *err = "symmetric form: run tile count out of range";
            return false;
        }
        ntiles = std::max(ntiles, run_t0[r] + run_nt[r]);
    }
    if (ntiles < 1) {
        *err = "symmetric form: no tiles";
        return false;
    }
    out.row_start.assign(ntiles + 1, 0);
    out.tile_t0.assign(ntiles, 0);
    out.tile_nt.assign(ntiles, 0);
    out.tile_pbase.assign(ntiles, 0);
    long long pbase = 0;
    int next_tile = 0;
    for (int r = 0; r < nruns; ++r) {
        const int t0 = run_t0[r], nt = run_nt[r];
        const int step = sym_item_tiles(nt, len);
        for (int T = next_tile; T < t0; ++T) out.row_start[T] = static_cast<int>(out.items.size());
        for (int I = t0; I < t0 + nt; ++I) {
            out.row_start[I] = static_cast<int>(out.items.size());
            out.tile_t0[I] = t0;
            out.tile_nt[I] = nt;
            out.tile_pbase[I] = pbase;
            // the diagonal tile is an item of its own; the row's other
            // tiles are cut into items of `step` tiles
            for (int J = I, first = 1; J < t0 + nt; J = first ? J + 1 : J + step, first = 0) {
                SymItem it{};
                it.I = I;
                it.jb = J;
                it.je = first ? J + 1 : std::min(J + step, t0 + nt);
                it.islot = static_cast<int>(out.items.size());
                it.pbase = pbase;
                it.t0 = t0;
                out.items.push_back(it);
            }
        }
        pbase += static_cast<long long>(nt) * (nt - 1) / 2;
        next_tile = std::max(next_tile, t0 + nt);
    }
    for (int T = next_tile; T <= ntiles; ++T) out.row_start[T] = static_cast<int>(out.items.size());
    // longest first; a diagonal item (half a tile pair in the symmetric
    // kernel) after the single-tile items of its length
    auto cost = [](const SymItem& a) { return 2 * (a.je - a.jb) - (a.jb == a.I ? 1 : 0); };
    std::stable_sort(out.items.begin(), out.items.end(),
                     [&](const SymItem& a, const SymItem& b) { return cost(a) > cost(b); });
    out.pairs = pbase;
    out.ntiles = ntiles;
    return true;
}

}  // namespace gc
