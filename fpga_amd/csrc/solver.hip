// Handle, device memory, drivers and the extern "C" boundary of the MI355X
// ILU0-BiCGSTAB solver.  See include/bda_bicgstab.h for the interface.
//
// Lifecycle mirrors the reference host library (fpga_functions_bicgstab.cpp):
// buffers are allocated once and grown only when a larger system arrives
// (:84-110, :502-505); each solve runs the whole iteration loop on the device
// (:1305-1325) and returns a decoded status record (:890-934).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bda_bicgstab.h"
#include "bda_host.hpp"
#include "kernels.hpp"

namespace bda {
namespace {

#define HIP_CHECK(expr)                                                                           \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw std::runtime_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));   \
    } while (0)

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void ensure(size_t cnt) {
        if (cnt <= n && p) return;
        release();
        const size_t bytes = std::max<size_t>(cnt, 1) * sizeof(T);
        HIP_CHECK(hipMalloc(&p, bytes));
        HIP_CHECK(hipMemset(p, 0, bytes));
        n = std::max<size_t>(cnt, 1);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct DevMat {
    int n = 0;
    int64_t nnzb = 0;
    int ntiles = 0;
    DBuf<int32_t> rowptr, colidx, src;
    // tile descriptors {first row, first block, end row, end block}: grouped
    // by colour (per-colour launches address a contiguous range), and all
    // tiles in one list for launches over the whole matrix; both in the
    // launch order of upload_tiles
    DBuf<int4> tile, tile_sp;
    DBuf<int4> tile_wide, tile_wide_sp;  // the wide tiling
    DBuf<int4> tile_tall;                // the tall tiling (sweeps: per-colour list only)
    // reduction slot of every entry of the lists above: the tile's place in
    // the colour-ordered list (within its colour for the per-colour lists)
    DBuf<int32_t> slot, slot_sp, slot_wide, slot_wide_sp, slot_tall;
    std::vector<int32_t> colour_tile, colour_tile_wide, colour_tile_tall;
    int ntiles_wide = 0;
    // rows of colour C-2 that carry blocks (for the merged tail launch of a
    // second sweep), when there are at most kTailRows of them; else -1
    DBuf<int32_t> tail_dep;
    std::vector<int32_t> tail_dep_rows;  // host copy
    int tail_ndep = -1;
    int tail_row0 = 0, tail_nrows = 0;
    // the extended merged tail (k_tail3), when also at most kTailRows rows of
    // colour C-3 have blocks into colour C-1 or into a row of tail_dep, and
    // every row it packs has at most kTail3Blocks blocks; else tail3_n = -1.
    // tail3_row / tail3_blk / tail3_src: the packed rows (k_tail3's
    // Tail3Args); tail3_n: the special rows of colour C-3; tail3_tile: tiles
    // of the other colour C-3 rows, tall when that saves a quarter of the
    // workgroups.
    DBuf<int32_t> tail3_row, tail3_blk, tail3_src;
    DBuf<int4> tail3_tile;
    int tail3_n = -1, tail3_ntiles = 0;
    bool tail3_tall = false;
    // the first sweep's colours C-3 and C-2 in one launch (k_lower2), for a
    // first-sweep pattern: low2_n tiles of colour C-2 rows (at most 64 rows,
    // at most 64 colour C-3 rows read, each side's blocks within a tile's
    // capacity), -1 when not built.  low2_tile: those tiles; low2_vtile:
    // their dependency tiles, rows t*64 + k of a packed pattern (low2_vrowptr,
    // low2_vcol, low2_vmap: the matrix row, low2_vsrc: the pattern's block
    // index of each packed block, for the value copies); low2_col: colidx
    // with the colour C-3 columns of colour C-2 rows replaced by -1 - their
    // place in the tile's dependency tile.  Both lists in launch order.
    DBuf<int4> low2_tile, low2_vtile;
    DBuf<int32_t> low2_vrowptr, low2_vcol, low2_vmap, low2_vsrc, low2_col;
    int low2_n = -1;
    int64_t low2_nvblk = 0;

    // slices per layer in the launch order, one per XCD (measured on a
    // 128x128x64 grid: 4 slices no better, 16 and 32 slower, FP64 iteration
    // 530 and 544 us against 513 us)
    static constexpr int64_t kLayerSlices = 8;
    // Typical coupling distance in the original numbering: the median over
    // rows of the row's widest coupling (for a structured grid in natural
    // order, the distance to the neighbour in the next layer).
    static int64_t coupling_span(const MatPattern &p, const std::vector<int32_t> &o) {
        std::vector<int32_t> w;
        w.reserve(p.n);
        for (int k = 0; k < p.n; ++k) {
            int32_t m = 0;
            for (int32_t q = p.rowptr[k]; q < p.rowptr[k + 1]; ++q) m = std::max(m, std::abs(o[k] - o[p.colidx[q]]));
            if (m > 0) w.push_back(m);
        }
        if (w.empty()) return 0;
        std::nth_element(w.begin(), w.begin() + w.size() / 2, w.end());
        return w[w.size() / 2];
    }

    // Launch-order key of a tile whose first row is r (below).
    static int64_t tile_key(const std::vector<int32_t> &o, int64_t n, int64_t L, int r) {
        const int64_t orow = o[r];
        const int64_t slice = (L >= kLayerSlices && 2 * L <= n) ? ((orow % L) * kLayerSlices) / L : 0;
        return slice * (n + 1) + orow;
    }

    // Launch order of the tiles.  Every tile is independent and its
    // reduction partials go to a fixed slot, so the order has no effect on
    // results.  Whole-matrix launches take all tiles, per-colour launches
    // their colour's tiles, sorted by the original index of the first row
    // within eighths of a layer (the typical coupling distance).  A launch's
    // XCDs each work through one contiguous eighth of the list, so each XCD
    // sweeps one slice of every layer, from the first layer to the last, and
    // the tiles running at one time on all XCDs cover the same few layers of
    // every colour.  Measured on a 128x128x64 grid against plain original
    // order (each XCD one slab of whole layers): FP64 SpMV 99.7 -> 94.9 us,
    // FP32 SpMV 56.4 -> 54.7 us, FP64 iteration 537 -> 514 us, with 1-2%
    // more fetched bytes (the gain is not from x reuse).
    static void upload_tiles(const MatPattern &p, const std::vector<int32_t> &tile_row,
                             const std::vector<int32_t> &colour_tile, const std::vector<int32_t> *orig_row,
                             int64_t L, DBuf<int4> &out, DBuf<int32_t> &out_slot, DBuf<int4> *out_sp,
                             DBuf<int32_t> *out_sp_slot, hipStream_t s) {
        const int nt = (int)tile_row.size() - 1;
        std::vector<int4> d(std::max(nt, 1), make_int4(0, 0, 0, 0));
        for (int t = 0; t < nt; ++t)
            d[t] = make_int4(tile_row[t], p.rowptr[tile_row[t]], tile_row[t + 1], p.rowptr[tile_row[t + 1]]);
        std::vector<int4> ds = d;
        std::vector<int32_t> sl(d.size(), 0), sls(d.size(), 0);
        for (int t = 0; t < nt; ++t) sls[t] = t;
        for (size_t c = 0; c + 1 < colour_tile.size(); ++c)
            for (int t = colour_tile[c]; t < colour_tile[c + 1]; ++t) sl[t] = t - colour_tile[c];
        if (orig_row && nt > 0) {
            const std::vector<int32_t> &o = *orig_row;
            const int64_t n = p.n;
            // key: slice of the layer (an eighth of L), then original row
            std::vector<int64_t> key(nt);
            for (int t = 0; t < nt; ++t) key[t] = tile_key(o, n, L, tile_row[t]);
            auto by_key = [&](int a, int b) { return key[a] < key[b]; };
            std::vector<int32_t> ord(nt);
            // whole-matrix list
            for (int t = 0; t < nt; ++t) ord[t] = t;
            std::stable_sort(ord.begin(), ord.end(), by_key);
            for (int t = 0; t < nt; ++t) {
                ds[t] = d[ord[t]];
                sls[t] = ord[t];
            }
            // per-colour launches: the same order within each colour's range
            for (int t = 0; t < nt; ++t) ord[t] = t;
            for (size_t c = 0; c + 1 < colour_tile.size(); ++c)
                std::stable_sort(ord.begin() + colour_tile[c], ord.begin() + colour_tile[c + 1], by_key);
            std::vector<int4> dc(d.size());
            std::vector<int32_t> slc(d.size(), 0);
            for (int t = 0; t < nt; ++t) {
                dc[t] = d[ord[t]];
                slc[t] = sl[ord[t]];
            }
            d.swap(dc);
            sl.swap(slc);
        }
        out.ensure(d.size());
        out_slot.ensure(sl.size());
        HIP_CHECK(hipMemcpyAsync(out.p, d.data(), d.size() * sizeof(int4), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(out_slot.p, sl.data(), sl.size() * 4, hipMemcpyHostToDevice, s));
        if (out_sp) {
            out_sp->ensure(ds.size());
            out_sp_slot->ensure(sls.size());
            HIP_CHECK(hipMemcpyAsync(out_sp->p, ds.data(), ds.size() * sizeof(int4), hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(out_sp_slot->p, sls.data(), sls.size() * 4, hipMemcpyHostToDevice, s));
        }
        HIP_CHECK(hipStreamSynchronize(s));
    }

    // orig_row: original row of every (permuted) row, and span: the
    // matrix's coupling_span, for the launch order of the tiles
    void upload(const MatPattern &p, hipStream_t s, const std::vector<int32_t> *orig_row = nullptr,
                int64_t span = 0) {
        n = p.n;
        nnzb = p.nnzb();
        ntiles = p.ntiles();
        colour_tile = p.colour_tile;
        colour_tile_wide = p.colour_tile_wide;
        colour_tile_tall = p.colour_tile_tall;
        ntiles_wide = (int)p.tile_row_wide.size() - 1;
        rowptr.ensure(p.rowptr.size());
        colidx.ensure(p.colidx.size() + kPadBlocks);
        src.ensure(std::max<size_t>(p.src.size(), 1));
        HIP_CHECK(hipMemcpyAsync(rowptr.p, p.rowptr.data(), p.rowptr.size() * 4, hipMemcpyHostToDevice, s));
        if (!p.colidx.empty())
            HIP_CHECK(hipMemcpyAsync(colidx.p, p.colidx.data(), p.colidx.size() * 4, hipMemcpyHostToDevice, s));
        if (!p.src.empty())
            HIP_CHECK(hipMemcpyAsync(src.p, p.src.data(), p.src.size() * 4, hipMemcpyHostToDevice, s));
        upload_tiles(p, p.tile_row, p.colour_tile, orig_row, span, tile, slot, &tile_sp, &slot_sp, s);
        upload_tiles(p, p.tile_row_wide, p.colour_tile_wide, orig_row, span, tile_wide, slot_wide, &tile_wide_sp,
                     &slot_wide_sp, s);
        upload_tiles(p, p.tile_row_tall, p.colour_tile_tall, orig_row, span, tile_tall, slot_tall, nullptr, nullptr,
                     s);
    }
    // Tiles over the row ranges runs (as build_tiles does for a colour), in
    // launch order; false when a row has more blocks than a tile holds.
    static bool tiles_for_runs(const MatPattern &p, const std::vector<std::pair<int, int>> &runs, int cap,
                               int max_rows, const std::vector<int32_t> &o, int64_t L, std::vector<int4> &out) {
        std::vector<std::pair<int64_t, int4>> t;
        for (const auto &run : runs) {
            int r = run.first;
            while (r < run.second) {
                const int first = r;
                int rows = 0, blocks = 0;
                while (r < run.second && rows < max_rows) {
                    const int nb = p.rowptr[r + 1] - p.rowptr[r];
                    if (nb > cap) return false;
                    if (blocks + nb > cap) break;
                    blocks += nb;
                    ++rows;
                    ++r;
                }
                t.push_back({tile_key(o, p.n, L, first), make_int4(first, p.rowptr[first], r, p.rowptr[r])});
            }
        }
        std::stable_sort(t.begin(), t.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        out.resize(t.size());
        for (size_t i = 0; i < t.size(); ++i) out[i] = t[i].second;
        return true;
    }

    // The rows k_tail3 computes in its first workgroup (the colour C-2 rows
    // of tail_dep, then the special rows of colour C-3), packed with their
    // blocks and each block's z source, and tiles of the other colour C-3 rows.
    void set_tail3(const MatPattern &p, const std::vector<int32_t> &colour_ptr, const std::vector<int32_t> &o,
                   int64_t span, int bs, hipStream_t s) {
        tail3_n = -1;
        const int nc = (int)colour_ptr.size() - 1;
        if (tail_ndep < 0 || nc < 3) return;
        const int c3_0 = colour_ptr[nc - 3], c3_1 = colour_ptr[nc - 2];
        const int last0 = colour_ptr[nc - 1];
        const std::vector<int32_t> &dep = tail_dep_rows;
        std::vector<int32_t> dslot(tail_nrows, -1);
        for (int j = 0; j < tail_ndep; ++j) dslot[dep[j] - tail_row0] = j;
        constexpr int MB = kTail3Blocks;
        std::vector<int32_t> row(2 * kTailRows, 0), blk(2 * kTailRows * MB, -1), src(2 * kTailRows * MB, -1);
        // one packed row; false when it has too many blocks or a block that
        // points before colour C-2
        auto pack = [&](int slot, int r) {
            if (p.rowptr[r + 1] - p.rowptr[r] > MB) return false;
            row[slot] = r;
            int m = 0;
            for (int32_t q = p.rowptr[r]; q < p.rowptr[r + 1]; ++q, ++m) {
                const int col = p.colidx[q];
                if (col < tail_row0) return false;
                blk[slot * MB + m] = q;
                if (col >= last0) src[slot * MB + m] = -1 - (col - last0);
                else if (dslot[col - tail_row0] >= 0) src[slot * MB + m] = -1 - (kTailRows + dslot[col - tail_row0]);
                else src[slot * MB + m] = col;
            }
            return true;
        };
        for (int j = 0; j < tail_ndep; ++j)
            if (!pack(j, dep[j])) return;
        int n3 = 0;
        std::vector<std::pair<int, int>> runs;
        int run0 = c3_0;
        for (int r = c3_0; r < c3_1; ++r) {
            bool special = false;
            for (int32_t q = p.rowptr[r]; q < p.rowptr[r + 1]; ++q) {
                const int col = p.colidx[q];
                if (col < tail_row0) return;  // second-sweep blocks only point to later colours
                if (col >= last0 || dslot[col - tail_row0] >= 0) special = true;
            }
            if (!special) continue;
            if (n3 == kTailRows || !pack(kTailRows + n3, r)) return;
            ++n3;
            if (r > run0) runs.push_back({run0, r});
            run0 = r + 1;
        }
        if (c3_1 > run0) runs.push_back({run0, c3_1});
        std::vector<int4> narrow, tall;
        if (!tiles_for_runs(p, runs, tile_cap(bs), kTileRows, o, span, narrow) ||
            !tiles_for_runs(p, runs, tile_cap(bs), kTileRowsTall, o, span, tall))
            return;
        tail3_tall = 4 * tall.size() <= 3 * narrow.size();
        const std::vector<int4> &tl = tail3_tall ? tall : narrow;
        tail3_ntiles = (int)tl.size();
        tail3_tile.ensure(std::max<size_t>(tl.size(), 1));
        if (!tl.empty())
            HIP_CHECK(hipMemcpyAsync(tail3_tile.p, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice, s));
        tail3_row.ensure(row.size());
        tail3_blk.ensure(blk.size());
        tail3_src.ensure(src.size());
        HIP_CHECK(hipMemcpyAsync(tail3_row.p, row.data(), row.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(tail3_blk.p, blk.data(), blk.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(tail3_src.p, src.data(), src.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
        tail3_n = n3;
    }

    static constexpr int kLow2MaxTiles = 2048;

    // Tiles of colour C-2 of a first-sweep pattern for k_lower2, each with the
    // packed copy of the colour C-3 rows it reads.
    void set_low2(const MatPattern &p, const std::vector<int32_t> &colour_ptr, const std::vector<int32_t> &o,
                  int64_t span, int bs, hipStream_t s) {
        low2_n = -1;
        const int nc = (int)colour_ptr.size() - 1;
        if (nc < 4) return;
        // only where colour C-3's own launch is latency-bound (fewer tiles than
        // the GPU holds at once): on the 4x4 256x256x64 grid (6,260 tiles) the
        // recomputed rows cost more than the saved launch (FP64 26.1 + 10.4 us
        // as two launches, 39.8 us merged)
        if (p.colour_tile[nc - 2] - p.colour_tile[nc - 3] > kLow2MaxTiles) return;
        const int c0 = colour_ptr[nc - 3], c1 = colour_ptr[nc - 2], c2 = colour_ptr[nc - 1];
        const int cap = tile_cap(bs);
        struct T {
            int first, end;
            std::vector<int32_t> deps;
        };
        std::vector<T> tiles;
        std::vector<int32_t> mark(c1 - c0, -1);  // slot of a colour C-3 row in the tile being built
        int r = c1;
        while (r < c2) {
            T t{r, r, {}};
            int blocks = 0, dblocks = 0;
            while (r < c2 && r - t.first < kTileRows) {
                const int nb = p.rowptr[r + 1] - p.rowptr[r];
                if (nb > cap) return;
                int nnew = 0, bnew = 0;
                for (int32_t q = p.rowptr[r]; q < p.rowptr[r + 1]; ++q) {
                    const int col = p.colidx[q];
                    if (col >= c1) return;  // a first sweep only reads earlier colours
                    if (col >= c0 && mark[col - c0] < 0) {
                        mark[col - c0] = -2;  // counted for this row
                        ++nnew;
                        bnew += p.rowptr[col + 1] - p.rowptr[col];
                    }
                }
                for (int32_t q = p.rowptr[r]; q < p.rowptr[r + 1]; ++q) {
                    const int col = p.colidx[q];
                    if (col >= c0 && mark[col - c0] == -2) mark[col - c0] = -1;
                }
                const bool fits = blocks + nb <= cap && (int)t.deps.size() + nnew <= kTileRows && dblocks + bnew <= cap;
                if (!fits) {
                    if (r == t.first) return;  // one row alone does not fit
                    break;
                }
                for (int32_t q = p.rowptr[r]; q < p.rowptr[r + 1]; ++q) {
                    const int col = p.colidx[q];
                    if (col >= c0 && mark[col - c0] < 0) {
                        mark[col - c0] = (int)t.deps.size();
                        t.deps.push_back(col);
                    }
                }
                blocks += nb;
                dblocks += bnew;
                ++r;
            }
            t.end = r;
            for (int32_t d : t.deps) mark[d - c0] = -1;
            tiles.push_back(std::move(t));
        }
        // launch order
        std::vector<int> ord(tiles.size());
        for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
            return tile_key(o, p.n, span, tiles[x].first) < tile_key(o, p.n, span, tiles[y].first);
        });
        const int nt = (int)tiles.size();
        std::vector<int4> td(std::max(nt, 1)), vd(std::max(nt, 1));
        std::vector<int32_t> vrp((size_t)nt * kTileRows + 1, 0), vmap((size_t)nt * kTileRows, 0), vcol, vsrc;
        std::vector<int32_t> col(p.colidx.begin(), p.colidx.end());
        for (int k = 0; k < nt; ++k) {
            const T &t = tiles[ord[k]];
            td[k] = make_int4(t.first, p.rowptr[t.first], t.end, p.rowptr[t.end]);
            const int vr0 = k * kTileRows;
            for (int j = 0; j < kTileRows; ++j) {
                if (j < (int)t.deps.size()) {
                    const int d = t.deps[j];
                    vmap[vr0 + j] = d;
                    for (int32_t q = p.rowptr[d]; q < p.rowptr[d + 1]; ++q) {
                        vcol.push_back(p.colidx[q]);
                        vsrc.push_back(q);
                    }
                }
                vrp[vr0 + j + 1] = (int32_t)vcol.size();
            }
            vd[k] = make_int4(vr0, vrp[vr0], vr0 + (int)t.deps.size(), vrp[vr0 + (int)t.deps.size()]);
            for (int32_t q = p.rowptr[t.first]; q < p.rowptr[t.end]; ++q) {
                const int c = p.colidx[q];
                if (c >= c0 && c < c1) {
                    const auto it = std::find(t.deps.begin(), t.deps.end(), c);
                    col[q] = -1 - (int32_t)(it - t.deps.begin());
                }
            }
        }
        low2_nvblk = (int64_t)vcol.size();
        for (int i = 0; i < kPadBlocks; ++i) vcol.push_back(0);
        if (vsrc.empty()) vsrc.push_back(0);
        for (int i = 0; i < kPadBlocks; ++i) col.push_back(0);
        low2_tile.ensure(td.size());
        low2_vtile.ensure(vd.size());
        low2_vrowptr.ensure(vrp.size());
        low2_vmap.ensure(std::max<size_t>(vmap.size(), 1));
        low2_vcol.ensure(vcol.size());
        low2_vsrc.ensure(vsrc.size());
        low2_col.ensure(col.size());
        HIP_CHECK(hipMemcpyAsync(low2_tile.p, td.data(), td.size() * sizeof(int4), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(low2_vtile.p, vd.data(), vd.size() * sizeof(int4), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(low2_vrowptr.p, vrp.data(), vrp.size() * 4, hipMemcpyHostToDevice, s));
        if (!vmap.empty())
            HIP_CHECK(hipMemcpyAsync(low2_vmap.p, vmap.data(), vmap.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(low2_vcol.p, vcol.data(), vcol.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(low2_vsrc.p, vsrc.data(), vsrc.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(low2_col.p, col.data(), col.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
        low2_n = nt;
    }

    void set_tail(const MatPattern &p, const std::vector<int32_t> &colour_ptr, hipStream_t s) {
        tail_ndep = -1;
        tail3_n = -1;
        const int nc = (int)colour_ptr.size() - 1;
        if (nc < 3) return;
        tail_row0 = colour_ptr[nc - 2];
        tail_nrows = colour_ptr[nc - 1] - tail_row0;
        const int last0 = colour_ptr[nc - 1], last1 = colour_ptr[nc];
        if (last1 - last0 > kTileRows) return;
        std::vector<int32_t> dep;
        for (int r = tail_row0; r < tail_row0 + tail_nrows; ++r) {
            if (p.rowptr[r + 1] == p.rowptr[r]) continue;
            for (int32_t q = p.rowptr[r]; q < p.rowptr[r + 1]; ++q)
                if (p.colidx[q] < last0 || p.colidx[q] >= last1) return;  // blocks only into colour C-1
            dep.push_back(r);
        }
        if ((int)dep.size() > kTailRows) return;
        tail_dep.ensure(std::max<size_t>(dep.size(), 1));
        if (!dep.empty())
            HIP_CHECK(hipMemcpyAsync(tail_dep.p, dep.data(), dep.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
        tail_ndep = (int)dep.size();
        tail_dep_rows = dep;
    }
    void release() {
        rowptr.release();
        colidx.release();
        tile.release();
        tile_sp.release();
        tile_wide.release();
        tile_wide_sp.release();
        tile_tall.release();
        for (DBuf<int32_t> *b : {&slot, &slot_sp, &slot_wide, &slot_wide_sp, &slot_tall}) b->release();
        tail_dep.release();
        for (DBuf<int32_t> *b : {&tail3_row, &tail3_blk, &tail3_src}) b->release();
        tail3_tile.release();
        tail3_n = -1;
        for (DBuf<int32_t> *b : {&low2_vrowptr, &low2_vcol, &low2_vmap, &low2_vsrc, &low2_col}) b->release();
        low2_tile.release();
        low2_vtile.release();
        low2_n = -1;
        src.release();
    }
};

template <typename TX>
struct Krylov {
    DBuf<TX> rt, r, p, phat, v, s, shat, t, y, x;
    void ensure(size_t n) {
        for (DBuf<TX> *b : {&rt, &r, &p, &phat, &v, &s, &shat, &t, &y, &x}) b->ensure(n);
    }
    void release() {
        for (DBuf<TX> *b : {&rt, &r, &p, &phat, &v, &s, &shat, &t, &y, &x}) b->release();
    }
};

// The matrices one BiCGSTAB run uses: the operator and the two sweeps of its
// preconditioner.
template <typename TV>
struct OpSet {
    const DevMat *A;
    const TV *Aval;
    const DevMat *S1;  // first sweep, colours ascending
    const TV *S1val;
    bool s1_dinv;
    const DevMat *S2;  // second sweep, colours descending
    const TV *S2val;
    bool s2_dinv;
    const double *dinv;
};

enum ProfKind { PK_SPMV = 0, PK_LOWER = 1, PK_UPPER = 2, PK_VEC = 3, PK_N = 4 };

}  // namespace
}  // namespace bda

// Build switches of the merged sweep launches (k_tail3: the last colour's
// first sweep with the second sweep of colours C-2 and C-3; k_lower2: the
// first sweep of colours C-3 and C-2).  1, the default, merges where the
// pattern allows it; 0 keeps the launches per colour (same results, for
// A/B timing).
#ifndef BDA_TAIL3
#define BDA_TAIL3 1
#endif
#ifndef BDA_LOW2
#define BDA_LOW2 1
#endif
// 1 (default): k_finalize_joint for 2,049 to 18,432 partials; 0: k_finalize
// (the same sums, for A/B timing)
#ifndef BDA_FIN_JOINT
#define BDA_FIN_JOINT 1
#endif

struct bda_handle_s {
    int bs = 0, n = 0;
    int64_t nnzb_max = 0;
    int device = 0;
    int cus = 0;  // compute units of the device
    int wall_khz = 0;  // rate of the device's wall clock (the watchdog's time base)
    hipStream_t stream = nullptr;
    std::string err;

    // host copies
    bda::Bcsr pattern;
    bool have_pattern = false, analysed = false, factored = false, have_values = false;
    bda::Analysis an;
    std::vector<int32_t> blkrow;  // block row of every permuted A block

    // device matrices
    bda::DevMat A, L, U, At, Ut, Lt;
    bda::DBuf<int32_t> d_perm, d_diag, d_blkrow, d_bad;
    bda::DBuf<int32_t> d_nonfinite;  // set by the factorisation when a matrix value is infinite or NaN
    bda::DBuf<double> d_orig, A64, L64, U64, LU, Dinv, DinvT, At64, Ut64, Lt64;
    bda::DBuf<float> A32, L32, U32;
    bda::DBuf<double> low2_v64;  // values of L's packed dependency rows (k_lower2)
    bda::DBuf<float> low2_v32;
    bda::DBuf<float> At32, Ut32, Lt32;  // FP32 transposes (component access only)
    uint64_t factor_gen = 0, mixed_gen = UINT64_MAX, adj_gen = UINT64_MAX, adj32_gen = UINT64_MAX;
    uint64_t values_gen = 0;  // changes with every upload, value update and factorisation
    // the extended merged tail launch (k_tail3) and the merged lower launch
    // (k_lower2) where the colouring allows them
    bool tail3 = BDA_TAIL3 != 0;
    bool low2 = BDA_LOW2 != 0;

    // vectors
    bda::Krylov<double> k64;
    bda::Krylov<float> k32;
    bda::DBuf<double> bp, xo, ro, vin, vout, vw, gl, gx, ggrad;
    bool last_solve_valid = false;
    int last_kind = 0;  // 0 plain / adjoint on k64, 1 mixed (residual in ro)

    // reductions and scalar state (0 = solve, 1 = outer loop or auxiliary, 2 = idle)
    bda::DBuf<double> partials;
    int32_t partial_cap = 0;
    bda::DBuf<bda::SolverState> st;
    bda::SolverState *h_st = nullptr;  // pinned
    int32_t *h_done = nullptr;          // mapped host flags, one per state: the id of the run that stopped
    int32_t *d_done = nullptr;          // their device address
    uint64_t *h_prog = nullptr;         // mapped progress words, one per state (run id << 32 | count)
    uint64_t *d_prog = nullptr;
    bda::DBuf<double> hist;             // norm history of the runs on state slot 0 (optional)
    int32_t hist_cap = 0;
    int32_t run_seq = 0;                // last run id handed out

    // options
    double timeout_s = 0.0;
    int abs_tol = 0;
    int poll_iters = 1;
    int refine_adaptive = 1;  // last refinement steps ask only what the outer target needs
    double refine_eta = 0.9;

    // profiling
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    size_t ev_used = 0;
    std::vector<int> ev_kind;
    std::vector<int> ev_tag;  // half-iteration a launch belongs to (-1: set-up)
    int prof_tag = -1;
    int64_t prof_count[bda::PK_N] = {};
    double prof_ms[bda::PK_N] = {};
    std::vector<float> prof_times[bda::PK_N];  // every counted launch, ms

    size_t N() const { return (size_t)n * bs; }
    size_t Npad() const { return (size_t)n * bs + bda::kVecPad; }
};

// Scratch of bda_spmv_dev_noalloc: the scalar state and the reduction
// partial sums of its SpMV, in the library's device code rather than in
// memory the library allocates.  Every GPU the library runs on holds its own
// copy at its own address until the process ends; every call on a GPU uses
// that GPU's copy, so calls that overlap in time on one GPU race.  On the
// host the name is only a stand-in: host code uses the device address
// hipGetSymbolAddress returns for the current device (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[size_t(1) << 20];

namespace bda {
namespace {

constexpr size_t kNoAllocScratchBytes = sizeof(scratch);
constexpr size_t kNoAllocStateBytes = 512;  // the SolverState, padded
static_assert(sizeof(SolverState) <= kNoAllocStateBytes, "state slot of the scratch array");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current (from bda_create).
void noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
}

}  // namespace
}  // namespace bda

namespace bda {
namespace {

static_assert(TileCap<3, false>::value == 168 && TileCap<4, false>::value == 128 &&
                  TileCap<3, true>::value == 336 && TileCap<4, true>::value == 256,
              "device tile capacity must match tile_cap()");

using H = bda_handle_s;

// A copy on the handle's stream, complete on return.  It is ordered with the
// handle's queued work like every other call; unlike hipMemcpy (the legacy
// default stream) it does not wait for the work of other handles' streams,
// so handles used from different threads do not serialise each other.
void copy_sync(H &h, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, h.stream));
    HIP_CHECK(hipStreamSynchronize(h.stream));
}

int blocks_for(int64_t work, int tpb) { return (int)((work + tpb - 1) / tpb); }

void prof_begin(H &h, int kind) {
    if (!h.profiling) return;
    if (h.ev_used == h.ev_pool.size()) {
        hipEvent_t a, b;
        HIP_CHECK(hipEventCreate(&a));
        HIP_CHECK(hipEventCreate(&b));
        h.ev_pool.emplace_back(a, b);
    }
    h.ev_kind.resize(h.ev_used + 1);
    h.ev_tag.resize(h.ev_used + 1);
    h.ev_kind[h.ev_used] = kind;
    h.ev_tag[h.ev_used] = h.prof_tag;
    HIP_CHECK(hipEventRecord(h.ev_pool[h.ev_used].first, h.stream));
}
void prof_end(H &h) {
    if (!h.profiling) return;
    HIP_CHECK(hipEventRecord(h.ev_pool[h.ev_used].second, h.stream));
    h.ev_used++;
}
// Launches queued after the device-side stop became no-ops; only those of
// half-iterations up to max_tag did work and are counted.
void prof_collect(H &h, int max_tag = INT_MAX) {
    if (!h.profiling || h.ev_used == 0) return;
    HIP_CHECK(hipStreamSynchronize(h.stream));
    for (size_t i = 0; i < h.ev_used; ++i) {
        if (h.ev_tag[i] > max_tag) continue;
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, h.ev_pool[i].first, h.ev_pool[i].second));
        h.prof_count[h.ev_kind[i]]++;
        h.prof_ms[h.ev_kind[i]] += ms;
        h.prof_times[h.ev_kind[i]].push_back(ms);
    }
    h.ev_used = 0;
}

void check_launch() { HIP_CHECK(hipGetLastError()); }

// State slots: 0 = the running BiCGSTAB, 1 = the outer refinement loop (or a
// standalone reduction), kIdleState = a state that is never stopped, for
// launches outside a solve (the streaming kernel always reads a state).
constexpr int kIdleState = 2;

RedArgs red(H &h, int state_idx, int fin) {
    RedArgs ra;
    ra.partials = h.partials.p;
    ra.cap = h.partial_cap;
    ra.st = h.st.p + state_idx;
    ra.fin = fin;
    return ra;
}

// grid-wide sum of the partials of the kernel just launched + scalar epilogue
void launch_finalize(H &h, int nred, const RedArgs &ra, int nparts) {
    const bool wide = nparts >= kFinWideFrom;
    if (nparts >= kFinTwoStageFrom) {
        if (nparts > ra.cap - kFinGroups) throw std::runtime_error("partials buffer too small for the two-stage sum");
        if (nred == 1) hipLaunchKernelGGL((k_finalize_stage1<1>), dim3(kFinGroups), dim3(kFinGroupThreads), 0, h.stream, ra, nparts);
        else hipLaunchKernelGGL((k_finalize_stage1<2>), dim3(kFinGroups), dim3(kFinGroupThreads), 0, h.stream, ra, nparts);
        check_launch();
        RedArgs r2 = ra;
        r2.partials = ra.partials + (ra.cap - kFinGroups);
        if (nred == 1) hipLaunchKernelGGL((k_finalize_one_pass<1, 2>), dim3(1), dim3(kFinThreads), 0, h.stream, r2, kFinGroups);
        else hipLaunchKernelGGL((k_finalize_one_pass<2, 2>), dim3(1), dim3(kFinThreads), 0, h.stream, r2, kFinGroups);
    } else if (nparts <= 2 * kFinThreads) {
        if (nred == 1) hipLaunchKernelGGL((k_finalize_one_pass<1, 2>), dim3(1), dim3(kFinThreads), 0, h.stream, ra, nparts);
        else hipLaunchKernelGGL((k_finalize_one_pass<2, 2>), dim3(1), dim3(kFinThreads), 0, h.stream, ra, nparts);
    } else if (BDA_FIN_JOINT && nparts <= kFinJointTo) {
        if (nred == 1) hipLaunchKernelGGL((k_finalize_joint<1, kFinJointU2>), dim3(1), dim3(kFinThreads), 0, h.stream, ra, nparts);
        else hipLaunchKernelGGL((k_finalize_joint<2, kFinJointU2>), dim3(1), dim3(kFinThreads), 0, h.stream, ra, nparts);
    } else if (nred == 1) {
        if (wide) hipLaunchKernelGGL((k_finalize<1, kFinUnrollWide>), dim3(1), dim3(kFinThreads), 0, h.stream, ra, nparts);
        else hipLaunchKernelGGL((k_finalize<1, kFinUnroll>), dim3(1), dim3(kFinThreads), 0, h.stream, ra, nparts);
    } else {
        if (wide) hipLaunchKernelGGL((k_finalize<2, kFinUnrollWide>), dim3(1), dim3(kFinThreads), 0, h.stream, ra, nparts);
        else hipLaunchKernelGGL((k_finalize<2, kFinUnroll>), dim3(1), dim3(kFinThreads), 0, h.stream, ra, nparts);
    }
    check_launch();
}

// Launches the streaming kernel, one workgroup per tile; returns the number
// of reduction partials written.
template <int BS, int MODE, int RED, typename TV, typename TX, int ROWS = kTileRows, bool ZX = false, bool PLAIN = false>
int stream_dispatch(H &h, const StreamArgs<TV, TX> &a) {
    hipLaunchKernelGGL((k_stream<BS, MODE, RED, TV, TX, ROWS, ZX, PLAIN>), dim3(a.ntiles), dim3(64 * BS), 0, h.stream, a);
    check_launch();
    return a.ntiles;
}

// --- streaming kernel launch over the tiles of one colour (or all) ---------
// prof_kind >= 0: with profiling on, the streaming kernel alone (not its
// finalize) is bracketed by an event pair of that kind.
// ZX: a step-0 residual that takes the short path when k_mark_nonzero left
// the state's full_residual clear (zero start vector, finite matrix).
template <int BS, int MODE, int RED, typename TV, typename TX, bool ZX = false>
void launch_stream(H &h, const DevMat &m, const TV *val, int colour, const TX *x, const TX *rhs, TX *out,
                   const double *dinv, const TX *w0, RedArgs ra, bool use_state, TX *out2 = nullptr,
                   TX *out3 = nullptr, const TX *x_lo = nullptr, int split = 0, int prof_kind = -1,
                   const int32_t *rowptr2 = nullptr, TX *zout = nullptr) {
    constexpr int nred = (RED & SR_W0 ? 1 : 0) + (RED & SR_SELF ? 1 : 0);
    StreamArgs<TV, TX> a;
    constexpr bool wide = wide_tiles<MODE>();
    const std::vector<int32_t> &ct = wide ? m.colour_tile_wide : m.colour_tile;
    a.tile = colour < 0 ? (wide ? m.tile_wide_sp.p : m.tile_sp.p) : (wide ? m.tile_wide.p : m.tile.p);
    a.slot = colour < 0 ? (wide ? m.slot_wide_sp.p : m.slot_sp.p) : (wide ? m.slot_wide.p : m.slot.p);
    if (colour < 0) {
        a.tile0 = 0;
        a.ntiles = wide ? m.ntiles_wide : m.ntiles;
    } else {
        a.tile0 = ct[colour];
        a.ntiles = ct[colour + 1] - ct[colour];
    }
    if (a.ntiles <= 0) return;
    // a sweep colour with few blocks per row takes the tall tiling when that
    // saves a quarter of the workgroups or more
    bool tall = false;
    if constexpr (MODE != SM_SPMV && RED == SR_NONE) {
        if (colour >= 0 && !wide) {
            const int nt = m.colour_tile_tall[colour + 1] - m.colour_tile_tall[colour];
            if (4 * nt <= 3 * a.ntiles) {
                tall = true;
                a.tile = m.tile_tall.p;
                a.slot = m.slot_tall.p;
                a.tile0 = m.colour_tile_tall[colour];
                a.ntiles = nt;
            }
        }
    }
    a.rowptr = m.rowptr.p;
    a.colidx = m.colidx.p;
    a.val = val;
    a.x = x;
    a.x_lo = x_lo ? x_lo : x;
    a.split = x_lo ? split : 0;
    a.rhs = rhs;
    a.out = out;
    a.out2 = out2;
    a.out3 = out3;
    a.dinv = dinv;
    a.w0 = w0;
    a.ra = ra;
    a.rowptr2 = rowptr2;
    a.zout = zout;
    if (!use_state) a.ra.st = h.st.p + kIdleState;
    if (MODE == SM_SUB_Z && (!rowptr2 || !zout || !dinv)) throw std::runtime_error("SUB_Z launch without z operands");
    if ((RED & SR_W0) && !w0) throw std::runtime_error("missing dot operand");
    if (nred > 0 && !use_state) throw std::runtime_error("a reduction needs the solver state");
    if (nred > 0 && a.ntiles > a.ra.cap) throw std::runtime_error("partials buffer too small");
    if (prof_kind >= 0) prof_begin(h, prof_kind);
    int nparts = 0;
    if constexpr (MODE != SM_SPMV && RED == SR_NONE) {
        bool done = false;
        if constexpr (plain_small<TV>() && !ZX) {
            if (a.ntiles < kPlainBelow) {
                if (tall) nparts = stream_dispatch<BS, MODE, RED, TV, TX, kTileRowsTall, false, true>(h, a);
                else nparts = stream_dispatch<BS, MODE, RED, TV, TX, kTileRows, false, true>(h, a);
                done = true;
            }
        }
        if (!done) {
            if (tall) nparts = stream_dispatch<BS, MODE, RED, TV, TX, kTileRowsTall>(h, a);
            else nparts = stream_dispatch<BS, MODE, RED, TV, TX, kTileRows, ZX>(h, a);
        }
    } else {
        nparts = stream_dispatch<BS, MODE, RED, TV, TX, kTileRows, ZX>(h, a);
    }
    if (prof_kind >= 0) prof_end(h);
    if (nred > 0) launch_finalize(h, nred, a.ra, nparts);
}

template <int BS, typename TV, typename TX>
void launch_stream_timed(H &h, const TV *val, const TX *x, TX *y, const TX *w0, RedArgs ra) {
    StreamArgs<TV, TX> a;
    std::memset(&a, 0, sizeof(a));
    constexpr bool wide = wide_tiles<SM_SPMV>();
    a.tile = wide ? h.A.tile_wide_sp.p : h.A.tile_sp.p;
    a.slot = wide ? h.A.slot_wide_sp.p : h.A.slot_sp.p;
    a.tile0 = 0;
    a.ntiles = wide ? h.A.ntiles_wide : h.A.ntiles;
    a.rowptr = h.A.rowptr.p;
    a.colidx = h.A.colidx.p;
    a.val = val;
    a.x = x;
    a.x_lo = x;
    a.split = 0;
    a.out = y;
    a.w0 = w0;
    a.ra = ra;
    if (a.ntiles > h.partial_cap) throw std::runtime_error("partials buffer too small");
    stream_dispatch<BS, SM_SPMV, SR_W0, TV, TX>(h, a);
}

int vec_blocks(size_t N) { return (int)((N + kVecElemsPerBlock - 1) / kVecElemsPerBlock); }

template <int BS, typename TV, typename TX>
void sweep_launch(H &h, const DevMat &m, const TV *val, int c, bool dinv_mode, const TX *x, const TX *x_lo, int split,
                  const TX *rhs, TX *out, const double *dinv, bool use_state) {
    RedArgs ra = red(h, 0, FIN_NONE);
    if (dinv_mode)
        launch_stream<BS, SM_SUB_DINV, SR_NONE, TV, TX>(h, m, val, c, x, rhs, out, dinv, nullptr, ra, use_state,
                                                        nullptr, nullptr, x_lo, split);
    else
        launch_stream<BS, SM_SUB, SR_NONE, TV, TX>(h, m, val, c, x, rhs, out, nullptr, nullptr, ra, use_state, nullptr,
                                                   nullptr, x_lo, split);
}

// The first sweep's last colour C-1 and the second sweep's colour C-2 in one
// launch (k_tail): workgroup 0 runs the single tile of colour C-1 and then
// the rows of colour C-2 that carry second-sweep blocks; the other
// workgroups, one row per thread, the rows of colour C-2 without blocks.
template <int BS, int MODEA, int MODEB, typename TV, typename TX>
void tail_dispatch(H &h, const StreamArgs<TV, TX> &a, const TailArgs<TV, TX> &b) {
    constexpr int NT = 64 * BS;
    const int grid = 1 + (b.nrows + NT - 1) / NT;
    hipLaunchKernelGGL((k_tail<BS, MODEA, MODEB, TV, TX>), dim3(grid), dim3(NT), 0, h.stream, a, b);
    check_launch();
}

template <int BS, typename TV, typename TX>
void tail_launch(H &h, const OpSet<TV> &op, const TX *y, const TX *lo, int split, const TX *p, TX *z,
                 bool use_state) {
    const int nc = h.an.ncolours;
    const DevMat &m1 = *op.S1, &m2 = *op.S2;
    StreamArgs<TV, TX> a;
    std::memset(&a, 0, sizeof(a));
    a.tile = m1.tile.p;
    a.slot = m1.slot.p;
    a.tile0 = m1.colour_tile[nc - 1];
    a.ntiles = 1;
    a.rowptr = m1.rowptr.p;
    a.colidx = m1.colidx.p;
    a.val = op.S1val;
    a.x = y;
    a.x_lo = lo ? lo : y;
    a.split = lo ? split : 0;
    a.rhs = p;
    a.out = z;
    a.dinv = op.dinv;
    a.ra = red(h, 0, FIN_NONE);
    if (!use_state) a.ra.st = h.st.p + kIdleState;
    TailArgs<TV, TX> b;
    b.rowptr = m2.rowptr.p;
    b.colidx = m2.colidx.p;
    b.val = op.S2val;
    b.rhs = y;
    b.out = z;
    b.dinv = op.dinv;
    b.row0 = m2.tail_row0;
    b.nrows = m2.tail_nrows;
    b.dep = m2.tail_dep.p;
    b.ndep = m2.tail_ndep;
    if (b.ndep < 0 || b.ndep > kTailRows || m1.colour_tile[nc] - m1.colour_tile[nc - 1] != 1 ||
        b.row0 != h.an.colour_ptr[nc - 2] || b.row0 + b.nrows != h.an.colour_ptr[nc - 1])
        throw std::runtime_error("merged tail launch does not fit the colouring");
    if (op.s2_dinv) tail_dispatch<BS, SM_SUB_DINV, SM_SUB_DINV, TV, TX>(h, a, b);
    else if (op.s1_dinv) tail_dispatch<BS, SM_SUB_DINV, SM_SUB, TV, TX>(h, a, b);
    else tail_dispatch<BS, SM_SUB, SM_SUB, TV, TX>(h, a, b);
}

template <int BS, int ROWSB, typename TV, typename TX>
void tail3_dispatch(H &h, const StreamArgs<TV, TX> &a, const Tail3Args<TV, TX> &b, const StreamArgs<TV, TX> &c) {
    hipLaunchKernelGGL((k_tail3<BS, ROWSB, TV, TX>), dim3(1 + c.ntiles), dim3(64 * BS), 0, h.stream, a, b, c);
    check_launch();
}

// The extended merged tail (k_tail3) of a forward application whose first
// sweep has a unit diagonal: the first sweep's colour C-1, the second
// sweep's colours C-2 and C-3; rhs3 is y, or p when colour C-3 is colour 0
// and the first sweep skipped it.
template <int BS, typename TV, typename TX>
void tail3_launch(H &h, const OpSet<TV> &op, const TX *y, const TX *lo, int split, const TX *p, TX *z,
                  const TX *rhs3, bool use_state) {
    const int nc = h.an.ncolours;
    const DevMat &m1 = *op.S1, &m2 = *op.S2;
    if (op.s1_dinv || !op.s2_dinv || m2.tail3_n < 0 || m2.tail_ndep < 0 || nc < 3 ||
        m1.colour_tile[nc] - m1.colour_tile[nc - 1] != 1 || m2.tail_row0 != h.an.colour_ptr[nc - 2])
        throw std::runtime_error("extended tail launch does not fit the colouring");
    StreamArgs<TV, TX> a;
    std::memset(&a, 0, sizeof(a));
    a.tile = m1.tile.p;
    a.slot = m1.slot.p;
    a.tile0 = m1.colour_tile[nc - 1];
    a.ntiles = 1;
    a.rowptr = m1.rowptr.p;
    a.colidx = m1.colidx.p;
    a.val = op.S1val;
    a.x = y;
    a.x_lo = lo ? lo : y;
    a.split = lo ? split : 0;
    a.rhs = p;
    a.out = z;
    a.dinv = op.dinv;
    a.ra = red(h, 0, FIN_NONE);
    if (!use_state) a.ra.st = h.st.p + kIdleState;
    Tail3Args<TV, TX> b;
    b.val = op.S2val;
    b.rhs2 = y;
    b.rhs3 = rhs3;
    b.out = z;
    b.dinv = op.dinv;
    b.n2 = m2.tail_ndep;
    b.n3 = m2.tail3_n;
    b.row = m2.tail3_row.p;
    b.blk = m2.tail3_blk.p;
    b.src = m2.tail3_src.p;
    StreamArgs<TV, TX> c;
    std::memset(&c, 0, sizeof(c));
    c.tile = m2.tail3_tile.p;
    c.slot = nullptr;  // no reductions
    c.tile0 = 0;
    c.ntiles = m2.tail3_ntiles;
    c.rowptr = m2.rowptr.p;
    c.colidx = m2.colidx.p;
    c.val = op.S2val;
    c.x = z;
    c.x_lo = z;
    c.split = 0;
    c.rhs = rhs3;
    c.out = z;
    c.dinv = op.dinv;
    c.ra = a.ra;
    if (m2.tail3_tall) tail3_dispatch<BS, kTileRowsTall, TV, TX>(h, a, b, c);
    else tail3_dispatch<BS, kTileRows, TV, TX>(h, a, b, c);
}

template <int BS, int MODEB, int ROWSA, typename TV, typename TX>
void lower2_dispatch(H &h, const StreamArgs<TV, TX> &a, const StreamArgs<TV, TX> &v, const StreamArgs<TV, TX> &g) {
    hipLaunchKernelGGL((k_lower2<BS, MODEB, ROWSA, TV, TX>), dim3(a.ntiles + g.ntiles), dim3(64 * BS), 0, h.stream, a,
                       v, g);
    check_launch();
}

// The first sweep's colours C-3 and C-2 in one launch (k_lower2) of a
// forward application: y for both; with zout (the extended tail follows),
// also z of colour C-2's rows without second-sweep blocks.
template <int BS, typename TV, typename TX>
void lower2_launch(H &h, const OpSet<TV> &op, const TV *vval, const TX *p, TX *y, const TX *lo, int split, TX *zout,
                   bool use_state) {
    const int nc = h.an.ncolours;
    const DevMat &m = *op.S1;
    const int c = nc - 3;
    if (op.s1_dinv || m.low2_n < 0 || nc < 4) throw std::runtime_error("merged lower launch does not fit the colouring");
    StreamArgs<TV, TX> a;
    std::memset(&a, 0, sizeof(a));
    a.rowptr = m.rowptr.p;
    a.colidx = m.colidx.p;
    a.val = op.S1val;
    a.x = y;
    a.x_lo = lo ? lo : y;
    a.split = lo ? split : 0;
    a.rhs = p;
    a.out = y;
    a.ra = red(h, 0, FIN_NONE);
    if (!use_state) a.ra.st = h.st.p + kIdleState;
    StreamArgs<TV, TX> v = a, g = a;
    // colour c's own tiles, tall when that saves a quarter of the workgroups
    a.tile0 = m.colour_tile[c];
    a.ntiles = m.colour_tile[c + 1] - m.colour_tile[c];
    a.tile = m.tile.p;
    const int ntall = m.colour_tile_tall[c + 1] - m.colour_tile_tall[c];
    const bool tall = 4 * ntall <= 3 * a.ntiles;
    if (tall) {
        a.tile = m.tile_tall.p;
        a.tile0 = m.colour_tile_tall[c];
        a.ntiles = ntall;
    }
    v.tile = m.low2_vtile.p;
    v.tile0 = 0;
    v.ntiles = m.low2_n;
    v.rowptr = m.low2_vrowptr.p;
    v.colidx = m.low2_vcol.p;
    v.val = vval;
    v.rowmap = m.low2_vmap.p;
    g.tile = m.low2_tile.p;
    g.tile0 = 0;
    g.ntiles = m.low2_n;
    g.colidx = m.low2_col.p;
    if (zout) {
        g.rowptr2 = op.S2->rowptr.p;
        g.zout = zout;
        g.dinv = op.dinv;
        if (tall) lower2_dispatch<BS, SM_SUB_Z, kTileRowsTall, TV, TX>(h, a, v, g);
        else lower2_dispatch<BS, SM_SUB_Z, kTileRows, TV, TX>(h, a, v, g);
    } else {
        if (tall) lower2_dispatch<BS, SM_SUB, kTileRowsTall, TV, TX>(h, a, v, g);
        else lower2_dispatch<BS, SM_SUB, kTileRows, TV, TX>(h, a, v, g);
    }
}

template <typename TV> const TV *low2_values(H &h);
template <> const double *low2_values<double>(H &h) { return h.low2_v64.p; }
template <> const float *low2_values<float>(H &h) { return h.low2_v32.p; }

// z = M^{-1} p through the two colour sweeps of op (first sweep colours
// ascending into y, second sweep colours descending into z).  Two launches
// are saved per application:
//  - when the first sweep has a unit diagonal, its colour-0 rows have no
//    blocks (nothing precedes colour 0) and y equals p there; that launch is
//    skipped and later gathers read p for colour-0 columns;
//  - the last colour has no blocks in the second sweep (nothing follows
//    it), so its first-sweep launch writes z directly, applying the second
//    sweep's diagonal inverse in the same epilogue when it has one, to the
//    first sweep's value before any rounding (that value is never stored).
// When the last colour is a single tile and few rows of colour C-2 have
// second-sweep blocks, that launch and the second sweep's colour C-2 launch
// are one (tail_launch).
template <int BS, typename TV, typename TX>
void precond(H &h, const OpSet<TV> &op, const TX *p, TX *y, TX *z, bool use_state) {
    const int nc = h.an.ncolours;
    if (op.s1_dinv && op.s2_dinv) throw std::runtime_error("at most one sweep may carry the diagonal inverse");
    const bool skip0 = !op.s1_dinv && nc > 1;
    const TX *lo = skip0 ? p : nullptr;
    const int split = skip0 ? h.an.colour_ptr[1] : 0;
    const bool merge = nc >= 3 && op.S2->tail_ndep >= 0 &&
                       op.S1->colour_tile[nc] - op.S1->colour_tile[nc - 1] == 1;
    const bool merge3 = merge && h.tail3 && !op.s1_dinv && op.s2_dinv && op.S2->tail3_n >= 0;
    const bool low2 = h.low2 && skip0 && op.S1->low2_n >= 0 && nc >= 4 && op.S1 == &h.L;
    prof_begin(h, PK_LOWER);
    for (int c = skip0 ? 1 : 0; c < nc - 1; ++c) {
        if (low2 && c == nc - 3) {
            lower2_launch<BS, TV, TX>(h, op, low2_values<TV>(h), p, y, lo, split, merge3 ? z : nullptr, use_state);
            ++c;  // colour C-2 is done too
        } else if (merge3 && c == nc - 2)
            launch_stream<BS, SM_SUB_Z, SR_NONE, TV, TX>(h, *op.S1, op.S1val, c, y, p, y, op.dinv, nullptr,
                                                         red(h, 0, FIN_NONE), use_state, nullptr, nullptr, lo, split,
                                                         -1, op.S2->rowptr.p, z);
        else
            sweep_launch<BS, TV, TX>(h, *op.S1, op.S1val, c, op.s1_dinv, y, lo, split, p, y, op.dinv, use_state);
    }
    if (!merge)
        sweep_launch<BS, TV, TX>(h, *op.S1, op.S1val, nc - 1, op.s1_dinv || op.s2_dinv, y, lo, split, p, z, op.dinv,
                                 use_state);
    prof_end(h);
    prof_begin(h, PK_UPPER);
    if (merge3) tail3_launch<BS, TV, TX>(h, op, y, lo, split, p, z, (nc - 3 == 0 && skip0) ? p : y, use_state);
    else if (merge) tail_launch<BS, TV, TX>(h, op, y, lo, split, p, z, use_state);
    for (int c = merge3 ? nc - 4 : merge ? nc - 3 : nc - 2; c >= 0; --c)
        sweep_launch<BS, TV, TX>(h, *op.S2, op.S2val, c, op.s2_dinv, z, nullptr, 0, (c == 0 && skip0) ? p : y, z,
                                 op.dinv, use_state);
    prof_end(h);
}

// Kernel launches of one application of op's preconditioner (as precond
// issues them).
template <typename TV>
int precond_launches(const H &h, const OpSet<TV> &op) {
    const int nc = h.an.ncolours;
    const bool skip0 = !op.s1_dinv && nc > 1;
    const bool merge = nc >= 3 && op.S2->tail_ndep >= 0 &&
                       op.S1->colour_tile[nc] - op.S1->colour_tile[nc - 1] == 1;
    const bool merge3 = merge && h.tail3 && !op.s1_dinv && op.s2_dinv && op.S2->tail3_n >= 0;
    const bool low2 = h.low2 && skip0 && op.S1->low2_n >= 0 && nc >= 4 && op.S1 == &h.L;
    auto nonempty = [&](const DevMat &m, int c) { return m.colour_tile[c + 1] > m.colour_tile[c] ? 1 : 0; };
    int k = 0;
    for (int c = skip0 ? 1 : 0; c < nc - 1; ++c) k += nonempty(*op.S1, c);
    if (low2) k -= 1;
    if (!merge) k += nonempty(*op.S1, nc - 1);
    if (merge || merge3) k += 1;
    for (int c = merge3 ? nc - 4 : merge ? nc - 3 : nc - 2; c >= 0; --c) k += nonempty(*op.S2, c);
    return k;
}

struct RunOut {
    SolverState st;
    double ms = 0.0;
    bool aborted = false;
};

// Fresh state for a run on state slot idx.  The state is written by a kernel
// on the stream, so no host synchronisation is needed; the run's id goes to
// the slot's mapped flag when the run stops.
SolverState new_state(H &h, int idx, double tol, int maxit, bool abs_mode) {
    SolverState s;
    std::memset(&s, 0, sizeof(s));
    s.tol = tol;
    s.maxit = maxit;
    s.abs_mode = abs_mode ? 1 : 0;
    s.norm_slot = 1;
    h.run_seq = h.run_seq == INT32_MAX ? 1 : h.run_seq + 1;
    s.run_id = h.run_seq;
    s.host_done = h.d_done + idx;
    s.host_prog = h.d_prog + idx;
    if (h.timeout_s > 0.0)
        s.timeout_ticks = (uint64_t)std::max(1.0, std::ceil(h.timeout_s * 1e3 * (double)h.wall_khz));
    if (idx == 0 && h.hist_cap > 0) {
        s.hist = h.hist.p;
        s.hist_cap = h.hist_cap;
    }
    return s;
}

bool run_stopped(const H &h, int idx, int32_t run_id) {
    return __atomic_load_n(h.h_done + idx, __ATOMIC_ACQUIRE) == run_id;
}

// Waits until run run_id on state slot idx has reported progress >= count,
// or has stopped; false if the handle's timeout (measured from t0) expired.
// The host polls the mapped words; a wait that has lasted longer than
// kSpinWait (the device is more than a short kernel sequence behind: large
// systems) polls every kSleepStep instead, so that the waiting thread gives
// its core back (a simulator runs one MPI rank per core).  Short waits, which
// decide how soon the next launches are queued, keep spinning.
constexpr std::chrono::microseconds kSpinWait{200}, kSleepStep{20};

bool wait_progress(const H &h, int idx, int32_t run_id, int32_t count,
                   std::chrono::steady_clock::time_point t0) {
    const auto w0 = std::chrono::steady_clock::now();
    for (uint32_t spin = 0;; ++spin) {
        const uint64_t v = __atomic_load_n(h.h_prog + idx, __ATOMIC_ACQUIRE);
        if ((int32_t)(v >> 32) == run_id && (int32_t)(uint32_t)v >= count) return true;
        if (run_stopped(h, idx, run_id)) return true;
        if ((spin & 255) == 255) {
            const auto now = std::chrono::steady_clock::now();
            if (h.timeout_s > 0.0 && std::chrono::duration<double>(now - t0).count() > h.timeout_s) return false;
            if (now - w0 > kSpinWait) std::this_thread::sleep_for(kSleepStep);
            else std::this_thread::yield();
        }
    }
}

int32_t init_state(H &h, int idx, double tol, int maxit, bool abs_mode) {
    const SolverState s = new_state(h, idx, tol, maxit, abs_mode);
    hipLaunchKernelGGL(k_set_state, dim3(1), dim3(1), 0, h.stream, h.st.p + idx, s);
    check_launch();
    return s.run_id;
}

// One BiCGSTAB run following the reference microprogram (solver_pkg.vhd:120-141).
// zero_x0: x0 = 0 and the initial residual is the right-hand side itself.
// outer != nullptr: an inner run of the precision form; its tolerance comes
// from the outer state on the device, and the run ends without reading its
// state back (the outer state collects the counters).
template <int BS, typename TV, typename TX>
RunOut run_bicgstab(H &h, const OpSet<TV> &op, Krylov<TX> &k, const double *rhs, bool zero_x0, double tol,
                    int maxit, bool abs_mode, const SolverState *outer = nullptr) {
    const size_t N = h.N();
    const int vb = vec_blocks(N);
    int32_t run_id;
    if (outer) {
        const SolverState s = new_state(h, 0, 0.0, maxit, false);
        run_id = s.run_id;
        hipLaunchKernelGGL(k_inner_start, dim3(1), dim3(1), 0, h.stream, h.st.p, outer, s);
        check_launch();
    } else {
        run_id = init_state(h, 0, tol, maxit, abs_mode);
    }
    RunOut out;
    const auto t0 = std::chrono::steady_clock::now();
    h.prof_tag = -1;
    if (zero_x0) {
        hipLaunchKernelGGL((k_init_zero_x0<TX, TX>), dim3(vb), dim3(kVecThreads), 0, h.stream, (int64_t)N, rhs,
                           k.rt.p, k.r.p, k.p.p, k.x.p, red(h, 0, FIN_INIT));
        check_launch();
        launch_finalize(h, 1, red(h, 0, FIN_INIT), vb);
    } else {
        if constexpr (std::is_same<TX, double>::value) {
            // r = b - A x0; with x0 = 0 (and a finite matrix) this is b, without the product
            hipLaunchKernelGGL((k_mark_nonzero<TX>), dim3(vb), dim3(kVecThreads), 0, h.stream, (int64_t)N,
                               (const TX *)k.x.p, (const int32_t *)h.d_nonfinite.p, h.st.p);
            check_launch();
            launch_stream<BS, SM_SUB, SR_SELF, TV, TX, true>(h, *op.A, op.Aval, -1, k.x.p, rhs, k.rt.p, nullptr,
                                                             nullptr, red(h, 0, FIN_INIT), true, k.r.p, k.p.p,
                                                             nullptr, 0, PK_SPMV);
        } else {
            throw std::runtime_error("non-zero start vector needs FP64 vectors");
        }
    }
    // Iterations are queued ahead of the device; the device stops itself
    // (every kernel returns at once after the stop) and raises a mapped host
    // flag.  The host checks its lead in the middle of every half-iteration,
    // after queueing that half's preconditioner: it waits until at most
    // 2 * ahead - 1 earlier half-iterations are unfinished (ahead =
    // bda_set_options' poll_iters, 1..3), on the mapped progress word (no
    // stream event), which the scalar step of every half-iteration advances,
    // and checks the stop flag each time.  What is still queued when the
    // device stops runs as no-op launches (a few microseconds each); with
    // the check in mid half-iteration that is the preconditioner's launches
    // of one half at most (ahead = 1), while the device still has that
    // preconditioner queued whenever the host waits.
    bool stop = false;
    const int ahead = std::min(3, std::max(1, h.poll_iters));
    const int ahead_h = 2 * ahead - 1;
    auto throttle = [&](int half) {
        if (half > ahead_h && !wait_progress(h, 0, run_id, half - ahead_h, t0)) {
            out.aborted = true;
            stop = true;
        }
        if (run_stopped(h, 0, run_id)) stop = true;
        if (h.timeout_s > 0.0) {
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el > h.timeout_s) {
                out.aborted = true;
                stop = true;
            }
        }
    };
    for (int it = 1; it <= maxit + 1 && !stop; ++it) {
        h.prof_tag = 2 * (it - 1);
        precond<BS, TV, TX>(h, op, k.p.p, k.y.p, k.phat.p, true);
        throttle(2 * it - 1);
        if (stop) break;
        launch_stream<BS, SM_SPMV, SR_W0, TV, TX>(h, *op.A, op.Aval, -1, k.phat.p, nullptr, k.v.p, nullptr, k.rt.p,
                                                  red(h, 0, FIN_ALPHA), true, nullptr, nullptr, nullptr, 0, PK_SPMV);
        prof_begin(h, PK_VEC);
        hipLaunchKernelGGL((k_update_s<TX>), dim3(vb), dim3(kVecThreads), 0, h.stream, (int64_t)N, k.r.p, k.v.p,
                           k.s.p, red(h, 0, FIN_HALF1));
        check_launch();
        launch_finalize(h, 1, red(h, 0, FIN_HALF1), vb);
        prof_end(h);
        h.prof_tag = 2 * (it - 1) + 1;
        precond<BS, TV, TX>(h, op, k.s.p, k.y.p, k.shat.p, true);
        throttle(2 * it);
        if (stop) break;
        launch_stream<BS, SM_SPMV, SR_BOTH, TV, TX>(h, *op.A, op.Aval, -1, k.shat.p, nullptr, k.t.p, nullptr, k.s.p,
                                                    red(h, 0, FIN_OMEGA), true, nullptr, nullptr, nullptr, 0, PK_SPMV);
        prof_begin(h, PK_VEC);
        hipLaunchKernelGGL((k_update_xr<TX, TX>), dim3(vb), dim3(kVecThreads), 0, h.stream, (int64_t)N, k.x.p,
                           k.phat.p, k.shat.p, k.s.p, k.t.p, k.r.p, k.rt.p, red(h, 0, FIN_HALF2));
        check_launch();
        launch_finalize(h, 2, red(h, 0, FIN_HALF2), vb);
        hipLaunchKernelGGL((k_update_p<TX>), dim3(vb), dim3(kVecThreads), 0, h.stream, (int64_t)N, k.p.p, k.r.p,
                           k.v.p, (const SolverState *)(h.st.p));
        check_launch();
        prof_end(h);
    }
    // x += alpha phat of a run that ended after a first half-iteration (the
    // x update of a first half is otherwise applied by the next k_update_xr)
    hipLaunchKernelGGL((k_finish_x<TX, TX>), dim3(vb), dim3(kVecThreads), 0, h.stream, (int64_t)N, k.x.p, k.phat.p,
                       (const SolverState *)(h.st.p));
    check_launch();
    if (outer) {
        h.prof_tag = -1;
        return out;
    }
    HIP_CHECK(hipStreamSynchronize(h.stream));
    out.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    copy_sync(h, h.h_st, h.st.p, sizeof(SolverState), hipMemcpyDeviceToHost);
    out.st = *h.h_st;
    prof_collect(h, out.st.no_results ? -1 : out.st.h);
    h.prof_tag = -1;
    return out;
}

void fill_result(const RunOut &o, bda_result *res) {
    if (!res) return;
    std::memset(res, 0, sizeof(*res));
    const SolverState &s = o.st;
    res->half_iterations = s.h;
    res->iterations = s.h / 2.0 + 0.5;
    res->last_norm_idx = s.last_norm_idx;
    for (int i = 0; i < 4; ++i) res->norms[i] = s.norms[i];
    res->final_norm = s.last_norm;
    res->rel_residual = s.norms[0] > 0 ? s.last_norm / s.norms[0] : 0.0;
    res->converged = s.converged;
    res->no_results = s.no_results;
    res->max_iter_reached = s.max_iter;
    res->breakdown = s.breakdown;
    res->nan_detected = s.nan_flag;
    res->aborted = (o.aborted || s.aborted) ? 1 : 0;
    res->solve_ms = o.ms;
}

// ----------------------------------------------------------------------------
void require_analysed(H &h) {
    if (!h.analysed) throw std::runtime_error("bda_analyse has not been called");
}
void require_factored(H &h) {
    require_analysed(h);
    if (!h.factored) throw std::runtime_error("bda_factor has not been called");
}

template <int BS, typename TD>
void gather_blocks(H &h, int64_t nblk, const int32_t *idx, const double *src, TD *dst, int trans) {
    if (nblk <= 0) return;
    const int64_t work = nblk * BS * BS;
    hipLaunchKernelGGL((k_gather_blocks<BS, TD>), dim3(blocks_for(work, 256)), dim3(256), 0, h.stream, nblk, idx,
                       src, dst, trans);
    check_launch();
}

template <int BS>
void permute_values(H &h) {
    gather_blocks<BS, double>(h, h.A.nnzb, h.A.src.p, h.d_orig.p, h.A64.p, 0);
    HIP_CHECK(hipStreamSynchronize(h.stream));
}

// Device side of an analysis held in h.an: patterns, tilings, permutation
// and the buffers sized by them.
void log_analysis(H &h);

void finish_analyse(H &h) {
    const Analysis &an = h.an;
    const int64_t span = DevMat::coupling_span(an.A, an.perm);
    h.A.upload(an.A, h.stream, &an.perm, span);
    h.L.upload(an.L, h.stream, &an.perm, span);
    h.U.upload(an.U, h.stream, &an.perm, span);
    h.At.upload(an.At, h.stream, &an.perm, span);
    h.Ut.upload(an.Ut, h.stream, &an.perm, span);
    h.Lt.upload(an.Lt, h.stream, &an.perm, span);
    h.U.set_tail(an.U, an.colour_ptr, h.stream);
    h.U.set_tail3(an.U, an.colour_ptr, an.perm, span, h.bs, h.stream);
    h.L.set_low2(an.L, an.colour_ptr, an.perm, span, h.bs, h.stream);
    h.Lt.set_tail(an.Lt, an.colour_ptr, h.stream);
    h.d_perm.ensure(h.n);
    h.d_diag.ensure(h.n);
    copy_sync(h, h.d_perm.p, an.perm.data(), h.n * 4, hipMemcpyHostToDevice);
    copy_sync(h, h.d_diag.p, an.diag.data(), h.n * 4, hipMemcpyHostToDevice);
    h.blkrow.resize(an.A.nnzb());
    for (int k = 0; k < h.n; ++k)
        for (int32_t q = an.A.rowptr[k]; q < an.A.rowptr[k + 1]; ++q) h.blkrow[q] = k;
    h.d_blkrow.ensure(h.blkrow.size());
    copy_sync(h, h.d_blkrow.p, h.blkrow.data(), h.blkrow.size() * 4, hipMemcpyHostToDevice);

    const size_t bb = (size_t)h.bs * h.bs;
    const int64_t nnzb = an.A.nnzb();
    h.A64.ensure((nnzb + kPadBlocks) * bb);
    h.LU.ensure(nnzb * bb);
    h.L64.ensure((an.L.nnzb() + kPadBlocks) * bb);
    h.U64.ensure((an.U.nnzb() + kPadBlocks) * bb);
    h.Dinv.ensure((size_t)h.n * bb);
    h.DinvT.ensure((size_t)h.n * bb);
    h.d_bad.ensure(1);
    h.d_nonfinite.ensure(1);
    // reduction partials: one slot pair per workgroup of the largest reducing
    // launch, and room for the first stage of a two-stage sum
    const int cap = std::max({h.A.ntiles, h.At.ntiles, vec_blocks(h.N()), 1}) + kFinGroups;
    h.partial_cap = cap;
    h.partials.ensure((size_t)2 * cap);
    h.k64.ensure(h.Npad());
    h.bp.ensure(h.Npad());
    h.xo.ensure(h.Npad());
    h.ro.ensure(h.Npad());
    h.vin.ensure(h.Npad());
    h.vout.ensure(h.Npad());
    h.vw.ensure(h.Npad());
    h.analysed = true;
    h.factored = false;
    if (h.have_values) {
        if (h.bs == 3) permute_values<3>(h);
        else permute_values<4>(h);
    }
    log_analysis(h);
}

void do_analyse(H &h, bool level) {
    if (!h.have_pattern) throw std::runtime_error("bda_upload_bcsr has not been called");
    analyse(h.pattern, level, h.an);
    finish_analyse(h);
}

// Values of L's packed dependency rows (k_lower2) from L64.
template <int BS>
void gather_low2(H &h) {
    if (h.L.low2_n < 0) return;
    h.low2_v64.ensure((h.L.low2_nvblk + kPadBlocks) * BS * BS);
    gather_blocks<BS, double>(h, h.L.low2_nvblk, h.L.low2_vsrc.p, h.L64.p, h.low2_v64.p, 0);
}

template <int BS>
void do_factor(H &h) {
    require_analysed(h);
    if (!h.have_values) throw std::runtime_error("no matrix values uploaded");
    constexpr int BB = BS * BS;
    const Analysis &an = h.an;
    const int32_t big = INT_MAX, zero = 0;
    const auto t0 = std::chrono::steady_clock::now();
    HIP_CHECK(hipMemcpyAsync(h.d_bad.p, &big, 4, hipMemcpyHostToDevice, h.stream));
    HIP_CHECK(hipMemcpyAsync(h.d_nonfinite.p, &zero, 4, hipMemcpyHostToDevice, h.stream));
    {   // LU = A, recording whether A holds a non-finite value (for the step-0 residual's short path)
        const int64_t nval = (int64_t)h.A.nnzb * BB;
        const int grid = (int)std::min<int64_t>(std::max<int64_t>(blocks_for(nval, 256), 1), 8192);
        hipLaunchKernelGGL(k_copy_check_finite, dim3(grid), dim3(256), 0, h.stream, nval, (const double *)h.A64.p,
                           h.LU.p, h.d_nonfinite.p);
        check_launch();
    }
    for (int c = 0; c < an.ncolours; ++c) {
        const int r0 = an.colour_ptr[c], r1 = an.colour_ptr[c + 1];
        hipLaunchKernelGGL((k_bilu0_rows<BS>), dim3(blocks_for(r1 - r0, 128)), dim3(128), 0, h.stream, r0, r1,
                           h.A.rowptr.p, h.A.colidx.p, h.d_diag.p, h.LU.p, h.Dinv.p, h.d_bad.p);
        check_launch();
    }
    int32_t bad = 0;
    HIP_CHECK(hipMemcpyAsync(&bad, h.d_bad.p, 4, hipMemcpyDeviceToHost, h.stream));
    HIP_CHECK(hipStreamSynchronize(h.stream));
    if (bad != INT_MAX)
        throw std::runtime_error("singular diagonal block in ILU0 at colour-ordered row " + std::to_string(bad) +
                                 " (original row " + std::to_string(an.perm[bad]) + ")");
    bda::info(1, "INFO: factor: block ILU0 of %d block rows over %d colours, %.3f ms", h.n, an.ncolours,
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    gather_blocks<BS, double>(h, h.L.nnzb, h.L.src.p, h.LU.p, h.L64.p, 0);
    gather_blocks<BS, double>(h, h.U.nnzb, h.U.src.p, h.LU.p, h.U64.p, 0);
    gather_blocks<BS, double>(h, h.n, nullptr, h.Dinv.p, h.DinvT.p, 1);
    gather_low2<BS>(h);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    h.factored = true;
    h.factor_gen++;
    h.values_gen++;
}

// The caller's own factors of a colour-ordered system (identity
// permutation): L and U from lu_values on A's pattern, D^-1 as given.  The
// finiteness flag of the step-0 short path is set from A as bda_factor
// sets it.
template <int BS>
void load_factors(H &h, const double *lu_values, const double *dinv) {
    constexpr int BB = BS * BS;
    const int32_t zero = 0;
    HIP_CHECK(hipMemcpyAsync(h.d_nonfinite.p, &zero, 4, hipMemcpyHostToDevice, h.stream));
    const int64_t nval = (int64_t)h.A.nnzb * BB;
    const int grid = (int)std::min<int64_t>(std::max<int64_t>(blocks_for(nval, 256), 1), 8192);
    hipLaunchKernelGGL(k_copy_check_finite, dim3(grid), dim3(256), 0, h.stream, nval, (const double *)h.A64.p, h.LU.p,
                       h.d_nonfinite.p);
    check_launch();
    HIP_CHECK(hipMemcpyAsync(h.LU.p, lu_values, (size_t)nval * 8, hipMemcpyHostToDevice, h.stream));
    HIP_CHECK(hipMemcpyAsync(h.Dinv.p, dinv, (size_t)h.n * BB * 8, hipMemcpyHostToDevice, h.stream));
    gather_blocks<BS, double>(h, h.L.nnzb, h.L.src.p, h.LU.p, h.L64.p, 0);
    gather_blocks<BS, double>(h, h.U.nnzb, h.U.src.p, h.LU.p, h.U64.p, 0);
    gather_blocks<BS, double>(h, h.n, nullptr, h.Dinv.p, h.DinvT.p, 1);
    gather_low2<BS>(h);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    h.factored = true;
    h.factor_gen++;
    h.values_gen++;
}

template <int BS>
void ensure_mixed(H &h) {
    require_factored(h);
    if (h.mixed_gen == h.factor_gen) return;
    const size_t bb = (size_t)BS * BS;
    h.A32.ensure((h.A.nnzb + kPadBlocks) * bb);
    h.L32.ensure((h.L.nnzb + kPadBlocks) * bb);
    h.U32.ensure((h.U.nnzb + kPadBlocks) * bb);
    gather_blocks<BS, float>(h, h.A.nnzb, nullptr, h.A64.p, h.A32.p, 0);
    gather_blocks<BS, float>(h, h.L.nnzb, nullptr, h.L64.p, h.L32.p, 0);
    gather_blocks<BS, float>(h, h.U.nnzb, nullptr, h.U64.p, h.U32.p, 0);
    if (h.L.low2_n >= 0) {
        h.low2_v32.ensure((h.L.low2_nvblk + kPadBlocks) * bb);
        gather_blocks<BS, float>(h, h.L.low2_nvblk, nullptr, h.low2_v64.p, h.low2_v32.p, 0);
    }
    HIP_CHECK(hipStreamSynchronize(h.stream));
    h.mixed_gen = h.factor_gen;
}

template <int BS>
void ensure_adjoint(H &h) {
    require_factored(h);
    if (h.adj_gen == h.factor_gen) return;
    const size_t bb = (size_t)BS * BS;
    h.At64.ensure((h.At.nnzb + kPadBlocks) * bb);
    h.Ut64.ensure((h.Ut.nnzb + kPadBlocks) * bb);
    h.Lt64.ensure((h.Lt.nnzb + kPadBlocks) * bb);
    gather_blocks<BS, double>(h, h.At.nnzb, h.At.src.p, h.A64.p, h.At64.p, 1);
    gather_blocks<BS, double>(h, h.Ut.nnzb, h.Ut.src.p, h.U64.p, h.Ut64.p, 1);
    gather_blocks<BS, double>(h, h.Lt.nnzb, h.Lt.src.p, h.L64.p, h.Lt64.p, 1);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    h.adj_gen = h.factor_gen;
}

// FP32 copies of the transposed operators, demoted from the FP64 transposes
// (the same values the oracle's FP32 transposes hold).  Only the component
// access uses them: the adjoint solve runs in FP64.
template <int BS>
void ensure_adjoint32(H &h) {
    ensure_adjoint<BS>(h);
    if (h.adj32_gen == h.factor_gen) return;
    const size_t bb = (size_t)BS * BS;
    h.At32.ensure((h.At.nnzb + kPadBlocks) * bb);
    h.Ut32.ensure((h.Ut.nnzb + kPadBlocks) * bb);
    h.Lt32.ensure((h.Lt.nnzb + kPadBlocks) * bb);
    gather_blocks<BS, float>(h, h.At.nnzb, nullptr, h.At64.p, h.At32.p, 0);
    gather_blocks<BS, float>(h, h.Ut.nnzb, nullptr, h.Ut64.p, h.Ut32.p, 0);
    gather_blocks<BS, float>(h, h.Lt.nnzb, nullptr, h.Lt64.p, h.Lt32.p, 0);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    h.adj32_gen = h.factor_gen;
}

OpSet<double> op_fwd64(H &h) {
    return OpSet<double>{&h.A, h.A64.p, &h.L, h.L64.p, false, &h.U, h.U64.p, true, h.Dinv.p};
}
OpSet<float> op_fwd32(H &h) {
    return OpSet<float>{&h.A, h.A32.p, &h.L, h.L32.p, false, &h.U, h.U32.p, true, h.Dinv.p};
}
OpSet<double> op_adj64(H &h) {
    return OpSet<double>{&h.At, h.At64.p, &h.Ut, h.Ut64.p, true, &h.Lt, h.Lt64.p, false, h.DinvT.p};
}
OpSet<float> op_adj32(H &h) {
    return OpSet<float>{&h.At, h.At32.p, &h.Ut, h.Ut32.p, true, &h.Lt, h.Lt32.p, false, h.DinvT.p};
}

// Diagnostics of an analysed system: sizes (level 1), then per colour and
// the preconditioner's launch plan (level 2), the counterpart of the
// reference's per-colour size printouts (fpga_functions_bicgstab.cpp:146-167).
void log_analysis(H &h) {
    if (bda::debug_level() < 1) return;
    const Analysis &an = h.an;
    bda::info(1, "INFO: analyse: %d block rows of %dx%d, %lld blocks, %d colours (%s); L %lld and U %lld blocks",
              h.n, h.bs, h.bs, (long long)an.A.nnzb(), an.ncolours, an.level_scheduling ? "level sets" : "greedy",
              (long long)an.L.nnzb(), (long long)an.U.nnzb());
    if (bda::debug_level() < 2) return;
    for (int c = 0; c < an.ncolours; ++c) {
        const int r0 = an.colour_ptr[c], r1 = an.colour_ptr[c + 1];
        bda::info(2, "INFO: analyse: colour %d: %d rows; A %d, L %d, U %d blocks; L %d and U %d tiles", c, r1 - r0,
                  an.A.rowptr[r1] - an.A.rowptr[r0], an.L.rowptr[r1] - an.L.rowptr[r0],
                  an.U.rowptr[r1] - an.U.rowptr[r0], h.L.colour_tile[c + 1] - h.L.colour_tile[c],
                  h.U.colour_tile[c + 1] - h.U.colour_tile[c]);
    }
    bda::info(2, "INFO: analyse: %d SpMV tiles; preconditioner launches per application: %d, transposed %d",
              h.A.ntiles, precond_launches(h, op_fwd64(h)), precond_launches(h, op_adj64(h)));
}

// One line per solve (level 1) and its norm ring (level 2).
void log_result(const char *what, const bda_result &r) {
    if (bda::debug_level() < 1) return;
    bda::info(1,
              "INFO: %s: %d half-iterations (%d refinement steps), rel residual %.3e; converged %d, no results %d, "
              "breakdown %d, NaN %d, iteration cap %d, aborted %d; %.3f ms",
              what, r.half_iterations, r.outer_iterations, r.rel_residual, r.converged, r.no_results, r.breakdown,
              r.nan_detected, r.max_iter_reached, r.aborted, r.solve_ms);
    bda::info(2, "INFO: %s: ||r0|| %.6e, last norms %.6e %.6e %.6e (newest %d)", what, r.norms[0], r.norms[1],
              r.norms[2], r.norms[3], r.last_norm_idx);
}

void gather_vec2(H &h, const double *in1, double *out1, const double *in2, double *out2) {
    hipLaunchKernelGGL(k_gather_rows2, dim3(blocks_for((int64_t)h.N(), 256)), dim3(256), 0, h.stream, h.n, h.bs,
                       h.d_perm.p, in1, out1, in2, out2);
    check_launch();
}

void gather_vec(H &h, const double *in, double *out) {
    hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for((int64_t)h.N(), 256)), dim3(256), 0, h.stream, h.n, h.bs,
                       h.d_perm.p, in, out);
    check_launch();
}
void scatter_vec(H &h, const double *in, double *out) {
    hipLaunchKernelGGL(k_scatter_rows, dim3(blocks_for((int64_t)h.N(), 256)), dim3(256), 0, h.stream, h.n, h.bs,
                       h.d_perm.p, in, out);
    check_launch();
}

void check_solve_args(double tol, int maxit) {
    if (!(tol >= 0.0) || std::isinf(tol)) throw std::runtime_error("tolerance must be finite and >= 0");
    if (maxit < 0 || maxit > 32767) throw std::runtime_error("maxit must be in [0, 32767]");
}

// plain FP64 solve on device vectors in original ordering
template <int BS>
void solve_dev(H &h, const double *d_b, double *d_x, double tol, int maxit, bda_result *res, bool adjoint) {
    require_factored(h);
    check_solve_args(tol, maxit);
    if (adjoint) ensure_adjoint<BS>(h);
    gather_vec2(h, d_b, h.bp.p, d_x, h.k64.x.p);
    OpSet<double> op = adjoint ? op_adj64(h) : op_fwd64(h);
    RunOut o = run_bicgstab<BS, double, double>(h, op, h.k64, h.bp.p, false, tol, maxit, h.abs_tol != 0);
    scatter_vec(h, h.k64.x.p, d_x);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    bda_result r;
    fill_result(o, &r);
    if (res) *res = r;
    log_result(adjoint ? "adjoint solve" : "solve", r);
    h.last_solve_valid = true;
    h.last_kind = 0;
}

// precision form: FP64 refinement around FP32-value inner solves.  The outer
// loop runs on the device: the FP64 residual's epilogue (FIN_REFINE) makes
// the stop decisions and sets the next inner tolerance, k_inner_start and
// k_refine_step link the inner runs to the outer state.  The host waits once
// per outer step, after the residual, to learn whether another inner run is
// needed.
template <int BS, typename TX>
bool mixed_inner(H &h, Krylov<TX> &k, int inner_maxit, bool adjoint) {
    OpSet<float> op = adjoint ? op_adj32(h) : op_fwd32(h);
    const RunOut o = run_bicgstab<BS, float, TX>(h, op, k, h.ro.p, true, 0.0, inner_maxit, false, h.st.p + 1);
    hipLaunchKernelGGL((k_add_correction<TX>), dim3(blocks_for((int64_t)h.N(), kVecThreads)), dim3(kVecThreads), 0,
                       h.stream, (int64_t)h.N(), h.xo.p, (const TX *)k.x.p, (const SolverState *)h.st.p);
    check_launch();
    hipLaunchKernelGGL(k_refine_step, dim3(1), dim3(1), 0, h.stream, h.st.p + 1, (const SolverState *)h.st.p);
    check_launch();
    return o.aborted;
}

// adjoint: the same refinement for A^T lambda = w, with the transposed
// operators (FP64 A^T for the residual, FP32 A^T, U^T, L^T for the inner runs).
template <int BS>
void solve_mixed_dev(H &h, const double *d_b, double *d_x, double tol, int max_outer, double inner_tol,
                     int inner_maxit, int vec32, bda_result *res, bool adjoint = false) {
    require_factored(h);
    check_solve_args(tol, inner_maxit);
    check_solve_args(inner_tol, 0);
    if (max_outer < 0) throw std::runtime_error("max_outer must be >= 0");
    if (adjoint) ensure_adjoint32<BS>(h);
    else ensure_mixed<BS>(h);
    const DevMat &M = adjoint ? h.At : h.A;
    const double *Mval = adjoint ? h.At64.p : h.A64.p;
    if (vec32) h.k32.ensure(h.Npad());
    const auto t0 = std::chrono::steady_clock::now();
    gather_vec2(h, d_b, h.bp.p, d_x, h.xo.p);
    SolverState os = new_state(h, 1, tol, 0, h.abs_tol != 0);
    os.max_outer = max_outer;
    os.inner_tol = inner_tol;
    os.eta = h.refine_eta;
    os.adaptive = h.refine_adaptive;
    hipLaunchKernelGGL(k_set_state, dim3(1), dim3(1), 0, h.stream, h.st.p + 1, os);
    check_launch();
    bool aborted = false;
    for (int outer = 0;; ++outer) {
        // FP64 true residual ro = b - A x and the outer step's decisions; the
        // first one is b without the product when x = 0 and A is finite
        if (outer == 0) {
            hipLaunchKernelGGL((k_mark_nonzero<double>), dim3(vec_blocks(h.N())), dim3(kVecThreads), 0, h.stream,
                               (int64_t)h.N(), (const double *)h.xo.p, (const int32_t *)h.d_nonfinite.p, h.st.p + 1);
            check_launch();
            launch_stream<BS, SM_SUB, SR_SELF, double, double, true>(h, M, Mval, -1, h.xo.p, h.bp.p, h.ro.p,
                                                                     nullptr, nullptr, red(h, 1, FIN_REFINE), true);
        } else {
            launch_stream<BS, SM_SUB, SR_SELF, double, double>(h, M, Mval, -1, h.xo.p, h.bp.p, h.ro.p, nullptr,
                                                               nullptr, red(h, 1, FIN_REFINE), true);
        }
        if (outer == max_outer) break;
        // After the first residual the host does not wait for the outer
        // step's decision: the refinement stops there only when b is zero
        // (or within the absolute tolerance), and then the inner run's start
        // finds the outer state stopped and the whole inner run, the
        // correction included, does nothing.
        if (outer > 0) {
            if (!wait_progress(h, 1, os.run_id, outer + 1, t0)) {
                aborted = true;
                break;
            }
            if (run_stopped(h, 1, os.run_id)) break;
        }
        const bool ab = vec32 ? mixed_inner<BS, float>(h, h.k32, inner_maxit, adjoint)
                              : mixed_inner<BS, double>(h, h.k64, inner_maxit, adjoint);
        if (ab) {
            aborted = true;
            break;
        }
        if (h.timeout_s > 0.0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > h.timeout_s) {
            aborted = true;
            break;
        }
    }
    scatter_vec(h, h.xo.p, d_x);
    HIP_CHECK(hipStreamSynchronize(h.stream));
    prof_collect(h);  // inner runs are not filtered: their no-op launches count too
    copy_sync(h, h.h_st, h.st.p + 1, sizeof(SolverState), hipMemcpyDeviceToHost);
    const SolverState s = *h.h_st;
    bda_result r;
    std::memset(&r, 0, sizeof(r));
    for (int i = 0; i < 4; ++i) r.norms[i] = s.norms[i];
    r.last_norm_idx = s.last_norm_idx;
    r.final_norm = s.last_norm;
    r.converged = s.converged;
    r.no_results = s.no_results;
    r.max_iter_reached = s.max_iter;
    r.breakdown = s.breakdown;
    r.nan_detected = s.nan_flag;
    r.aborted = (aborted || s.aborted) ? 1 : 0;
    r.outer_iterations = s.outer;
    r.inner_half_iterations = s.inner_h;
    r.half_iterations = s.inner_h;
    r.iterations = s.iters;
    r.rel_residual = s.nrm0 > 0 ? s.last_norm / s.nrm0 : 0.0;
    r.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (res) *res = r;
    log_result(adjoint ? "precision-form adjoint solve" : "precision-form solve", r);
    h.last_solve_valid = true;
    h.last_kind = 1;
}

// Component access in the three kernel forms of the solves:
//   PREC_F64      FP64 values and vectors (bda_solve, the adjoint)
//   PREC_F32      FP32 values and vectors (bda_solve_mixed with vec32 != 0)
//   PREC_F32_X64  FP32 values, FP64 vectors (bda_solve_mixed with vec32 = 0)
enum Prec : int { PREC_F64 = 0, PREC_F32 = 1, PREC_F32_X64 = 2 };

// y = A x (or A^T x) with the solve's SpMV kernel instance; redm = 1 or 3
// launches the instance with the fused reductions the solve uses
// (v = A phat with v.rt; t = A shat with t.s and t.t), summed in the
// solve's fixed order into dots[].  Vectors are device buffers of the
// handle in the colour-ordered numbering.
// Where the fused reductions of a component SpMV go: the partial sums
// (2 x cap doubles) and the scalar state that adds them up.  nullptr: the
// handle's own (partials and state slot 1).
struct RedScratch {
    double *partials;
    int32_t cap;
    SolverState *st;
};

// The scratch array of a device as reduction space: the state in its first
// 512 bytes, then 2 x cap partial sums.
RedScratch noalloc_scratch(int device) {
    void *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (device >= 0 && device < kMaxDevices) p = g_scratch_ptr[device];
    }
    if (!p) throw std::runtime_error("no scratch address for this device (bda_create sets it up)");
    RedScratch sc;
    sc.st = reinterpret_cast<SolverState *>(p);
    sc.partials = reinterpret_cast<double *>(static_cast<unsigned char *>(p) + kNoAllocStateBytes);
    sc.cap = (int32_t)((kNoAllocScratchBytes - kNoAllocStateBytes) / (2 * sizeof(double)));
    return sc;
}

template <int BS, typename TV, typename TX>
void spmv_comp(H &h, const TX *x, TX *y, int transpose, const TX *w, int redm, double *dots,
               const RedScratch *sc = nullptr) {
    const DevMat &m = transpose ? h.At : h.A;
    const TV *val;
    if constexpr (std::is_same<TV, float>::value) val = transpose ? h.At32.p : h.A32.p;
    else val = transpose ? h.At64.p : h.A64.p;
    SolverState *st = sc ? sc->st : h.st.p + 1;
    if (redm == 0) {
        launch_stream<BS, SM_SPMV, SR_NONE, TV, TX>(h, m, val, -1, x, nullptr, y, nullptr, nullptr,
                                                    red(h, 1, FIN_NONE), false);
    } else {
        RedArgs ra = red(h, 1, FIN_STORE);
        if (sc) {
            // a fresh state (no host flags: FIN_STORE raises none), written
            // by a kernel on the stream
            SolverState v;
            std::memset(&v, 0, sizeof(v));
            v.norm_slot = 1;
            hipLaunchKernelGGL(k_set_state, dim3(1), dim3(1), 0, h.stream, sc->st, v);
            check_launch();
            ra.partials = sc->partials;
            ra.cap = sc->cap;
            ra.st = sc->st;
        } else {
            init_state(h, 1, 0.0, 0, false);
        }
        if (redm == 1)
            launch_stream<BS, SM_SPMV, SR_W0, TV, TX>(h, m, val, -1, x, nullptr, y, nullptr, w, ra, true);
        else
            launch_stream<BS, SM_SPMV, SR_BOTH, TV, TX>(h, m, val, -1, x, nullptr, y, nullptr, w, ra, true);
    }
    HIP_CHECK(hipStreamSynchronize(h.stream));
    if (redm != 0 && dots) {
        SolverState s;
        copy_sync(h, &s, st, sizeof(SolverState), hipMemcpyDeviceToHost);
        dots[0] = s.tot[0];
        dots[1] = redm == 3 ? s.tot[1] : 0.0;
    }
}

template <int BS>
void spmv_comp_prec(H &h, int prec, const void *x, void *y, int transpose, const void *w, int redm, double *dots,
                    const RedScratch *sc = nullptr) {
    if (redm != 0 && redm != 1 && redm != 3) throw std::runtime_error("red must be 0, 1 or 3");
    if (redm != 0 && transpose) throw std::runtime_error("fused reductions: forward operator only");
    if (redm != 0 && !w) throw std::runtime_error("the fused reductions need w");
    if (prec == PREC_F32)
        spmv_comp<BS, float, float>(h, (const float *)x, (float *)y, transpose, (const float *)w, redm, dots, sc);
    else if (prec == PREC_F32_X64)
        spmv_comp<BS, float, double>(h, (const double *)x, (double *)y, transpose, (const double *)w, redm, dots,
                                     sc);
    else
        spmv_comp<BS, double, double>(h, (const double *)x, (double *)y, transpose, (const double *)w, redm, dots,
                                      sc);
}

template <int BS>
void comp_prepare(H &h, int prec, int transpose) {
    require_analysed(h);
    if (!h.have_values) throw std::runtime_error("no matrix values uploaded");
    if (prec != PREC_F64 && prec != PREC_F32 && prec != PREC_F32_X64)
        throw std::runtime_error("precision must be 0 (FP64), 1 (FP32 vectors) or 2 (FP64 vectors)");
    if (prec != PREC_F64) {
        ensure_mixed<BS>(h);  // the FP32 copies are made from the factored system
        if (transpose) ensure_adjoint32<BS>(h);
    } else if (transpose) {
        require_factored(h);
        ensure_adjoint<BS>(h);
    }
}

template <int BS>
void spmv_perm(H &h, int prec, const void *x, void *y, int transpose, const void *w, int redm, double *dots) {
    comp_prepare<BS>(h, prec, transpose);
    const size_t es = prec == PREC_F32 ? 4 : 8;
    copy_sync(h, h.vin.p, x, h.N() * es, hipMemcpyHostToDevice);
    if (w) copy_sync(h, h.vw.p, w, h.N() * es, hipMemcpyHostToDevice);
    spmv_comp_prec<BS>(h, prec, h.vin.p, h.vout.p, transpose, w ? h.vw.p : nullptr, redm, dots);
    copy_sync(h, y, h.vout.p, h.N() * es, hipMemcpyDeviceToHost);
}

// The same on the caller's device vectors.
template <int BS>
void spmv_dev(H &h, int prec, const void *d_x, void *d_y, int transpose, const void *d_w, int redm, double *dots) {
    comp_prepare<BS>(h, prec, transpose);
    spmv_comp_prec<BS>(h, prec, d_x, d_y, transpose, d_w, redm, dots);
}

// bda_spmv_dev without any device allocation: the handle's copies of the
// matrix must already exist and be current, and the fused reductions use
// the per-device scratch array instead of the handle's buffers.
template <int BS>
void spmv_dev_noalloc(H &h, int prec, const void *d_x, void *d_y, int transpose, const void *d_w, int redm,
                      double *dots) {
    require_analysed(h);
    if (!h.have_values) throw std::runtime_error("no matrix values uploaded");
    if (prec != PREC_F64 && prec != PREC_F32 && prec != PREC_F32_X64)
        throw std::runtime_error("precision must be 0 (FP64), 1 (FP32 vectors) or 2 (FP64 vectors)");
    if (prec != PREC_F64) {
        require_factored(h);
        if (h.mixed_gen != h.factor_gen)
            throw std::runtime_error("the FP32 copies of the matrix are not current: call bda_spmv_dev or a "
                                     "precision-form solve once after bda_factor");
    }
    if (transpose) {
        require_factored(h);
        if (h.adj_gen != h.factor_gen || (prec != PREC_F64 && h.adj32_gen != h.factor_gen))
            throw std::runtime_error("the transposed copies of the matrix are not current: call bda_spmv_dev "
                                     "with transpose once after bda_factor");
    }
    const DevMat &m = transpose ? h.At : h.A;
    const int ntiles = wide_tiles<SM_SPMV>() ? m.ntiles_wide : m.ntiles;
    const RedScratch sc = noalloc_scratch(h.device);
    if (redm != 0 && ntiles > sc.cap)
        throw std::runtime_error("this system's SpMV writes " + std::to_string(ntiles) +
                                 " partial sums per reduction; the no-allocation entry point holds at most " +
                                 std::to_string(sc.cap) + " (" + std::to_string(kNoAllocScratchBytes >> 10) +
                                 " KiB of scratch)");
    spmv_comp_prec<BS>(h, prec, d_x, d_y, transpose, d_w, redm, dots, &sc);
}

template <int BS>
void precond_perm(H &h, int prec, const void *p, void *z, int transpose) {
    require_factored(h);
    comp_prepare<BS>(h, prec, transpose);
    const size_t es = prec == PREC_F32 ? 4 : 8;
    copy_sync(h, h.vin.p, p, h.N() * es, hipMemcpyHostToDevice);
    // the intermediate of the two sweeps goes to vw
    if (prec == PREC_F32) {
        OpSet<float> op = transpose ? op_adj32(h) : op_fwd32(h);
        precond<BS, float, float>(h, op, (const float *)h.vin.p, (float *)h.vw.p, (float *)h.vout.p, false);
    } else if (prec == PREC_F32_X64) {
        OpSet<float> op = transpose ? op_adj32(h) : op_fwd32(h);
        precond<BS, float, double>(h, op, h.vin.p, h.vw.p, h.vout.p, false);
    } else {
        OpSet<double> op = transpose ? op_adj64(h) : op_fwd64(h);
        precond<BS, double, double>(h, op, h.vin.p, h.vw.p, h.vout.p, false);
    }
    HIP_CHECK(hipStreamSynchronize(h.stream));
    prof_collect(h);
    copy_sync(h, z, h.vout.p, h.N() * es, hipMemcpyDeviceToHost);
}

template <int BS>
void grad_dev(H &h, const double *d_lam, const double *d_x, double *d_grad) {
    require_analysed(h);
    h.gl.ensure(h.Npad());
    h.gx.ensure(h.Npad());
    gather_vec(h, d_lam, h.gl.p);
    gather_vec(h, d_x, h.gx.p);
    const int64_t work = h.A.nnzb * BS * BS;
    hipLaunchKernelGGL((k_grad_blocks<BS>), dim3(blocks_for(work, 256)), dim3(256), 0, h.stream, h.A.nnzb,
                       h.d_blkrow.p, h.A.colidx.p, h.A.src.p, h.gl.p, h.gx.p, d_grad);
    check_launch();
    HIP_CHECK(hipStreamSynchronize(h.stream));
}

// Average duration of the SpMV launch the solve uses for v = A phat (with
// the fused v.rt partial sums), from reps back-to-back launches bracketed by
// two HIP events on the handle's stream.
template <int BS>
double time_spmv(H &h, int reps, int fp32) {
    require_analysed(h);
    if (fp32) ensure_mixed<BS>(h);
    if (fp32) h.k32.ensure(h.Npad());
    init_state(h, 1, 0.0, 0, false);
    RedArgs ra = red(h, 1, FIN_NONE);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    for (int rep = -1; rep < reps; ++rep) {
        if (rep == 0) HIP_CHECK(hipEventRecord(e0, h.stream));
        // no finalize: only the k_stream launch is timed
        if (fp32) {
            launch_stream_timed<BS, float, float>(h, h.A32.p, h.k32.phat.p, h.k32.v.p, h.k32.rt.p, ra);
        } else {
            launch_stream_timed<BS, double, double>(h, h.A64.p, h.k64.phat.p, h.k64.v.p, h.k64.rt.p, ra);
        }
    }
    HIP_CHECK(hipEventRecord(e1, h.stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return reps > 0 ? ms / reps : 0.0;
}

// Average duration of one preconditioner application z = M^-1 p (every
// launch of both sweeps) in the form the solves run, from reps back-to-back
// applications bracketed by two HIP events on the handle's stream.
template <int BS>
double time_precond(H &h, int reps, int fp32) {
    require_factored(h);
    if (fp32) {
        ensure_mixed<BS>(h);
        h.k32.ensure(h.Npad());
    }
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    for (int rep = -1; rep < reps; ++rep) {
        if (rep == 0) HIP_CHECK(hipEventRecord(e0, h.stream));
        if (fp32) precond<BS, float, float>(h, op_fwd32(h), h.k32.p.p, h.k32.y.p, h.k32.phat.p, false);
        else precond<BS, double, double>(h, op_fwd64(h), h.k64.p.p, h.k64.y.p, h.k64.phat.p, false);
    }
    HIP_CHECK(hipEventRecord(e1, h.stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return reps > 0 ? ms / reps : 0.0;
}

template <typename F>
int guarded(bda_handle h, F &&f) {
    if (!h) return 2;
    try {
        h->err.clear();
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && dev != h->device) HIP_CHECK(hipSetDevice(h->device));
        f();
        return 0;
    } catch (const std::exception &e) {
        h->err = e.what();
        bda::info(1, "ERROR: %s", e.what());
        return 1;
    } catch (...) {
        h->err = "unknown error";
        bda::info(1, "ERROR: unknown error");
        return 1;
    }
}

}  // namespace
}  // namespace bda

using namespace bda;

#define BS_CALL(h, fn, ...)                   \
    do {                                      \
        if ((h)->bs == 3) fn<3>(__VA_ARGS__); \
        else fn<4>(__VA_ARGS__);              \
    } while (0)

extern "C" {

int bda_create(int32_t bs, int32_t n, int32_t nnzb_max, bda_handle *out) {
    if (!out) return 2;
    *out = nullptr;
    if ((bs != 3 && bs != 4) || n <= 0 || nnzb_max < n) return 2;
    bda_handle h = new (std::nothrow) bda_handle_s();
    if (!h) return 1;
    h->bs = bs;
    h->n = n;
    h->nnzb_max = nnzb_max;
    try {
        HIP_CHECK(hipGetDevice(&h->device));
        HIP_CHECK(hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device));
        HIP_CHECK(hipDeviceGetAttribute(&h->wall_khz, hipDeviceAttributeWallClockRate, h->device));
        noalloc_scratch_setup(h->device);
        if (h->wall_khz <= 0) throw std::runtime_error("device reports no wall clock rate");
        // a blocking stream: it orders itself after work the caller queued on
        // the legacy default stream (torch's default), so device inputs
        // produced there are complete before the handle reads them
        HIP_CHECK(hipStreamCreate(&h->stream));
        h->st.ensure(kIdleState + 1);  // DBuf zero-fills: the idle slot never has done set
        HIP_CHECK(hipHostMalloc((void **)&h->h_st, sizeof(SolverState), hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc((void **)&h->h_done, 2 * sizeof(int32_t), hipHostMallocMapped));
        h->h_done[0] = h->h_done[1] = 0;
        HIP_CHECK(hipHostGetDevicePointer((void **)&h->d_done, h->h_done, 0));
        HIP_CHECK(hipHostMalloc((void **)&h->h_prog, 2 * sizeof(uint64_t), hipHostMallocMapped));
        h->h_prog[0] = h->h_prog[1] = 0;
        HIP_CHECK(hipHostGetDevicePointer((void **)&h->d_prog, h->h_prog, 0));
        h->d_orig.ensure((size_t)nnzb_max * bs * bs);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ERROR: bda_create: %s\n", e.what());
        bda_destroy(h);
        return 1;
    }
    bda::info(1, "INFO: bda_create: %dx%d blocks, %d block rows, room for %d blocks; device %d, %d compute units", bs,
              bs, n, nnzb_max, h->device, h->cus);
    *out = h;
    return 0;
}

int bda_destroy(bda_handle h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (DevMat *m : {&h->A, &h->L, &h->U, &h->At, &h->Ut, &h->Lt}) m->release();
    for (DBuf<int32_t> *b : {&h->d_perm, &h->d_diag, &h->d_blkrow, &h->d_bad}) b->release();
    for (DBuf<double> *b : {&h->d_orig, &h->A64, &h->L64, &h->U64, &h->LU, &h->Dinv, &h->DinvT, &h->At64, &h->Ut64,
                            &h->Lt64, &h->bp, &h->xo, &h->ro, &h->vin, &h->vout, &h->vw, &h->gl, &h->gx, &h->ggrad,
                            &h->partials})
        b->release();
    for (DBuf<float> *b : {&h->A32, &h->L32, &h->U32, &h->At32, &h->Ut32, &h->Lt32}) b->release();
    h->k64.release();
    h->k32.release();
    h->st.release();
    if (h->h_st) (void)hipHostFree(h->h_st);
    if (h->h_done) (void)hipHostFree(h->h_done);
    if (h->h_prog) (void)hipHostFree(h->h_prog);
    for (auto &pr : h->ev_pool) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

const char *bda_last_error(bda_handle h) { return h ? h->err.c_str() : "null handle"; }

int bda_query(bda_handle h, bda_limits *lim) {
    return guarded(h, [&] {
        if (!lim) throw std::runtime_error("null limits");
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, h->device));
        lim->max_block_size = 4;
        lim->max_blocks_per_row = tile_cap(h->bs);
        lim->max_iterations = 32767;
        lim->rows_per_tile = kTileRows;
        lim->device_memory_bytes = (int64_t)prop.totalGlobalMem;
        lim->compute_units = prop.multiProcessorCount;
        lim->wavefront_size = prop.warpSize;
    });
}

int bda_upload_bcsr(bda_handle h, const int32_t *rowptr, const int32_t *colidx, const double *values) {
    return guarded(h, [&] {
        if (!rowptr || !colidx || !values) throw std::runtime_error("null array");
        // the pattern is checked before anything is copied: rowptr
        // non-decreasing from 0, within nnzb_max, and every column in range
        if (rowptr[0] != 0) throw std::runtime_error("rowptr[0] must be 0");
        for (int i = 0; i < h->n; ++i)
            if (rowptr[i + 1] < rowptr[i] || rowptr[i + 1] > h->nnzb_max)
                throw std::runtime_error("rowptr[" + std::to_string(i + 1) + "] = " + std::to_string(rowptr[i + 1]) +
                                         " is decreasing or exceeds nnzb_max " + std::to_string(h->nnzb_max));
        const int64_t nnzb = rowptr[h->n];
        for (int64_t q = 0; q < nnzb; ++q)
            if (colidx[q] < 0 || colidx[q] >= h->n)
                throw std::runtime_error("colidx[" + std::to_string(q) + "] = " + std::to_string(colidx[q]) +
                                         " out of range");
        h->pattern.bs = h->bs;
        h->pattern.n = h->n;
        h->pattern.rowptr.assign(rowptr, rowptr + h->n + 1);
        h->pattern.colidx.assign(colidx, colidx + nnzb);
        h->have_pattern = true;
        h->analysed = false;
        h->factored = false;
        copy_sync(*h, h->d_orig.p, values, (size_t)nnzb * h->bs * h->bs * 8, hipMemcpyHostToDevice);
        h->have_values = true;
        h->values_gen++;
    });
}

int bda_update_values(bda_handle h, const double *values) {
    return guarded(h, [&] {
        if (!h->have_pattern) throw std::runtime_error("bda_upload_bcsr has not been called");
        const int64_t nnzb = h->pattern.nnzb();
        copy_sync(*h, h->d_orig.p, values, (size_t)nnzb * h->bs * h->bs * 8, hipMemcpyHostToDevice);
        h->have_values = true;
        h->factored = false;
        h->values_gen++;
        if (h->analysed) BS_CALL(h, permute_values, *h);
    });
}

int bda_update_values_dev(bda_handle h, const double *d_values) {
    return guarded(h, [&] {
        if (!h->have_pattern) throw std::runtime_error("bda_upload_bcsr has not been called");
        const int64_t nnzb = h->pattern.nnzb();
        copy_sync(*h, h->d_orig.p, d_values, (size_t)nnzb * h->bs * h->bs * 8, hipMemcpyDeviceToDevice);
        h->have_values = true;
        h->factored = false;
        h->values_gen++;
        if (h->analysed) BS_CALL(h, permute_values, *h);
    });
}

int bda_upload_factored(bda_handle h, int32_t ncolours, const int32_t *colour_ptr, const int32_t *rowptr,
                        const int32_t *colidx, const double *values, const double *lu_values, const double *dinv) {
    return guarded(h, [&] {
        if (!colour_ptr || !rowptr || !colidx || !values || !lu_values || !dinv)
            throw std::runtime_error("null array");
        if (ncolours < 1 || ncolours > h->n) throw std::runtime_error("ncolours must be in [1, n]");
        if (rowptr[0] != 0) throw std::runtime_error("rowptr[0] must be 0");
        for (int i = 0; i < h->n; ++i)
            if (rowptr[i + 1] < rowptr[i] || rowptr[i + 1] > h->nnzb_max)
                throw std::runtime_error("rowptr[" + std::to_string(i + 1) + "] is decreasing or exceeds nnzb_max");
        const int64_t nnzb = rowptr[h->n];
        Bcsr m;
        m.bs = h->bs;
        m.n = h->n;
        m.rowptr.assign(rowptr, rowptr + h->n + 1);
        m.colidx.assign(colidx, colidx + nnzb);
        const std::vector<int32_t> cptr(colour_ptr, colour_ptr + ncolours + 1);
        Analysis an;
        analyse_given_order(m, cptr, an);   // validates the pattern and the colouring before any copy
        h->pattern = m;
        h->have_pattern = true;
        h->an = std::move(an);
        copy_sync(*h, h->d_orig.p, values, (size_t)nnzb * h->bs * h->bs * 8, hipMemcpyHostToDevice);
        h->have_values = true;
        h->values_gen++;
        finish_analyse(*h);   // identity permutation: A's values are copied as given
        BS_CALL(h, load_factors, *h, lu_values, dinv);
    });
}

int bda_values_generation(bda_handle h, uint64_t *gen) {
    return guarded(h, [&] {
        if (!gen) throw std::runtime_error("null pointer");
        *gen = h->values_gen;
    });
}

int bda_analyse(bda_handle h, int32_t level_scheduling) {
    return guarded(h, [&] { do_analyse(*h, level_scheduling != 0); });
}

int bda_num_colours(bda_handle h, int32_t *nc) {
    return guarded(h, [&] {
        require_analysed(*h);
        *nc = h->an.ncolours;
    });
}

int bda_get_colouring(bda_handle h, int32_t *colour, int32_t *perm, int32_t *colour_ptr) {
    return guarded(h, [&] {
        require_analysed(*h);
        if (colour) std::memcpy(colour, h->an.colour.data(), h->n * 4);
        if (perm) std::memcpy(perm, h->an.perm.data(), h->n * 4);
        if (colour_ptr) std::memcpy(colour_ptr, h->an.colour_ptr.data(), (h->an.ncolours + 1) * 4);
    });
}

int bda_factor(bda_handle h) {
    return guarded(h, [&] { BS_CALL(h, do_factor, *h); });
}

int bda_prepare_mixed(bda_handle h) {
    return guarded(h, [&] { BS_CALL(h, ensure_mixed, *h); });
}

int bda_set_options(bda_handle h, double timeout_s, int32_t abs_tol, int32_t poll_iters) {
    return guarded(h, [&] {
        h->timeout_s = timeout_s > 0 ? timeout_s : 0.0;
        h->abs_tol = abs_tol != 0;
        h->poll_iters = poll_iters > 0 ? poll_iters : 1;
    });
}

int bda_set_refinement(bda_handle h, int32_t adaptive, double eta) {
    return guarded(h, [&] {
        if (!(eta > 0.0 && eta <= 1.0)) throw std::runtime_error("eta must be in (0, 1]");
        h->refine_adaptive = adaptive != 0;
        h->refine_eta = eta;
    });
}

int bda_solve_dev(bda_handle h, const double *d_b, double *d_x, double tol, int32_t maxit, bda_result *res) {
    return guarded(h, [&] { BS_CALL(h, solve_dev, *h, d_b, d_x, tol, maxit, res, false); });
}

int bda_solve(bda_handle h, const double *b, double *x, double tol, int32_t maxit, bda_result *res) {
    return guarded(h, [&] {
        if (!b || !x) throw std::runtime_error("null vector");
        require_factored(*h);
        copy_sync(*h, h->vin.p, b, h->N() * 8, hipMemcpyHostToDevice);
        copy_sync(*h, h->vout.p, x, h->N() * 8, hipMemcpyHostToDevice);
        BS_CALL(h, solve_dev, *h, h->vin.p, h->vout.p, tol, maxit, res, false);
        copy_sync(*h, x, h->vout.p, h->N() * 8, hipMemcpyDeviceToHost);
    });
}

int bda_get_residual(bda_handle h, double *r) {
    return guarded(h, [&] {
        if (!h->last_solve_valid) throw std::runtime_error("no solve has run");
        const double *src;
        if (h->last_kind == 1) {
            src = h->ro.p;
        } else {
            copy_sync(*h, h->h_st, h->st.p, sizeof(SolverState), hipMemcpyDeviceToHost);
            const int lh = h->h_st->last_half;
            src = lh == 1 ? h->k64.s.p : lh == 2 ? h->k64.r.p : h->k64.rt.p;
        }
        scatter_vec(*h, src, h->vout.p);
        HIP_CHECK(hipStreamSynchronize(h->stream));
        copy_sync(*h, r, h->vout.p, h->N() * 8, hipMemcpyDeviceToHost);
    });
}

int bda_solve_mixed_dev(bda_handle h, const double *d_b, double *d_x, double tol, int32_t max_outer,
                        double inner_tol, int32_t inner_maxit, int32_t vec32, bda_result *res) {
    return guarded(h, [&] {
        BS_CALL(h, solve_mixed_dev, *h, d_b, d_x, tol, max_outer, inner_tol, inner_maxit, vec32, res);
    });
}

int bda_solve_mixed(bda_handle h, const double *b, double *x, double tol, int32_t max_outer, double inner_tol,
                    int32_t inner_maxit, int32_t vec32, bda_result *res) {
    return guarded(h, [&] {
        if (!b || !x) throw std::runtime_error("null vector");
        require_factored(*h);
        copy_sync(*h, h->vin.p, b, h->N() * 8, hipMemcpyHostToDevice);
        copy_sync(*h, h->vout.p, x, h->N() * 8, hipMemcpyHostToDevice);
        BS_CALL(h, solve_mixed_dev, *h, h->vin.p, h->vout.p, tol, max_outer, inner_tol, inner_maxit, vec32, res);
        copy_sync(*h, x, h->vout.p, h->N() * 8, hipMemcpyDeviceToHost);
    });
}

int bda_adjoint_solve_dev(bda_handle h, const double *d_w, double *d_lambda, double tol, int32_t maxit,
                          bda_result *res) {
    return guarded(h, [&] { BS_CALL(h, solve_dev, *h, d_w, d_lambda, tol, maxit, res, true); });
}

int bda_adjoint_solve(bda_handle h, const double *w, double *lambda, double tol, int32_t maxit, bda_result *res) {
    return guarded(h, [&] {
        if (!w || !lambda) throw std::runtime_error("null vector");
        require_factored(*h);
        copy_sync(*h, h->vin.p, w, h->N() * 8, hipMemcpyHostToDevice);
        copy_sync(*h, h->vout.p, lambda, h->N() * 8, hipMemcpyHostToDevice);
        BS_CALL(h, solve_dev, *h, h->vin.p, h->vout.p, tol, maxit, res, true);
        copy_sync(*h, lambda, h->vout.p, h->N() * 8, hipMemcpyDeviceToHost);
    });
}

int bda_adjoint_solve_mixed_dev(bda_handle h, const double *d_w, double *d_lambda, double tol, int32_t max_outer,
                                double inner_tol, int32_t inner_maxit, int32_t vec32, bda_result *res) {
    return guarded(h, [&] {
        BS_CALL(h, solve_mixed_dev, *h, d_w, d_lambda, tol, max_outer, inner_tol, inner_maxit, vec32, res, true);
    });
}

int bda_adjoint_solve_mixed(bda_handle h, const double *w, double *lambda, double tol, int32_t max_outer,
                            double inner_tol, int32_t inner_maxit, int32_t vec32, bda_result *res) {
    return guarded(h, [&] {
        if (!w || !lambda) throw std::runtime_error("null vector");
        require_factored(*h);
        copy_sync(*h, h->vin.p, w, h->N() * 8, hipMemcpyHostToDevice);
        copy_sync(*h, h->vout.p, lambda, h->N() * 8, hipMemcpyHostToDevice);
        BS_CALL(h, solve_mixed_dev, *h, h->vin.p, h->vout.p, tol, max_outer, inner_tol, inner_maxit, vec32, res,
                true);
        copy_sync(*h, lambda, h->vout.p, h->N() * 8, hipMemcpyDeviceToHost);
    });
}

int bda_grad_values_dev(bda_handle h, const double *d_lambda, const double *d_x, double *d_grad) {
    return guarded(h, [&] { BS_CALL(h, grad_dev, *h, d_lambda, d_x, d_grad); });
}

int bda_grad_values(bda_handle h, const double *lambda, const double *x, double *grad) {
    return guarded(h, [&] {
        if (!lambda || !x || !grad) throw std::runtime_error("null vector");
        require_analysed(*h);
        const size_t nv = (size_t)h->A.nnzb * h->bs * h->bs;
        h->ggrad.ensure(nv);
        copy_sync(*h, h->vin.p, lambda, h->N() * 8, hipMemcpyHostToDevice);
        copy_sync(*h, h->vout.p, x, h->N() * 8, hipMemcpyHostToDevice);
        BS_CALL(h, grad_dev, *h, h->vin.p, h->vout.p, h->ggrad.p);
        copy_sync(*h, grad, h->ggrad.p, nv * 8, hipMemcpyDeviceToHost);
    });
}

int bda_spmv_perm(bda_handle h, const double *x, double *y, int32_t transpose) {
    return guarded(h, [&] {
        if (!x || !y) throw std::runtime_error("null vector");
        BS_CALL(h, spmv_perm, *h, PREC_F64, x, y, transpose, nullptr, 0, nullptr);
    });
}

int bda_precond_apply_perm(bda_handle h, const double *p, double *z, int32_t transpose) {
    return guarded(h, [&] {
        if (!p || !z) throw std::runtime_error("null vector");
        BS_CALL(h, precond_perm, *h, PREC_F64, p, z, transpose);
    });
}

int bda_spmv_perm_ex(bda_handle h, int32_t prec, const void *x, void *y, int32_t transpose, const void *w,
                     int32_t red, double dots[2]) {
    return guarded(h, [&] {
        if (!x || !y) throw std::runtime_error("null vector");
        BS_CALL(h, spmv_perm, *h, prec, x, y, transpose, w, red, dots);
    });
}

int bda_spmv_dev(bda_handle h, int32_t prec, const void *d_x, void *d_y, int32_t transpose, const void *d_w,
                 int32_t red, double dots[2]) {
    return guarded(h, [&] {
        if (!d_x || !d_y) throw std::runtime_error("null vector");
        BS_CALL(h, spmv_dev, *h, prec, d_x, d_y, transpose, d_w, red, dots);
    });
}

int bda_spmv_dev_noalloc(bda_handle h, int32_t prec, const void *d_x, void *d_y, int32_t transpose,
                         const void *d_w, int32_t red, double dots[2]) {
    return guarded(h, [&] {
        if (!d_x || !d_y) throw std::runtime_error("null vector");
        BS_CALL(h, spmv_dev_noalloc, *h, prec, d_x, d_y, transpose, d_w, red, dots);
    });
}

int bda_precond_apply_perm_ex(bda_handle h, int32_t prec, const void *p, void *z, int32_t transpose) {
    return guarded(h, [&] {
        if (!p || !z) throw std::runtime_error("null vector");
        BS_CALL(h, precond_perm, *h, prec, p, z, transpose);
    });
}

int bda_precond_launches(bda_handle h, int32_t transpose, int32_t *launches) {
    return guarded(h, [&] {
        require_analysed(*h);
        if (!launches) throw std::runtime_error("null output");
        *launches = transpose ? precond_launches(*h, op_adj64(*h)) : precond_launches(*h, op_fwd64(*h));
    });
}

int bda_factor_sizes(bda_handle h, int32_t *nl, int32_t *nu) {
    return guarded(h, [&] {
        require_analysed(*h);
        if (nl) *nl = (int32_t)h->an.L.nnzb();
        if (nu) *nu = (int32_t)h->an.U.nnzb();
    });
}

int bda_get_factors(bda_handle h, int32_t *l_rowptr, int32_t *l_colidx, double *l_values, int32_t *u_rowptr,
                    int32_t *u_colidx, double *u_values, double *dinv) {
    return guarded(h, [&] {
        require_factored(*h);
        const Analysis &an = h->an;
        const size_t bb = (size_t)h->bs * h->bs;
        if (l_rowptr) std::memcpy(l_rowptr, an.L.rowptr.data(), (h->n + 1) * 4);
        if (l_colidx) std::memcpy(l_colidx, an.L.colidx.data(), an.L.colidx.size() * 4);
        if (u_rowptr) std::memcpy(u_rowptr, an.U.rowptr.data(), (h->n + 1) * 4);
        if (u_colidx) std::memcpy(u_colidx, an.U.colidx.data(), an.U.colidx.size() * 4);
        if (l_values && an.L.nnzb() > 0)
            copy_sync(*h, l_values, h->L64.p, an.L.nnzb() * bb * 8, hipMemcpyDeviceToHost);
        if (u_values && an.U.nnzb() > 0)
            copy_sync(*h, u_values, h->U64.p, an.U.nnzb() * bb * 8, hipMemcpyDeviceToHost);
        if (dinv) copy_sync(*h, dinv, h->Dinv.p, (size_t)h->n * bb * 8, hipMemcpyDeviceToHost);
    });
}

// Reference wire format of the handle's analysed and factored system
// (host arrays; the device copies of A, L, U and D^-1 are read back).
namespace {
struct WireInputs {
    std::vector<double> a, l, u, dinv;
};
void wire_inputs(H &h, WireInputs &w) {
    require_factored(h);
    const Analysis &an = h.an;
    const size_t bb = (size_t)h.bs * h.bs;
    w.a.resize(an.A.nnzb() * bb);
    w.l.resize(std::max<size_t>(an.L.nnzb() * bb, 1));
    w.u.resize(std::max<size_t>(an.U.nnzb() * bb, 1));
    w.dinv.resize((size_t)h.n * bb);
    copy_sync(h, w.a.data(), h.A64.p, w.a.size() * 8, hipMemcpyDeviceToHost);
    if (an.L.nnzb()) copy_sync(h, w.l.data(), h.L64.p, an.L.nnzb() * bb * 8, hipMemcpyDeviceToHost);
    if (an.U.nnzb()) copy_sync(h, w.u.data(), h.U64.p, an.U.nnzb() * bb * 8, hipMemcpyDeviceToHost);
    copy_sync(h, w.dinv.data(), h.Dinv.p, w.dinv.size() * 8, hipMemcpyDeviceToHost);
}
std::vector<double> to_colour_order(const H &h, const double *v) {
    std::vector<double> out((size_t)h.n * h.bs, 0.0);
    if (!v) return out;
    for (int k = 0; k < h.n; ++k)
        for (int a = 0; a < h.bs; ++a) out[(size_t)k * h.bs + a] = v[(size_t)h.an.perm[k] * h.bs + a];
    return out;
}
}  // namespace

int bda_wire_export_sizes(bda_handle h, bda_wire_sizes *ws) {
    return guarded(h, [&] {
        require_factored(*h);
        const Analysis &an = h->an;
        if (bda_wire_build_sizes(h->bs, h->n, an.colour_ptr.data(), an.ncolours, an.A.rowptr.data(),
                                 an.A.colidx.data(), an.L.rowptr.data(), an.L.colidx.data(), an.U.rowptr.data(),
                                 an.U.colidx.data(), ws) != 0)
            throw std::runtime_error("wire format sizes failed");
    });
}

int bda_wire_export(bda_handle h, const double *b, const double *x0, void *vector_pointers[21]) {
    return guarded(h, [&] {
        WireInputs w;
        wire_inputs(*h, w);
        const Analysis &an = h->an;
        const std::vector<double> bp = to_colour_order(*h, b), xp = to_colour_order(*h, x0);
        if (bda_wire_build(h->bs, h->n, an.colour_ptr.data(), an.ncolours, an.A.rowptr.data(), an.A.colidx.data(),
                           w.a.data(), an.L.rowptr.data(), an.L.colidx.data(), w.l.data(), an.U.rowptr.data(),
                           an.U.colidx.data(), w.u.data(), w.dinv.data(), bp.data(), xp.data(),
                           vector_pointers) != 0)
            throw std::runtime_error("wire format export failed");
    });
}

int bda_set_history(bda_handle h, int32_t capacity) {
    return guarded(h, [&] {
        if (capacity < 0) throw std::runtime_error("negative history capacity");
        h->hist.release();
        h->hist_cap = 0;
        if (capacity > 0) {
            h->hist.ensure((size_t)capacity);
            h->hist_cap = capacity;
        }
    });
}

int bda_get_history(bda_handle h, double *norms, int32_t capacity, int32_t *count) {
    return guarded(h, [&] {
        if (!h->last_solve_valid) throw std::runtime_error("no solve has run");
        HIP_CHECK(hipStreamSynchronize(h->stream));
        SolverState s;
        copy_sync(*h, &s, h->st.p, sizeof(SolverState), hipMemcpyDeviceToHost);
        const int32_t avail = std::min(s.nhist, h->hist_cap);
        const int32_t n = std::max(0, std::min(avail, capacity));
        if (n > 0 && norms) copy_sync(*h, norms, h->hist.p, (size_t)n * 8, hipMemcpyDeviceToHost);
        if (count) *count = s.nhist;
    });
}

int bda_set_profiling(bda_handle h, int32_t on) {
    return guarded(h, [&] {
        h->profiling = on != 0;
        for (int k = 0; k < PK_N; ++k) {
            h->prof_count[k] = 0;
            h->prof_ms[k] = 0.0;
            h->prof_times[k].clear();
        }
        h->ev_used = 0;
    });
}

int bda_get_kernel_stats(bda_handle h, int32_t kind, int64_t *launches, double *total_ms) {
    return guarded(h, [&] {
        if (kind < 0 || kind >= PK_N) throw std::runtime_error("bad kernel kind");
        if (launches) *launches = h->prof_count[kind];
        if (total_ms) *total_ms = h->prof_ms[kind];
    });
}

int bda_get_kernel_times(bda_handle h, int32_t kind, float *ms, int32_t capacity, int32_t *count) {
    return guarded(h, [&] {
        if (kind < 0 || kind >= PK_N) throw std::runtime_error("bad kernel kind");
        if (capacity < 0) throw std::runtime_error("negative capacity");
        const std::vector<float> &v = h->prof_times[kind];
        const int32_t n = (int32_t)std::min<size_t>(v.size(), (size_t)capacity);
        if (n > 0 && ms) std::memcpy(ms, v.data(), (size_t)n * sizeof(float));
        if (count) *count = (int32_t)v.size();
    });
}

int bda_time_spmv(bda_handle h, int32_t reps, int32_t fp32, double *avg_ms) {
    return guarded(h, [&] {
        double v = 0.0;
        if (h->bs == 3) v = time_spmv<3>(*h, reps, fp32);
        else v = time_spmv<4>(*h, reps, fp32);
        if (avg_ms) *avg_ms = v;
    });
}

int bda_time_precond(bda_handle h, int32_t reps, int32_t fp32, double *avg_ms) {
    return guarded(h, [&] {
        double v = 0.0;
        if (h->bs == 3) v = time_precond<3>(*h, reps, fp32);
        else v = time_precond<4>(*h, reps, fp32);
        if (avg_ms) *avg_ms = v;
    });
}

int bda_measure_bandwidth(bda_handle h, int64_t bytes, int32_t reps, double *read_gbs, double *copy_gbs) {
    return guarded(h, [&] {
        if (bytes < (1 << 20) || reps < 1) throw std::runtime_error("bytes >= 1 MiB and reps >= 1");
        constexpr int64_t chunk = (int64_t)kBwThreads * kBwUnroll;   // 16-byte words per workgroup
        const int64_t n4 = bytes / 16 / chunk * chunk;
        DBuf<double2> a, b;
        DBuf<double> part;
        a.ensure((size_t)n4);
        b.ensure((size_t)n4);
        const int grid = (int)(n4 / chunk);
        part.ensure((size_t)grid * (kBwThreads / 64));
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        auto timed = [&](bool copy) {
            for (int rep = -1; rep < reps; ++rep) {
                if (rep == 0) HIP_CHECK(hipEventRecord(e0, h->stream));
                if (copy)
                    hipLaunchKernelGGL(k_bw_copy, dim3(grid), dim3(kBwThreads), 0, h->stream, (const double2 *)a.p,
                                       b.p);
                else
                    hipLaunchKernelGGL(k_bw_read, dim3(grid), dim3(kBwThreads), 0, h->stream, (const double2 *)a.p,
                                       part.p);
                check_launch();
            }
            HIP_CHECK(hipEventRecord(e1, h->stream));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / reps;
        };
        const double ms_r = timed(false), ms_c = timed(true);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (read_gbs) *read_gbs = (double)n4 * 16 / (ms_r * 1e-3) / 1e9;
        if (copy_gbs) *copy_gbs = 2.0 * (double)n4 * 16 / (ms_c * 1e-3) / 1e9;
        a.release();
        b.release();
        part.release();
    });
}

int bda_algorithmic_bytes(bda_handle h, int32_t kind, int32_t fp32, double *bytes) {
    return guarded(h, [&] {
        require_analysed(*h);
        const double bb = (double)h->bs * h->bs;
        const double sv = fp32 ? 4.0 : 8.0, sx = fp32 ? 4.0 : 8.0;
        const double N = (double)h->N(), n = (double)h->n;
        double v = 0.0;
        switch (kind) {
        case 0: v = (double)h->A.nnzb * (bb * sv + 4.0) + 4.0 * (n + 1) + 2.0 * N * sx; break;
        case 1: v = (double)h->L.nnzb * (bb * sv + 4.0) + 4.0 * (n + 1) + 2.0 * N * sx; break;
        case 2: v = (double)h->U.nnzb * (bb * sv + 4.0) + n * bb * 8.0 + 4.0 * (n + 1) + 2.0 * N * sx; break;
        default: throw std::runtime_error("bad kernel kind");
        }
        if (bytes) *bytes = v;
    });
}

}  // extern "C"
