// The reference's per-colour device format ("wire format"): the arrays the
// FPGA kernel reads (fpga_functions_bicgstab.cpp:526-545,606-629), their
// sizes (vectorSizes, :527-544), and the five data buffers plus setup block
// they are packed into (:226-289,369-477,558-629).  Host-only code: the GPU
// path never uses this format; it exists so that matrices can be exchanged
// with the reference's FPGA path (export, import, buffer images, dumps).
//
// Element semantics, read from the RTL that consumes the arrays:
//  - A "colour" of the format is a partition of one colour of the colour
//    ordering, bounded by the device's per-colour limits (constants.vhd:
//    80-98): at most 2048 rows (ROW_INDEX_WIDTH 11), 8192 gathered vector
//    entries (COL_INDEX_WIDTH 13), 32768 non-zeros (MAT_PART_ADDR_WIDTH 15).
//  - Per partition: {row, arow, col, val} (rw_pkg.vhd:79-84).  arow is the
//    number of scalar rows, written 0..arow-1 after the rows of the partitions
//    before it (write_unit.vhd:225-240,379,419-421); col is the number of
//    gathered vector entries; val the number of stream entries.
//  - P indices: the colour-ordered scalar indices of the gathered entries
//    (int_read_unit.vhd:346-360).  For L and U the partition's own rows take
//    the last arow slots (ilu0_start_addr = col - arow, int_read_unit.vhd:
//    277-284).  Column index (uint16) = slot in this list.
//  - NR (uint8): row advance of each stream entry; the row of entry k is
//    sum_{m<=k} NR_m - 1 (NR2offset.vhd:32-41).  Rows without entries are
//    zero-filled (write_unit.vhd:225-240).
//  - Value and column-index streams are read in whole lines of 8 and 32
//    (ext_read_unit.vhd:147-152,282-287, sizes val>>3 and val>>5), so val is a
//    multiple of 32; NRs are read in ceil(val/64) lines (:246) and each
//    partition's NR block is padded to 64, P blocks to 16 (:97), D^-1 blocks
//    (4 doubles per scalar row, :114-116) to 8.
//  - L is swept over partitions in colour order, U in reverse, with rows
//    descending (write_unit.vhd:200-203; int_read_unit.vhd:295-302).
//
// Choices of this build where the reference leaves the producer free (the
// producer is opm-simulators, not part of the reference):
//  - partitions are cut greedily along block rows in processing order;
//  - stream entries of a row are in ascending column order, every entry of
//    every stored block (zeros included);
//  - filler entries hold -0.0 (x + -0.0 == x for every x): padding to a
//    multiple of 32 repeats the last entry's row and column (NR 0); a gap of
//    more than 255 empty rows is bridged by fillers at slot 0 of the rows in
//    between (only L and U have empty rows);
//  - row == arow; the 8 header words of each colour-size array are 0.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bda_host.hpp"
#include "../../include/bda_bicgstab.h"

namespace bda {
namespace wire {

constexpr int kMaxRows = 2048;
constexpr int kMaxCols = 8192;
constexpr int kMaxPartVals = 32768;
constexpr int kMaxNR = 255;
constexpr int kMaxParts = 256;              // MAX_COLORS_DEPTH 8 (constants.vhd:107-108)
constexpr int64_t kMaxMatrixVals = 1 << 21; // MATRIX_ADDR_WIDTH 21 (constants.vhd:99-101)
constexpr int kVectorElems = 163840;        // VECTOR_SIZE_ELEM (constants.vhd:76)
constexpr int kMaxRowVals = 8 * 16 - 8 + 1; // MAX_NNZS_PER_ROW (constants.vhd:150)
constexpr int kValAlign = 32, kPAlign = 16, kNRAlign = 64, kDiagAlign = 8, kDiagPerRow = 4;
constexpr int kHeaderWords = 8;             // colour sizes start 8 ints in (fpga_functions_bicgstab.cpp:606-608)

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

enum Kind { KA = 0, KL = 1, KU = 2 };

struct Part {
    std::vector<int32_t> rows;   // scalar rows (colour order) in processing order
    std::vector<uint32_t> P;     // gathered entries
    std::vector<double> val;     // stream (padded to kValAlign)
    std::vector<uint16_t> col;
    std::vector<uint8_t> nr;
    int32_t nreal = 0;           // real entries before padding
};

struct Csr {
    const int32_t *rowptr, *colidx;
    const double *vals;
};

// One scalar row's stream entries: (global scalar column, value), ascending.
void row_entries(const Csr &m, int bs, int32_t r, std::vector<std::pair<int32_t, double>> &out) {
    out.clear();
    const int i = r / bs, a = r % bs;
    for (int32_t q = m.rowptr[i]; q < m.rowptr[i + 1]; ++q)
        for (int b = 0; b < bs; ++b)
            out.emplace_back(m.colidx[q] * bs + b, m.vals[((size_t)q * bs + a) * bs + b]);
}

// Greedy partitioning of one matrix in processing order, and its streams.
std::vector<Part> build_matrix(Kind kind, const Csr &m, int bs, int n, const std::vector<int32_t> &colour_ptr) {
    const int nc = (int)colour_ptr.size() - 1;
    std::vector<Part> parts;
    std::vector<int32_t> stamp((size_t)n * bs, -1);
    std::vector<std::pair<int32_t, double>> ent;
    const bool own = kind != KA;
    const bool desc = kind == KU;

    // block rows of each colour in processing order
    for (int cc = 0; cc < nc; ++cc) {
        const int c = desc ? nc - 1 - cc : cc;
        std::vector<int32_t> brows;
        for (int i = colour_ptr[c]; i < colour_ptr[c + 1]; ++i) brows.push_back(i);
        if (desc) std::reverse(brows.begin(), brows.end());

        size_t k = 0;
        while (k < brows.size()) {
            Part p;
            const int pid = (int)parts.size();
            int ncols = 0;
            int64_t nval = 0;      // real + gap fillers
            int last_pos = -1;     // position of the last row with entries
            // try to add block rows while the limits hold
            while (k < brows.size()) {
                const int i = brows[k];
                // entries of the block row's bs scalar rows (processing order)
                int add_cols = 0;
                int64_t add_val = 0;
                std::vector<int32_t> newcols;
                int pos = (int)p.rows.size();
                int lp = last_pos;
                for (int aa = 0; aa < bs; ++aa) {
                    const int a = desc ? bs - 1 - aa : aa;
                    row_entries(m, bs, i * bs + a, ent);
                    if ((int)ent.size() > kMaxRowVals)
                        throw std::runtime_error("wire format: a scalar row has more than 121 entries");
                    if (!ent.empty()) {
                        add_val += (pos - lp - 1) / kMaxNR;  // gap fillers
                        add_val += (int64_t)ent.size();
                        lp = pos;
                    }
                    for (auto &e : ent)
                        if (stamp[e.first] != pid) {
                            stamp[e.first] = pid;
                            newcols.push_back(e.first);
                        }
                    ++pos;
                }
                add_cols = (int)newcols.size() + (own ? bs : 0);
                const bool fits = (int)p.rows.size() + bs <= kMaxRows && ncols + add_cols <= kMaxCols &&
                                  round_up(nval + add_val, kValAlign) <= kMaxPartVals;
                if (!fits) {
                    for (int32_t cidx : newcols) stamp[cidx] = -1;  // not taken
                    if (p.rows.empty())
                        throw std::runtime_error("wire format: one block row exceeds the per-colour limits");
                    break;
                }
                for (int aa = 0; aa < bs; ++aa) p.rows.push_back(i * bs + (desc ? bs - 1 - aa : aa));
                ncols += add_cols;
                nval += add_val;
                last_pos = lp;
                ++k;
            }
            // gathered list: distinct columns ascending, then own rows ascending
            std::vector<int32_t> cols;
            for (int32_t r : p.rows) {
                row_entries(m, bs, r, ent);
                for (auto &e : ent)
                    if (stamp[e.first] == pid) {
                        stamp[e.first] = -2 - pid;  // listed
                        cols.push_back(e.first);
                    }
            }
            std::sort(cols.begin(), cols.end());
            for (int32_t c2 : cols) p.P.push_back((uint32_t)c2);
            if (own) {
                std::vector<int32_t> mine(p.rows);
                std::sort(mine.begin(), mine.end());
                for (int32_t r : mine) p.P.push_back((uint32_t)r);
            }
            // streams
            int prev = -1;
            for (int pos = 0; pos < (int)p.rows.size(); ++pos) {
                row_entries(m, bs, p.rows[pos], ent);
                if (ent.empty()) continue;
                while (pos - prev > kMaxNR) {
                    prev += kMaxNR;
                    p.val.push_back(-0.0);
                    p.col.push_back(0);
                    p.nr.push_back((uint8_t)kMaxNR);
                }
                bool first = true;
                for (auto &e : ent) {
                    const auto it = std::lower_bound(cols.begin(), cols.end(), e.first);
                    p.val.push_back(e.second);
                    p.col.push_back((uint16_t)(it - cols.begin()));
                    p.nr.push_back((uint8_t)(first ? pos - prev : 0));
                    first = false;
                }
                prev = pos;
            }
            p.nreal = (int32_t)p.val.size();
            if (!p.val.empty()) {
                const uint16_t lc = p.col.back();
                while (p.val.size() % kValAlign) {
                    p.val.push_back(-0.0);
                    p.col.push_back(lc);
                    p.nr.push_back(0);
                }
            }
            parts.push_back(std::move(p));
        }
    }
    return parts;
}

struct Built {
    int bs = 0, n = 0;
    std::vector<Part> m[3];
    std::vector<double> blkd;
};

void build_all(int bs, int n, const Csr &A, const Csr &L, const Csr &U, const double *dinv,
               const std::vector<int32_t> &colour_ptr, Built &out) {
    out.bs = bs;
    out.n = n;
    out.m[KA] = build_matrix(KA, A, bs, n, colour_ptr);
    out.m[KL] = build_matrix(KL, L, bs, n, colour_ptr);
    out.m[KU] = build_matrix(KU, U, bs, n, colour_ptr);
    out.blkd.clear();
    for (const Part &p : out.m[KU]) {
        for (int32_t r : p.rows) {
            const int i = r / bs, a = r % bs;
            for (int b = 0; b < kDiagPerRow; ++b)
                out.blkd.push_back(b < bs ? dinv[((size_t)i * bs + a) * bs + b] : 0.0);
        }
        while (out.blkd.size() % kDiagAlign) out.blkd.push_back(0.0);
    }
}

void sizes_of(const Built &b, bda_wire_sizes *ws) {
    std::memset(ws, 0, sizeof(*ws));
    const int64_t N = (int64_t)b.n * b.bs;
    bool fits = N <= kVectorElems;
    for (int k = 0; k < 3; ++k) {
        int64_t val = 0, cols = 0, nrs = 0;
        for (const Part &p : b.m[k]) {
            val += (int64_t)p.val.size();
            cols += round_up((int64_t)p.P.size(), kPAlign);
            nrs += round_up((int64_t)p.val.size(), kNRAlign);
        }
        int32_t *s = ws->vector_sizes + 6 * k;
        s[0] = (int32_t)N;
        s[1] = (int32_t)val;
        s[2] = (int32_t)b.m[k].size();
        s[3] = (int32_t)cols;
        s[4] = (int32_t)nrs;
        s[5] = (int32_t)b.blkd.size();
        ws->nnz_val_sizes[k] = (int32_t)val;
        ws->partitions[k] = (int32_t)b.m[k].size();
        fits = fits && (int)b.m[k].size() <= kMaxParts && val <= kMaxMatrixVals;
    }
    ws->device_fits = fits ? 1 : 0;
}

void write_all(const Built &b, const double *b_perm, const double *x0_perm, void *vp[21]) {
    for (int k = 0; k < 3; ++k) {
        const int base = 6 * k;
        int32_t *cs = (int32_t *)vp[base + 0];
        uint32_t *P = (uint32_t *)vp[base + 1];
        double *V = ((double **)vp[base + 2])[0];
        uint16_t *Ci = (uint16_t *)vp[base + 3];
        uint8_t *NR = (uint8_t *)vp[base + 4];
        if (!cs || !P || !V || !Ci || !NR) throw std::runtime_error("wire format: missing output array");
        std::memset(cs, 0, sizeof(int32_t) * kHeaderWords);
        size_t op = 0, ov = 0, onr = 0;
        for (size_t j = 0; j < b.m[k].size(); ++j) {
            const Part &p = b.m[k][j];
            int32_t *e = cs + kHeaderWords + 4 * j;
            e[0] = (int32_t)p.rows.size();
            e[1] = (int32_t)p.rows.size();
            e[2] = (int32_t)p.P.size();
            e[3] = (int32_t)p.val.size();
            std::copy(p.P.begin(), p.P.end(), P + op);
            std::fill(P + op + p.P.size(), P + op + round_up(p.P.size(), kPAlign), 0u);
            op += round_up(p.P.size(), kPAlign);
            std::copy(p.val.begin(), p.val.end(), V + ov);
            std::copy(p.col.begin(), p.col.end(), Ci + ov);
            ov += p.val.size();
            std::copy(p.nr.begin(), p.nr.end(), NR + onr);
            std::fill(NR + onr + p.nr.size(), NR + onr + round_up(p.nr.size(), kNRAlign), (uint8_t)0);
            onr += round_up(p.nr.size(), kNRAlign);
        }
    }
    if (vp[18]) std::copy(b.blkd.begin(), b.blkd.end(), (double *)vp[18]);
    const size_t N = (size_t)b.n * b.bs;
    if (vp[19]) {
        if (b_perm) std::memcpy(vp[19], b_perm, N * sizeof(double));
        else std::memset(vp[19], 0, N * sizeof(double));
    }
    if (vp[20]) {
        if (x0_perm) std::memcpy(vp[20], x0_perm, N * sizeof(double));
        else std::memset(vp[20], 0, N * sizeof(double));
    }
}

// ---- data buffers (fpga_setup_host_datamem / fpga_copy_host_datamem) ------
constexpr int kLine = 64;
constexpr int kSetupLines = 5;
constexpr int kSetupWords = kSetupLines * kLine / 8;
constexpr uint64_t kUnused = 0xDEADC0DEDEADC0DEull;

struct Layout {
    uint32_t pos[5][10] = {};
    uint32_t total[5] = {};
};

Layout layout_of(const int32_t s[18], const int32_t nv[3], uint32_t dbg_bytes) {
    (void)dbg_bytes;
    auto rd = [](int64_t x, int64_t elem) { return (uint32_t)(elem * round_up(x, kLine / elem)); };
    const int rowSize = s[0];
    Layout L;
    const uint32_t vec = rd(rowSize, 8);
    // buffer 0: setup, values of A/L/U, colour sizes of A/L/U, D^-1
    L.pos[0][0] = 0;
    L.pos[0][1] = L.pos[0][0] + kSetupLines * kLine;
    L.pos[0][2] = L.pos[0][1] + rd(nv[0], 8);
    L.pos[0][3] = L.pos[0][2] + rd(nv[1], 8);
    L.pos[0][4] = L.pos[0][3] + rd(nv[2], 8);
    L.pos[0][5] = L.pos[0][4] + rd(4 * (int64_t)s[2], 4);
    L.pos[0][6] = L.pos[0][5] + rd(4 * (int64_t)s[8], 4);
    L.pos[0][7] = L.pos[0][6] + rd(4 * (int64_t)s[14], 4);
    L.pos[0][8] = L.pos[0][7] + rd(s[5], 8);
    L.total[0] = L.pos[0][8];
    // buffer 1: P indices, column indices, NR offsets of A/L/U
    L.pos[1][0] = 0;
    L.pos[1][1] = L.pos[1][0] + rd(s[3], 4);
    L.pos[1][2] = L.pos[1][1] + rd(s[9], 4);
    L.pos[1][3] = L.pos[1][2] + rd(s[15], 4);
    L.pos[1][4] = L.pos[1][3] + rd(s[1], 2);
    L.pos[1][5] = L.pos[1][4] + rd(s[7], 2);
    L.pos[1][6] = L.pos[1][5] + rd(s[13], 2);
    L.pos[1][7] = L.pos[1][6] + rd(s[4], 1);
    L.pos[1][8] = L.pos[1][7] + rd(s[10], 1);
    L.pos[1][9] = L.pos[1][8] + rd(s[16], 1);
    L.total[1] = L.pos[1][9];
    // buffer 2: X2, R1; buffer 3: X1, R2, P1, P2, RT; buffer 4: T, V, Lres, Ures
    L.pos[2][1] = vec;
    L.total[2] = 2 * vec;
    for (int j = 1; j <= 4; ++j) L.pos[3][j] = j * vec;
    L.total[3] = 5 * vec;
    for (int j = 1; j <= 3; ++j) L.pos[4][j] = j * vec;
    L.total[4] = 4 * vec;
    return L;
}

void setup_of(const Layout &L, const int32_t s[18], uint32_t config_bits, uint64_t setup[kSetupWords]) {
    for (int i = 0; i < kSetupWords; ++i) setup[i] = kUnused;
    auto cl = [&](int b, int j) { return (uint64_t)(L.pos[b][j] / kLine); };
    auto pack = [](int32_t hi, int32_t lo) { return ((uint64_t)(uint32_t)hi << 32) + (uint64_t)(uint32_t)lo; };
    setup[0] = pack(s[1], s[0]);
    setup[1] = pack((int32_t)config_bits, s[2]);
    setup[2] = cl(2, 1);   // R1
    setup[3] = cl(3, 1);   // R2
    setup[4] = cl(3, 0);   // X1
    setup[5] = cl(2, 0);   // X2
    setup[6] = cl(3, 2);   // P1
    setup[7] = cl(3, 3);   // P2
    setup[8] = pack(s[7], s[6]);
    setup[9] = (uint64_t)(uint32_t)s[8];
    setup[10] = cl(0, 4);  // colour sizes
    setup[11] = cl(1, 0);  // P indices
    setup[12] = cl(0, 1);  // values
    setup[13] = cl(1, 3);  // column indices
    setup[14] = cl(1, 6);  // NR offsets
    setup[15] = cl(3, 4);  // RT
    setup[16] = pack(s[13], s[12]);
    setup[17] = (uint64_t)(uint32_t)s[14];
    setup[18] = cl(0, 5);  // L colour sizes
    setup[19] = cl(1, 1);
    setup[20] = cl(0, 2);
    setup[21] = cl(1, 4);
    setup[22] = cl(1, 7);
    setup[23] = cl(0, 7);  // D^-1
    setup[24] = cl(0, 6);  // U colour sizes
    setup[25] = cl(1, 2);
    setup[26] = cl(0, 3);
    setup[27] = cl(1, 5);
    setup[28] = cl(1, 8);
    setup[29] = cl(4, 0);  // T
    setup[30] = cl(4, 1);  // V
}

}  // namespace wire
}  // namespace bda

using namespace bda;
using namespace bda::wire;

namespace {
template <typename F>
int wire_guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ERROR: wire format: %s\n", e.what());
        return 1;
    }
}

std::vector<int32_t> colour_vec(const int32_t *colour_ptr, int32_t ncolours, int32_t n) {
    if (!colour_ptr || ncolours <= 0 || colour_ptr[0] != 0 || colour_ptr[ncolours] != n)
        throw std::runtime_error("bad colour_ptr");
    return std::vector<int32_t>(colour_ptr, colour_ptr + ncolours + 1);
}
}  // namespace

extern "C" {

int bda_wire_build_sizes(int32_t bs, int32_t n, const int32_t *colour_ptr, int32_t ncolours, const int32_t *a_rowptr,
                         const int32_t *a_colidx, const int32_t *l_rowptr, const int32_t *l_colidx,
                         const int32_t *u_rowptr, const int32_t *u_colidx, bda_wire_sizes *ws) {
    return wire_guard([&] {
        if ((bs != 3 && bs != 4) || n <= 0 || !ws) throw std::runtime_error("bad arguments");
        const auto cp = colour_vec(colour_ptr, ncolours, n);
        // values are irrelevant for the sizes: a zero block per entry
        const int64_t mx = std::max<int64_t>({a_rowptr[n], l_rowptr[n], u_rowptr[n]});
        std::vector<double> zero((size_t)mx * bs * bs, 0.0);
        std::vector<double> dz((size_t)n * bs * bs, 0.0);
        Built b;
        build_all(bs, n, Csr{a_rowptr, a_colidx, zero.data()}, Csr{l_rowptr, l_colidx, zero.data()},
                  Csr{u_rowptr, u_colidx, zero.data()}, dz.data(), cp, b);
        sizes_of(b, ws);
    });
}

int bda_wire_build(int32_t bs, int32_t n, const int32_t *colour_ptr, int32_t ncolours, const int32_t *a_rowptr,
                   const int32_t *a_colidx, const double *a_values, const int32_t *l_rowptr,
                   const int32_t *l_colidx, const double *l_values, const int32_t *u_rowptr,
                   const int32_t *u_colidx, const double *u_values, const double *dinv, const double *b_perm,
                   const double *x0_perm, void *vector_pointers[21]) {
    return wire_guard([&] {
        if ((bs != 3 && bs != 4) || n <= 0 || !vector_pointers) throw std::runtime_error("bad arguments");
        const auto cp = colour_vec(colour_ptr, ncolours, n);
        Built b;
        build_all(bs, n, Csr{a_rowptr, a_colidx, a_values}, Csr{l_rowptr, l_colidx, l_values},
                  Csr{u_rowptr, u_colidx, u_values}, dinv, cp, b);
        write_all(b, b_perm, x0_perm, vector_pointers);
    });
}

int bda_wire_layout(const int32_t vector_sizes[18], const int32_t nnz_val_sizes[3], uint32_t config_bits,
                    uint32_t total_size[5], uint64_t setup[40]) {
    return wire_guard([&] {
        const Layout L = layout_of(vector_sizes, nnz_val_sizes, 0);
        for (int b = 0; b < 5; ++b) total_size[b] = L.total[b];
        if (setup) setup_of(L, vector_sizes, config_bits, setup);
    });
}

int bda_wire_pack(const int32_t vector_sizes[18], const int32_t nnz_val_sizes[3], uint32_t config_bits,
                  void *const vector_pointers[21], uint8_t *buffers[5]) {
    return wire_guard([&] {
        const int32_t *s = vector_sizes;
        const Layout L = layout_of(s, nnz_val_sizes, 0);
        for (int b = 0; b < 5; ++b) std::memset(buffers[b], 0, L.total[b]);
        setup_of(L, s, config_bits, (uint64_t *)buffers[0]);
        auto put = [&](int b, int j, const void *src, size_t bytes) {
            if (bytes && !src) throw std::runtime_error("missing input array");
            if (bytes) std::memcpy(buffers[b] + L.pos[b][j], src, bytes);
        };
        void *const *vp = vector_pointers;
        put(0, 4, (const int32_t *)vp[0] + kHeaderWords, 16 * (size_t)s[2]);
        put(0, 5, (const int32_t *)vp[6] + kHeaderWords, 16 * (size_t)s[8]);
        put(0, 6, (const int32_t *)vp[12] + kHeaderWords, 16 * (size_t)s[14]);
        put(1, 0, vp[1], 4 * (size_t)s[3]);
        put(1, 1, vp[7], 4 * (size_t)s[9]);
        put(1, 2, vp[13], 4 * (size_t)s[15]);
        put(0, 1, ((double *const *)vp[2])[0], 8 * (size_t)nnz_val_sizes[0]);
        put(0, 2, ((double *const *)vp[8])[0], 8 * (size_t)nnz_val_sizes[1]);
        put(0, 3, ((double *const *)vp[14])[0], 8 * (size_t)nnz_val_sizes[2]);
        put(1, 3, vp[3], 2 * (size_t)s[1]);
        put(1, 4, vp[9], 2 * (size_t)s[7]);
        put(1, 5, vp[15], 2 * (size_t)s[13]);
        put(1, 6, vp[4], (size_t)s[4]);
        put(1, 7, vp[10], (size_t)s[10]);
        put(1, 8, vp[16], (size_t)s[16]);
        put(2, 1, vp[19], 8 * (size_t)s[0]);   // R1 = b
        put(3, 0, vp[20], 8 * (size_t)s[0]);   // X1 = x0
        put(0, 7, vp[18], 8 * (size_t)s[5]);   // D^-1
    });
}

// Arrays of a packed image: pointers into the buffers, sizes from the setup
// block and the colour-size arrays.
int bda_wire_unpack(uint8_t *const buffers[5], int32_t vector_sizes[18], int32_t nnz_val_sizes[3],
                    void *vector_pointers[21], double *value_arrays[3]) {
    return wire_guard([&] {
        const uint64_t *setup = (const uint64_t *)buffers[0];
        auto lo = [](uint64_t w) { return (int32_t)(uint32_t)(w & 0xFFFFFFFFu); };
        auto hi = [](uint64_t w) { return (int32_t)(uint32_t)(w >> 32); };
        int32_t *s = vector_sizes;
        const int wr[3] = {0, 8, 16}, wc[3] = {1, 9, 17};
        const int wcs[3] = {10, 18, 24}, wp[3] = {11, 19, 25}, wv[3] = {12, 20, 26}, wci[3] = {13, 21, 27},
                  wnr[3] = {14, 22, 28};
        int64_t blkd = 0;
        for (int k = 0; k < 3; ++k) {
            s[6 * k + 0] = lo(setup[wr[k]]);
            s[6 * k + 1] = hi(setup[wr[k]]);
            s[6 * k + 2] = lo(setup[wc[k]]);
            const int32_t *cs = (const int32_t *)(buffers[0] + setup[wcs[k]] * kLine);
            int64_t cols = 0, nrs = 0;
            for (int j = 0; j < s[6 * k + 2]; ++j) {
                cols += round_up(cs[4 * j + 2], kPAlign);
                nrs += round_up(cs[4 * j + 3], kNRAlign);
                if (k == KU) blkd += round_up((int64_t)kDiagPerRow * cs[4 * j + 1], kDiagAlign);
            }
            s[6 * k + 3] = (int32_t)cols;
            s[6 * k + 4] = (int32_t)nrs;
            nnz_val_sizes[k] = s[6 * k + 1];
            // the header words before the colour sizes are not part of the image
            vector_pointers[6 * k + 0] = (void *)(cs - kHeaderWords);
            vector_pointers[6 * k + 1] = buffers[1] + setup[wp[k]] * kLine;
            value_arrays[k] = (double *)(buffers[0] + setup[wv[k]] * kLine);
            vector_pointers[6 * k + 2] = (void *)&value_arrays[k];
            vector_pointers[6 * k + 3] = buffers[1] + setup[wci[k]] * kLine;
            vector_pointers[6 * k + 4] = buffers[1] + setup[wnr[k]] * kLine;
        }
        for (int k = 0; k < 3; ++k) s[6 * k + 5] = (int32_t)blkd;
        vector_pointers[18] = buffers[0] + setup[23] * kLine;
        vector_pointers[19] = buffers[2] + setup[2] * kLine;
        vector_pointers[20] = buffers[3] + setup[4] * kLine;
    });
}

// Decoding into colour-ordered BCSR.  A's block pattern is the union of the
// blocks its stream touches; L and U values are read on A's strictly lower
// and upper blocks (entries outside them must be -0.0 or 0.0 fillers).
}  // extern "C"

namespace {
// Checks every stream of a foreign buffer image before anything is decoded:
// partition sizes non-negative, each entry's row inside its partition, its
// slot inside the partition's gathered list, every gathered index inside the
// vector, the running offsets inside the arrays vector_sizes declares, the
// partitions' rows within N (covering it exactly for A), and D^-1's rows and
// values inside their array.
void validate_streams(int32_t bs, const int32_t vector_sizes[18], void *const vector_pointers[21]) {
    if (bs != 3 && bs != 4) throw std::runtime_error("bad block size");
    const int32_t N = vector_sizes[0];
    if (N <= 0 || N % bs) throw std::runtime_error("row count is not a positive multiple of the block size");
    for (int k = 0; k < 3; ++k) {
        const int base = 6 * k;
        const int32_t parts = vector_sizes[base + 2];
        const int64_t n_slots = vector_sizes[base + 1], n_p = vector_sizes[base + 3], n_nr = vector_sizes[base + 4];
        if (parts < 0 || n_slots < 0 || n_p < 0 || n_nr < 0) throw std::runtime_error("negative array size");
        if (parts > 0 && (!vector_pointers[base] || !vector_pointers[base + 1] || !vector_pointers[base + 3] ||
                          !vector_pointers[base + 4]))
            throw std::runtime_error("missing stream array");
        const int32_t *cs = parts > 0 ? (const int32_t *)vector_pointers[base] + kHeaderWords : nullptr;
        const uint32_t *P = (const uint32_t *)vector_pointers[base + 1];
        const uint16_t *Ci = (const uint16_t *)vector_pointers[base + 3];
        const uint8_t *NR = (const uint8_t *)vector_pointers[base + 4];
        int64_t row0 = 0, op = 0, ov = 0, onr = 0;
        for (int j = 0; j < parts; ++j) {
            const int32_t arow = cs[4 * j + 1], col = cs[4 * j + 2], val = cs[4 * j + 3];
            if (arow < 0 || col < 0 || val < 0) throw std::runtime_error("negative partition size");
            if (op + col > n_p || ov + val > n_slots || onr + val > n_nr)
                throw std::runtime_error("partition beyond its stream arrays");
            if (row0 + arow > N) throw std::runtime_error("partition rows beyond the vector");
            for (int32_t q = 0; q < col; ++q)
                if (P[op + q] >= (uint32_t)N) throw std::runtime_error("P index outside the vector");
            int64_t rr = -1;
            for (int32_t q = 0; q < val; ++q) {
                rr += NR[onr + q];
                if (rr < 0 || rr >= arow) throw std::runtime_error("stream row outside the partition");
                if (Ci[ov + q] >= col) throw std::runtime_error("column slot outside the gathered list");
            }
            row0 += arow;
            op += round_up(col, kPAlign);
            ov += val;
            onr += round_up(val, kNRAlign);
        }
        if (k == KA && row0 != N) throw std::runtime_error("partition rows do not cover the vector");
        if (k == KU) {
            int64_t od = 0;
            for (int j = 0; j < parts; ++j) od += round_up((int64_t)kDiagPerRow * cs[4 * j + 1], kDiagAlign);
            if (od > vector_sizes[5]) throw std::runtime_error("D^-1 rows beyond their array");
        }
    }
}
}  // namespace

extern "C" {

int bda_wire_import_sizes(int32_t bs, const int32_t vector_sizes[18], void *const vector_pointers[21], int32_t *n,
                          int64_t *nnzb, int32_t *num_parts) {
    return wire_guard([&] {
        validate_streams(bs, vector_sizes, vector_pointers);
        const int32_t N = vector_sizes[0];
        const int nb = N / bs;
        const int32_t *cs = (const int32_t *)vector_pointers[0] + kHeaderWords;
        const uint32_t *P = (const uint32_t *)vector_pointers[1];
        const uint16_t *Ci = (const uint16_t *)vector_pointers[3];
        const uint8_t *NR = (const uint8_t *)vector_pointers[4];
        std::vector<std::vector<int32_t>> cols(nb);
        int64_t row0 = 0, op = 0, ov = 0, onr = 0;
        for (int j = 0; j < vector_sizes[2]; ++j) {
            const int32_t arow = cs[4 * j + 1], col = cs[4 * j + 2], val = cs[4 * j + 3];
            int64_t rr = -1;
            for (int32_t k = 0; k < val; ++k) {
                rr += NR[onr + k];
                if (rr < 0 || rr >= arow) throw std::runtime_error("stream row outside the partition");
                if (Ci[ov + k] >= col) throw std::runtime_error("column slot outside the gathered list");
                const int64_t r = row0 + rr, c = P[op + Ci[ov + k]];
                if (c >= N) throw std::runtime_error("P index outside the vector");
                cols[r / bs].push_back((int32_t)(c / bs));
            }
            row0 += arow;
            op += round_up(col, kPAlign);
            ov += val;
            onr += round_up(val, kNRAlign);
        }
        if (row0 != N) throw std::runtime_error("partition rows do not cover the vector");
        int64_t tot = 0;
        for (auto &v : cols) {
            std::sort(v.begin(), v.end());
            tot += std::unique(v.begin(), v.end()) - v.begin();
        }
        *n = nb;
        *nnzb = tot;
        *num_parts = vector_sizes[2];
    });
}

int bda_wire_import(int32_t bs, const int32_t vector_sizes[18], void *const vector_pointers[21], int32_t *rowptr,
                    int32_t *colidx, double *values, double *lu_values, double *dinv, int32_t *part_rows,
                    double *b, double *x0) {
    return wire_guard([&] {
        validate_streams(bs, vector_sizes, vector_pointers);
        const int32_t N = vector_sizes[0];
        const int nb = N / bs;
        const int bb = bs * bs;
        // pattern of A
        {
            const int32_t *cs = (const int32_t *)vector_pointers[0] + kHeaderWords;
            const uint32_t *P = (const uint32_t *)vector_pointers[1];
            const uint16_t *Ci = (const uint16_t *)vector_pointers[3];
            const uint8_t *NR = (const uint8_t *)vector_pointers[4];
            std::vector<std::vector<int32_t>> cols(nb);
            int64_t row0 = 0, op = 0, ov = 0, onr = 0;
            for (int j = 0; j < vector_sizes[2]; ++j) {
                const int32_t arow = cs[4 * j + 1], col = cs[4 * j + 2], val = cs[4 * j + 3];
                int64_t rr = -1;
                for (int32_t k = 0; k < val; ++k) {
                    rr += NR[onr + k];
                    cols[(row0 + rr) / bs].push_back((int32_t)(P[op + Ci[ov + k]] / bs));
                }
                if (part_rows) part_rows[j] = (int32_t)row0;
                row0 += arow;
                op += round_up(col, kPAlign);
                ov += val;
                onr += round_up(val, kNRAlign);
            }
            if (part_rows) part_rows[vector_sizes[2]] = (int32_t)row0;
            rowptr[0] = 0;
            for (int i = 0; i < nb; ++i) {
                auto &v = cols[i];
                std::sort(v.begin(), v.end());
                v.erase(std::unique(v.begin(), v.end()), v.end());
                std::copy(v.begin(), v.end(), colidx + rowptr[i]);
                rowptr[i + 1] = rowptr[i] + (int32_t)v.size();
            }
        }
        auto find_block = [&](int i, int j) -> int64_t {
            const int32_t *b0 = colidx + rowptr[i], *b1 = colidx + rowptr[i + 1];
            const int32_t *it = std::lower_bound(b0, b1, j);
            return (it != b1 && *it == j) ? (int64_t)(it - colidx) : -1;
        };
        // values: A into values, L (strictly lower) and U (strictly upper) into lu_values
        const int64_t nnzb = rowptr[nb];
        std::fill(values, values + nnzb * bb, 0.0);
        if (lu_values) std::fill(lu_values, lu_values + nnzb * bb, 0.0);
        for (int k = 0; k < 3; ++k) {
            if (k > 0 && !lu_values) break;
            const int base = 6 * k;
            const int32_t *cs = (const int32_t *)vector_pointers[base] + kHeaderWords;
            const uint32_t *P = (const uint32_t *)vector_pointers[base + 1];
            const double *V = ((double *const *)vector_pointers[base + 2])[0];
            const uint16_t *Ci = (const uint16_t *)vector_pointers[base + 3];
            const uint8_t *NR = (const uint8_t *)vector_pointers[base + 4];
            double *dst = k == 0 ? values : lu_values;
            int64_t op = 0, ov = 0, onr = 0;
            std::vector<int32_t> rows;
            // rows of each partition in processing order: A and L ascending
            // from the running start; U descending from the top
            int64_t up = N;
            int64_t row0 = 0;
            for (int j = 0; j < vector_sizes[base + 2]; ++j) {
                const int32_t arow = cs[4 * j + 1], col = cs[4 * j + 2], val = cs[4 * j + 3];
                int64_t rr = -1;
                for (int32_t q = 0; q < val; ++q) {
                    rr += NR[onr + q];
                    const int64_t r = k == KU ? up - 1 - rr : row0 + rr;
                    const int64_t c = P[op + Ci[ov + q]];
                    const double v = V[ov + q];
                    const int i = (int)(r / bs), a = (int)(r % bs), jb = (int)(c / bs), bcol = (int)(c % bs);
                    const bool ok = k == KA || (k == KL ? jb < i : jb > i);
                    const int64_t blk = ok ? find_block(i, jb) : -1;
                    if (blk < 0) {
                        if (v != 0.0) throw std::runtime_error("factor entry outside the pattern of A");
                        continue;
                    }
                    const size_t e = (size_t)blk * bb + a * bs + bcol;
                    dst[e] = dst[e] + v;  // fillers are -0.0
                }
                if (k == KU) up -= arow;
                else row0 += arow;
                op += round_up(col, kPAlign);
                ov += val;
                onr += round_up(val, kNRAlign);
            }
        }
        if (dinv) {
            const double *D = (const double *)vector_pointers[18];
            const int32_t *cs = (const int32_t *)vector_pointers[12] + kHeaderWords;
            int64_t up = N, od = 0;
            for (int j = 0; j < vector_sizes[14]; ++j) {
                const int32_t arow = cs[4 * j + 1];
                for (int32_t rr = 0; rr < arow; ++rr) {
                    const int64_t r = up - 1 - rr;
                    const int i = (int)(r / bs), a = (int)(r % bs);
                    for (int c = 0; c < bs; ++c) dinv[((size_t)i * bs + a) * bs + c] = D[od + (size_t)rr * kDiagPerRow + c];
                }
                up -= arow;
                od += round_up((int64_t)kDiagPerRow * arow, kDiagAlign);
            }
        }
        if (b && vector_pointers[19]) std::memcpy(b, vector_pointers[19], (size_t)N * 8);
        if (x0 && vector_pointers[20]) std::memcpy(x0, vector_pointers[20], (size_t)N * 8);
    });
}

}  // extern "C"
