// Host analysis of a block-CSR pattern: colouring, colour-ordered
// permutation, the permuted pattern, its strictly-lower (L) and
// strictly-upper (U) parts, their transposes (for the adjoint) and the
// tiling used by the streaming kernels.
//
// The reference consumes this information ready-made: colour sizes and
// P-index arrays per colour (fpga_functions_bicgstab.cpp:606-613, read per
// colour by sparstition.vhd:673-717); rows of one colour are processed with
// no intra-colour dependency (sparstition.vhd:33-37), L is swept over colours
// ascending and U descending (int_read_unit.vhd:287-302).
#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>

#include "bda_host.hpp"

namespace bda {
namespace {

// CSR transpose of a square pattern; out.src[k] = index of the source entry.
void transpose_pattern(const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colidx, int n,
                       MatPattern &out) {
    const int64_t nnz = rowptr[n];
    out.n = n;
    out.rowptr.assign(n + 1, 0);
    for (int64_t q = 0; q < nnz; ++q) out.rowptr[colidx[q] + 1]++;
    for (int i = 0; i < n; ++i) out.rowptr[i + 1] += out.rowptr[i];
    out.colidx.assign((size_t)nnz, 0);
    out.src.assign((size_t)nnz, 0);
    std::vector<int32_t> fill(out.rowptr.begin(), out.rowptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            const int32_t j = colidx[q];
            out.colidx[fill[j]] = i;  // rows visited ascending => sorted columns
            out.src[fill[j]] = q;
            fill[j]++;
        }
}

}  // namespace

int colour_rows(const Bcsr &m, bool level_scheduling, std::vector<int32_t> &colour) {
    const int n = m.n;
    // symmetrised neighbour lists: the pattern and its transpose
    MatPattern t;
    transpose_pattern(m.rowptr, m.colidx, n, t);
    colour.assign(n, -1);
    int ncol = 0;
    if (level_scheduling) {
        for (int i = 0; i < n; ++i) {
            int lvl = 0;
            for (int32_t q = m.rowptr[i]; q < m.rowptr[i + 1]; ++q) {
                const int32_t j = m.colidx[q];
                if (j < i) lvl = std::max(lvl, colour[j] + 1);
            }
            for (int32_t q = t.rowptr[i]; q < t.rowptr[i + 1]; ++q) {
                const int32_t j = t.colidx[q];
                if (j < i) lvl = std::max(lvl, colour[j] + 1);
            }
            colour[i] = lvl;
            ncol = std::max(ncol, lvl + 1);
        }
        return ncol;
    }
    std::vector<int32_t> mark;  // mark[c] == i  <=> colour c is taken by a neighbour of i
    for (int i = 0; i < n; ++i) {
        for (int32_t q = m.rowptr[i]; q < m.rowptr[i + 1]; ++q) {
            const int32_t j = m.colidx[q];
            if (j != i && colour[j] >= 0) mark[colour[j]] = i;
        }
        for (int32_t q = t.rowptr[i]; q < t.rowptr[i + 1]; ++q) {
            const int32_t j = t.colidx[q];
            if (j != i && colour[j] >= 0) mark[colour[j]] = i;
        }
        int c = 0;
        while (c < (int)mark.size() && mark[c] == i) ++c;
        if (c == (int)mark.size()) mark.push_back(-1);
        colour[i] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

void build_tiles(const MatPattern &p, const std::vector<int32_t> &colour_ptr, int cap,
                 std::vector<int32_t> &tile_row, std::vector<int32_t> &colour_tile, int max_rows) {
    const int ncol = (int)colour_ptr.size() - 1;
    tile_row.clear();
    colour_tile.assign(ncol + 1, 0);
    for (int c = 0; c < ncol; ++c) {
        colour_tile[c] = (int)tile_row.size();
        int r = colour_ptr[c];
        const int end = colour_ptr[c + 1];
        while (r < end) {
            tile_row.push_back(r);
            int rows = 0, blocks = 0;
            while (r < end && rows < max_rows) {
                const int nb = p.rowptr[r + 1] - p.rowptr[r];
                if (nb > cap)
                    throw std::runtime_error("block row " + std::to_string(r) + " has " + std::to_string(nb) +
                                             " blocks; the streaming kernels accept at most " +
                                             std::to_string(cap));
                if (blocks + nb > cap) break;
                blocks += nb;
                ++rows;
                ++r;
            }
        }
    }
    colour_tile[ncol] = (int)tile_row.size();
    tile_row.push_back(colour_ptr[ncol]);
}

void build_tiles(MatPattern &p, const std::vector<int32_t> &colour_ptr, int bs) {
    build_tiles(p, colour_ptr, tile_cap(bs, false), p.tile_row, p.colour_tile);
    build_tiles(p, colour_ptr, tile_cap(bs, true), p.tile_row_wide, p.colour_tile_wide);
    build_tiles(p, colour_ptr, tile_cap(bs, false), p.tile_row_tall, p.colour_tile_tall, kTileRowsTall);
}

void validate_pattern(const Bcsr &m) {
    const int n = m.n;
    if (n <= 0) throw std::runtime_error("empty matrix");
    if ((int)m.rowptr.size() != n + 1 || m.rowptr[0] != 0) throw std::runtime_error("bad rowptr");
    // the whole rowptr first, so that no column read below leaves colidx
    for (int i = 0; i < n; ++i)
        if (m.rowptr[i + 1] < m.rowptr[i]) throw std::runtime_error("rowptr not monotone");
    if ((size_t)m.rowptr[n] != m.colidx.size()) throw std::runtime_error("rowptr[n] is not the block count");
    for (int i = 0; i < n; ++i) {
        for (int32_t q = m.rowptr[i]; q < m.rowptr[i + 1]; ++q) {
            const int32_t j = m.colidx[q];
            if (j < 0 || j >= n) throw std::runtime_error("column index out of range in row " + std::to_string(i));
            if (q > m.rowptr[i] && m.colidx[q - 1] >= j)
                throw std::runtime_error("column indices not strictly increasing in row " + std::to_string(i));
        }
    }
}

void analyse(const Bcsr &m, bool level_scheduling, Analysis &an) {
    const int n = m.n;
    validate_pattern(m);
    an.n = n;
    an.bs = m.bs;
    an.level_scheduling = level_scheduling;
    an.ncolours = colour_rows(m, level_scheduling, an.colour);
    analyse_coloured(m, an);
}

void analyse_given_order(const Bcsr &m, const std::vector<int32_t> &colour_ptr, Analysis &an) {
    const int n = m.n;
    validate_pattern(m);
    const int nc = (int)colour_ptr.size() - 1;
    if (nc < 1 || colour_ptr[0] != 0 || colour_ptr[nc] != n)
        throw std::runtime_error("colour_ptr must run from 0 to n");
    for (int c = 0; c < nc; ++c)
        if (colour_ptr[c + 1] <= colour_ptr[c]) throw std::runtime_error("every colour must hold at least one row");
    an.n = n;
    an.bs = m.bs;
    an.level_scheduling = false;
    an.ncolours = nc;
    an.colour.assign(n, 0);
    for (int c = 0; c < nc; ++c)
        for (int i = colour_ptr[c]; i < colour_ptr[c + 1]; ++i) an.colour[i] = c;
    analyse_coloured(m, an);   // the colours are non-decreasing in row order: the permutation is the identity
}

void analyse_coloured(const Bcsr &m, Analysis &an) {
    const int n = m.n;
    const int ncol = an.ncolours;

    // stable counting sort by colour
    an.colour_ptr.assign(ncol + 1, 0);
    for (int i = 0; i < n; ++i) an.colour_ptr[an.colour[i] + 1]++;
    for (int c = 0; c < ncol; ++c) an.colour_ptr[c + 1] += an.colour_ptr[c];
    an.perm.assign(n, 0);
    an.iperm.assign(n, 0);
    {
        std::vector<int32_t> fill(an.colour_ptr.begin(), an.colour_ptr.end() - 1);
        for (int i = 0; i < n; ++i) {
            const int k = fill[an.colour[i]]++;
            an.perm[k] = i;
            an.iperm[i] = k;
        }
    }

    // permuted A, columns sorted
    MatPattern &A = an.A;
    A.n = n;
    A.rowptr.assign(n + 1, 0);
    for (int k = 0; k < n; ++k) {
        const int i = an.perm[k];
        A.rowptr[k + 1] = A.rowptr[k] + (m.rowptr[i + 1] - m.rowptr[i]);
    }
    const int64_t nnzb = A.rowptr[n];
    A.colidx.assign((size_t)nnzb, 0);
    A.src.assign((size_t)nnzb, 0);
    an.diag.assign(n, -1);
    an.max_row_blocks = 0;
    {
        std::vector<std::pair<int32_t, int32_t>> tmp;
        for (int k = 0; k < n; ++k) {
            const int i = an.perm[k];
            tmp.clear();
            for (int32_t q = m.rowptr[i]; q < m.rowptr[i + 1]; ++q) tmp.emplace_back(an.iperm[m.colidx[q]], q);
            std::sort(tmp.begin(), tmp.end());
            int32_t o = A.rowptr[k];
            for (auto &pr : tmp) {
                A.colidx[o] = pr.first;
                A.src[o] = pr.second;
                if (pr.first == k) an.diag[k] = o;
                ++o;
            }
            if (an.diag[k] < 0) throw std::runtime_error("missing diagonal block in row " + std::to_string(i));
            an.max_row_blocks = std::max(an.max_row_blocks, (int)tmp.size());
        }
    }
    // a valid colouring never couples two rows of one colour
    for (int k = 0; k < n; ++k)
        for (int32_t q = A.rowptr[k]; q < A.rowptr[k + 1]; ++q) {
            const int32_t j = A.colidx[q];
            if (j != k && an.colour[an.perm[j]] == an.colour[an.perm[k]])
                throw std::runtime_error("colouring invalid: coupled rows share a colour");
        }

    // L (strictly lower) and U (strictly upper) of the permuted matrix
    auto split = [&](MatPattern &P, bool lower) {
        P.n = n;
        P.rowptr.assign(n + 1, 0);
        for (int k = 0; k < n; ++k) {
            int cnt = 0;
            for (int32_t q = A.rowptr[k]; q < A.rowptr[k + 1]; ++q)
                if (lower ? A.colidx[q] < k : A.colidx[q] > k) ++cnt;
            P.rowptr[k + 1] = P.rowptr[k] + cnt;
        }
        P.colidx.assign((size_t)P.rowptr[n], 0);
        P.src.assign((size_t)P.rowptr[n], 0);
        for (int k = 0; k < n; ++k) {
            int32_t o = P.rowptr[k];
            for (int32_t q = A.rowptr[k]; q < A.rowptr[k + 1]; ++q)
                if (lower ? A.colidx[q] < k : A.colidx[q] > k) {
                    P.colidx[o] = A.colidx[q];
                    P.src[o] = q;
                    ++o;
                }
        }
    };
    split(an.L, true);
    split(an.U, false);
    transpose_pattern(A.rowptr, A.colidx, n, an.At);
    transpose_pattern(an.U.rowptr, an.U.colidx, n, an.Ut);
    transpose_pattern(an.L.rowptr, an.L.colidx, n, an.Lt);

    for (MatPattern *p : {&an.A, &an.L, &an.U, &an.At, &an.Ut, &an.Lt}) build_tiles(*p, an.colour_ptr, m.bs);
}

}  // namespace bda
