// Synthetic block systems of a reservoir-like 3D grid (7-point stencil plus
// optional non-neighbour connections), used as benchmark and test input.
//
// Random numbers: std::mt19937_64, one independent engine per purpose,
// seeded with seed ^ (0x9E3779B97F4A7C15 * (stream + 1)):
//   stream 0  per-layer normals           (nz values)
//   stream 1  per-cell normals            (natural cell order)
//   stream 2  inactive-cell draw          (natural cell order)
//   stream 3  non-neighbour connections   (active cells, natural order)
//   stream 4  off-diagonal perturbations  (connection order, bs*bs per block)
//   stream 5  diagonal noise              (row order, bs*bs per row)
//   stream 6  right-hand side             (n*bs values)
// uniform() = (next() >> 11) * 2^-53 in [0,1); normal() = Box-Muller
// sqrt(-2 ln(1-u1)) cos(2 pi u2) from two uniforms.
//
// Cells are numbered naturally, c = i + nx*(j + ny*k); removing inactive
// cells keeps that order.  Faces are visited per active cell in the order
// +x, +y, +z; non-neighbour connections follow, one per drawn cell.
//   log K   = sigma * (sqrt(corr) * z_layer + sqrt(1-corr) * z_cell)
//   T_ab    = 2 K_a K_b / (K_a + K_b)                (harmonic mean)
//   A[a][b] = -T_ab (I + P_ab)                        (a < b)
//   A[b][a] = -T_ab (I + P_ba) - upwind * T_ab * I    (higher-index side)
//   A[a][a] = diag(rowsum_r |offdiag| + 1) + E,  E_rc ~ N(0, diag_noise)
//   b       ~ U[-1, 1]
#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>
#include <unordered_set>

#include "bda_host.hpp"

namespace bda {
namespace {

struct Rng {
    std::mt19937_64 e;
    Rng(uint64_t seed, int stream)
        : e(seed ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(stream + 1))) {}
    double uniform() { return (double)(e() >> 11) * 0x1.0p-53; }
    double normal() {
        const double u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Conn {
    int32_t a, b;  // compressed indices, a < b
    double t;
};

}  // namespace

void generate_grid(const GridSpec &s, Grid &g) {
    if (s.nx <= 0 || s.ny <= 0 || s.nz <= 0) throw std::runtime_error("grid dimensions must be positive");
    if (s.bs != 3 && s.bs != 4) throw std::runtime_error("block size must be 3 or 4");
    const int64_t ncell = (int64_t)s.nx * s.ny * s.nz;
    if (ncell > (int64_t)1 << 30) throw std::runtime_error("grid too large");
    const int bs = s.bs, bb = bs * bs;

    // permeability
    std::vector<double> perm((size_t)ncell);
    {
        Rng rl(s.seed, 0), rc(s.seed, 1);
        std::vector<double> zl(s.nz);
        for (int k = 0; k < s.nz; ++k) zl[k] = rl.normal();
        const double wl = std::sqrt(s.layer_corr), wc = std::sqrt(1.0 - s.layer_corr);
        for (int64_t c = 0; c < ncell; ++c) {
            const int k = (int)(c / ((int64_t)s.nx * s.ny));
            perm[c] = std::exp(s.perm_sigma * (wl * zl[k] + wc * rc.normal()));
        }
    }
    // active cells
    std::vector<int32_t> cidx((size_t)ncell);
    int32_t n = 0;
    {
        Rng ra(s.seed, 2);
        for (int64_t c = 0; c < ncell; ++c) {
            const bool inactive = ra.uniform() < s.inactive_frac;
            cidx[c] = inactive ? -1 : n++;
        }
    }
    if (n == 0) throw std::runtime_error("no active cells");
    std::vector<int64_t> cell_of(n);
    for (int64_t c = 0; c < ncell; ++c)
        if (cidx[c] >= 0) cell_of[cidx[c]] = c;

    auto harm = [&](int64_t c1, int64_t c2) { return 2.0 * perm[c1] * perm[c2] / (perm[c1] + perm[c2]); };
    std::vector<Conn> conns;
    conns.reserve((size_t)n * 3 + n / 50 + 16);
    const int64_t sx = 1, sy = s.nx, sz = (int64_t)s.nx * s.ny;
    for (int32_t a = 0; a < n; ++a) {
        const int64_t c = cell_of[a];
        const int i = (int)(c % s.nx), j = (int)((c / s.nx) % s.ny), k = (int)(c / sz);
        if (i + 1 < s.nx && cidx[c + sx] >= 0) conns.push_back({a, cidx[c + sx], harm(c, c + sx)});
        if (j + 1 < s.ny && cidx[c + sy] >= 0) conns.push_back({a, cidx[c + sy], harm(c, c + sy)});
        if (k + 1 < s.nz && cidx[c + sz] >= 0) conns.push_back({a, cidx[c + sz], harm(c, c + sz)});
    }
    if (s.nnc_frac > 0.0) {
        Rng rn(s.seed, 3);
        std::unordered_set<uint64_t> seen;
        const int d = s.nnc_dist, w = 2 * d + 1;
        for (int32_t a = 0; a < n; ++a) {
            if (!(rn.uniform() < s.nnc_frac)) continue;
            const int64_t c = cell_of[a];
            const int i = (int)(c % s.nx), j = (int)((c / s.nx) % s.ny), k = (int)(c / sz);
            for (int attempt = 0; attempt < 16; ++attempt) {
                const int dx = (int)(rn.e() % (uint64_t)w) - d;
                const int dy = (int)(rn.e() % (uint64_t)w) - d;
                const int dz = (int)(rn.e() % (uint64_t)w) - d;
                const int man = std::abs(dx) + std::abs(dy) + std::abs(dz);
                if (man <= 1) continue;  // itself or a face neighbour
                const int ii = i + dx, jj = j + dy, kk = k + dz;
                if (ii < 0 || ii >= s.nx || jj < 0 || jj >= s.ny || kk < 0 || kk >= s.nz) continue;
                const int64_t c2 = ii + sy * (jj + (int64_t)s.ny * kk);
                const int32_t b2 = cidx[c2];
                if (b2 < 0) continue;
                const uint64_t key = ((uint64_t)std::min(a, b2) << 32) | (uint32_t)std::max(a, b2);
                if (!seen.insert(key).second) continue;
                conns.push_back({std::min(a, b2), std::max(a, b2), harm(c, c2)});
                break;
            }
        }
    }

    // pattern
    Bcsr &m = g.m;
    m.bs = bs;
    m.n = n;
    m.rowptr.assign(n + 1, 0);
    for (int32_t r = 0; r < n; ++r) m.rowptr[r + 1] = 1;  // diagonal
    for (const Conn &cn : conns) {
        m.rowptr[cn.a + 1]++;
        m.rowptr[cn.b + 1]++;
    }
    for (int32_t r = 0; r < n; ++r) m.rowptr[r + 1] += m.rowptr[r];
    const int64_t nnzb64 = m.rowptr[n];
    if (nnzb64 * bb > ((int64_t)1 << 40)) throw std::runtime_error("too many blocks");
    m.colidx.assign((size_t)nnzb64, 0);
    {
        std::vector<int32_t> fill(m.rowptr.begin(), m.rowptr.end() - 1);
        for (int32_t r = 0; r < n; ++r) m.colidx[fill[r]++] = r;
        for (const Conn &cn : conns) {
            m.colidx[fill[cn.a]++] = cn.b;
            m.colidx[fill[cn.b]++] = cn.a;
        }
        for (int32_t r = 0; r < n; ++r) std::sort(m.colidx.begin() + m.rowptr[r], m.colidx.begin() + m.rowptr[r + 1]);
    }
    auto pos = [&](int32_t r, int32_t col) -> int64_t {
        auto beg = m.colidx.begin() + m.rowptr[r], end = m.colidx.begin() + m.rowptr[r + 1];
        auto it = std::lower_bound(beg, end, col);
        return it - m.colidx.begin();
    };

    g.values.assign((size_t)nnzb64 * bb, 0.0);
    {
        Rng rp(s.seed, 4);
        for (const Conn &cn : conns) {
            double *ab = &g.values[(size_t)pos(cn.a, cn.b) * bb];
            double *ba = &g.values[(size_t)pos(cn.b, cn.a) * bb];
            for (int e = 0; e < bb; ++e) {
                const double pert = s.pert_max * rp.uniform();
                ab[e] = -cn.t * (((e / bs) == (e % bs) ? 1.0 : 0.0) + pert);
            }
            for (int e = 0; e < bb; ++e) {
                const double pert = s.pert_max * rp.uniform();
                ba[e] = -cn.t * (((e / bs) == (e % bs) ? 1.0 : 0.0) + pert);
            }
            for (int r = 0; r < bs; ++r) ba[r * bs + r] -= s.upwind * cn.t;
        }
    }
    {
        Rng rd(s.seed, 5);
        for (int32_t r = 0; r < n; ++r) {
            double rowsum[4] = {0, 0, 0, 0};
            int64_t dpos = -1;
            for (int32_t q = m.rowptr[r]; q < m.rowptr[r + 1]; ++q) {
                if (m.colidx[q] == r) {
                    dpos = q;
                    continue;
                }
                const double *blk = &g.values[(size_t)q * bb];
                for (int rr = 0; rr < bs; ++rr)
                    for (int cc = 0; cc < bs; ++cc) rowsum[rr] += std::fabs(blk[rr * bs + cc]);
            }
            double *d = &g.values[(size_t)dpos * bb];
            for (int e = 0; e < bb; ++e) {
                const int rr = e / bs, cc = e % bs;
                d[e] = (rr == cc ? rowsum[rr] + 1.0 : 0.0) + s.diag_noise * rd.normal();
            }
        }
    }
    g.b.resize((size_t)n * bs);
    {
        Rng rb(s.seed, 6);
        for (auto &v : g.b) v = 2.0 * rb.uniform() - 1.0;
    }
}

}  // namespace bda
