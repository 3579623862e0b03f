// extern "C" access to the host-only parts: the synthetic grid generator and
// the colouring / colour-ordered permutation (no device needed).
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "../../include/bda_bicgstab.h"
#include "bda_host.hpp"

struct bda_grid_s {
    bda::Grid g;
};

extern "C" {

int bda_grid_create(const bda_grid_spec *spec, bda_grid *out) {
    if (!spec || !out) return 2;
    *out = nullptr;
    bda_grid g = new (std::nothrow) bda_grid_s();
    if (!g) return 1;
    bda::GridSpec s;
    s.nx = spec->nx;
    s.ny = spec->ny;
    s.nz = spec->nz;
    s.bs = spec->bs;
    s.seed = spec->seed;
    s.perm_sigma = spec->perm_sigma;
    s.layer_corr = spec->layer_corr;
    s.inactive_frac = spec->inactive_frac;
    s.nnc_frac = spec->nnc_frac;
    s.nnc_dist = spec->nnc_dist;
    s.pert_max = spec->pert_max;
    s.upwind = spec->upwind;
    s.diag_noise = spec->diag_noise;
    try {
        bda::generate_grid(s, g->g);
    } catch (const std::exception &) {
        delete g;
        return 1;
    }
    *out = g;
    return 0;
}

int bda_grid_sizes(bda_grid g, int32_t *n, int32_t *nnzb, int32_t *bs) {
    if (!g) return 2;
    if (n) *n = g->g.m.n;
    if (nnzb) *nnzb = g->g.m.nnzb();
    if (bs) *bs = g->g.m.bs;
    return 0;
}

int bda_grid_copy(bda_grid g, int32_t *rowptr, int32_t *colidx, double *values, double *b) {
    if (!g) return 2;
    const bda::Grid &G = g->g;
    if (rowptr) std::memcpy(rowptr, G.m.rowptr.data(), G.m.rowptr.size() * sizeof(int32_t));
    if (colidx) std::memcpy(colidx, G.m.colidx.data(), G.m.colidx.size() * sizeof(int32_t));
    if (values) std::memcpy(values, G.values.data(), G.values.size() * sizeof(double));
    if (b) std::memcpy(b, G.b.data(), G.b.size() * sizeof(double));
    return 0;
}

int bda_grid_destroy(bda_grid g) {
    delete g;
    return 0;
}

int bda_colour_pattern(int32_t n, const int32_t *rowptr, const int32_t *colidx, int32_t level_scheduling,
                       int32_t *colour, int32_t *perm, int32_t *colour_ptr, int32_t *num_colours) {
    if (n <= 0 || !rowptr || !colidx) return 2;
    try {
        bda::Bcsr m;
        m.bs = 3;
        m.n = n;
        m.rowptr.assign(rowptr, rowptr + n + 1);
        m.colidx.assign(colidx, colidx + rowptr[n]);
        std::vector<int32_t> col;
        const int nc = bda::colour_rows(m, level_scheduling != 0, col);
        if (num_colours) *num_colours = nc;
        if (colour) std::memcpy(colour, col.data(), n * sizeof(int32_t));
        std::vector<int32_t> ptr(nc + 1, 0);
        for (int i = 0; i < n; ++i) ptr[col[i] + 1]++;
        for (int c = 0; c < nc; ++c) ptr[c + 1] += ptr[c];
        if (colour_ptr) std::memcpy(colour_ptr, ptr.data(), (nc + 1) * sizeof(int32_t));
        if (perm) {
            std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
            for (int i = 0; i < n; ++i) perm[fill[col[i]]++] = i;
        }
    } catch (const std::exception &) {
        return 1;
    }
    return 0;
}

}  // extern "C"
