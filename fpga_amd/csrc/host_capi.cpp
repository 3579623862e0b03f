// extern "C" access to the host-only parts: the synthetic grid generator and
// the colouring / colour-ordered permutation (no device needed).
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "../../include/bda_bicgstab.h"
#include "bda_host.hpp"

struct bda_grid_s {
    bda::Grid g;
};

// Run-time counterpart of the reference's compile-time BDA_DEBUG_LEVEL
// printouts (bda_utils.hpp:46-57): the same variable name, read from the
// environment once, and a setter, so that a built library can be asked for
// its diagnostics without a rebuild.
namespace bda {
namespace {
std::atomic<int> g_debug_level{-1};
}  // namespace

int debug_level() {
    int v = g_debug_level.load(std::memory_order_relaxed);
    if (v < 0) {
        const char *e = std::getenv("BDA_DEBUG_LEVEL");
        v = e ? std::atoi(e) : 0;
        v = v < 0 ? 0 : v;
        int expected = -1;
        if (!g_debug_level.compare_exchange_strong(expected, v)) v = expected;
    }
    return v;
}

void set_debug_level(int level) { g_debug_level.store(level < 0 ? 0 : level, std::memory_order_relaxed); }

void info(int level, const char *fmt, ...) {
    if (debug_level() < level) return;
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    std::fprintf(stderr, "%s\n", buf);
    std::fflush(stderr);
}
}  // namespace bda

extern "C" {

int bda_set_debug_level(int32_t level) {
    if (level < 0) return 2;
    bda::set_debug_level(level);
    return 0;
}

int bda_get_debug_level(void) { return bda::debug_level(); }

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
        if (rowptr[0] != 0 || rowptr[n] < 0) return 2;
        for (int32_t i = 0; i < n; ++i)
            if (rowptr[i + 1] < rowptr[i]) return 2;
        m.colidx.assign(colidx, colidx + rowptr[n]);
        bda::validate_pattern(m);
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

// ---------------------------------------------------------------------------
// Status and limits lines in the layout of the reference kernel's debug
// port, so that its host decoder reads them unchanged.
//   status line 0 (bicgstab_solver_wrapper_2r_3r3w_simple-ddr_v1.vhd:75-82):
//     bit 0 aborted, bit 1 no results, bits 95..64 cycles, bits 511..488 "BDA"
//   status line 1 (decoded at bicgstab_utils.cpp:138-259): word 0 overflow
//     flags, word 2 bits 63..48 solver state, word 3 bits 31..16 line count
//     and 47..32 half-iteration count, words 4..7 the four norms
//   limits line (wrapper vhd:51-73, decoded at bicgstab_utils.cpp:282-369)
namespace {
constexpr uint64_t kSignature = 0x414442ULL;   // "BDA"
constexpr double kKernelClockHz = 250e6;      // the reference kernel clock (ooc.xdc:8)
uint64_t dbits(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    return u;
}
}  // namespace

int bda_encode_status(const bda_result *r, uint64_t line[16]) {
    if (!r || !line) return 2;
    std::memset(line, 0, 16 * sizeof(uint64_t));
    line[0] = (r->aborted ? 1ULL : 0ULL) | (r->no_results ? 2ULL : 0ULL);
    double cyc = r->solve_ms * 1e-3 * kKernelClockHz;
    if (!(cyc >= 0.0)) cyc = 0.0;
    line[1] = cyc >= 4294967295.0 ? 0xFFFFFFFFULL : (uint64_t)std::llround(cyc);
    line[7] = kSignature << 40;
    line[8] = 0;                                   // no overflow / underflow flags
    line[10] = 0xCULL << 48;                       // solver state write_debug (solver_pkg.vhd:69)
    line[11] = (1ULL << 16) | ((uint64_t)(uint16_t)r->half_iterations << 32);
    for (int i = 0; i < 4; ++i) line[12 + i] = dbits(r->norms[i]);
    return 0;
}

int bda_encode_query(const bda_limits *lim, int32_t n, int64_t nnzb_max, int32_t bs, uint64_t line[8]) {
    if (!lim || !line || n <= 0 || nnzb_max < 0 || (bs != 3 && bs != 4)) return 2;
    std::memset(line, 0, 8 * sizeof(uint64_t));
    const uint64_t rows = (uint64_t)n * bs;
    auto sat32 = [](uint64_t v) { return v > 0xFFFFFFFFULL ? 0xFFFFFFFFULL : v; };
    line[0] = sat32(rows) | (sat32(rows) << 32);                 // x vector elements, max rows
    line[1] = sat32(rows) | (sat32((uint64_t)n) << 32);          // max gathered columns, max colours
    const uint64_t nnz_row = (uint64_t)lim->max_blocks_per_row * bs;
    line[2] = (nnz_row > 0xFFFF ? 0xFFFFULL : nnz_row) | (sat32((uint64_t)nnzb_max * bs * bs) << 16);
    line[6] = 0ULL                                               // no URAM
              | (1ULL << 1)                                      // ILU0 results written
              | ((uint64_t)512 << 16)                            // data path width in bits
              | ((uint64_t)(lim->wavefront_size & 0xFF) << 56);  // parallel multipliers per wave
    line[7] = 1ULL | (1ULL << 4) | (kSignature << 40);           // one read / one write stream
    return 0;
}
