// The host-only parts of the C ABI under AddressSanitizer and
// UndefinedBehaviorSanitizer: the grid generator, the colouring, the
// reference's per-colour device format (build, layout, pack, unpack,
// import) and the status / limits encoders, on small systems, followed by
// malformed inputs that must come back as error codes without any access
// outside the caller's arrays.  Compiled with g++ from the library's host
// sources (no device code) and run by tests/test_host_sanitize.py; prints
// one line per failed check and exits with the number of failures (any
// sanitizer report aborts the program instead).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/bda_bicgstab.h"

namespace {

int g_fails = 0;
#define CHECK(c)                                                                   \
    do {                                                                           \
        if (!(c)) {                                                                \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);      \
            ++g_fails;                                                             \
        }                                                                          \
    } while (0)

struct Sys {
    int32_t n = 0, nnzb = 0, bs = 0;
    std::vector<int32_t> rowptr, colidx;
    std::vector<double> val, b;
};

Sys make_grid(int nx, int ny, int nz, int bs, uint64_t seed, double inactive, double nnc) {
    bda_grid_spec s;
    std::memset(&s, 0, sizeof(s));
    s.nx = nx;
    s.ny = ny;
    s.nz = nz;
    s.bs = bs;
    s.seed = seed;
    s.perm_sigma = 2.0;
    s.layer_corr = 0.5;
    s.inactive_frac = inactive;
    s.nnc_frac = nnc;
    s.nnc_dist = 3;
    s.pert_max = 0.2;
    s.upwind = 0.3;
    s.diag_noise = 0.05;
    bda_grid g = nullptr;
    Sys m;
    CHECK(bda_grid_create(&s, &g) == 0 && g);
    if (!g) return m;
    CHECK(bda_grid_sizes(g, &m.n, &m.nnzb, &m.bs) == 0);
    m.rowptr.resize(m.n + 1);
    m.colidx.resize(m.nnzb);
    m.val.resize((size_t)m.nnzb * bs * bs);
    m.b.resize((size_t)m.n * bs);
    CHECK(bda_grid_copy(g, m.rowptr.data(), m.colidx.data(), m.val.data(), m.b.data()) == 0);
    CHECK(bda_grid_destroy(g) == 0);
    return m;
}

struct Colouring {
    std::vector<int32_t> colour, perm, cptr;
    int32_t nc = 0;
};

Colouring colour_of(const Sys &m, int level) {
    Colouring c;
    c.colour.resize(m.n);
    c.perm.resize(m.n);
    c.cptr.resize(m.n + 1);
    CHECK(bda_colour_pattern(m.n, m.rowptr.data(), m.colidx.data(), level, c.colour.data(), c.perm.data(),
                             c.cptr.data(), &c.nc) == 0);
    c.cptr.resize(c.nc + 1);
    // perm is a permutation grouped by colour, and coupled rows differ in colour
    std::vector<int32_t> seen(m.n, 0);
    for (int32_t k = 0; k < m.n; ++k) {
        const int32_t i = c.perm[k];
        CHECK(i >= 0 && i < m.n && seen[i] == 0);
        if (i < 0 || i >= m.n) continue;
        seen[i] = 1;
        const int32_t col = c.colour[i];
        CHECK(col >= 0 && col < c.nc && k >= c.cptr[col] && k < c.cptr[col + 1]);
    }
    for (int32_t i = 0; i < m.n; ++i)
        for (int32_t q = m.rowptr[i]; q < m.rowptr[i + 1]; ++q)
            if (m.colidx[q] != i) CHECK(c.colour[m.colidx[q]] != c.colour[i]);
    return c;
}

struct Csr {
    std::vector<int32_t> rowptr, colidx;
    std::vector<double> val;
};

// A in colour order, and L / U: its strictly lower / upper blocks (values
// drawn at random: the format does not depend on them being factors)
void colour_ordered(const Sys &m, const Colouring &c, std::mt19937_64 &rng, Csr &A, Csr &L, Csr &U) {
    const int bb = m.bs * m.bs;
    std::vector<int32_t> iperm(m.n);
    for (int32_t k = 0; k < m.n; ++k) iperm[c.perm[k]] = k;
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (Csr *x : {&A, &L, &U}) {
        x->rowptr.assign(1, 0);
        x->colidx.clear();
        x->val.clear();
    }
    for (int32_t k = 0; k < m.n; ++k) {
        const int32_t i = c.perm[k];
        std::vector<std::pair<int32_t, int32_t>> row;  // (new column, old position)
        for (int32_t q = m.rowptr[i]; q < m.rowptr[i + 1]; ++q) row.push_back({iperm[m.colidx[q]], q});
        std::sort(row.begin(), row.end());
        for (const auto &e : row) {
            A.colidx.push_back(e.first);
            A.val.insert(A.val.end(), m.val.begin() + (size_t)e.second * bb, m.val.begin() + (size_t)(e.second + 1) * bb);
            Csr *t = e.first < k ? &L : e.first > k ? &U : nullptr;
            if (t) {
                t->colidx.push_back(e.first);
                for (int j = 0; j < bb; ++j) t->val.push_back(u(rng));
            }
        }
        A.rowptr.push_back((int32_t)A.colidx.size());
        L.rowptr.push_back((int32_t)L.colidx.size());
        U.rowptr.push_back((int32_t)U.colidx.size());
    }
}

// the 21 arrays of the format, sized from bda_wire_sizes
struct Wire {
    std::vector<int32_t> sizes, nv;
    std::vector<std::vector<uint8_t>> arr;   // slot -> bytes
    std::vector<void *> vp;
    double *vals[3] = {nullptr, nullptr, nullptr};

    void allocate(const int32_t s[18], const int32_t nnz_val[3]) {
        sizes.assign(s, s + 18);
        nv.assign(nnz_val, nnz_val + 3);
        arr.assign(21, {});
        for (int k = 0; k < 3; ++k) {
            const int b = 6 * k;
            arr[b + 0].assign((size_t)(8 + 4 * s[b + 2]) * 4, 0);
            arr[b + 1].assign((size_t)std::max(s[b + 3], 1) * 4, 0);
            arr[b + 2].assign((size_t)std::max(nv[k], 1) * 8, 0);
            arr[b + 3].assign((size_t)std::max(s[b + 1], 1) * 2, 0);
            arr[b + 4].assign((size_t)std::max(s[b + 4], 1), 0);
        }
        arr[18].assign((size_t)std::max(s[5], 1) * 8, 0);
        arr[19].assign((size_t)s[0] * 8, 0);
        arr[20].assign((size_t)s[0] * 8, 0);
        pointers();
    }
    void pointers() {
        vp.assign(21, nullptr);
        for (int j = 0; j < 21; ++j) vp[j] = arr[j].data();
        for (int k = 0; k < 3; ++k) {
            vals[k] = reinterpret_cast<double *>(arr[6 * k + 2].data());
            vp[6 * k + 2] = &vals[k];   // the value slots hold double** (one array each)
        }
    }
};

struct Imported {
    std::vector<int32_t> rowptr, colidx, part_rows;
    std::vector<double> val, lu, dinv, b, x0;
};

int import_of(int bs, Wire &w, Imported &o, int32_t n_hint = -1, int64_t nnzb_hint = -1, int32_t parts_hint = -1) {
    int32_t n = n_hint, parts = parts_hint;
    int64_t nnzb = nnzb_hint;
    if (n < 0) {
        const int rc = bda_wire_import_sizes(bs, w.sizes.data(), w.vp.data(), &n, &nnzb, &parts);
        if (rc != 0) return rc;
    }
    const int bb = bs * bs;
    o.rowptr.assign(n + 1, 0);
    o.colidx.assign(std::max<int64_t>(nnzb, 1), 0);
    o.part_rows.assign(parts + 1, 0);
    o.val.assign(std::max<int64_t>(nnzb, 1) * bb, 0.0);
    o.lu.assign(std::max<int64_t>(nnzb, 1) * bb, 0.0);
    o.dinv.assign((size_t)n * bb, 0.0);
    o.b.assign((size_t)n * bs, 0.0);
    o.x0.assign((size_t)n * bs, 0.0);
    return bda_wire_import(bs, w.sizes.data(), w.vp.data(), o.rowptr.data(), o.colidx.data(), o.val.data(),
                           o.lu.data(), o.dinv.data(), o.part_rows.data(), o.b.data(), o.x0.data());
}

void check_case(const char *name, int nx, int ny, int nz, int bs, uint64_t seed, double inactive, double nnc) {
    std::fprintf(stderr, "case %s\n", name);
    const Sys m = make_grid(nx, ny, nz, bs, seed, inactive, nnc);
    if (m.n == 0) return;
    colour_of(m, 1);   // level sets: the same checks
    const Colouring c = colour_of(m, 0);
    std::mt19937_64 rng(seed + 17);
    Csr A, L, U;
    colour_ordered(m, c, rng, A, L, U);
    const int bb = bs * bs;
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> dinv((size_t)m.n * bb), bp((size_t)m.n * bs), xp((size_t)m.n * bs);
    for (double &v : dinv) v = u(rng);
    for (double &v : bp) v = u(rng);
    for (double &v : xp) v = u(rng);

    bda_wire_sizes ws;
    CHECK(bda_wire_build_sizes(bs, m.n, c.cptr.data(), c.nc, A.rowptr.data(), A.colidx.data(), L.rowptr.data(),
                               L.colidx.data(), U.rowptr.data(), U.colidx.data(), &ws) == 0);
    Wire w;
    w.allocate(ws.vector_sizes, ws.nnz_val_sizes);
    CHECK(bda_wire_build(bs, m.n, c.cptr.data(), c.nc, A.rowptr.data(), A.colidx.data(), A.val.data(),
                         L.rowptr.data(), L.colidx.data(), L.val.data(), U.rowptr.data(), U.colidx.data(),
                         U.val.data(), dinv.data(), bp.data(), xp.data(), w.vp.data()) == 0);

    // import gives back the colour-ordered system bit for bit
    Imported o;
    CHECK(import_of(bs, w, o) == 0);
    CHECK(o.rowptr == A.rowptr);
    CHECK(std::vector<int32_t>(o.colidx.begin(), o.colidx.begin() + A.colidx.size()) == A.colidx);
    CHECK(std::memcmp(o.val.data(), A.val.data(), A.val.size() * 8) == 0);
    CHECK(std::memcmp(o.dinv.data(), dinv.data(), dinv.size() * 8) == 0);
    CHECK(std::memcmp(o.b.data(), bp.data(), bp.size() * 8) == 0);
    CHECK(std::memcmp(o.x0.data(), xp.data(), xp.size() * 8) == 0);
    {
        // L and U share A's block positions in lu_values
        std::vector<double> lu(A.val.size(), 0.0);
        for (const Csr *t : {&L, &U})
            for (int32_t i = 0; i < m.n; ++i)
                for (int32_t q = t->rowptr[i]; q < t->rowptr[i + 1]; ++q) {
                    int32_t p = A.rowptr[i];
                    while (A.colidx[p] != t->colidx[q]) ++p;
                    std::memcpy(&lu[(size_t)p * bb], &t->val[(size_t)q * bb], bb * 8);
                }
        CHECK(std::memcmp(o.lu.data(), lu.data(), lu.size() * 8) == 0);
    }

    // pack, unpack, and the unpacked arrays equal the built ones
    uint32_t tot[5];
    uint64_t setup[40];
    CHECK(bda_wire_layout(w.sizes.data(), w.nv.data(), 0, tot, setup) == 0);
    std::vector<std::vector<uint8_t>> bufs(5);
    uint8_t *bp5[5];
    for (int j = 0; j < 5; ++j) {
        bufs[j].assign(std::max<uint32_t>(tot[j], 1), 0xA5);
        bp5[j] = bufs[j].data();
    }
    CHECK(bda_wire_pack(w.sizes.data(), w.nv.data(), 0, w.vp.data(), bp5) == 0);
    int32_t s2[18], nv2[3];
    void *vp2[21];
    double *va2[3];
    CHECK(bda_wire_unpack(bp5, s2, nv2, vp2, va2) == 0);
    CHECK(std::memcmp(s2, w.sizes.data(), sizeof(s2)) == 0);
    CHECK(std::memcmp(nv2, w.nv.data(), sizeof(nv2)) == 0);
    for (int k = 0; k < 3; ++k) {
        const int b = 6 * k;
        // the colour sizes (the 8 header words before them are not part of the packed data)
        CHECK(std::memcmp((const int32_t *)vp2[b + 0] + 8, w.arr[b + 0].data() + 32, (size_t)s2[b + 2] * 16) == 0);
        CHECK(std::memcmp(vp2[b + 1], w.arr[b + 1].data(), (size_t)s2[b + 3] * 4) == 0);
        CHECK(std::memcmp(va2[k], w.arr[b + 2].data(), (size_t)nv2[k] * 8) == 0);
        CHECK(std::memcmp(vp2[b + 3], w.arr[b + 3].data(), (size_t)s2[b + 1] * 2) == 0);
        CHECK(std::memcmp(vp2[b + 4], w.arr[b + 4].data(), (size_t)s2[b + 4]) == 0);
    }
    CHECK(std::memcmp(vp2[18], w.arr[18].data(), (size_t)s2[5] * 8) == 0);
    CHECK(std::memcmp(vp2[19], w.arr[19].data(), (size_t)s2[0] * 8) == 0);
    CHECK(std::memcmp(vp2[20], w.arr[20].data(), (size_t)s2[0] * 8) == 0);

    // malformed streams: each one an error, from import_sizes and from
    // import itself (called with the good system's output sizes)
    int32_t n0, parts0;
    int64_t nnzb0;
    CHECK(bda_wire_import_sizes(bs, w.sizes.data(), w.vp.data(), &n0, &nnzb0, &parts0) == 0);
    auto expect_bad = [&](const char *what, auto &&mutate) {
        Wire x = w;
        x.pointers();
        mutate(x);
        Imported t;
        int32_t n1, parts1;
        int64_t nnzb1;
        const int rs = bda_wire_import_sizes(bs, x.sizes.data(), x.vp.data(), &n1, &nnzb1, &parts1);
        const int ri = import_of(bs, x, t, n0, nnzb0, parts0);
        if (rs == 0 || ri == 0) std::fprintf(stderr, "  accepted: %s (sizes %d, import %d)\n", what, rs, ri);
        CHECK(rs != 0 && ri != 0);
    };
    for (int k = 0; k < 3; ++k) {
        const int b = 6 * k;
        if (w.sizes[b + 2] == 0) continue;
        int32_t *hdr = reinterpret_cast<int32_t *>(w.arr[b].data());
        const int32_t col0 = hdr[8 + 2], val0 = hdr[8 + 3];
        if (col0 > 0)
            expect_bad("P index past N", [&](Wire &x) { reinterpret_cast<uint32_t *>(x.arr[b + 1].data())[0] = (uint32_t)m.n * bs; });
        if (val0 > 0) {
            expect_bad("column slot past the gathered list",
                       [&](Wire &x) { reinterpret_cast<uint16_t *>(x.arr[b + 3].data())[0] = (uint16_t)col0; });
            expect_bad("row before the partition", [&](Wire &x) { x.arr[b + 4][0] = 0; });
            expect_bad("row past the partition", [&](Wire &x) { x.arr[b + 4][val0 - 1] = 255; });
        }
        expect_bad("partition rows past N", [&](Wire &x) { reinterpret_cast<int32_t *>(x.arr[b].data())[8 + 1] = m.n * bs + 1; });
        expect_bad("partition values past the stream",
                   [&](Wire &x) { reinterpret_cast<int32_t *>(x.arr[b].data())[8 + 3] = x.sizes[b + 1] + 1; });
        expect_bad("negative partition size", [&](Wire &x) { reinterpret_cast<int32_t *>(x.arr[b].data())[8 + 2] = -1; });
    }
    expect_bad("D^-1 array too short", [&](Wire &x) { x.sizes[5] = 0; });
    expect_bad("N not a multiple of the block size", [&](Wire &x) { x.sizes[0] -= 1; });
    expect_bad("A's partitions short of N", [&](Wire &x) { x.sizes[2] -= 1; });
    {
        // read with the other block size: rejected when N is not a multiple
        // of it, else decoded into arrays sized by that reading's own sizes
        Imported t;
        const int other = bs == 3 ? 4 : 3;
        const int rc = import_of(other, w, t);
        CHECK(rc != 0 || (m.n * bs) % other == 0);
    }
}

void check_colour_inputs() {
    std::fprintf(stderr, "case malformed patterns\n");
    int32_t colour[4], perm[4], cptr[5], nc;
    const int32_t rp_ok[4] = {0, 2, 4, 5};
    const int32_t ci_ok[5] = {0, 1, 0, 1, 2};
    CHECK(bda_colour_pattern(3, rp_ok, ci_ok, 0, colour, perm, cptr, &nc) == 0);
    const int32_t rp_dec[4] = {0, 5, 2, 5};       // rowptr decreasing inside
    CHECK(bda_colour_pattern(3, rp_dec, ci_ok, 0, colour, perm, cptr, &nc) != 0);
    const int32_t rp_neg[4] = {0, 2, 4, -1};      // negative block count
    CHECK(bda_colour_pattern(3, rp_neg, ci_ok, 0, colour, perm, cptr, &nc) != 0);
    const int32_t rp_off[4] = {1, 2, 4, 5};       // rowptr[0] != 0
    CHECK(bda_colour_pattern(3, rp_off, ci_ok, 0, colour, perm, cptr, &nc) != 0);
    const int32_t ci_range[5] = {0, 1, 0, 7, 2};  // column past n
    CHECK(bda_colour_pattern(3, rp_ok, ci_range, 1, colour, perm, cptr, &nc) != 0);
    const int32_t ci_negc[5] = {0, 1, -3, 1, 2};  // negative column
    CHECK(bda_colour_pattern(3, rp_ok, ci_negc, 0, colour, perm, cptr, &nc) != 0);
    const int32_t ci_order[5] = {1, 0, 0, 1, 2};  // columns not increasing
    CHECK(bda_colour_pattern(3, rp_ok, ci_order, 0, colour, perm, cptr, &nc) != 0);
    CHECK(bda_colour_pattern(0, rp_ok, ci_ok, 0, colour, perm, cptr, &nc) != 0);

    bda_grid_spec s;
    std::memset(&s, 0, sizeof(s));
    bda_grid g = nullptr;
    s.nx = s.ny = s.nz = 4;
    s.bs = 5;                                     // unsupported block size
    CHECK(bda_grid_create(&s, &g) != 0 && g == nullptr);
    s.bs = 3;
    s.nx = 0;                                     // empty grid
    CHECK(bda_grid_create(&s, &g) != 0 && g == nullptr);
}

void check_encoders() {
    std::fprintf(stderr, "case status and limits lines\n");
    bda_result r;
    std::memset(&r, 0, sizeof(r));
    r.half_iterations = 37;
    r.iterations = 19.0;
    r.last_norm_idx = 2;
    for (int i = 0; i < 4; ++i) r.norms[i] = 1.0 / (1 + i);
    r.converged = 1;
    r.solve_ms = 12.5;
    uint64_t line[16];
    CHECK(bda_encode_status(&r, line) == 0);
    CHECK((line[7] >> 40) == 0x414442ULL);   // "BDA" in bits 511..488
    CHECK(bda_encode_status(nullptr, line) != 0);
    bda_limits lim;
    std::memset(&lim, 0, sizeof(lim));
    lim.max_block_size = 4;
    lim.max_blocks_per_row = 64;
    lim.max_iterations = 32767;
    lim.rows_per_tile = 64;
    lim.device_memory_bytes = 288LL << 30;
    lim.compute_units = 256;
    lim.wavefront_size = 64;
    uint64_t q[8];
    CHECK(bda_encode_query(&lim, 1000, 7000, 3, q) == 0);
    CHECK(bda_encode_query(nullptr, 1000, 7000, 3, q) != 0);
}

}  // namespace

int main() {
    check_case("grid12x10x8 3x3 with fault connections", 12, 10, 8, 3, 3, 0.1, 0.1);
    check_case("grid9x8x6 4x4 with fault connections", 9, 8, 6, 4, 5, 0.1, 0.1);
    check_case("grid6x5x4 3x3 plain", 6, 5, 4, 3, 7, 0.0, 0.0);
    check_colour_inputs();
    check_encoders();
    std::fprintf(stderr, "%d failed checks\n", g_fails);
    return g_fails == 0 ? 0 : 1;
}
