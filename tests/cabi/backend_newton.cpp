// A Newton-style driver of bda_backend::Mi355xBackend (include/bda_backend.hpp),
// compiled as a plain C++ program against the C ABI, as a simulator's solver
// backend would be.  Used by tests/test_gpu_backend.py.
//
//   backend_newton nx ny nz bs out_dir
//
// Step 1: the synthetic grid's system A1 x1 = b1 (upload, colour, factor, solve).
// Step 2: new values on the same pattern, A2 = 1.25 A1 + 0.5 I, and b2 = b1 + 0.1
//         (values copied, factor, solve; no new analysis).
// Step 3: the same system handed over pre-coloured and pre-factorised, as the
//         reference's caller hands it to its kernel (colour-ordered A, L, U and
//         D^-1 read back from step 2's handle): solve_prefactored must give
//         step 2's solution bit for bit.
// Step 4: a larger grid (nx + 2): a new pattern on the same backend object.
// x1, x2, x4 (caller's ordering) go to out_dir/x1.bin, x2.bin, x4.bin; one
// line per step is printed.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bda_backend.hpp"

namespace {

struct System {
    int n = 0, nnzb = 0, bs = 0;
    std::vector<int32_t> rowptr, colidx;
    std::vector<double> values, b;
};

bool make_grid(int nx, int ny, int nz, int bs, System &s) {
    bda_grid_spec spec{};
    spec.nx = nx;
    spec.ny = ny;
    spec.nz = nz;
    spec.bs = bs;
    spec.seed = 0;
    spec.perm_sigma = 2.0;
    spec.layer_corr = 0.5;
    spec.inactive_frac = 0.1;
    spec.nnc_frac = 0.01;
    spec.nnc_dist = 3;
    spec.pert_max = 0.2;
    spec.upwind = 0.3;
    spec.diag_noise = 0.05;
    bda_grid g = nullptr;
    if (bda_grid_create(&spec, &g) != 0) return false;
    bda_grid_sizes(g, &s.n, &s.nnzb, &s.bs);
    s.rowptr.resize(s.n + 1);
    s.colidx.resize(s.nnzb);
    s.values.resize((size_t)s.nnzb * s.bs * s.bs);
    s.b.resize((size_t)s.n * s.bs);
    bda_grid_copy(g, s.rowptr.data(), s.colidx.data(), s.values.data(), s.b.data());
    bda_grid_destroy(g);
    return true;
}

bool write(const std::string &path, const std::vector<double> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(v.data(), sizeof(double), v.size(), f) == v.size();
    std::fclose(f);
    return ok;
}

int report(const char *step, bda_backend::SolverStatus st, const bda_backend::Stats &s,
           const bda_backend::Mi355xBackend &be) {
    std::printf("%s status=%d converged=%d iterations=%d half=%d reduction=%.3e elapsed=%.6f%s%s\n", step, (int)st,
                (int)s.converged, s.iterations, s.half_iterations, s.reduction, s.elapsed,
                st == bda_backend::SolverStatus::Success ? "" : " error=", be.last_error().c_str());
    std::fflush(stdout);
    return st == bda_backend::SolverStatus::Success && s.converged ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s nx ny nz bs out_dir\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]), bs = std::atoi(argv[4]);
    const std::string out = argv[5];
    System s;
    if (!make_grid(nx, ny, nz, bs, s)) return 2;
    const size_t N = (size_t)s.n * bs, bb = (size_t)bs * bs;
    bda_backend::Mi355xBackend be(bs, 1000, 1e-10);
    bda_backend::Stats st;
    std::vector<double> x(N);
    int bad = 0;

    // step 1
    bad |= report("step1", be.solve_system(s.n, s.nnzb, s.rowptr.data(), s.colidx.data(), s.values.data(),
                                           s.b.data(), st), st, be);
    be.get_result(x.data());
    if (!write(out + "/x1.bin", x)) return 2;

    // step 2: new values and right-hand side on the same pattern
    std::vector<double> v2(s.values), b2(s.b);
    for (auto &v : v2) v *= 1.25;
    for (int i = 0; i < s.n; ++i)
        for (int32_t q = s.rowptr[i]; q < s.rowptr[i + 1]; ++q)
            if (s.colidx[q] == i)
                for (int r = 0; r < bs; ++r) v2[(size_t)q * bb + r * bs + r] += 0.5;
    for (auto &v : b2) v += 0.1;
    bda_handle h_before = be.handle();
    bad |= report("step2", be.solve_system(s.n, s.nnzb, s.rowptr.data(), s.colidx.data(), v2.data(), b2.data(), st),
                  st, be);
    if (be.handle() != h_before) {
        std::printf("step2 reallocated the handle\n");
        bad = 1;
    }
    std::vector<double> x2(N);
    be.get_result(x2.data());
    if (!write(out + "/x2.bin", x2)) return 2;

    // step 3: step 2's system, coloured and factorised by the caller
    bda_handle h = be.handle();
    int32_t nc = 0, nl = 0, nu = 0;
    if (bda_num_colours(h, &nc) != 0 || bda_factor_sizes(h, &nl, &nu) != 0) return 2;
    std::vector<int32_t> colour(s.n), perm(s.n), cptr(nc + 1), iperm(s.n);
    bda_get_colouring(h, colour.data(), perm.data(), cptr.data());
    for (int k = 0; k < s.n; ++k) iperm[perm[k]] = k;
    std::vector<int32_t> lr(s.n + 1), lc(nl), ur(s.n + 1), uc(nu);
    std::vector<double> lv((size_t)nl * bb), uv((size_t)nu * bb), dinv((size_t)s.n * bb);
    bda_get_factors(h, lr.data(), lc.data(), lv.data(), ur.data(), uc.data(), uv.data(), dinv.data());
    // colour-ordered A (columns sorted) and L | U on its pattern
    std::vector<int32_t> pr(s.n + 1, 0), pc;
    std::vector<double> pv, plu;
    for (int k = 0; k < s.n; ++k) {
        const int i = perm[k];
        std::vector<std::pair<int32_t, int32_t>> row;
        for (int32_t q = s.rowptr[i]; q < s.rowptr[i + 1]; ++q) row.emplace_back(iperm[s.colidx[q]], q);
        std::sort(row.begin(), row.end());
        int32_t ql = lr[k], qu = ur[k];
        for (auto &e : row) {
            pc.push_back(e.first);
            pv.insert(pv.end(), &v2[(size_t)e.second * bb], &v2[(size_t)e.second * bb] + bb);
            if (e.first < k) {
                plu.insert(plu.end(), &lv[(size_t)ql * bb], &lv[(size_t)ql * bb] + bb);
                ++ql;
            } else if (e.first > k) {
                plu.insert(plu.end(), &uv[(size_t)qu * bb], &uv[(size_t)qu * bb] + bb);
                ++qu;
            } else {
                plu.insert(plu.end(), bb, 0.0);
            }
        }
        pr[k + 1] = (int32_t)pc.size();
    }
    std::vector<double> b3(N);
    for (int k = 0; k < s.n; ++k)
        for (int r = 0; r < bs; ++r) b3[(size_t)k * bs + r] = b2[(size_t)perm[k] * bs + r];
    bad |= report("step3", be.solve_prefactored(s.n, s.nnzb, nc, cptr.data(), pr.data(), pc.data(), pv.data(),
                                                plu.data(), dinv.data(), b3.data(), st), st, be);
    std::vector<double> x3(N);
    be.get_result(x3.data());
    size_t diff = 0;
    for (int k = 0; k < s.n; ++k)
        for (int r = 0; r < bs; ++r)
            if (std::memcmp(&x3[(size_t)k * bs + r], &x2[(size_t)perm[k] * bs + r], sizeof(double)) != 0) ++diff;
    std::printf("step3 entries differing from step2: %zu\n", diff);
    if (diff) bad = 1;

    // step 4: a larger grid, a new pattern on the same object
    System s4;
    if (!make_grid(nx + 2, ny, nz, bs, s4)) return 2;
    bad |= report("step4", be.solve_system(s4.n, s4.nnzb, s4.rowptr.data(), s4.colidx.data(), s4.values.data(),
                                           s4.b.data(), st), st, be);
    std::vector<double> x4((size_t)s4.n * bs);
    be.get_result(x4.data());
    if (!write(out + "/x4.bin", x4)) return 2;
    std::printf("done bad=%d\n", bad);
    return bad;
}
