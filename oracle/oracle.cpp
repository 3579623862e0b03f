// CPU oracle of the ILU0-BiCGSTAB path -- TEST INFRASTRUCTURE ONLY.
//
// Restates, in plain C++ (compiled with -ffp-contract=off, OpenMP over the
// rows of a colour), what the reference computes.  Paths are relative to
// linearalgebra/ilu0bicgstab/xilinx/src/.
//
//   BiCGSTAB sequence   rtl/bicgstab_solver_2r-3r3w/pkg/solver_pkg.vhd:120-141 (PROGRAM),
//                       rtl/bicgstab_solver_2r-3r3w/solver.vhd:552-558,785-965
//   vector ops          rtl/bicgstab_solver_2r-3r3w/vector_ops.vhd:290-458,
//                       dot_axpy.vhd:249-275 (out = vect1*scale + vect2, sign by bit flip)
//   L / U sweeps        rtl/bicgstab_solver_2r-3r3w/ilu0.vhd:385-416,
//                       int_read_unit.vhd:287-302 (L colours ascending, U descending)
//   already-solved exit alveo_u280/rtl/pkg/constants.vhd:59 (1e-30)
//   iteration count     sda_app/common/fpga_functions_bicgstab.cpp:922-924 (h/2 + 0.5)
//   norm ring           sda_app/common/bicgstab_utils.cpp:248-258
//
// Colouring, the colour-ordered permutation and the block ILU0 have no source
// in the reference (its caller, opm-simulators, is not vendored); they follow
// the definitions this solver documents: first-fit greedy colouring in
// natural order on the symmetrised block graph, a stable sort by (colour,
// index), IKJ block ILU0 in colour order with Gauss-Jordan (partial pivoting)
// inverses of the diagonal blocks.
//
// Summation orders (the FPGA's adder-tree order cannot be reproduced, see
// rtl/bicgstab_solver_2r-3r3w/dot_reduce.vhd:91-118): a block row's product
// adds, per block, the BS products left to right and then the block sums in
// column order starting from 0.  Dots are summed in fixed chunks of 4096
// elements, chunk sums added in order.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace {

template <typename TV>
struct BMat {
    int n = 0, bs = 0;
    std::vector<int32_t> ptr, col;
    std::vector<TV> val;  // bs*bs row-major per block
};

constexpr double kAlreadySolved = 1e-30;
constexpr int64_t kChunk = 4096;

template <typename A, typename B>
double dot(int64_t N, const A *a, const B *b) {
    const int64_t nch = (N + kChunk - 1) / kChunk;
    std::vector<double> part(nch);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nch; ++c) {
        double s = 0.0;
        const int64_t e = std::min(N, (c + 1) * kChunk);
        for (int64_t i = c * kChunk; i < e; ++i) s += (double)a[i] * (double)b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (double v : part) s += v;
    return s;
}

template <typename TV, typename TX>
inline double mulp(TV a, TX x) {
    if constexpr (std::is_same<TV, float>::value) return (double)(a * (float)x);
    else return (double)a * (double)x;
}

// sum over blocks of row i of (A_ij x_j)[r]
template <typename TV, typename TX>
inline void row_product(const BMat<TV> &m, int i, const TX *x, double *out) {
    const int bs = m.bs, bb = bs * bs;
    for (int r = 0; r < bs; ++r) out[r] = 0.0;
    for (int q = m.ptr[i]; q < m.ptr[i + 1]; ++q) {
        const TV *a = &m.val[(size_t)q * bb];
        const TX *xj = x + (size_t)m.col[q] * bs;
        for (int r = 0; r < bs; ++r) {
            double t = mulp(a[r * bs], xj[0]);
            for (int c = 1; c < bs; ++c) t = t + mulp(a[r * bs + c], xj[c]);
            out[r] += t;
        }
    }
}

template <typename TV, typename TX>
void spmv(const BMat<TV> &m, const TX *x, TX *y) {
    const int bs = m.bs;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m.n; ++i) {
        double acc[4];
        row_product(m, i, x, acc);
        for (int r = 0; r < bs; ++r) y[(size_t)i * bs + r] = (TX)acc[r];
    }
}

// one colour-ordered sweep: out_i = [D_i] (rhs_i - sum_j S_ij out_j) over the
// rows of each colour, colours in the given order
template <typename TV, typename TX>
void sweep(const BMat<TV> &S, const std::vector<int32_t> &cptr, bool ascending, const double *dinv, const TX *rhs,
           TX *out) {
    const int bs = S.bs, bb = bs * bs, nc = (int)cptr.size() - 1;
    for (int ci = 0; ci < nc; ++ci) {
        const int c = ascending ? ci : nc - 1 - ci;
#pragma omp parallel for schedule(static)
        for (int i = cptr[c]; i < cptr[c + 1]; ++i) {
            double acc[4], t[4];
            row_product(S, i, out, acc);
            for (int r = 0; r < bs; ++r) t[r] = (double)rhs[(size_t)i * bs + r] - acc[r];
            if (dinv) {
                const double *D = dinv + (size_t)i * bb;
                for (int r = 0; r < bs; ++r) {
                    double z = 0.0;
                    for (int cc = 0; cc < bs; ++cc) z += D[r * bs + cc] * t[cc];
                    out[(size_t)i * bs + r] = (TX)z;
                }
            } else {
                for (int r = 0; r < bs; ++r) out[(size_t)i * bs + r] = (TX)t[r];
            }
        }
    }
}

// z = M^-1 p with M the two sweeps (S1 colours ascending into y, S2
// descending into z; at most one of them carries D^-1).  A stored vector is
// rounded to TX.  When only the second sweep carries D^-1, the last colour's
// first-sweep value is not stored: that colour has no second-sweep blocks
// (they point to higher colours), so its second-sweep step is the diagonal
// solve alone and is applied to the unrounded value, z_i = D_i (p_i - S1 y)_i.
template <typename TV, typename TX>
void precond(const BMat<TV> &S1, const BMat<TV> &S2, const std::vector<int32_t> &cptr, const double *d1,
             const double *d2, const TX *p, TX *y, TX *z) {
    const int nc = (int)cptr.size() - 1;
    if (d1 || !d2) {
        sweep(S1, cptr, true, d1, p, y);
        sweep(S2, cptr, false, d2, y, z);
        return;
    }
    const std::vector<int32_t> head(cptr.begin(), cptr.end() - 1);   // colours 0..C-2
    sweep(S1, head, true, (const double *)nullptr, p, y);
    const int bs = S1.bs, bb = bs * bs;
#pragma omp parallel for schedule(static)
    for (int i = cptr[nc - 1]; i < cptr[nc]; ++i) {
        double acc[4], t[4];
        row_product(S1, i, y, acc);
        for (int r = 0; r < bs; ++r) t[r] = (double)p[(size_t)i * bs + r] - acc[r];
        const double *D = d2 + (size_t)i * bb;
        for (int r = 0; r < bs; ++r) {
            double v = 0.0;
            for (int c = 0; c < bs; ++c) v += D[r * bs + c] * t[c];
            z[(size_t)i * bs + r] = (TX)v;
        }
    }
    sweep(S2, head, false, d2, y, z);
}

// Gauss-Jordan inverse with partial pivoting (first maximal |pivot|)
bool invert(int bs, const double *a, double *inv) {
    double m[4][4], v[4][4];
    for (int r = 0; r < bs; ++r)
        for (int c = 0; c < bs; ++c) {
            m[r][c] = a[r * bs + c];
            v[r][c] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < bs; ++c) {
        int piv = c;
        double best = std::fabs(m[c][c]);
        for (int r = c + 1; r < bs; ++r)
            if (std::fabs(m[r][c]) > best) {
                best = std::fabs(m[r][c]);
                piv = r;
            }
        if (!(best > 0.0)) return false;
        if (piv != c)
            for (int q = 0; q < bs; ++q) {
                std::swap(m[c][q], m[piv][q]);
                std::swap(v[c][q], v[piv][q]);
            }
        const double p = m[c][c];
        for (int q = 0; q < bs; ++q) {
            m[c][q] = m[c][q] / p;
            v[c][q] = v[c][q] / p;
        }
        for (int r = 0; r < bs; ++r) {
            if (r == c) continue;
            const double f = m[r][c];
            for (int q = 0; q < bs; ++q) {
                m[r][q] = m[r][q] - f * m[c][q];
                v[r][q] = v[r][q] - f * v[c][q];
            }
        }
    }
    for (int r = 0; r < bs; ++r)
        for (int c = 0; c < bs; ++c) inv[r * bs + c] = v[r][c];
    return true;
}

void bmul(int bs, const double *a, const double *b, double *o) {
    for (int r = 0; r < bs; ++r)
        for (int c = 0; c < bs; ++c) {
            double s = a[r * bs] * b[c];
            for (int m = 1; m < bs; ++m) s = s + a[r * bs + m] * b[m * bs + c];
            o[r * bs + c] = s;
        }
}

template <typename TV>
BMat<TV> transpose(const BMat<double> &m) {
    BMat<TV> t;
    const int bs = m.bs, bb = bs * bs;
    t.n = m.n;
    t.bs = bs;
    t.ptr.assign(m.n + 1, 0);
    for (int q = 0; q < m.ptr[m.n]; ++q) t.ptr[m.col[q] + 1]++;
    for (int i = 0; i < m.n; ++i) t.ptr[i + 1] += t.ptr[i];
    t.col.resize(m.ptr[m.n]);
    t.val.resize((size_t)m.ptr[m.n] * bb);
    std::vector<int32_t> fill(t.ptr.begin(), t.ptr.end() - 1);
    for (int i = 0; i < m.n; ++i)
        for (int q = m.ptr[i]; q < m.ptr[i + 1]; ++q) {
            const int j = m.col[q], o = fill[j]++;
            t.col[o] = i;
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) t.val[(size_t)o * bb + c * bs + r] = (TV)m.val[(size_t)q * bb + r * bs + c];
        }
    return t;
}

template <typename TV>
BMat<TV> convert(const BMat<double> &m) {
    BMat<TV> t;
    t.n = m.n;
    t.bs = m.bs;
    t.ptr = m.ptr;
    t.col = m.col;
    t.val.assign(m.val.begin(), m.val.end());
    return t;
}

struct RunState {
    int h = 0, norm_slot = 1, last_idx = 0;
    double norms[4] = {0, 0, 0, 0};
    double last = 0.0;
    bool converged = false, no_results = false, max_iter = false, breakdown = false, nan = false;
    std::vector<double> hist;  // ||r0||, then every half-iteration's norm, in order
    void push(double v) {
        hist.push_back(v);
        norms[norm_slot] = v;
        last_idx = norm_slot;
        norm_slot = norm_slot == 3 ? 1 : norm_slot + 1;
        last = v;
    }
};

template <typename TV>
struct OpSet {
    const BMat<TV> *A, *S1, *S2;
    const double *d1, *d2;  // dinv of sweep 1 / 2 (or null)
    const std::vector<int32_t> *cptr;
};

template <typename T>
inline T rnd(double v) { return (T)v; }
inline bool finite(double v) { return std::isfinite(v); }

// The microprogram of solver_pkg.vhd:120-141 (see Appendix-A form in the
// header comment).  x holds x0 (ignored when zero_x0) and the result.
template <typename TV, typename TX>
RunState bicgstab(const OpSet<TV> &op, int64_t N, const double *rhs, TX *x, bool zero_x0, double tol, int maxit,
                  bool abs_mode) {
    RunState st;
    std::vector<TX> rt(N), r(N), p(N), phat(N), v(N), s(N), shat(N), t(N), y(N);
    // steps 0-1: v = A x0, rt = (-1) v + b, rho = rt.rt
    if (zero_x0) {
#pragma omp parallel for
        for (int64_t i = 0; i < N; ++i) {
            rt[i] = rnd<TX>(rhs[i]);
            x[i] = (TX)0;
        }
    } else {
        spmv(*op.A, x, v.data());
#pragma omp parallel for
        for (int64_t i = 0; i < N; ++i) rt[i] = rnd<TX>(-1.0 * (double)v[i] + rhs[i]);
    }
    double rho = dot(N, rt.data(), rt.data());
    const double n0 = std::sqrt(rho);
    st.norms[0] = n0;
    st.last = n0;
    st.hist.push_back(n0);
    if (!(n0 > (abs_mode ? tol : kAlreadySolved))) {
        st.no_results = true;
        st.nan = !finite(n0);
        return st;
    }
    const double thr = abs_mode ? tol : tol * n0;
    r = rt;
    p = rt;
    double alpha = 0, omega = 0, beta = 0;
    for (;;) {
        // steps 2-3: phat = M^-1 p
        precond(*op.S1, *op.S2, *op.cptr, op.d1, op.d2, p.data(), y.data(), phat.data());
        spmv(*op.A, phat.data(), v.data());                 // step 4
        const double vr = dot(N, v.data(), rt.data());       // step 5
        if (vr == 0.0 || !finite(vr)) {
            (vr == 0.0 ? st.breakdown : st.nan) = true;
            return st;
        }
        alpha = rho / vr;
        const double ma = -alpha;
#pragma omp parallel for
        for (int64_t i = 0; i < N; ++i) {                     // steps 6-8
            x[i] = rnd<TX>(alpha * (double)phat[i] + (double)x[i]);
            s[i] = rnd<TX>(ma * (double)v[i] + (double)r[i]);
        }
        const double ns = std::sqrt(dot(N, s.data(), s.data()));
        st.push(ns);
        if (!finite(ns)) { st.nan = true; return st; }
        if (!(ns > thr)) { st.converged = true; return st; }
        if ((st.h >> 1) == maxit) { st.max_iter = true; return st; }
        st.h += 1;
        precond(*op.S1, *op.S2, *op.cptr, op.d1, op.d2, s.data(), y.data(), shat.data());   // shat = M^-1 s
        spmv(*op.A, shat.data(), t.data());                  // step 9
        const double ts = dot(N, t.data(), s.data());        // step 10
        const double tt = dot(N, t.data(), t.data());
        if (tt == 0.0 || !finite(ts) || !finite(tt)) {
            (tt == 0.0 ? st.breakdown : st.nan) = true;
            return st;
        }
        omega = ts / tt;
        const double mo = -omega;
#pragma omp parallel for
        for (int64_t i = 0; i < N; ++i) {                     // steps 11-12
            x[i] = rnd<TX>(omega * (double)shat[i] + (double)x[i]);
            r[i] = rnd<TX>(mo * (double)t[i] + (double)s[i]);
        }
        const double nr = std::sqrt(dot(N, r.data(), r.data()));
        st.push(nr);
        if (!finite(nr)) { st.nan = true; return st; }
        if (!(nr > thr)) { st.converged = true; return st; }
        st.h += 1;
        const double rho_prev = rho;                          // steps 13-14
        rho = dot(N, rt.data(), r.data());
        if (rho == 0.0 || omega == 0.0) { st.breakdown = true; return st; }
        beta = (rho / rho_prev) * (alpha / omega);
#pragma omp parallel for
        for (int64_t i = 0; i < N; ++i) {                     // steps 15-17
            const double tmp = mo * (double)v[i] + (double)p[i];
            p[i] = rnd<TX>(beta * tmp + (double)r[i]);
        }
    }
}

void fill(const RunState &s, double ms, orc_result *res) {
    if (!res) return;
    std::memset(res, 0, sizeof(*res));
    res->half_iterations = s.h;
    res->iterations = s.h / 2.0 + 0.5;
    res->last_norm_idx = s.last_idx;
    for (int i = 0; i < 4; ++i) res->norms[i] = s.norms[i];
    res->final_norm = s.last;
    res->rel_residual = s.norms[0] > 0 ? s.last / s.norms[0] : 0.0;
    res->converged = s.converged;
    res->no_results = s.no_results;
    res->max_iter_reached = s.max_iter;
    res->breakdown = s.breakdown;
    res->nan_detected = s.nan;
    res->solve_ms = ms;
}

}  // namespace

struct orc_system {
    int bs = 0, n = 0, ncol = 0;
    BMat<double> orig;                 // caller's ordering
    std::vector<int32_t> colour, perm, iperm, cptr;
    BMat<double> A;                    // colour-ordered
    std::vector<int32_t> srcblk;       // A block -> original block
    BMat<double> L, U;                 // strictly lower / upper factors
    std::vector<double> dinv, dinvT;
    bool factored = false;
    BMat<double> At, Ut, Lt;           // adjoint operators (built on demand)
    bool adj_ready = false;
    BMat<float> A32, L32, U32;
    bool mixed_ready = false;
    BMat<float> At32, Ut32, Lt32;      // FP32 transposes (component checks)
    bool adj32_ready = false;
    std::vector<double> last_hist;     // norm history of the last orc_solve / orc_adjoint_solve
};

extern "C" {

void orc_set_threads(int32_t t) { omp_set_num_threads(t > 0 ? t : 1); }
int32_t orc_get_threads(void) { return omp_get_max_threads(); }

int32_t orc_colour(int32_t n, const int32_t *rowptr, const int32_t *colidx, int32_t level_scheduling,
                   int32_t *colour) {
    // symmetrised adjacency as explicit neighbour lists
    std::vector<std::vector<int32_t>> nb(n);
    for (int i = 0; i < n; ++i)
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            const int j = colidx[q];
            if (j < 0 || j >= n) return -1;
            if (j == i) continue;
            nb[i].push_back(j);
            nb[j].push_back(i);
        }
    int ncol = 0;
    if (level_scheduling) {
        for (int i = 0; i < n; ++i) {
            int lvl = 0;
            for (int j : nb[i])
                if (j < i) lvl = std::max(lvl, colour[j] + 1);
            colour[i] = lvl;
            ncol = std::max(ncol, lvl + 1);
        }
        return ncol;
    }
    std::vector<char> used;
    for (int i = 0; i < n; ++i) {
        used.assign(ncol + 1, 0);
        for (int j : nb[i])
            if (j < i) used[colour[j]] = 1;
        int c = 0;
        while (used[c]) ++c;
        colour[i] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

int orc_create(int32_t bs, int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *values,
               int32_t level_scheduling, orc_handle *out) {
    try {
        auto h = std::make_unique<orc_system>();
        h->bs = bs;
        h->n = n;
        const int bb = bs * bs;
        const int64_t nnzb = rowptr[n];
        h->orig.n = n;
        h->orig.bs = bs;
        h->orig.ptr.assign(rowptr, rowptr + n + 1);
        h->orig.col.assign(colidx, colidx + nnzb);
        h->orig.val.assign(values, values + nnzb * bb);
        h->colour.resize(n);
        h->ncol = orc_colour(n, rowptr, colidx, level_scheduling, h->colour.data());
        if (h->ncol <= 0) return 1;
        // stable sort of rows by colour
        h->perm.resize(n);
        std::iota(h->perm.begin(), h->perm.end(), 0);
        std::stable_sort(h->perm.begin(), h->perm.end(),
                         [&](int a, int b) { return h->colour[a] < h->colour[b]; });
        h->iperm.resize(n);
        for (int k = 0; k < n; ++k) h->iperm[h->perm[k]] = k;
        h->cptr.assign(h->ncol + 1, 0);
        for (int i = 0; i < n; ++i) h->cptr[h->colour[i] + 1]++;
        for (int c = 0; c < h->ncol; ++c) h->cptr[c + 1] += h->cptr[c];
        // permuted pattern
        BMat<double> &A = h->A;
        A.n = n;
        A.bs = bs;
        A.ptr.assign(1, 0);
        for (int k = 0; k < n; ++k) {
            const int i = h->perm[k];
            std::vector<std::pair<int, int>> row;
            for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) row.emplace_back(h->iperm[colidx[q]], q);
            std::sort(row.begin(), row.end());
            for (auto &pr : row) {
                A.col.push_back(pr.first);
                h->srcblk.push_back(pr.second);
            }
            A.ptr.push_back((int)A.col.size());
        }
        A.val.resize((size_t)nnzb * bb);
        *out = h.release();
        return orc_update_values(*out, values);
    } catch (...) {
        return 1;
    }
}

int orc_destroy(orc_handle h) {
    delete h;
    return 0;
}

int orc_num_colours(orc_handle h, int32_t *nc) {
    *nc = h->ncol;
    return 0;
}

int orc_get_colouring(orc_handle h, int32_t *colour, int32_t *perm, int32_t *colour_ptr) {
    if (colour) std::memcpy(colour, h->colour.data(), h->n * 4);
    if (perm) std::memcpy(perm, h->perm.data(), h->n * 4);
    if (colour_ptr) std::memcpy(colour_ptr, h->cptr.data(), (h->ncol + 1) * 4);
    return 0;
}

int orc_update_values(orc_handle h, const double *values) {
    const int bb = h->bs * h->bs;
    const int64_t nnzb = h->A.ptr[h->n];
    std::memcpy(h->orig.val.data(), values, nnzb * bb * 8);
    for (int64_t q = 0; q < nnzb; ++q)
        std::memcpy(&h->A.val[q * bb], &values[(int64_t)h->srcblk[q] * bb], bb * 8);
    h->factored = h->adj_ready = h->mixed_ready = h->adj32_ready = false;
    return 0;
}

// IKJ block ILU0 on the colour-ordered matrix; rows of one colour in parallel
int orc_factor(orc_handle h) {
    const int n = h->n, bs = h->bs, bb = bs * bs;
    BMat<double> LU = h->A;
    std::vector<int32_t> diag(n, -1);
    for (int i = 0; i < n; ++i)
        for (int q = LU.ptr[i]; q < LU.ptr[i + 1]; ++q)
            if (LU.col[q] == i) diag[i] = q;
    for (int i = 0; i < n; ++i)
        if (diag[i] < 0) return 1;
    h->dinv.assign((size_t)n * bb, 0.0);
    int bad = 0;
    for (int c = 0; c < h->ncol; ++c) {
#pragma omp parallel
        {
            std::vector<int32_t> pos(n, -1);
#pragma omp for schedule(static) reduction(| : bad)
            for (int i = h->cptr[c]; i < h->cptr[c + 1]; ++i) {
                for (int q = LU.ptr[i]; q < LU.ptr[i + 1]; ++q) pos[LU.col[q]] = q;
                for (int q = LU.ptr[i]; q < diag[i]; ++q) {
                    const int k = LU.col[q];
                    double l[16];
                    bmul(bs, &LU.val[(size_t)q * bb], &h->dinv[(size_t)k * bb], l);
                    std::memcpy(&LU.val[(size_t)q * bb], l, bb * 8);
                    for (int pk = diag[k] + 1; pk < LU.ptr[k + 1]; ++pk) {
                        const int qi = pos[LU.col[pk]];
                        if (qi < 0) continue;
                        double t[16];
                        bmul(bs, l, &LU.val[(size_t)pk * bb], t);
                        for (int e = 0; e < bb; ++e) LU.val[(size_t)qi * bb + e] = LU.val[(size_t)qi * bb + e] - t[e];
                    }
                }
                if (!invert(bs, &LU.val[(size_t)diag[i] * bb], &h->dinv[(size_t)i * bb])) bad |= 1;
                for (int q = LU.ptr[i]; q < LU.ptr[i + 1]; ++q) pos[LU.col[q]] = -1;
            }
        }
    }
    if (bad) return 1;
    auto split = [&](BMat<double> &P, bool lower) {
        P.n = n;
        P.bs = bs;
        P.ptr.assign(1, 0);
        P.col.clear();
        P.val.clear();
        for (int i = 0; i < n; ++i) {
            for (int q = LU.ptr[i]; q < LU.ptr[i + 1]; ++q) {
                const int j = LU.col[q];
                if (lower ? j < i : j > i) {
                    P.col.push_back(j);
                    P.val.insert(P.val.end(), &LU.val[(size_t)q * bb], &LU.val[(size_t)q * bb] + bb);
                }
            }
            P.ptr.push_back((int)P.col.size());
        }
    };
    split(h->L, true);
    split(h->U, false);
    h->dinvT.resize(h->dinv.size());
    for (int i = 0; i < n; ++i)
        for (int r = 0; r < bs; ++r)
            for (int c = 0; c < bs; ++c) h->dinvT[(size_t)i * bb + c * bs + r] = h->dinv[(size_t)i * bb + r * bs + c];
    h->factored = true;
    h->adj_ready = h->mixed_ready = h->adj32_ready = false;
    return 0;
}

int orc_factor_sizes(orc_handle h, int32_t *nl, int32_t *nu) {
    if (!h->factored) return 1;
    *nl = h->L.ptr[h->n];
    *nu = h->U.ptr[h->n];
    return 0;
}

int orc_get_factors(orc_handle h, int32_t *lr, int32_t *lc, double *lv, int32_t *ur, int32_t *uc, double *uv,
                    double *dinv) {
    if (!h->factored) return 1;
    std::memcpy(lr, h->L.ptr.data(), (h->n + 1) * 4);
    std::memcpy(lc, h->L.col.data(), h->L.col.size() * 4);
    std::memcpy(lv, h->L.val.data(), h->L.val.size() * 8);
    std::memcpy(ur, h->U.ptr.data(), (h->n + 1) * 4);
    std::memcpy(uc, h->U.col.data(), h->U.col.size() * 4);
    std::memcpy(uv, h->U.val.data(), h->U.val.size() * 8);
    std::memcpy(dinv, h->dinv.data(), h->dinv.size() * 8);
    return 0;
}

static void ensure_adj(orc_handle h) {
    if (h->adj_ready) return;
    h->At = transpose<double>(h->A);
    h->Ut = transpose<double>(h->U);
    h->Lt = transpose<double>(h->L);
    h->adj_ready = true;
}

static void ensure_mixed(orc_handle h) {
    if (h->mixed_ready) return;
    h->A32 = convert<float>(h->A);
    h->L32 = convert<float>(h->L);
    h->U32 = convert<float>(h->U);
    h->mixed_ready = true;
}

static void ensure_adj32(orc_handle h) {
    if (h->adj32_ready) return;
    ensure_adj(h);
    h->At32 = convert<float>(h->At);
    h->Ut32 = convert<float>(h->Ut);
    h->Lt32 = convert<float>(h->Lt);
    h->adj32_ready = true;
}

// The precision form's operators: FP32 values, vectors of type TX
extern "C++" {
template <typename TX>
static void spmv32(orc_handle h, const TX *x, TX *y, bool transpose) {
    ensure_mixed(h);
    if (transpose) ensure_adj32(h);
    spmv(transpose ? h->At32 : h->A32, x, y);
}

template <typename TX>
static void precond32(orc_handle h, const TX *p, TX *z, bool transpose) {
    ensure_mixed(h);
    std::vector<TX> y((size_t)h->n * h->bs);
    if (transpose) {
        ensure_adj32(h);
        precond(h->Ut32, h->Lt32, h->cptr, h->dinvT.data(), (const double *)nullptr, p, y.data(), z);
    } else {
        precond(h->L32, h->U32, h->cptr, (const double *)nullptr, h->dinv.data(), p, y.data(), z);
    }
}
}  // extern "C++"

int orc_spmv_perm(orc_handle h, const double *x, double *y, int32_t transpose) {
    if (transpose) {
        if (!h->factored) return 1;
        ensure_adj(h);
        spmv(h->At, x, y);
    } else {
        spmv(h->A, x, y);
    }
    return 0;
}

int orc_precond_perm(orc_handle h, const double *p, double *z, int32_t transpose) {
    if (!h->factored) return 1;
    std::vector<double> y((size_t)h->n * h->bs);
    if (transpose) {
        ensure_adj(h);
        precond(h->Ut, h->Lt, h->cptr, h->dinvT.data(), (const double *)nullptr, p, y.data(), z);
    } else {
        precond(h->L, h->U, h->cptr, (const double *)nullptr, h->dinv.data(), p, y.data(), z);
    }
    return 0;
}

// prec 0: FP64 (double vectors); 1: FP32 values and float vectors; 2: FP32
// values and double vectors (each x entry rounded to FP32 in the product)
int orc_spmv_perm_ex(orc_handle h, int32_t prec, const void *x, void *y, int32_t transpose) {
    if (!h->factored && (prec != 0 || transpose)) return 1;
    switch (prec) {
    case 0: return orc_spmv_perm(h, (const double *)x, (double *)y, transpose);
    case 1: spmv32<float>(h, (const float *)x, (float *)y, transpose != 0); return 0;
    case 2: spmv32<double>(h, (const double *)x, (double *)y, transpose != 0); return 0;
    default: return 2;
    }
}

int orc_precond_perm_ex(orc_handle h, int32_t prec, const void *p, void *z, int32_t transpose) {
    if (!h->factored) return 1;
    switch (prec) {
    case 0: return orc_precond_perm(h, (const double *)p, (double *)z, transpose);
    case 1: precond32<float>(h, (const float *)p, (float *)z, transpose != 0); return 0;
    case 2: precond32<double>(h, (const double *)p, (double *)z, transpose != 0); return 0;
    default: return 2;
    }
}

static void permute_in(orc_handle h, const double *in, double *out) {
    const int bs = h->bs;
    for (int k = 0; k < h->n; ++k)
        for (int r = 0; r < bs; ++r) out[(size_t)k * bs + r] = in[(size_t)h->perm[k] * bs + r];
}
static void permute_out(orc_handle h, const double *in, double *out) {
    const int bs = h->bs;
    for (int k = 0; k < h->n; ++k)
        for (int r = 0; r < bs; ++r) out[(size_t)h->perm[k] * bs + r] = in[(size_t)k * bs + r];
}

static int solve_common(orc_handle h, const double *b, double *x, double tol, int32_t maxit, int32_t abs_tol,
                        bool adjoint, orc_result *res) {
    if (!h->factored) return 1;
    if (maxit < 0 || maxit > 32767 || !(tol >= 0.0)) return 2;
    const int64_t N = (int64_t)h->n * h->bs;
    std::vector<double> bp(N), xp(N);
    permute_in(h, b, bp.data());
    permute_in(h, x, xp.data());
    OpSet<double> op;
    if (adjoint) {
        ensure_adj(h);
        op = OpSet<double>{&h->At, &h->Ut, &h->Lt, h->dinvT.data(), nullptr, &h->cptr};
    } else {
        op = OpSet<double>{&h->A, &h->L, &h->U, nullptr, h->dinv.data(), &h->cptr};
    }
    const auto t0 = std::chrono::steady_clock::now();
    RunState st = bicgstab<double, double>(op, N, bp.data(), xp.data(), false, tol, maxit, abs_tol != 0);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    permute_out(h, xp.data(), x);
    fill(st, ms, res);
    h->last_hist = st.hist;
    return 0;
}

int orc_get_history(orc_handle h, double *norms, int32_t capacity, int32_t *count) {
    const int32_t n = (int32_t)h->last_hist.size();
    for (int32_t i = 0; i < std::min(n, capacity); ++i) norms[i] = h->last_hist[i];
    if (count) *count = n;
    return 0;
}

int orc_solve(orc_handle h, const double *b, double *x, double tol, int32_t maxit, int32_t abs_tol,
              orc_result *res) {
    return solve_common(h, b, x, tol, maxit, abs_tol, false, res);
}

int orc_adjoint_solve(orc_handle h, const double *w, double *lam, double tol, int32_t maxit, orc_result *res) {
    return solve_common(h, w, lam, tol, maxit, 0, true, res);
}

}  // extern "C"

template <typename TX>
static RunState mixed_inner(orc_handle h, const double *rhs, std::vector<double> &xo, double inner_tol,
                            int inner_maxit, bool adjoint) {
    const int64_t N = (int64_t)h->n * h->bs;
    const OpSet<float> op = adjoint ? OpSet<float>{&h->At32, &h->Ut32, &h->Lt32, h->dinvT.data(), nullptr, &h->cptr}
                                    : OpSet<float>{&h->A32, &h->L32, &h->U32, nullptr, h->dinv.data(), &h->cptr};
    std::vector<TX> d(N);
    RunState st = bicgstab<float, TX>(op, N, rhs, d.data(), true, inner_tol, inner_maxit, false);
    for (int64_t i = 0; i < N; ++i) xo[i] = xo[i] + (double)d[i];
    return st;
}

// The precision form's refinement for A x = b, or (adjoint) A^T x = b with
// the transposed operators: FP64 A^T for the residual, FP32 transposes for
// the inner runs.
static int solve_mixed_common(orc_handle h, const double *b, double *x, double tol, int32_t max_outer,
                              double inner_tol, int32_t inner_maxit, int32_t vec32, int32_t adaptive, double eta,
                              bool adjoint, orc_result *res) {
    if (!h->factored) return 1;
    if (adjoint) ensure_adj32(h);
    else ensure_mixed(h);
    const BMat<double> &M = adjoint ? h->At : h->A;
    const int64_t N = (int64_t)h->n * h->bs;
    std::vector<double> bp(N), xo(N), ro(N), v(N);
    permute_in(h, b, bp.data());
    permute_in(h, x, xo.data());
    orc_result r;
    std::memset(&r, 0, sizeof(r));
    const auto t0 = std::chrono::steady_clock::now();
    double nrm0 = 0, thr = 0;
    int slot = 1;
    for (int outer = 0;; ++outer) {
        spmv(M, xo.data(), v.data());
        for (int64_t i = 0; i < N; ++i) ro[i] = bp[i] - v[i];
        const double nrm = std::sqrt(dot(N, ro.data(), ro.data()));
        if (outer == 0) {
            nrm0 = nrm;
            r.norms[0] = nrm;
            if (!(nrm > kAlreadySolved)) {
                r.no_results = 1;
                r.nan_detected = std::isnan(nrm);
                break;
            }
            thr = tol * nrm;
        } else {
            r.norms[slot] = nrm;
            r.last_norm_idx = slot;
            slot = slot == 3 ? 1 : slot + 1;
        }
        r.final_norm = nrm;
        if (!std::isfinite(nrm)) { r.nan_detected = 1; break; }
        if (outer > 0 && !(nrm > thr)) { r.converged = 1; break; }
        if (outer == max_outer) { r.max_iter_reached = 1; break; }
        // adaptive: ask only the reduction the outer target still needs
        const double itol = adaptive ? std::min(0.5, std::max(inner_tol, eta * thr / nrm)) : inner_tol;
        RunState st = vec32 ? mixed_inner<float>(h, ro.data(), xo, itol, inner_maxit, adjoint)
                            : mixed_inner<double>(h, ro.data(), xo, itol, inner_maxit, adjoint);
        r.outer_iterations++;
        r.inner_half_iterations += st.h;
        r.iterations += st.h / 2.0 + 0.5;
        if (st.breakdown) r.breakdown = 1;
        if (st.nan) r.nan_detected = 1;
        if (st.no_results) break;
        // a breakdown or NaN before the first half-iteration completed left x
        // unchanged: the next outer step would repeat the same inner run
        if ((st.breakdown || st.nan) && st.h == 0 && st.hist.size() == 1) break;
    }
    r.rel_residual = nrm0 > 0 ? r.final_norm / nrm0 : 0.0;
    r.half_iterations = r.inner_half_iterations;
    r.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    permute_out(h, xo.data(), x);
    if (res) *res = r;
    return 0;
}

extern "C" {

int orc_solve_mixed(orc_handle h, const double *b, double *x, double tol, int32_t max_outer, double inner_tol,
                    int32_t inner_maxit, int32_t vec32, int32_t adaptive, double eta, orc_result *res) {
    return solve_mixed_common(h, b, x, tol, max_outer, inner_tol, inner_maxit, vec32, adaptive, eta, false, res);
}

int orc_adjoint_solve_mixed(orc_handle h, const double *w, double *lam, double tol, int32_t max_outer,
                            double inner_tol, int32_t inner_maxit, int32_t vec32, int32_t adaptive, double eta,
                            orc_result *res) {
    return solve_mixed_common(h, w, lam, tol, max_outer, inner_tol, inner_maxit, vec32, adaptive, eta, true, res);
}

// dL/dA on the pattern, original numbering: grad_ij = -lambda_i x_j^T
int orc_grad(orc_handle h, const double *lam, const double *x, double *grad) {
    const int bs = h->bs, bb = bs * bs;
    for (int i = 0; i < h->n; ++i)
        for (int q = h->orig.ptr[i]; q < h->orig.ptr[i + 1]; ++q) {
            const int j = h->orig.col[q];
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c)
                    grad[(size_t)q * bb + r * bs + c] = -(lam[(size_t)i * bs + r] * x[(size_t)j * bs + c]);
        }
    return 0;
}

}  // extern "C"
