// HIP kernels of the ILU0-BiCGSTAB hot path, written for gfx950 (CDNA4,
// wave64).  Everything here is HBM-bound integer/FP64 streaming work; no
// MFMA is used.
//
// Kernel map (reference op -> kernel):
//   SpMV mode, sparstition.vhd:655-878 / spmvp.vhd:240-383  -> k_stream<SPMV>
//   steps 0-1 (r = b - A x0, rho)  solver_pkg.vhd:122-123     -> k_stream<SUB> + FIN_INIT
//   ILU0_L_fs, ilu0.vhd:412-416                               -> k_stream<SUB> per colour
//   ILU0_U_bs, ilu0.vhd:385-410                               -> k_stream<SUB_DINV> per colour
//   dot1 -> alpha, solver.vhd:793-798                         -> SpMV epilogue + FIN_ALPHA
//   axpy1 + axpy2/norm, solver_pkg.vhd:128-130                -> k_update_xs + FIN_HALF1
//   dot2 -> omega, solver.vhd:831-842                         -> SpMV epilogue + FIN_OMEGA
//   axpy1 + axpy2/norm + dot1 -> beta, solver_pkg.vhd:133-136 -> k_update_xr + FIN_HALF2
//   calc_p, vector_ops.vhd:319-323,429-458                    -> k_update_p
//
// Arithmetic follows the reference's separately rounded multiply-then-add
// (dot_axpy.vhd:249-275): this file is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bda {

// ---------------------------------------------------------------------------
// Device-resident solver state.  The scalar unit of the reference
// (solver.vhd:785-936) lives here so that no host round trip is needed
// between the steps of an iteration.
struct SolverState {
    double rho, rho_prev, alpha, omega, beta;
    double tol, thr, last_norm;
    double norms[4];
    double tot[2];          // raw totals of the last FIN_STORE reduction
    int32_t h;              // half-iteration counter (starts at 0 after step 1)
    int32_t maxit;
    int32_t abs_mode;
    int32_t norm_slot;      // ring slot (1..3) the next norm goes to
    int32_t last_norm_idx;
    int32_t done;
    int32_t converged, no_results, max_iter, breakdown, nan_flag;
    int32_t last_half;      // 1: s is the latest residual, 2: r is
};

enum Fin : int32_t {
    FIN_NONE = 0,
    FIN_STORE = 1,   // tot[] = totals
    FIN_INIT = 2,    // rho = r.r, ||r0||, threshold, "already solved" exit
    FIN_ALPHA = 3,   // alpha = rho / (v.rt)
    FIN_HALF1 = 4,   // ||s||, compare, final-iteration exit
    FIN_OMEGA = 5,   // omega = (t.s) / (t.t)
    FIN_HALF2 = 6,   // ||r||, compare, rho_new = rt.r, beta
};

// Threshold below which the initial residual counts as already solved in
// relative mode (alveo_u280/rtl/pkg/constants.vhd:59).
constexpr double kAlreadySolved = 1e-30;

struct RedArgs {
    double *partials;   // [2][cap]
    int32_t cap;
    SolverState *st;
    int32_t fin;
};

__device__ inline void push_norm(SolverState *st, double nrm) {
    const int slot = st->norm_slot;
    st->norms[slot] = nrm;
    st->last_norm_idx = slot;
    st->norm_slot = slot == 3 ? 1 : slot + 1;
    st->last_norm = nrm;
}

__device__ inline bool finite(double v) { return !__builtin_isnan(v) && !__builtin_isinf(v); }

// Scalar epilogues; run by one thread after a grid-wide reduction.
__device__ inline void finalize(SolverState *st, int fin, double t0, double t1) {
    switch (fin) {
    case FIN_STORE:
        st->tot[0] = t0;
        st->tot[1] = t1;
        break;
    case FIN_INIT: {   // solver.vhd:867-936 with dot_res = rho_new
        const double nrm = sqrt(t0);
        st->rho = t0;
        st->norms[0] = nrm;
        st->norm_slot = 1;
        st->last_norm_idx = 0;
        st->last_norm = nrm;
        st->h = 0;
        st->last_half = 0;
        const double precision = st->abs_mode ? st->tol : kAlreadySolved;
        if (!(nrm > precision)) {
            st->no_results = 1;
            if (!finite(nrm)) st->nan_flag = 1;
            st->done = 1;
        } else {
            st->thr = st->abs_mode ? st->tol : st->tol * nrm;
        }
        break;
    }
    case FIN_ALPHA:    // step 5: divide rho_new by the dot result
        if (t0 == 0.0 || !finite(t0)) {
            if (t0 == 0.0) st->breakdown = 1; else st->nan_flag = 1;
            st->done = 1;
        } else {
            st->alpha = st->rho / t0;
        }
        break;
    case FIN_HALF1: {  // steps 7/8
        const double nrm = sqrt(t0);
        push_norm(st, nrm);
        st->last_half = 1;
        if (!finite(nrm)) { st->nan_flag = 1; st->done = 1; }
        else if (!(nrm > st->thr)) { st->converged = 1; st->done = 1; }
        else if ((st->h >> 1) == st->maxit) { st->max_iter = 1; st->done = 1; }
        else st->h += 1;
        break;
    }
    case FIN_OMEGA:    // step 10
        if (t1 == 0.0 || !finite(t0) || !finite(t1)) {
            if (t1 == 0.0) st->breakdown = 1; else st->nan_flag = 1;
            st->done = 1;
        } else {
            st->omega = t0 / t1;
        }
        break;
    case FIN_HALF2: {  // step 12, then steps 13/14
        const double nrm = sqrt(t0);
        push_norm(st, nrm);
        st->last_half = 2;
        if (!finite(nrm)) { st->nan_flag = 1; st->done = 1; }
        else if (!(nrm > st->thr)) { st->converged = 1; st->done = 1; }
        else {
            st->h += 1;
            st->rho_prev = st->rho;
            st->rho = t1;
            if (t1 == 0.0 || st->omega == 0.0) { st->breakdown = 1; st->done = 1; }
            else st->beta = (st->rho / st->rho_prev) * (st->alpha / st->omega);
        }
        break;
    }
    default:
        break;
    }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Workgroup reduction of NRED values into one partial per workgroup.  The
// grid-wide sum and the scalar epilogue run in k_finalize, launched right
// after the producing kernel: the kernel boundary makes the partials
// visible without any device-scope fence (an agent-scope release on CDNA
// writes back the whole L2 of the XCD, far too costly per workgroup).  The
// partials are summed in a fixed order, so results do not depend on
// workgroup scheduling.
template <int NT, int NRED>
__device__ inline void block_partials(double (&v)[NRED], const RedArgs &ra) {
    __shared__ double sred[NT / 64][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
        v[j] = wave_sum(v[j]);
        if (lane == 0) sred[wave][j] = v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NRED; ++j) {
            double t = sred[0][j];
            for (int w = 1; w < NT / 64; ++w) t += sred[w][j];
            ra.partials[(size_t)j * ra.cap + blockIdx.x] = t;
        }
    }
}

constexpr int kFinThreads = 1024;

template <int NRED>
__global__ __launch_bounds__(kFinThreads) void k_finalize(RedArgs ra, int nparts) {
    __shared__ double sred[kFinThreads / 64][2];
    if (ra.st->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[NRED];
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
        acc[j] = 0.0;
        for (int i = threadIdx.x; i < nparts; i += kFinThreads) acc[j] += ra.partials[(size_t)j * ra.cap + i];
        acc[j] = wave_sum(acc[j]);
        if (lane == 0) sred[wave][j] = acc[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[2] = {0.0, 0.0};
        for (int j = 0; j < NRED; ++j) {
            t[j] = sred[0][j];
            for (int w = 1; w < kFinThreads / 64; ++w) t[j] += sred[w][j];
        }
        finalize(ra.st, ra.fin, t[0], t[1]);
    }
}

// ---------------------------------------------------------------------------
// Streaming block-row kernel.
//
// One workgroup (64*BS threads, BS waves) handles one tile: up to 64
// consecutive block rows of one colour whose blocks are contiguous in the
// value array.  Phase 1 streams the tile's values with one lane per matrix
// entry (7 blocks of 3x3 = 63 lanes, or 4 blocks of 4x4 = 64 lanes per wave
// step), so every wave load instruction reads consecutive memory; each lane
// gathers its x entry, multiplies, and the BS lanes of one block row combine
// their products left to right with cross-lane shuffles.  The per-block row
// partials go to LDS.  Phase 2 gives one thread per scalar row, which adds
// its block partials in column order and applies the epilogue.
//
// Summation order per scalar row r of block row i:
//   sum = 0;  for blocks j in column order: sum += ((a_r0 x_0 + a_r1 x_1) + a_r2 x_2)
enum StreamMode : int { SM_SPMV = 0, SM_SUB = 1, SM_SUB_DINV = 2 };

template <typename TV, typename TX>
struct StreamArgs {
    const int32_t *tile_row;
    int32_t tile0, ntiles;
    const int32_t *rowptr;
    const int32_t *colidx;
    const TV *val;
    const TX *x;        // vector multiplied (gathered through colidx)
    const TX *rhs;      // SUB modes: out = rhs - A x
    TX *out;
    TX *out2;           // optional copies of out (initial residual -> r, p)
    TX *out3;
    const double *dinv; // SUB_DINV: out = dinv * (rhs - A x)
    const TX *w0;       // reduction 0: sum out * w0   (if non-null)
    int32_t red_self;   // reduction 1 (or 0 if w0 null): sum out * out
    int32_t nred;
    RedArgs ra;
};

template <typename TV, typename TX>
__device__ inline double prod(TV a, TX x);
template <>
__device__ inline double prod<double, double>(double a, double x) { return a * x; }
template <>
__device__ inline double prod<float, float>(float a, float x) { return (double)(a * x); }
template <>
__device__ inline double prod<float, double>(float a, double x) { return (double)(a * (float)x); }

template <typename T> __device__ inline T from_double(double v);
template <> __device__ inline double from_double<double>(double v) { return v; }
template <> __device__ inline float from_double<float>(double v) { return (float)v; }

// Map workgroups to tiles so that each XCD (workgroups are dispatched to the
// 8 XCDs round-robin) works on one contiguous range of tiles and the x
// gathers of neighbouring rows hit the same L2.
__device__ inline int xcd_tile(int b, int G) {
    const int q = G >> 3, rem = G & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + (x < rem ? x : rem) + i;
}

// Lane layouts of phase 1:
//   VAR 0  one lane per matrix entry: a wave step reads 7 (3x3) or 4 (4x4)
//          whole blocks as one contiguous run of 63/64 values; the BS lanes
//          of a block row combine their products with cross-lane shuffles.
//   VAR 1  one lane per block row segment (BS consecutive values): a wave
//          step covers 21 (3x3) or 16 (4x4) blocks, each lane multiplies its
//          BS values by the BS gathered x entries on its own.
// Both add the products of a block row segment left to right, so the
// result is the same to the bit.
template <int BS, int MODE, typename TV, typename TX, int VAR>
__global__ __launch_bounds__(64 * BS) void k_stream(StreamArgs<TV, TX> a) {
    constexpr int BB = BS * BS;
    constexpr int NT = 64 * BS;            // threads per workgroup
    constexpr int NW = BS;                 // waves per workgroup
    constexpr int CAP = BS == 3 ? 640 : 480;
    constexpr int UNR = VAR == 0 ? 4 : 2;  // wave steps in flight per lane

    __shared__ double part[CAP * BS];
    __shared__ int32_t scol[CAP];
    __shared__ double tv[MODE == SM_SUB_DINV ? NT : 1];

    if (a.ra.st != nullptr && a.ra.st->done) return;

    const int tile = a.tile0 + xcd_tile(blockIdx.x, a.ntiles);
    const int r0 = a.tile_row[tile], r1 = a.tile_row[tile + 1];
    const int b0 = a.rowptr[r0], b1 = a.rowptr[r1];
    const int nblk = b1 - b0;

    // phase-2 operands that do not depend on phase 1, fetched early
    const int t = threadIdx.x;
    const int k = t / BS, r = t - k * BS;
    const int row = r0 + k;
    const bool has_row = row < r1;
    int q1 = 0, q2 = 0;
    double rhs_v = 0.0;
    double dv[BS];
    if (has_row) {
        q1 = a.rowptr[row] - b0;
        q2 = a.rowptr[row + 1] - b0;
        if constexpr (MODE != SM_SPMV) rhs_v = (double)a.rhs[(size_t)row * BS + r];
        if constexpr (MODE == SM_SUB_DINV) {
#pragma unroll
            for (int c = 0; c < BS; ++c) dv[c] = a.dinv[(size_t)row * BB + r * BS + c];
        }
    }

    for (int i = t; i < nblk; i += NT) scol[i] = a.colidx[b0 + i];
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const TV *val = a.val + (size_t)b0 * BB;
    if constexpr (VAR == 0) {
        constexpr int BPW = 64 / BB;       // blocks per wave step (7 or 4)
        constexpr int VPW = BPW * BB;      // lanes carrying a value (63 or 64)
        const int lb = lane / BB, e = lane - lb * BB, er = e / BS, ec = e - er * BS;
        const bool active_lane = lane < VPW;
        for (int q0 = wave * BPW; q0 < nblk; q0 += NW * BPW * UNR) {
            TV av[UNR];
            TX xv[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int bl = q0 + u * NW * BPW + lb;
                const bool ok = active_lane && bl < nblk;
                av[u] = ok ? val[(size_t)bl * BB + e] : (TV)0;
                xv[u] = ok ? a.x[(size_t)scol[bl] * BS + ec] : (TX)0;
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const double p = prod<TV, TX>(av[u], xv[u]);
                double sm = p;
                const double p1 = __shfl_down(p, 1);
                const double p2 = __shfl_down(p, 2);
                sm = sm + p1;
                sm = sm + p2;
                if constexpr (BS == 4) {
                    const double p3 = __shfl_down(p, 3);
                    sm = sm + p3;
                }
                const int bl = q0 + u * NW * BPW + lb;
                if (active_lane && ec == 0 && bl < nblk) part[bl * BS + er] = sm;
            }
        }
    } else {
        constexpr int SPW = 64 / BS;       // blocks per wave step (21 or 16)
        const int lb = lane / BS, lr = lane - lb * BS;
        const bool active_lane = lane < SPW * BS;
        for (int q0 = wave * SPW; q0 < nblk; q0 += NW * SPW * UNR) {
            TV av[UNR][BS];
            TX xv[UNR][BS];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int bl = q0 + u * NW * SPW + lb;
                const bool ok = active_lane && bl < nblk;
                const int cb = ok ? scol[bl] : 0;
                const TV *ap = val + ((size_t)bl * BS + lr) * BS;
                const TX *xp = a.x + (size_t)cb * BS;
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    av[u][c] = ok ? ap[c] : (TV)0;
                    xv[u][c] = ok ? xp[c] : (TX)0;
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                double sm = prod<TV, TX>(av[u][0], xv[u][0]);
#pragma unroll
                for (int c = 1; c < BS; ++c) sm = sm + prod<TV, TX>(av[u][c], xv[u][c]);
                const int bl = q0 + u * NW * SPW + lb;
                if (active_lane && bl < nblk) part[bl * BS + lr] = sm;
            }
        }
    }
    __syncthreads();

    double res = 0.0;
    if (has_row) {
        double sum = 0.0;
        for (int q = q1; q < q2; ++q) sum += part[q * BS + r];
        if constexpr (MODE == SM_SPMV) res = sum;
        else res = rhs_v - sum;
    }
    if constexpr (MODE == SM_SUB_DINV) {
        tv[t] = res;
        __syncthreads();
        if (has_row) {
            double z = 0.0;
#pragma unroll
            for (int c = 0; c < BS; ++c) z += dv[c] * tv[k * BS + c];
            res = z;
        }
    }
    double red[2] = {0.0, 0.0};
    if (has_row) {
        const size_t gi = (size_t)row * BS + r;
        const TX o = from_double<TX>(res);
        a.out[gi] = o;
        if (a.out2) a.out2[gi] = o;
        if (a.out3) a.out3[gi] = o;
        const double od = (double)o;
        int j = 0;
        if (a.w0) red[j++] = od * (double)a.w0[gi];
        if (a.red_self) red[j] = od * od;
    }
    if (a.nred == 1) {
        double v1[1] = {red[0]};
        block_partials<NT, 1>(v1, a.ra);
    } else if (a.nred == 2) {
        block_partials<NT, 2>(red, a.ra);
    }
}

// ---------------------------------------------------------------------------
// Fused BiCGSTAB vector updates.  Each thread handles pairs of elements;
// vectors are allocated with at least one element of zero padding so that a
// pair load at the odd tail stays inside the allocation.
constexpr int kVecThreads = 256;
constexpr int kVecPairsPerThread = 4;
constexpr int kVecElemsPerBlock = kVecThreads * kVecPairsPerThread * 2;

template <typename T> struct Pair;
template <> struct Pair<double> { using type = double2; };
template <> struct Pair<float> { using type = float2; };

template <typename T>
__device__ inline void ld2(const T *p, size_t i, double &a, double &b) {
    const typename Pair<T>::type v = *reinterpret_cast<const typename Pair<T>::type *>(p + i);
    a = (double)v.x;
    b = (double)v.y;
}
template <typename T>
__device__ inline void st2(T *p, size_t i, double a, double b, bool both) {
    if (both) {
        typename Pair<T>::type v;
        v.x = (T)a;
        v.y = (T)b;
        *reinterpret_cast<typename Pair<T>::type *>(p + i) = v;
    } else {
        p[i] = (T)a;
    }
}
template <typename T> __device__ inline double rnd(double v) { return (double)(T)v; }

// steps 6-8: x += alpha phat; s = r - alpha v; ||s||^2
template <typename TX, typename TXX>
__global__ __launch_bounds__(kVecThreads) void k_update_xs(int64_t N, TXX *x, const TX *phat, const TX *r,
                                                          const TX *v, TX *s, RedArgs ra) {
    if (ra.st->done) return;
    const double alpha = ra.st->alpha, malpha = -alpha;
    double acc = 0.0;
    const int64_t base = (int64_t)blockIdx.x * (kVecElemsPerBlock / 2);
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const int64_t pi = base + threadIdx.x + (int64_t)j * kVecThreads;
        const int64_t i = 2 * pi;
        if (i < N) {
            const bool both = i + 1 < N;
            double x0, x1, p0, p1, r0, r1, v0, v1;
            ld2(x, i, x0, x1);
            ld2(phat, i, p0, p1);
            ld2(r, i, r0, r1);
            ld2(v, i, v0, v1);
            const double nx0 = alpha * p0 + x0, nx1 = alpha * p1 + x1;
            const double s0 = rnd<TX>(malpha * v0 + r0), s1 = rnd<TX>(malpha * v1 + r1);
            st2(x, i, nx0, nx1, both);
            st2(s, i, s0, s1, both);
            acc += s0 * s0;
            if (both) acc += s1 * s1;
        }
    }
    double vv[1] = {acc};
    block_partials<kVecThreads, 1>(vv, ra);
}

// steps 11-14: x += omega shat; r = s - omega t; ||r||^2 and rt.r
template <typename TX, typename TXX>
__global__ __launch_bounds__(kVecThreads) void k_update_xr(int64_t N, TXX *x, const TX *shat, const TX *s,
                                                          const TX *t, TX *r, const TX *rt, RedArgs ra) {
    if (ra.st->done) return;
    const double omega = ra.st->omega, momega = -omega;
    double acc[2] = {0.0, 0.0};
    const int64_t base = (int64_t)blockIdx.x * (kVecElemsPerBlock / 2);
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const int64_t pi = base + threadIdx.x + (int64_t)j * kVecThreads;
        const int64_t i = 2 * pi;
        if (i < N) {
            const bool both = i + 1 < N;
            double x0, x1, h0, h1, s0, s1, t0, t1, q0, q1;
            ld2(x, i, x0, x1);
            ld2(shat, i, h0, h1);
            ld2(s, i, s0, s1);
            ld2(t, i, t0, t1);
            ld2(rt, i, q0, q1);
            const double nx0 = omega * h0 + x0, nx1 = omega * h1 + x1;
            const double r0 = rnd<TX>(momega * t0 + s0), r1 = rnd<TX>(momega * t1 + s1);
            st2(x, i, nx0, nx1, both);
            st2(r, i, r0, r1, both);
            acc[0] += r0 * r0;
            acc[1] += q0 * r0;
            if (both) {
                acc[0] += r1 * r1;
                acc[1] += q1 * r1;
            }
        }
    }
    block_partials<kVecThreads, 2>(acc, ra);
}

// steps 15-17: p = r + beta (p - omega v), as tmp = (-omega) v + p; p = beta tmp + r
template <typename TX>
__global__ __launch_bounds__(kVecThreads) void k_update_p(int64_t N, TX *p, const TX *r, const TX *v,
                                                         const SolverState *st) {
    if (st->done) return;
    const double beta = st->beta, momega = -st->omega;
    const int64_t base = (int64_t)blockIdx.x * (kVecElemsPerBlock / 2);
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const int64_t pi = base + threadIdx.x + (int64_t)j * kVecThreads;
        const int64_t i = 2 * pi;
        if (i < N) {
            const bool both = i + 1 < N;
            double p0, p1, r0, r1, v0, v1;
            ld2(p, i, p0, p1);
            ld2(r, i, r0, r1);
            ld2(v, i, v0, v1);
            const double t0 = momega * v0 + p0, t1 = momega * v1 + p1;
            st2(p, i, beta * t0 + r0, beta * t1 + r1, both);
        }
    }
}

// Steps 0-1 for x0 = 0 (A*0 = 0, so rt = b exactly): rt = r = p = rhs,
// x = 0, rho = rt.rt.  Used by the inner solves of the precision form.
template <typename TX, typename TXX>
__global__ __launch_bounds__(kVecThreads) void k_init_zero_x0(int64_t N, const double *rhs, TX *rt, TX *r, TX *p,
                                                             TXX *x, RedArgs ra) {
    if (ra.st->done) return;
    double acc = 0.0;
    const int64_t base = (int64_t)blockIdx.x * (kVecElemsPerBlock / 2);
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const int64_t pi = base + threadIdx.x + (int64_t)j * kVecThreads;
        const int64_t i = 2 * pi;
        if (i < N) {
            const bool both = i + 1 < N;
            double b0, b1;
            ld2(rhs, i, b0, b1);
            b0 = rnd<TX>(b0);
            b1 = rnd<TX>(b1);
            st2(rt, i, b0, b1, both);
            st2(r, i, b0, b1, both);
            st2(p, i, b0, b1, both);
            st2(x, i, 0.0, 0.0, both);
            acc += b0 * b0;
            if (both) acc += b1 * b1;
        }
    }
    double vv[1] = {acc};
    block_partials<kVecThreads, 1>(vv, ra);
}

// x64 += d (outer refinement update of the precision form)
template <typename TXX>
__global__ __launch_bounds__(kVecThreads) void k_add_correction(int64_t N, double *x, const TXX *d) {
    const int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x;
    if (i < N) x[i] = x[i] + (double)d[i];
}

// ---------------------------------------------------------------------------
// Permutations between the caller's ordering and the colour ordering.
__global__ void k_gather_rows(int n, int bs, const int32_t *perm, const double *in, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * bs) return;
    const int64_t k = i / bs, c = i - k * bs;
    out[i] = in[(int64_t)perm[k] * bs + c];
}
__global__ void k_scatter_rows(int n, int bs, const int32_t *perm, const double *in, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * bs) return;
    const int64_t k = i / bs, c = i - k * bs;
    out[(int64_t)perm[k] * bs + c] = in[i];
}

// dst block k = src block idx[k] (idx null: identity), optionally transposed
// and converted (K17 demotion when TD = float).
template <int BS, typename TD>
__global__ void k_gather_blocks(int64_t nblk, const int32_t *idx, const double *src, TD *dst, int trans) {
    constexpr int BB = BS * BS;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nblk * BB) return;
    const int64_t k = i / BB;
    const int e = (int)(i - k * BB);
    const int er = e / BS, ec = e - er * BS;
    const int se = trans ? ec * BS + er : e;
    const int64_t s = idx ? (int64_t)idx[k] : k;
    dst[i] = (TD)src[s * BB + se];
}

// ---------------------------------------------------------------------------
// Block ILU0, one thread per block row of one colour (rows of a colour are
// not coupled, so they factorise independently once all earlier colours are
// done).  IKJ order on the colour-ordered matrix:
//   for each L block (i,k), k ascending:  L_ik = A_ik * Dinv_k
//       for each U block (k,j) of row k:   A_ij -= L_ik * U_kj   if (i,j) in pattern
//   Dinv_i = inverse of the updated A_ii (Gauss-Jordan, partial pivoting)
template <int BS>
__device__ inline bool invert_block(const double *a, double *inv) {
    double m[BS][BS], v[BS][BS];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            m[r][c] = a[r * BS + c];
            v[r][c] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < BS; ++c) {
        int piv = c;
        double best = fabs(m[c][c]);
        for (int r = c + 1; r < BS; ++r)
            if (fabs(m[r][c]) > best) { best = fabs(m[r][c]); piv = r; }
        if (!(best > 0.0)) return false;
        if (piv != c)
            for (int q = 0; q < BS; ++q) {
                double t = m[c][q]; m[c][q] = m[piv][q]; m[piv][q] = t;
                t = v[c][q]; v[c][q] = v[piv][q]; v[piv][q] = t;
            }
        const double p = m[c][c];
        for (int q = 0; q < BS; ++q) { m[c][q] = m[c][q] / p; v[c][q] = v[c][q] / p; }
        for (int r = 0; r < BS; ++r) {
            if (r == c) continue;
            const double f = m[r][c];
            for (int q = 0; q < BS; ++q) {
                m[r][q] = m[r][q] - f * m[c][q];
                v[r][q] = v[r][q] - f * v[c][q];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) inv[r * BS + c] = v[r][c];
    return true;
}

// out = a * b for BS x BS row-major blocks, entries summed over m ascending
template <int BS>
__device__ inline void block_mul(const double *a, const double *b, double *out) {
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            double s = a[r * BS + 0] * b[0 * BS + c];
#pragma unroll
            for (int m = 1; m < BS; ++m) s = s + a[r * BS + m] * b[m * BS + c];
            out[r * BS + c] = s;
        }
}

template <int BS>
__global__ void k_bilu0_rows(int row0, int row1, const int32_t *rowptr, const int32_t *colidx,
                             const int32_t *diag, double *LU, double *dinv, int32_t *bad_row) {
    constexpr int BB = BS * BS;
    const int i = row0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= row1) return;
    const int rend = rowptr[i + 1], di = diag[i];
    for (int q = rowptr[i]; q < di; ++q) {
        const int kk = colidx[q];
        double a[BB], d[BB], l[BB];
#pragma unroll
        for (int e = 0; e < BB; ++e) { a[e] = LU[(size_t)q * BB + e]; d[e] = dinv[(size_t)kk * BB + e]; }
        block_mul<BS>(a, d, l);
#pragma unroll
        for (int e = 0; e < BB; ++e) LU[(size_t)q * BB + e] = l[e];
        int qi = q + 1;
        for (int pk = diag[kk] + 1; pk < rowptr[kk + 1]; ++pk) {
            const int j = colidx[pk];
            while (qi < rend && colidx[qi] < j) ++qi;
            if (qi >= rend) break;
            if (colidx[qi] != j) continue;
            double u[BB], t[BB];
#pragma unroll
            for (int e = 0; e < BB; ++e) u[e] = LU[(size_t)pk * BB + e];
            block_mul<BS>(l, u, t);
#pragma unroll
            for (int e = 0; e < BB; ++e) LU[(size_t)qi * BB + e] = LU[(size_t)qi * BB + e] - t[e];
        }
    }
    double dd[BB], inv[BB];
#pragma unroll
    for (int e = 0; e < BB; ++e) dd[e] = LU[(size_t)di * BB + e];
    if (!invert_block<BS>(dd, inv)) {
        atomicMin(bad_row, i);
#pragma unroll
        for (int e = 0; e < BB; ++e) inv[e] = 0.0;
    }
#pragma unroll
    for (int e = 0; e < BB; ++e) dinv[(size_t)i * BB + e] = inv[e];
}

// ---------------------------------------------------------------------------
// Adjoint gradient on the block pattern: grad_orig[src[q]] = -lambda_i x_j^T
// for permuted block q = (i, j).
template <int BS>
__global__ void k_grad_blocks(int64_t nnzb, const int32_t *blkrow, const int32_t *colidx, const int32_t *src,
                              const double *lam, const double *x, double *grad) {
    constexpr int BB = BS * BS;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nnzb * BB) return;
    const int64_t q = idx / BB;
    const int e = (int)(idx - q * BB);
    const int r = e / BS, c = e - r * BS;
    const int64_t i = blkrow[q], j = colidx[q];
    grad[(int64_t)src[q] * BB + e] = -(lam[i * BS + r] * x[j * BS + c]);
}

}  // namespace bda
