// HIP kernels of the ILU0-BiCGSTAB hot path, written for gfx950 (CDNA4,
// wave64).  Everything here is HBM-bound integer/FP64 streaming work; no
// MFMA is used.
//
// Kernel map (reference op -> kernel):
//   SpMV mode, sparstition.vhd:655-878 / spmvp.vhd:240-383  -> k_stream<SPMV>
//   steps 0-1 (r = b - A x0, rho)  solver_pkg.vhd:122-123     -> k_mark_nonzero + k_stream<SUB, ZX> + FIN_INIT
//                                                                 (r = b without the product when x0 = 0)
//   ILU0_L_fs, ilu0.vhd:412-416                               -> k_stream<SUB> per colour
//   ILU0_U_bs, ilu0.vhd:385-410                               -> k_stream<SUB_DINV> per colour
//   dot1 -> alpha, solver.vhd:793-798                         -> SpMV epilogue + FIN_ALPHA
//   axpy1 + axpy2/norm, solver_pkg.vhd:128-130                -> k_update_s + FIN_HALF1 (axpy1,
//                                                                 x += alpha phat, deferred into
//                                                                 k_update_xr / k_finish_x)
//   dot2 -> omega, solver.vhd:831-842                         -> SpMV epilogue + FIN_OMEGA
//   axpy1 + axpy2/norm + dot1 -> beta, solver_pkg.vhd:133-136 -> k_update_xr + FIN_HALF2
//   calc_p, vector_ops.vhd:319-323,429-458                    -> k_update_p
//
// Arithmetic follows the reference's separately rounded multiply-then-add
// (dot_axpy.vhd:249-275): this file is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "bda_host.hpp"  // kTileRows

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
    int32_t run_id;         // written to host_done when the run stops
    int32_t *host_done;     // mapped host memory the host polls
    uint64_t *host_prog;    // mapped: run_id << 32 | progress (half iterations, or outer steps)
    double *hist;           // optional norm history: ||r0||, then every half-iteration's norm
    int32_t hist_cap, nhist;
    // outer refinement of the precision form (used on the outer state)
    double nrm0;            // FP64 residual norm of the start vector
    double itol;            // inner tolerance for the next inner run
    double inner_tol, eta;  // adaptive rule: min(0.5, max(inner_tol, eta thr / ||r_k||))
    double iters;           // sum of the inner runs' iteration counts
    int32_t outer;          // inner runs completed
    int32_t max_outer;
    int32_t adaptive;
    int32_t inner_h;        // sum of the inner runs' half iterations
    int32_t full_residual;  // set by k_mark_nonzero: the step-0 residual needs the product
    // watchdog (the counterpart of the reference kernel's abort cycle count,
    // param0): timeout in ticks of the device's constant-rate wall clock
    // (0 = off); the run's deadline is set on the device when its state is
    // initialised, and every scalar step past it stops the run as aborted
    uint64_t timeout_ticks;
    uint64_t deadline;
    int32_t aborted;
};

enum Fin : int32_t {
    FIN_NONE = 0,
    FIN_STORE = 1,   // tot[] = totals
    FIN_INIT = 2,    // rho = r.r, ||r0||, threshold, "already solved" exit
    FIN_ALPHA = 3,   // alpha = rho / (v.rt)
    FIN_HALF1 = 4,   // ||s||, compare, final-iteration exit
    FIN_OMEGA = 5,   // omega = (t.s) / (t.t)
    FIN_HALF2 = 6,   // ||r||, compare, rho_new = rt.r, beta
    FIN_REFINE = 7,  // outer refinement: ||b - A x||, stop tests, next inner tolerance
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

__device__ inline void record_norm(SolverState *st, double nrm) {
    if (st->hist != nullptr && st->nhist < st->hist_cap) st->hist[st->nhist] = nrm;
    st->nhist += 1;
}

__device__ inline void push_norm(SolverState *st, double nrm) {
    record_norm(st, nrm);
    const int slot = st->norm_slot;
    st->norms[slot] = nrm;
    st->last_norm_idx = slot;
    st->norm_slot = slot == 3 ? 1 : slot + 1;
    st->last_norm = nrm;
}

__device__ inline bool finite(double v) { return !__builtin_isnan(v) && !__builtin_isinf(v); }

// The flag receives the run's id, so the host never has to clear it between
// runs (a clear could race a late write of the previous run).
__device__ inline void signal_done(SolverState *st) {
    st->done = 1;
    if (st->host_done) __hip_atomic_store(st->host_done, st->run_id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Progress for the host's throttling (it keeps a bounded number of
// half-iterations queued ahead of the device): a mapped word, so no event
// has to be recorded in the stream.
__device__ inline void signal_progress(SolverState *st, int32_t count) {
    if (st->host_prog)
        __hip_atomic_store(st->host_prog, ((uint64_t)(uint32_t)st->run_id << 32) | (uint32_t)count, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

// Scalar epilogues; run by one thread after a grid-wide reduction.  The
// watchdog is checked first: a run past its deadline stops here, and every
// launch queued after it returns at once.
__device__ inline void finalize(SolverState *st, int fin, double t0, double t1) {
    if (fin != FIN_STORE && st->deadline != 0 && wall_clock64() > st->deadline) {
        st->aborted = 1;
        signal_done(st);
        return;
    }
    switch (fin) {
    case FIN_STORE:
        st->tot[0] = t0;
        st->tot[1] = t1;
        break;
    case FIN_INIT: {   // solver.vhd:867-936 with dot_res = rho_new
        const double nrm = sqrt(t0);
        st->nhist = 0;
        record_norm(st, nrm);
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
            signal_done(st);
        } else {
            st->thr = st->abs_mode ? st->tol : st->tol * nrm;
        }
        break;
    }
    case FIN_ALPHA:    // step 5: divide rho_new by the dot result
        if (t0 == 0.0 || !finite(t0)) {
            if (t0 == 0.0) st->breakdown = 1; else st->nan_flag = 1;
            signal_done(st);
        } else {
            st->alpha = st->rho / t0;
        }
        break;
    case FIN_HALF1: {  // steps 7/8
        const double nrm = sqrt(t0);
        push_norm(st, nrm);
        st->last_half = 1;
        if (!finite(nrm)) { st->nan_flag = 1; signal_done(st); }
        else if (!(nrm > st->thr)) { st->converged = 1; signal_done(st); }
        else if ((st->h >> 1) == st->maxit) { st->max_iter = 1; signal_done(st); }
        else {
            st->h += 1;
            signal_progress(st, st->h);
        }
        break;
    }
    case FIN_OMEGA:    // step 10
        if (t1 == 0.0 || !finite(t0) || !finite(t1)) {
            if (t1 == 0.0) st->breakdown = 1; else st->nan_flag = 1;
            signal_done(st);
        } else {
            st->omega = t0 / t1;
        }
        break;
    case FIN_HALF2: {  // step 12, then steps 13/14
        const double nrm = sqrt(t0);
        push_norm(st, nrm);
        st->last_half = 2;
        if (!finite(nrm)) { st->nan_flag = 1; signal_done(st); }
        else if (!(nrm > st->thr)) { st->converged = 1; signal_done(st); }
        else {
            st->h += 1;
            st->rho_prev = st->rho;
            st->rho = t1;
            if (t1 == 0.0 || st->omega == 0.0) { st->breakdown = 1; signal_done(st); }
            else {
                st->beta = (st->rho / st->rho_prev) * (st->alpha / st->omega);
                signal_progress(st, st->h);
            }
        }
        break;
    }
    case FIN_REFINE: {  // one outer step of the precision form
        const double nrm = sqrt(t0);
        if (st->outer == 0) {
            st->nrm0 = nrm;
            st->norms[0] = nrm;
            const double precision = st->abs_mode ? st->tol : kAlreadySolved;
            if (!(nrm > precision)) {
                st->no_results = 1;
                if (__builtin_isnan(nrm)) st->nan_flag = 1;
                signal_done(st);
                break;
            }
            st->thr = st->abs_mode ? st->tol : st->tol * nrm;
            st->last_norm = nrm;
        } else {
            push_norm(st, nrm);
        }
        if (!finite(nrm)) { st->nan_flag = 1; signal_done(st); }
        else if (st->outer > 0 && !(nrm > st->thr)) { st->converged = 1; signal_done(st); }
        else if (st->outer == st->max_outer) { st->max_iter = 1; signal_done(st); }
        else if (st->adaptive) {
            const double want = st->eta * st->thr / nrm;
            const double m = st->inner_tol < want ? want : st->inner_tol;
            st->itol = m < 0.5 ? m : 0.5;
        } else {
            st->itol = st->inner_tol;
        }
        signal_progress(st, st->outer + 1);
        break;
    }
    default:
        break;
    }
}

// Solver state set-up on the stream (no host copy, no synchronisation); the
// watchdog's deadline counts from here.
__global__ void k_set_state(SolverState *st, SolverState v) {
    if (v.timeout_ticks != 0) v.deadline = wall_clock64() + v.timeout_ticks;
    *st = v;
}

// Start of an inner run of the precision form: tolerance and deadline from
// the outer state; an outer state that has already stopped stops the inner
// run.
__global__ void k_inner_start(SolverState *in, const SolverState *out, SolverState v) {
    v.deadline = out->deadline;
    *in = v;
    if (out->done) {
        in->no_results = 1;
        signal_done(in);
    } else {
        in->tol = out->itol;
    }
}

// End of an inner run: counters and flags into the outer state.  An inner
// run whose right-hand side was already zero ends the refinement, and so
// does one that broke down or met a NaN before completing a half-iteration:
// it changed nothing, so the next outer step would repeat it exactly.
__global__ void k_refine_step(SolverState *out, const SolverState *in) {
    if (out->done) return;
    out->outer += 1;
    out->inner_h += in->h;
    out->iters += in->h / 2.0 + 0.5;
    if (in->breakdown || in->nan_flag) {
        out->breakdown = in->breakdown;
        out->nan_flag = in->nan_flag;
    }
    if (in->aborted) out->aborted = 1;
    if (in->no_results || in->aborted || ((in->breakdown || in->nan_flag) && in->h == 0 && in->last_half == 0))
        signal_done(out);
}

// Sum over the 64 lanes of a wave as the butterfly v += v[lane ^ o] for
// o = 32, 16, 8, 4, 2, 1, with VALU lane moves instead of LDS-crossbar
// shuffles: the permlane swaps exchange the wave's halves and the 16-lane
// rows; DPP moves the rest.  row_ror:8 is lane ^ 8 exactly; at the o = 4
// step every lane holds a sum that depends only on lane mod 8, and
// row_ror:4 reads lane - 4 within the row, whose residue is (lane mod 8) ^ 4,
// so each step adds the same two values as the shuffle butterfly and the
// result is the same bit for bit (checked on the GPU against __shfl_xor over
// 262,144 random sums).
__device__ inline double lane_swap32(double v) {
    const int lane = threadIdx.x & 63;
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
    return lane < 32 ? __hiloint2double(hi[1], lo[1]) : __hiloint2double(hi[0], lo[0]);
}
__device__ inline double lane_swap16(double v) {
    const int lane = threadIdx.x & 63;
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
    return (lane & 16) == 0 ? __hiloint2double(hi[1], lo[1]) : __hiloint2double(hi[0], lo[0]);
}
template <int CTRL>
__device__ inline double lane_dpp(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double wave_sum(double v) {
    v += lane_swap32(v);
    v += lane_swap16(v);
    v += lane_dpp<0x128>(v);  // row_ror:8
    v += lane_dpp<0x124>(v);  // row_ror:4
    v += lane_dpp<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
    v += lane_dpp<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
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
__device__ inline void block_partials(double (&v)[NRED], const RedArgs &ra, int slot) {
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
            ra.partials[(size_t)j * ra.cap + slot] = t;
        }
    }
}

constexpr int kFinThreads = 1024;
// Loads in flight per thread: 16 (two passes cover 32,768 partials), or 32
// for larger partial counts (the 4x4 256x256x64 grid's SpMV writes 95,044:
// three passes instead of six); 16 running sums per thread either way, the
// u-th load of a pass added to sum u mod 16.
constexpr int kFinUnroll = 16, kFinUnrollWide = 32;
constexpr int kFinWideFrom = 2 * kFinThreads * kFinUnroll + 1;

template <int NRED, int U>
__global__ __launch_bounds__(kFinThreads) void k_finalize(RedArgs ra, int nparts) {
    __shared__ double sred[kFinThreads / 64][2];
    if (ra.st->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int A = kFinUnroll;
    double acc[NRED];
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
        const double *pp = ra.partials + (size_t)j * ra.cap;
        double a8[A];
#pragma unroll
        for (int u = 0; u < A; ++u) a8[u] = 0.0;
        for (int i0 = 0; i0 < nparts; i0 += kFinThreads * U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kFinThreads + threadIdx.x;
                v[u] = i < nparts ? pp[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) a8[u % A] += v[u];
        }
        double t = a8[0];
#pragma unroll
        for (int u = 1; u < A; ++u) t += a8[u];
        acc[j] = wave_sum(t);
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

// k_finalize for at most kFinThreads * U partials (U < kFinUnroll): one pass
// that issues the loads of all NRED reductions together (k_finalize sums
// them one reduction after the other), with the same sums: k_finalize's
// running sums past the first U stay +0.0, and adding them changes the
// total only as one final + 0.0 does.
template <int NRED, int U>
__global__ __launch_bounds__(kFinThreads) void k_finalize_one_pass(RedArgs ra, int nparts) {
    static_assert(U < kFinUnroll, "one pass of fewer loads than k_finalize's running sums");
    __shared__ double sred[kFinThreads / 64][2];
    if (ra.st->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double v[NRED][U];
    // clamped loads, all in flight together, and a select afterwards (a
    // load under a branch gets a wait of its own before the branch joins)
#pragma unroll
    for (int j = 0; j < NRED; ++j)
#pragma unroll
        for (int u = 0; u < U; ++u)
            v[j][u] = ra.partials[(size_t)j * ra.cap + min(u * kFinThreads + (int)threadIdx.x, max(nparts - 1, 0))];
#pragma unroll
    for (int j = 0; j < NRED; ++j)
#pragma unroll
        for (int u = 0; u < U; ++u) v[j][u] = u * kFinThreads + (int)threadIdx.x < nparts ? v[j][u] : 0.0;
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
        double t = 0.0 + v[j][0];
#pragma unroll
        for (int u = 1; u < U; ++u) t += 0.0 + v[j][u];
        t += 0.0;
        const double acc = wave_sum(t);
        if (lane == 0) sred[wave][j] = acc;
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

// k_finalize<NRED, kFinUnroll> for at most kFinThreads * (kFinUnroll + U2)
// partials (the 3x3 128x128x64 grid's SpMV writes 18,019): the loads of both
// passes and of all NRED reductions are issued together (k_finalize waits
// for each pass of each reduction in turn), then summed exactly as
// k_finalize sums them.  Its second-pass loads past the first U2 read +0.0
// for every thread, and a running sum that started as 0.0 + v is never -0.0,
// so leaving those additions out leaves every sum bit for bit the same.
template <int NRED, int U2>
__global__ __launch_bounds__(kFinThreads) void k_finalize_joint(RedArgs ra, int nparts) {
    static_assert(U2 < kFinUnroll, "a short second pass");
    __shared__ double sred[kFinThreads / 64][2];
    if (ra.st->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int A = kFinUnroll;
    double v0[NRED][A], v1[NRED][U2];
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
        const double *pp = ra.partials + (size_t)j * ra.cap;
        // clamped loads and a select afterwards: a load under a branch
        // gets a wait of its own before the branch joins
#pragma unroll
        for (int u = 0; u < A; ++u) v0[j][u] = pp[min(u * kFinThreads + (int)threadIdx.x, nparts - 1)];
#pragma unroll
        for (int u = 0; u < U2; ++u) v1[j][u] = pp[min((A + u) * kFinThreads + (int)threadIdx.x, nparts - 1)];
    }
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
#pragma unroll
        for (int u = 0; u < A; ++u) v0[j][u] = u * kFinThreads + (int)threadIdx.x < nparts ? v0[j][u] : 0.0;
#pragma unroll
        for (int u = 0; u < U2; ++u) v1[j][u] = (A + u) * kFinThreads + (int)threadIdx.x < nparts ? v1[j][u] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
        double a8[A];
#pragma unroll
        for (int u = 0; u < A; ++u) a8[u] = 0.0 + v0[j][u];
#pragma unroll
        for (int u = 0; u < U2; ++u) a8[u] += v1[j][u];
        double t = a8[0];
#pragma unroll
        for (int u = 1; u < A; ++u) t += a8[u];
        const double acc = wave_sum(t);
        if (lane == 0) sred[wave][j] = acc;
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
constexpr int kFinJointU2 = 2;
constexpr int kFinJointTo = kFinThreads * (kFinUnroll + kFinJointU2);

// Above kFinTwoStageFrom partials (the 4x4 256x256x64 grid's SpMV writes
// 95,044) one workgroup reads them at about 55 GB/s; a first stage of
// kFinGroups workgroups sums contiguous chunks instead (each chunk in a fixed
// order), into the last kFinGroups slots of each reduction's row of the
// partials buffer, and k_finalize_one_pass adds those up.
constexpr int kFinGroups = 256, kFinGroupThreads = 256, kFinGroupUnroll = 4;
constexpr int kFinTwoStageFrom = 2 * kFinThreads * kFinUnrollWide + 1;

template <int NRED>
__global__ __launch_bounds__(kFinGroupThreads) void k_finalize_stage1(RedArgs ra, int nparts) {
    __shared__ double sred[kFinGroupThreads / 64][2];
    if (ra.st->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int chunk = (nparts + kFinGroups - 1) / kFinGroups;
    const int i0 = blockIdx.x * chunk, i1 = min(i0 + chunk, nparts);
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
        const double *pp = ra.partials + (size_t)j * ra.cap;
        double t = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += kFinGroupThreads * kFinGroupUnroll) {
            double v[kFinGroupUnroll];
#pragma unroll
            for (int u = 0; u < kFinGroupUnroll; ++u) {
                const int k = i + u * kFinGroupThreads;
                v[u] = k < i1 ? pp[k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kFinGroupUnroll; ++u) t += v[u];
        }
        t = wave_sum(t);
        if (lane == 0) sred[wave][j] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NRED; ++j) {
            double t = sred[0][j];
            for (int w = 1; w < kFinGroupThreads / 64; ++w) t += sred[w][j];
            ra.partials[(size_t)j * ra.cap + ra.cap - kFinGroups + blockIdx.x] = t;
        }
    }
}

// dst = src (n doubles); *nonfinite = 1 when a value is infinite or NaN.
// Every writer stores the same value.
__global__ __launch_bounds__(256) void k_copy_check_finite(int64_t n, const double *src, double *dst,
                                                           int32_t *nonfinite) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = src[i];
        dst[i] = v;
        bad = bad || !finite(v);
    }
    if (bad) __hip_atomic_store(nonfinite, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Streaming block-row kernel.
//
// One workgroup (64*BS threads, BS waves) handles one tile: up to 64 (or,
// in the tall tiling, 128) consecutive block rows of one colour whose blocks
// are contiguous in the value array.  Phase 1 streams the tile's values,
// with FP32 values one lane per block row (BS values, one vector load), with
// FP64 values one lane per two consecutive entries (one 16-byte load; 63
// lanes per wave step for 3x3, 64 for 4x4); either way every wave load
// instruction reads consecutive memory.  The tile's x entries are gathered
// into LDS (column indices staged there first); a row lane adds its block
// row's products left to right, pair lanes add theirs with the next lane's
// in the same order through cross-lane shifts.  The per-block row partials
// go to LDS.  Phase 2 gives
// one thread per scalar row (two in the tall tiling), which adds its block
// partials in column order and applies the epilogue.
//
// Summation order per scalar row r of block row i:
//   sum = 0;  for blocks j in column order: sum += ((a_r0 x_0 + a_r1 x_1) + a_r2 x_2)
//
// SM_SUB_Z is SM_SUB for the first sweep's colour C-2 when the extended
// merged tail (k_tail3) follows: rows without blocks in the second sweep
// also get z = D^-1 y there, computed from y as stored (rounded to the
// vector type), which is what the second sweep would compute for them.
enum StreamMode : int { SM_SPMV = 0, SM_SUB = 1, SM_SUB_DINV = 2, SM_SUB_Z = 3 };

template <typename TV, typename TX>
struct StreamArgs {
    const int4 *tile;   // tile descriptors {first row, first block, end row, end block}
    const int32_t *slot;// reduction slot of each tile: its place in the colour-ordered tile list
    int32_t tile0, ntiles;
    const int32_t *rowptr;
    const int32_t *colidx;
    const TV *val;
    const TX *x;        // vector multiplied (gathered through colidx) ...
    const TX *x_lo;     // ... except columns below split, which come from x_lo
    int32_t split;
    const TX *rhs;      // SUB modes: out = rhs - A x
    TX *out;
    TX *out2;           // optional copies of out (initial residual -> r, p)
    TX *out3;
    const double *dinv; // SUB_DINV: out = dinv * (rhs - A x); SUB_Z: the second sweep's inverse
    const TX *w0;       // SR_W0 reductions: sum out * w0
    RedArgs ra;
    const int32_t *rowptr2 = nullptr;  // SUB_Z: the second sweep's row pointers
    TX *zout = nullptr;                // SUB_Z: z of the rows without second-sweep blocks
    const int32_t *rowmap = nullptr;   // dependency tiles (k_lower2): matrix row of each packed row
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

// Value of lane i + 1 (lane 63 receives 0): a DPP wavefront shift, a VALU
// operation, where __shfl_down would go through the LDS crossbar.
__device__ inline double lane_down1(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130 /* wave_shl:1 */, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Matrix values are loaded non-temporally.  For FP32 values this looks 25%
// slower when one SpMV is timed back to back on itself (the 216 MB of values
// then stay in the 256 MB last-level cache between launches), but inside a
// solve, where the values cannot stay cached, it makes the precision form 4%
// faster (measured on MI355X).
__device__ inline double load_stream(const double *p) { return __builtin_nontemporal_load(p); }
__device__ inline float load_stream(const float *p) { return __builtin_nontemporal_load(p); }
template <bool PLAIN, typename T> __device__ inline T load_mat(const T *p) {
    if constexpr (PLAIN) return *p;
    else return __builtin_nontemporal_load(p);
}

// Blocks of padding behind every matrix value and column-index array, so
// that the streaming kernel's clamped loads never leave the allocation.
constexpr int kPadBlocks = 8;

// The BS x entries of one block, read with one vector access (12 bytes for
// FP32 3x3, 16 for FP32 4x4; 24 and 32 bytes for FP64).
template <typename T, int BS> struct __attribute__((packed, aligned(sizeof(T)))) XVec { T v[BS]; };

// Tile capacity in blocks, narrow and wide tiling; the host tiler uses the
// same numbers.
template <int BS, bool WIDE> struct TileCap;
template <> struct TileCap<3, false> { static constexpr int value = 168; };  // 8 workgroup steps of 21 blocks
template <> struct TileCap<4, false> { static constexpr int value = 128; };  // 8 workgroup steps of 16 blocks
template <> struct TileCap<3, true> { static constexpr int value = 336; };   // 16 steps
template <> struct TileCap<4, true> { static constexpr int value = 256; };

// Which tiling a kernel instance uses: the wide one in the modes whose bit is
// set in BDA_WIDE_MODES (the SpMV, in both precisions), the narrow one
// otherwise (the sweeps: with wide tiles their register and LDS use cuts the
// resident workgroups, and the FP32 sweeps lose more than they gain).
#ifndef BDA_WIDE_MODES
#define BDA_WIDE_MODES 1
#endif
template <int MODE> constexpr bool wide_tiles() { return ((BDA_WIDE_MODES >> MODE) & 1) != 0; }
template <int BS, int MODE> constexpr int stream_cap() { return TileCap<BS, wide_tiles<MODE>()>::value; }

// Matrix values and column indices are read once per launch and use
// non-temporal loads.  (The measured fetch volume is the same with default,
// non-temporal and system-scope value loads.)
//
// Reductions fused into the epilogue: bit 0 = sum out*w0, bit 1 = sum out*out.
enum StreamRed : int { SR_NONE = 0, SR_W0 = 1, SR_SELF = 2, SR_BOTH = 3 };
template <int RED> struct NRedOf { static constexpr int value = (RED & 1) + ((RED >> 1) & 1); };

// No occupancy request is made of the register allocator.  The FP32 3x3
// product kernel lands just above 64 VGPRs (7 waves per SIMD); forcing 8
// waves was measured slower in the current launch order (54.4-54.5 against
// 53.8-54.0 us per SpMV on a 128x128x64 grid).

// Phase 1 has no loop: every lane owns a fixed set of S value slots of the
// tile (S = CAP / blocks per workgroup step) and issues all their loads at
// once.  The x entries a tile needs are gathered cooperatively into LDS,
// one block's BS entries per lane with one vector load, instead of once
// per matrix entry.
//
// All loads are issued without intervening branches (indices are clamped
// into range and results masked afterwards; the value and column arrays
// carry a few blocks of padding for this): the compiler's wait-count
// insertion falls back to waiting for every outstanding load at a branch,
// which would serialise the memory round trips.  Column indices and row
// pointers are issued first, so the first barrier waits for them and not
// for the value stream issued after them.  A tile costs three dependent
// memory round trips: descriptor; columns/values/row operands; x gathers.
//
// ROWS = 128 (the tall tiling, used by sweep colours with few blocks per row)
// gives every thread two rows of the epilogue, 64 rows apart.
//
// Value layout of the streaming kernel's lanes.  Row lanes: one lane per
// block row, its BS values one vector load, the row's products added in the
// lane.  Pair lanes: one lane per two consecutive matrix entries (one
// 16-byte load, a wave's loads contiguous), a block row's products added
// with the next lane's through DPP shifts.  Entry lanes: one lane per matrix
// entry, the products of a block row added across its BS lanes with DPP
// shifts.  All add the same products in the same order, so the results are
// the same bit for bit.  Measured on the 128x128x64 3x3 grid (A/B, two
// rounds): with FP32 values row lanes take the SpMV from 53.7 to 50.1 us and
// the preconditioner application from 70.3 to 65.7 us (49 instead of 65
// VGPRs).  With FP64 values row lanes are slower than entry lanes (SpMV 93 ->
// 110.5 us: a lane's 24 bytes take two or three strided loads) and pair
// lanes faster (SpMV 93.5 -> 91.0 us, application 110 -> 104.5 us, 91
// instead of 102 VGPRs; the FP64 solve 16.87 -> 16.40 ms).  BDA_LANES32 /
// BDA_LANES64 select the layout for FP32 / FP64 values.
enum LaneLayout : int { LAY_ENTRY = 0, LAY_ROW = 1, LAY_PAIR = 2 };
#ifndef BDA_LANES32
#define BDA_LANES32 LAY_ROW
#endif
#ifndef BDA_LANES64
#define BDA_LANES64 LAY_PAIR
#endif
template <typename TV> constexpr int lane_layout() {
    return std::is_same<TV, float>::value ? BDA_LANES32 : BDA_LANES64;
}

template <typename TX, int BS, int CAP>
struct TileLds {
    double part[CAP * BS];
    TX sx[CAP * BS];
    int32_t scol[(CAP + 64 * BS - 1) / (64 * BS) * 64 * BS];
    int32_t srow[64 * BS];
};
template <typename TX, int BS, int CAP>
__device__ inline TileLds<TX, BS, CAP> &tile_lds() {
    __shared__ TileLds<TX, BS, CAP> l;
    return l;
}
template <int N>
__device__ inline double *tile_tv() {
    __shared__ double tv[N];
    return tv;
}

// stream_tile is the work of one workgroup on the tile with descriptor d;
// with KEEP it also leaves the tile's outputs (rounded to TX, (row - first
// row) * BS + component) in keep[], an LDS array of 64 * BS entries.  Its
// fused reductions write partial-sum slot `slot`.
//
// V selects the two tile forms of the merged lower-sweep launch (k_lower2):
// 1, a dependency tile: packed copies of matrix rows (a.rowmap gives each
// packed row's matrix row, for its right-hand side); its outputs only go to
// keep[]; 2, a tile whose column indices below zero name a row of the
// dependency tile: -1 - c is the row's place there, and its x comes from
// ldsx[] (that tile's keep[]) instead of memory.
template <int BS, int MODE, int RED, typename TV, typename TX, int ROWS, bool KEEP, bool PLAIN = false, int V = 0>
__device__ __attribute__((always_inline)) inline void stream_tile(const StreamArgs<TV, TX> &a, const int4 d, TX *keep,
                                                                   int slot, const TX *ldsx = nullptr) {
    static_assert(V != 1 || (KEEP && MODE == SM_SUB && RED == SR_NONE), "dependency tiles keep plain sweep rows");
    constexpr int BB = BS * BS;
    constexpr int NT = 64 * BS;            // threads per workgroup
    constexpr int NW = BS;                 // waves per workgroup
    constexpr int CAP = stream_cap<BS, MODE>();
    constexpr int BPW = 64 / BB;           // blocks per wave step (7 or 4)
    constexpr int VPW = BPW * BB;          // lanes carrying a value (63 or 64)
    constexpr int S = CAP / (NW * BPW);    // value slots per lane
    constexpr int CJ = (CAP + NT - 1) / NT;       // column loads per thread
    constexpr int XV = (CAP + NT - 1) / NT;       // x gathers (whole blocks) per thread
    constexpr int NRED = (RED & 1) + ((RED >> 1) & 1);
    constexpr int RS = ROWS / 64;                 // epilogue rows per thread
    static_assert(ROWS == 64 || ROWS == 128, "tile rows");
    static_assert(ROWS + 1 <= NT, "one row pointer per thread");

    // LDS arrays are sized to whole multiples of the workgroup so that their
    // fills are unconditional (a conditional store lets the compiler sink the
    // feeding load next to it, where it is waited for at once).  They belong
    // to the tile shape, not to this instance, so that the instances one
    // kernel runs one after the other share them.
    TileLds<TX, BS, CAP> &lds = tile_lds<TX, BS, CAP>();
    double *part = lds.part;
    TX *sx = lds.sx;
    int32_t *scol = lds.scol;
    int32_t *srow = lds.srow;
    double *tv = tile_tv<(MODE == SM_SUB_DINV || MODE == SM_SUB_Z) ? RS * NT : 1>();

    static_assert(!KEEP || ROWS == 64, "kept outputs: one row per thread");
    const int r0 = d.x, r1 = d.z, b0 = d.y;
    const int nblk = d.w - b0;
    const int nrow = r1 - r0;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;

    // 1. column indices and row pointers
    int ci[CJ];
#pragma unroll
    for (int j = 0; j < CJ; ++j) ci[j] = load_mat<PLAIN>(&a.colidx[b0 + min(t + j * NT, nblk)]);
    const int rp = a.rowptr[r0 + min(t, nrow)];

    // 2. values.  Row lanes: slot u of thread t is block row j = u*NT + t of
    // the tile (block j / BS, row j % BS), its BS values one vector load;
    // the tile's block rows are contiguous, so a wave reads 64 consecutive
    // rows.  Pair lanes: slot u of lane l < PPW is entry pair q = (u*NW +
    // wave)*PPW + l, entries 2q and 2q + 1 of the tile.  Entry lanes: slot u
    // of this lane is block (u*NW + wave)*BPW + lb, entry e.
    constexpr int LAY = lane_layout<TV>();
    constexpr int SR = (CAP * BS + NT - 1) / NT;  // row lanes: block rows per thread
    constexpr int PPW = BS == 3 ? 63 : 64;        // pair lanes: pairs per wave step
    constexpr int SP = CAP * BB / 2 / (NW * PPW); // pair lanes: pairs per lane
    static_assert(LAY != LAY_PAIR || (CAP * BB) % (2 * NW * PPW) == 0, "whole pair slots");
    constexpr int NS = LAY == LAY_ROW ? SR : LAY == LAY_PAIR ? SP : S;  // value slots per thread
    constexpr int NC = LAY == LAY_ROW ? BS : LAY == LAY_PAIR ? 2 : 1;   // values per slot
    const int lb = lane / BB, e = lane - lb * BB, er = e / BS, ec = e - er * BS;
    const bool active_lane = lane < VPW;
    TV av[NS][NC];
    if constexpr (LAY == LAY_ROW) {
        const TV *valr = a.val + (size_t)b0 * BB;
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int jc = min(u * NT + t, nblk * BS);
#pragma unroll
            for (int c = 0; c < NC; ++c) av[u][c] = load_mat<PLAIN>(&valr[(size_t)jc * BS + c]);
        }
    } else if constexpr (LAY == LAY_PAIR) {
        const TV *valr = a.val + (size_t)b0 * BB;
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int qc = min((u * NW + wave) * PPW + lane, nblk * BB / 2);
#pragma unroll
            for (int c = 0; c < NC; ++c) av[u][c] = load_mat<PLAIN>(&valr[(size_t)qc * 2 + c]);
        }
    } else {
        const TV *val = a.val + (size_t)b0 * BB + e;
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int bl = (u * NW + wave) * BPW + lb;
            av[u][0] = load_mat<PLAIN>(&val[(size_t)min(bl, nblk) * BB]);
        }
    }

    // 3. per-row operands of the epilogue: thread t serves scalar row r of
    // block rows k0 and (ROWS = 128) k0 + 64
    const int k0 = t / BS, r = t - k0 * BS;
    int kk[RS];
    bool has_row[RS];
    size_t gi[RS];
    double rhs_v[RS], w_v[RS];
    double dv[RS][BS];
    bool zrow[RS];
#pragma unroll
    for (int s = 0; s < RS; ++s) {
        kk[s] = k0 + 64 * s;
        has_row[s] = kk[s] < nrow;
        gi[s] = (size_t)(r0 + min(kk[s], nrow - 1)) * BS + r;
        rhs_v[s] = 0.0;
        w_v[s] = 0.0;
        zrow[s] = false;
        if constexpr (V == 1) {
            // (a dependency tile may have no rows)
            const size_t gm = (size_t)a.rowmap[r0 + max(min(kk[s], nrow - 1), 0)] * BS + r;
            rhs_v[s] = (double)a.rhs[gm];
        } else if constexpr (MODE != SM_SPMV) {
            rhs_v[s] = (double)a.rhs[gi[s]];
        }
        if constexpr ((RED & SR_W0) != 0) w_v[s] = (double)a.w0[gi[s]];
        if constexpr (MODE == SM_SUB_DINV || MODE == SM_SUB_Z) {
            const double *D = a.dinv + gi[s] * BS;
#pragma unroll
            for (int c = 0; c < BS; ++c) dv[s][c] = D[c];
        }
        if constexpr (MODE == SM_SUB_Z) {
            const int i = r0 + min(kk[s], nrow - 1);
            zrow[s] = a.rowptr2[i + 1] == a.rowptr2[i];
        }
    }

#pragma unroll
    for (int j = 0; j < CJ; ++j) scol[t + j * NT] = ci[j];
    srow[t] = rp - b0;
    __syncthreads();

    // 4. x entries of the tile, one block (BS entries, one vector load) per
    // lane, stored (block, component) like part[]; lanes past the last block
    // repeat it (the same values to the same place).  A lane per entry needed
    // BS times the address work: the FP32 3x3 SpMV took 59.7 us instead of
    // 56.0 us with it (measured).
    XVec<TX, BS> xg[XV];
    int xq[XV];
    int xl[XV];
#pragma unroll
    for (int j = 0; j < XV; ++j) {
        const int q = min(t + j * NT, nblk > 0 ? nblk - 1 : 0);
        int cb = nblk > 0 ? scol[q] : 0;
        xl[j] = -1;
        if constexpr (V == 2) {
            xl[j] = cb < 0 ? -1 - cb : -1;
            cb = max(cb, 0);
        }
        xq[j] = q;
        xg[j] = *reinterpret_cast<const XVec<TX, BS> *>((cb < a.split ? a.x_lo : a.x) + (size_t)cb * BS);
    }
    if constexpr (V == 2) {
#pragma unroll
        for (int j = 0; j < XV; ++j)
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const TX lv = ldsx[max(xl[j], 0) * BS + c];
                if (xl[j] >= 0) xg[j].v[c] = lv;
            }
    }
#pragma unroll
    for (int j = 0; j < XV; ++j)
#pragma unroll
        for (int c = 0; c < BS; ++c) sx[xq[j] * BS + c] = xg[j].v[c];
    __syncthreads();

    // all LDS reads first (clamped, unconditional), then the products
    TX xs[NS][NC];
    if constexpr (LAY == LAY_ROW) {
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int bl = min((u * NT + t) / BS, CAP - 1);
#pragma unroll
            for (int c = 0; c < NC; ++c) xs[u][c] = sx[bl * BS + c];
        }
        // a block row's products added left to right, as the entry lanes do
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int j = u * NT + t;
            const bool ok = j / BS < nblk;
            double sm = prod<TV, TX>(ok ? av[u][0] : (TV)0, ok ? xs[u][0] : (TX)0);
#pragma unroll
            for (int c = 1; c < NC; ++c) sm = sm + prod<TV, TX>(ok ? av[u][c] : (TV)0, ok ? xs[u][c] : (TX)0);
            if (ok) part[j] = sm;
        }
    } else if constexpr (LAY == LAY_PAIR) {
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int q = (u * NW + wave) * PPW + lane;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const int en = 2 * q + c;
                xs[u][c] = sx[min(en / BB, CAP - 1) * BS + en % BS];
            }
        }
        // 3x3: lanes 3k, 3k+1, 3k+2 hold entries 6k'..6k'+5, two block rows;
        // lane 3k adds row 2k' (its two products, then the next lane's
        // first), lane 3k+1 row 2k'+1 (its second product, then the next
        // lane's two).  4x4: lane 2k adds its two products, then the next
        // lane's two.  Each is ((p0 + p1) + p2) [+ p3], the entry lanes' order.
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int q = (u * NW + wave) * PPW + lane;
            const double pa = prod<TV, TX>(av[u][0], xs[u][0]);
            const double pb = prod<TV, TX>(av[u][1], xs[u][1]);
            const double na = lane_down1(pa);
            const double nb = lane_down1(pb);
            int row;
            double sm;
            if constexpr (BS == 3) {
                const int m = lane % 3;
                row = m == 0 ? 2 * q / 3 : m == 1 ? (2 * q + 1) / 3 : -1;
                sm = m == 0 ? (pa + pb) + na : (pb + na) + nb;
            } else {
                row = (lane & 1) == 0 ? q / 2 : -1;
                sm = ((pa + pb) + na) + nb;
            }
            if (lane < PPW && row >= 0 && row < nblk * BS) part[row] = sm;
        }
    } else {
#pragma unroll
        for (int u = 0; u < NS; ++u) xs[u][0] = sx[min((u * NW + wave) * BPW + lb, CAP - 1) * BS + ec];
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int bl = (u * NW + wave) * BPW + lb;
            const bool ok = active_lane && bl < nblk;
            const double p = prod<TV, TX>(ok ? av[u][0] : (TV)0, ok ? xs[u][0] : (TX)0);
            double sm = p;
            const double p1 = lane_down1(p);
            const double p2 = lane_down1(p1);
            sm = sm + p1;
            sm = sm + p2;
            if constexpr (BS == 4) {
                const double p3 = lane_down1(p2);
                sm = sm + p3;
            }
            if (ok && ec == 0) part[bl * BS + er] = sm;
        }
    }
    __syncthreads();

    double res[RS];
#pragma unroll
    for (int s = 0; s < RS; ++s) {
        res[s] = 0.0;
        if (has_row[s]) {
            const int q1 = srow[kk[s]], q2 = srow[kk[s] + 1];
            double sum = 0.0;
            for (int q = q1; q < q2; ++q) sum += part[q * BS + r];
            if constexpr (MODE == SM_SPMV) res[s] = sum;
            else res[s] = rhs_v[s] - sum;
        }
    }
    if constexpr (MODE == SM_SUB_DINV) {
#pragma unroll
        for (int s = 0; s < RS; ++s) tv[s * NT + t] = res[s];
        __syncthreads();
#pragma unroll
        for (int s = 0; s < RS; ++s) {
            if (has_row[s]) {
                double z = 0.0;
#pragma unroll
                for (int c = 0; c < BS; ++c) z += dv[s][c] * tv[s * NT + k0 * BS + c];
                res[s] = z;
            }
        }
    }
    if constexpr (MODE == SM_SUB_Z) {
        // z = D^-1 (y - 0) from the stored y, with the second sweep's
        // arithmetic (y - 0.0 is y exactly, signed zeros included)
#pragma unroll
        for (int s = 0; s < RS; ++s) tv[s * NT + t] = (double)from_double<TX>(res[s]);
        __syncthreads();
#pragma unroll
        for (int s = 0; s < RS; ++s) {
            if (has_row[s] && zrow[s]) {
                double z = 0.0;
#pragma unroll
                for (int c = 0; c < BS; ++c) z += dv[s][c] * tv[s * NT + k0 * BS + c];
                a.zout[gi[s]] = from_double<TX>(z);
            }
        }
    }
    double red[2] = {0.0, 0.0};
#pragma unroll
    for (int s = 0; s < RS; ++s) {
        if (!has_row[s]) continue;
        const TX o = from_double<TX>(res[s]);
        if constexpr (KEEP) keep[t] = o;
        if constexpr (V != 1) {
            a.out[gi[s]] = o;
            if (a.out2) a.out2[gi[s]] = o;
            if (a.out3) a.out3[gi[s]] = o;
        }
        const double od = (double)o;
        double v0 = 0.0, v1 = 0.0;
        if constexpr (RED == SR_W0) v0 = od * w_v[s];
        if constexpr (RED == SR_SELF) v0 = od * od;
        if constexpr (RED == SR_BOTH) {
            v0 = od * w_v[s];
            v1 = od * od;
        }
        red[0] = s == 0 ? v0 : red[0] + v0;
        red[1] = s == 0 ? v1 : red[1] + v1;
    }
    if constexpr (NRED == 1) {
        double v1[1] = {red[0]};
        block_partials<NT, 1>(v1, a.ra, slot);
    } else if constexpr (NRED == 2) {
        block_partials<NT, 2>(red, a.ra, slot);
    }
}

// The tile's rows of rhs - A x for a start vector x that is all zero and a
// matrix whose values are all finite.  The full kernel's products are then
// signed zeros and its row sum, which starts from +0.0, is +0.0, so
// rhs - sum is rhs exactly; this path writes the same outputs and the same
// reduction partials without streaming the matrix.
template <int BS, int MODE, int RED, typename TV, typename TX>
__device__ inline void stream_rows_zero_x(const StreamArgs<TV, TX> &a, const int4 d, int slot) {
    static_assert(MODE == SM_SUB && (RED & SR_W0) == 0, "zero start vector: the plain residual only");
    constexpr int NT = 64 * BS;
    constexpr int NRED = NRedOf<RED>::value;
    const int r0 = d.x, nrow = d.z - d.x;
    const int t = threadIdx.x;
    const int k0 = t / BS, r = t - k0 * BS;
    double red[2] = {0.0, 0.0};
    if (k0 < nrow) {
        const size_t gi = (size_t)(r0 + k0) * BS + r;
        const double sum = 0.0;
        const TX o = from_double<TX>((double)a.rhs[gi] - sum);
        a.out[gi] = o;
        if (a.out2) a.out2[gi] = o;
        if (a.out3) a.out3[gi] = o;
        const double od = (double)o;
        if constexpr (RED == SR_SELF) red[0] = od * od;
    }
    if constexpr (NRED == 1) {
        double v1[1] = {red[0]};
        block_partials<NT, 1>(v1, a.ra, slot);
    }
}

template <int BS, int MODE, int RED, typename TV, typename TX, int ROWS = kTileRows, bool ZX = false, bool PLAIN = false>
__global__ __launch_bounds__(64 * BS) void k_stream(StreamArgs<TV, TX> a) {
    // the stop flag and the tile descriptor are loaded together (one scalar
    // round trip instead of two before the first vector load): the exit test
    // also reads the descriptor (a negative entry never occurs), so the
    // compiler cannot sink its load below the branch
    const int tile = a.tile0 + xcd_tile(blockIdx.x, a.ntiles);
    const int4 d = a.tile[tile];
    if ((a.ra.st->done | ((d.x | d.y | d.z | d.w) < 0 ? 1 : 0)) != 0) return;
    // the partial sums go to the tile's own slot, so that the fixed-order sum
    // over them does not depend on the order in which tiles are launched
    const int slot = NRedOf<RED>::value > 0 ? a.slot[tile] : 0;
    static_assert(!ZX || ROWS == 64, "the zero-start path serves one row per thread");
    if constexpr (ZX) {
        if (a.ra.st->full_residual == 0) {
            stream_rows_zero_x<BS, MODE, RED, TV, TX>(a, d, slot);
            return;
        }
    }
    stream_tile<BS, MODE, RED, TV, TX, ROWS, false, PLAIN>(a, d, nullptr, slot);
}

// ---------------------------------------------------------------------------
// The two sweep launches around the last colour C-1, merged.  When the
// first sweep's last colour is one tile and at most 64 rows of colour C-2
// carry blocks of the second sweep (their blocks can only point into colour
// C-1), one launch does:
//  - workgroup 0: the first sweep's tile of colour C-1 (which writes z
//    there), then, from those z values kept in LDS, the second sweep's rows
//    of colour C-2 that carry blocks;
//  - the other workgroups: the second sweep's rows of colour C-2 without
//    blocks, z_i = D_i^-1 y_i (or y_i without a diagonal inverse).
// Every row is computed with the arithmetic of k_stream, in the same order.
template <typename TV, typename TX>
struct TailArgs {
    const int32_t *rowptr;   // second-sweep pattern
    const int32_t *colidx;
    const TV *val;
    const TX *rhs;           // second sweep: out = [D^-1] (rhs - S2 out)
    TX *out;
    const double *dinv;
    int32_t row0, nrows;     // the rows of colour C-2
    const int32_t *dep;      // its rows that carry blocks (ndep <= 64)
    int32_t ndep;
};

// Sweep launches of fewer than kPlainBelow tiles read FP32 matrix values and
// column indices with plain loads instead of non-temporal ones (the merged
// tail launch always does): their few MB then stay in the last-level cache
// from one preconditioner application to the next, while the big launches'
// non-temporal streams do not displace them.  On a 128x128x64 grid this
// takes 0.5 us from the second lower colour, 0.4 us from the tall upper
// colour 1 and 0.2 us from the tail per application.  FP64 sweeps keep
// non-temporal loads: with FP64 vectors twice as large, keeping their matrix
// data cached slowed the FP64 SpMV by 1.6 us.
#ifndef BDA_PLAIN_BELOW
#define BDA_PLAIN_BELOW 8192
#endif
constexpr int kPlainBelow = BDA_PLAIN_BELOW;
template <typename TV> constexpr bool plain_small() { return std::is_same<TV, float>::value && kPlainBelow > 0; }

template <int BS, int MODEA, int MODEB, typename TV, typename TX>
__global__ __launch_bounds__(64 * BS) void k_tail(StreamArgs<TV, TX> a, TailArgs<TV, TX> b) {
    constexpr int BB = BS * BS;
    constexpr int NT = 64 * BS;
    __shared__ TX zk[64 * BS];
    __shared__ double tvb[NT];
    if (a.ra.st->done) return;
    const int t = threadIdx.x;
    if (blockIdx.x == 0) {
        const int4 d = a.tile[a.tile0];
        stream_tile<BS, MODEA, SR_NONE, TV, TX, 64, true, plain_small<TV>()>(a, d, zk, 0);
        __syncthreads();
        const int r0 = d.x;
        const int k = t / BS, r = t - k * BS;
        const bool has = k < b.ndep;
        const int i = b.ndep > 0 ? b.dep[min(k, b.ndep - 1)] : b.row0;
        const size_t gi = (size_t)i * BS + r;
        double res = 0.0;
        if (has) {
            double sum = 0.0;
            for (int q = b.rowptr[i]; q < b.rowptr[i + 1]; ++q) {
                const int j = b.colidx[q] - r0;
                const TV *av = b.val + (size_t)q * BB + r * BS;
                double sm = prod<TV, TX>(av[0], zk[j * BS]);
                sm = sm + prod<TV, TX>(av[1], zk[j * BS + 1]);
                sm = sm + prod<TV, TX>(av[2], zk[j * BS + 2]);
                if constexpr (BS == 4) sm = sm + prod<TV, TX>(av[3], zk[j * BS + 3]);
                sum += sm;
            }
            res = (double)b.rhs[gi] - sum;
        }
        if constexpr (MODEB == SM_SUB_DINV) {
            tvb[t] = res;
            __syncthreads();
            if (has) {
                const double *D = b.dinv + gi * BS;
                double z = 0.0;
#pragma unroll
                for (int c = 0; c < BS; ++c) z += D[c] * tvb[k * BS + c];
                res = z;
            }
        }
        if (has) b.out[gi] = from_double<TX>(res);
    } else {
        const int i = b.row0 + (int)(blockIdx.x - 1) * NT + t;
        if (i < b.row0 + b.nrows && b.rowptr[i + 1] == b.rowptr[i]) {
            double res[BS];
#pragma unroll
            for (int c = 0; c < BS; ++c) res[c] = (double)b.rhs[(size_t)i * BS + c] - 0.0;
#pragma unroll
            for (int rr = 0; rr < BS; ++rr) {
                double o = res[rr];
                if constexpr (MODEB == SM_SUB_DINV) {
                    const double *D = b.dinv + ((size_t)i * BS + rr) * BS;
                    double z = 0.0;
#pragma unroll
                    for (int c = 0; c < BS; ++c) z += D[c] * res[c];
                    o = z;
                }
                b.out[(size_t)i * BS + rr] = from_double<TX>(o);
            }
        }
    }
}

// The extended merged tail: k_tail's work on colours C-1 and C-2 and the
// second sweep's colour C-3 in one launch, for a forward preconditioner
// whose first sweep has a unit diagonal.  The first sweep's launch of colour
// C-2 (SM_SUB_Z) has already written z for the rows of colour C-2 without
// second-sweep blocks.  Of colour C-3, the rows whose blocks reach colour
// C-1 or a colour C-2 row with blocks ("special" rows, at most 64) depend
// on this launch's own results; the others only on z written before it.
//  - workgroup 0: the first sweep's tile of colour C-1 (z kept in LDS), the
//    colour C-2 rows with blocks (z kept in LDS), then the special rows of
//    colour C-3;
//  - the other workgroups: tiles of the other colour C-3 rows, as k_stream.
// Workgroup 0 loads the operands of both row sets (packed lists of at most
// kTail3Blocks blocks per row, with the source of each block's z) before the
// tile's own loads, so that its chain stays at the tile's three memory
// round trips.  Every row is computed with k_stream's arithmetic, in the
// same order.
template <typename TV, typename TX>
struct Tail3Args {
    const TV *val;           // second-sweep values
    const TX *rhs2;          // y of colour C-2
    const TX *rhs3;          // y of colour C-3 (p where the first sweep skipped colour 0)
    TX *out;                 // z
    const double *dinv;
    int32_t n2, n3;          // colour C-2 rows with blocks; special rows of colour C-3
    const int32_t *row;      // those rows: row[k] (k < n2), row[64 + k] (k < n3)
    const int32_t *blk;      // their blocks, kTail3Blocks per row (the same layout), -1 past the last
    const int32_t *src;      // z of the block's column: >= 0 that row, in memory; -1 - s: LDS slot s
                             // (s < 64: row s of colour C-1; 64 + j: the j-th colour C-2 row)
};

template <int BS, typename TV, typename TX>
struct Tail3Rows {
    int32_t blk[kTail3Blocks], src[kTail3Blocks];
    TV v[kTail3Blocks][BS];
    TX zg[kTail3Blocks][BS];
    double rhs, d[BS];
};

// operands of packed row k (slot base + k) for thread component r; GZ: some
// blocks may take z from memory
template <int BS, bool GZ, typename TV, typename TX>
__device__ inline void tail3_load(const Tail3Args<TV, TX> &b, const TX *rhs, int base, int k, int n, int r,
                                  Tail3Rows<BS, TV, TX> &o) {
    constexpr int BB = BS * BS;
    const int kc = min(k, max(n - 1, 0));
    const int i = n > 0 ? b.row[base + kc] : 0;
#pragma unroll
    for (int m = 0; m < kTail3Blocks; ++m) {
        o.blk[m] = n > 0 ? b.blk[(base + kc) * kTail3Blocks + m] : -1;
        o.src[m] = n > 0 ? b.src[(base + kc) * kTail3Blocks + m] : -1;
    }
    const size_t gi = (size_t)i * BS + r;
    o.rhs = (double)rhs[gi];
#pragma unroll
    for (int c = 0; c < BS; ++c) o.d[c] = b.dinv[gi * BS + c];
#pragma unroll
    for (int m = 0; m < kTail3Blocks; ++m) {
        const TV *av = b.val + (size_t)max(o.blk[m], 0) * BB + r * BS;
        const TX *zp = b.out + (size_t)max(o.src[m], 0) * BS;
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            o.v[m][c] = av[c];
            o.zg[m][c] = GZ ? zp[c] : (TX)0;
        }
    }
}

// rhs - sum over the row's blocks, in column order, of the block products
template <int BS, bool GZ, typename TV, typename TX>
__device__ inline double tail3_row(const Tail3Rows<BS, TV, TX> &o, const TX *zk, const TX *zd) {
    double sum = 0.0;
#pragma unroll
    for (int m = 0; m < kTail3Blocks; ++m) {
        if (o.blk[m] < 0) break;
        TX zc[BS];
        const bool mem = GZ && o.src[m] >= 0;
        const int sl = mem ? 0 : -1 - o.src[m];
        const TX *zl = sl < 64 ? zk + sl * BS : zd + (sl - 64) * BS;
#pragma unroll
        for (int c = 0; c < BS; ++c) zc[c] = mem ? o.zg[m][c] : zl[c];
        double sm = prod<TV, TX>(o.v[m][0], zc[0]);
        sm = sm + prod<TV, TX>(o.v[m][1], zc[1]);
        sm = sm + prod<TV, TX>(o.v[m][2], zc[2]);
        if constexpr (BS == 4) sm = sm + prod<TV, TX>(o.v[m][3], zc[3]);
        sum += sm;
    }
    return o.rhs - sum;
}

template <int BS, int ROWSB, typename TV, typename TX>
__global__ __launch_bounds__(64 * BS) void k_tail3(StreamArgs<TV, TX> a, Tail3Args<TV, TX> b,
                                                   StreamArgs<TV, TX> c) {
    constexpr int NT = 64 * BS;
    if (a.ra.st->done) return;
    const int t = threadIdx.x;
    if (blockIdx.x > 0) {
        const int4 d = c.tile[c.tile0 + xcd_tile(blockIdx.x - 1, c.ntiles)];
        stream_tile<BS, SM_SUB_DINV, SR_NONE, TV, TX, ROWSB, false, plain_small<TV>()>(c, d, nullptr, 0);
        return;
    }
    __shared__ TX zk[64 * BS];
    __shared__ TX zd[64 * BS];
    __shared__ double tvb[NT];
    const int k = t / BS, r = t - k * BS;
    Tail3Rows<BS, TV, TX> o2, o3;
    tail3_load<BS, false, TV, TX>(b, b.rhs2, 0, k, b.n2, r, o2);
    tail3_load<BS, true, TV, TX>(b, b.rhs3, 64, k, b.n3, r, o3);
    const int4 d = a.tile[a.tile0];
    stream_tile<BS, SM_SUB_DINV, SR_NONE, TV, TX, 64, true, plain_small<TV>()>(a, d, zk, 0);
    __syncthreads();
    // colour C-2 rows with blocks (only into colour C-1)
    {
        const bool has = k < b.n2;
        tvb[t] = has ? tail3_row<BS, false, TV, TX>(o2, zk, zd) : 0.0;
        __syncthreads();
        if (has) {
            double z = 0.0;
#pragma unroll
            for (int cc = 0; cc < BS; ++cc) z += o2.d[cc] * tvb[k * BS + cc];
            const TX zo = from_double<TX>(z);
            b.out[(size_t)b.row[k] * BS + r] = zo;
            zd[t] = zo;
        }
    }
    __syncthreads();
    // the special rows of colour C-3
    {
        const bool has = k < b.n3;
        tvb[t] = has ? tail3_row<BS, true, TV, TX>(o3, zk, zd) : 0.0;
        __syncthreads();
        if (has) {
            double z = 0.0;
#pragma unroll
            for (int cc = 0; cc < BS; ++cc) z += o3.d[cc] * tvb[k * BS + cc];
            b.out[(size_t)b.row[64 + k] * BS + r] = from_double<TX>(z);
        }
    }
}

// The first sweep's colours c and c + 1 in one launch (forward
// preconditioner, unit diagonal).  The first a.ntiles workgroups run colour
// c's tiles.  Each of the others runs one tile of colour c + 1 ("own" rows)
// after recomputing, into LDS, y of the colour c rows its own rows read: a
// dependency tile v (packed copies of those rows, built at analysis, with
// values refreshed at each factorisation) whose results stay in LDS, then
// the own tile g, whose column indices of those rows are LDS places.  Both
// compute every row with k_stream's arithmetic, so the recomputed y equal
// the ones colour c's tiles store, bit for bit.  MODEB: SM_SUB, or SM_SUB_Z
// when colour c + 1 is colour C-2 and the extended tail follows.
template <int BS, int MODEB, int ROWSA, typename TV, typename TX>
__global__ __launch_bounds__(64 * BS) void k_lower2(StreamArgs<TV, TX> a, StreamArgs<TV, TX> v,
                                                    StreamArgs<TV, TX> g) {
    if (a.ra.st->done) return;
    if ((int)blockIdx.x < a.ntiles) {
        const int4 d = a.tile[a.tile0 + xcd_tile(blockIdx.x, a.ntiles)];
        stream_tile<BS, SM_SUB, SR_NONE, TV, TX, ROWSA, false, plain_small<TV>()>(a, d, nullptr, 0);
        return;
    }
    __shared__ TX ydep[64 * BS];
    const int j = xcd_tile(blockIdx.x - a.ntiles, g.ntiles);
    const int4 dv = v.tile[j];
    const int4 dg = g.tile[j];
    stream_tile<BS, SM_SUB, SR_NONE, TV, TX, 64, true, plain_small<TV>(), 1>(v, dv, ydep, 0);
    __syncthreads();
    stream_tile<BS, MODEB, SR_NONE, TV, TX, 64, false, plain_small<TV>(), 2>(g, dg, nullptr, 0, ydep);
}

// ---------------------------------------------------------------------------
// Fused BiCGSTAB vector updates.  Each thread handles kVecPairsPerThread
// element pairs.  Vectors carry kVecPad elements of padding: the pair at an
// odd tail has padding as its second element, and threads past the end all
// work on one dummy pair just after the last pair.  Padding is computed like
// the rest, never read by the matrix kernels and left out of every
// reduction.  Loads and stores are therefore unconditional, and all loads of
// a thread are in flight together.
constexpr int kVecThreads = 256;
constexpr int kVecPairsPerThread = 4;
constexpr int kVecElemsPerBlock = kVecThreads * kVecPairsPerThread * 2;
constexpr int kVecPad = 4;

template <typename T> struct Pair;
template <> struct Pair<double> { using type = double2; };
template <> struct Pair<float> { using type = float2; };

template <typename T>
__device__ inline void ld2(const T *p, int64_t pi, double &a, double &b) {
    const typename Pair<T>::type v = reinterpret_cast<const typename Pair<T>::type *>(p)[pi];
    a = (double)v.x;
    b = (double)v.y;
}
template <typename T>
__device__ inline void st2(T *p, int64_t pi, double a, double b) {
    typename Pair<T>::type v;
    v.x = (T)a;
    v.y = (T)b;
    reinterpret_cast<typename Pair<T>::type *>(p)[pi] = v;
}
template <typename T> __device__ inline double rnd(double v) { return (double)(T)v; }

struct VecIdx {
    int64_t pi[kVecPairsPerThread];   // pair index (the dummy pair past the end)
    bool ok[kVecPairsPerThread];      // pair exists
    bool second[kVecPairsPerThread];  // its second element is < N
};
__device__ inline VecIdx vec_index(int64_t N) {
    VecIdx v;
    const int64_t npairs = (N + 1) / 2;
    const int64_t base = (int64_t)blockIdx.x * (kVecElemsPerBlock / 2) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const int64_t p = base + (int64_t)j * kVecThreads;
        v.ok[j] = p < npairs;
        v.pi[j] = p < npairs ? p : npairs;
        v.second[j] = 2 * p + 1 < N;
    }
    return v;
}

// Check before a step-0 residual: sets st->full_residual when an entry of x
// is not zero (a NaN counts as nonzero) or the matrix holds a non-finite
// value (*nonfinite, recorded when the factorisation copied it).  Every
// writer stores the same value.
template <typename TX>
__global__ __launch_bounds__(kVecThreads) void k_mark_nonzero(int64_t N, const TX *x, const int32_t *nonfinite,
                                                              SolverState *st) {
    const VecIdx ix = vec_index(N);
    bool nz = *nonfinite != 0;
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        double a, b;
        ld2(x, ix.pi[j], a, b);
        nz = nz || (ix.ok[j] && !(a == 0.0)) || (ix.second[j] && !(b == 0.0));
    }
    if (nz) __hip_atomic_store(&st->full_residual, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// steps 6-8 without the x update: s = r - alpha v; ||s||^2.  Step 6,
// x += alpha phat, is deferred: k_update_xr applies it together with step 11
// (the same two roundings, one read and one write of x fewer), and
// k_finish_x applies it alone when a run ends after a first half-iteration.
template <typename TX>
__global__ __launch_bounds__(kVecThreads) void k_update_s(int64_t N, const TX *r, const TX *v, TX *s, RedArgs ra) {
    if (ra.st->done) return;
    const double malpha = -ra.st->alpha;
    const VecIdx ix = vec_index(N);
    double r0[kVecPairsPerThread], r1[kVecPairsPerThread], v0[kVecPairsPerThread], v1[kVecPairsPerThread];
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        ld2(r, ix.pi[j], r0[j], r1[j]);
        ld2(v, ix.pi[j], v0[j], v1[j]);
    }
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const double s0 = rnd<TX>(malpha * v0[j] + r0[j]), s1 = rnd<TX>(malpha * v1[j] + r1[j]);
        st2(s, ix.pi[j], s0, s1);
        acc += ix.ok[j] ? s0 * s0 : 0.0;
        acc += ix.second[j] ? s1 * s1 : 0.0;
    }
    double vv[1] = {acc};
    block_partials<kVecThreads, 1>(vv, ra, blockIdx.x);
}

// steps 6 and 11-14: x = (x + alpha phat) + omega shat, rounded to the type
// of x after each addition as two separate updates would be; r = s - omega
// t; ||r||^2 and rt.r
template <typename TX, typename TXX>
__global__ __launch_bounds__(kVecThreads) void k_update_xr(int64_t N, TXX *x, const TX *phat, const TX *shat,
                                                          const TX *s, const TX *t, TX *r, const TX *rt, RedArgs ra) {
    if (ra.st->done) return;
    const double alpha = ra.st->alpha, omega = ra.st->omega, momega = -omega;
    const VecIdx ix = vec_index(N);
    double x0[kVecPairsPerThread], x1[kVecPairsPerThread], h0[kVecPairsPerThread], h1[kVecPairsPerThread];
    double s0[kVecPairsPerThread], s1[kVecPairsPerThread], t0[kVecPairsPerThread], t1[kVecPairsPerThread];
    double q0[kVecPairsPerThread], q1[kVecPairsPerThread], p0[kVecPairsPerThread], p1[kVecPairsPerThread];
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        ld2(x, ix.pi[j], x0[j], x1[j]);
        ld2(phat, ix.pi[j], p0[j], p1[j]);
        ld2(shat, ix.pi[j], h0[j], h1[j]);
        ld2(s, ix.pi[j], s0[j], s1[j]);
        ld2(t, ix.pi[j], t0[j], t1[j]);
        ld2(rt, ix.pi[j], q0[j], q1[j]);
    }
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const double ra0 = rnd<TX>(momega * t0[j] + s0[j]), ra1 = rnd<TX>(momega * t1[j] + s1[j]);
        const double xa0 = rnd<TXX>(alpha * p0[j] + x0[j]), xa1 = rnd<TXX>(alpha * p1[j] + x1[j]);
        st2(x, ix.pi[j], omega * h0[j] + xa0, omega * h1[j] + xa1);
        st2(r, ix.pi[j], ra0, ra1);
        acc[0] += ix.ok[j] ? ra0 * ra0 : 0.0;
        acc[1] += ix.ok[j] ? q0[j] * ra0 : 0.0;
        acc[0] += ix.second[j] ? ra1 * ra1 : 0.0;
        acc[1] += ix.second[j] ? q1[j] * ra1 : 0.0;
    }
    block_partials<kVecThreads, 2>(acc, ra, blockIdx.x);
}

// End of a run: the deferred step 6, x += alpha phat, when the last
// half-iteration that ran was a first half (its ||s|| test stopped the run,
// or the run ended there); it runs whether or not the run has stopped.
template <typename TX, typename TXX>
__global__ __launch_bounds__(kVecThreads) void k_finish_x(int64_t N, TXX *x, const TX *phat, const SolverState *st) {
    if (st->last_half != 1) return;
    const double alpha = st->alpha;
    const VecIdx ix = vec_index(N);
    double x0[kVecPairsPerThread], x1[kVecPairsPerThread], p0[kVecPairsPerThread], p1[kVecPairsPerThread];
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        ld2(x, ix.pi[j], x0[j], x1[j]);
        ld2(phat, ix.pi[j], p0[j], p1[j]);
    }
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) st2(x, ix.pi[j], alpha * p0[j] + x0[j], alpha * p1[j] + x1[j]);
}

// steps 15-17: p = r + beta (p - omega v), as tmp = (-omega) v + p; p = beta tmp + r
template <typename TX>
__global__ __launch_bounds__(kVecThreads) void k_update_p(int64_t N, TX *p, const TX *r, const TX *v,
                                                         const SolverState *st) {
    if (st->done) return;
    const double beta = st->beta, momega = -st->omega;
    const VecIdx ix = vec_index(N);
    double p0[kVecPairsPerThread], p1[kVecPairsPerThread], r0[kVecPairsPerThread], r1[kVecPairsPerThread];
    double v0[kVecPairsPerThread], v1[kVecPairsPerThread];
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        ld2(p, ix.pi[j], p0[j], p1[j]);
        ld2(r, ix.pi[j], r0[j], r1[j]);
        ld2(v, ix.pi[j], v0[j], v1[j]);
    }
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const double t0 = momega * v0[j] + p0[j], t1 = momega * v1[j] + p1[j];
        st2(p, ix.pi[j], beta * t0 + r0[j], beta * t1 + r1[j]);
    }
}

// Steps 0-1 for x0 = 0 (A*0 = 0, so rt = b exactly): rt = r = p = rhs,
// x = 0, rho = rt.rt.  Used by the inner solves of the precision form.
template <typename TX, typename TXX>
__global__ __launch_bounds__(kVecThreads) void k_init_zero_x0(int64_t N, const double *rhs, TX *rt, TX *r, TX *p,
                                                             TXX *x, RedArgs ra) {
    if (ra.st->done) return;
    const VecIdx ix = vec_index(N);
    double b0[kVecPairsPerThread], b1[kVecPairsPerThread];
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) ld2(rhs, ix.pi[j], b0[j], b1[j]);
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kVecPairsPerThread; ++j) {
        const double c0 = rnd<TX>(b0[j]), c1 = rnd<TX>(b1[j]);
        st2(rt, ix.pi[j], c0, c1);
        st2(r, ix.pi[j], c0, c1);
        st2(p, ix.pi[j], c0, c1);
        st2(x, ix.pi[j], 0.0, 0.0);
        acc += ix.ok[j] ? c0 * c0 : 0.0;
        acc += ix.second[j] ? c1 * c1 : 0.0;
    }
    double vv[1] = {acc};
    block_partials<kVecThreads, 1>(vv, ra, blockIdx.x);
}

// x64 += d (outer refinement update of the precision form), unless the inner
// run produced no results: it then either never started (the outer loop had
// already stopped when its start was queued) or had a zero right-hand side,
// and d is not its output
template <typename TXX>
__global__ __launch_bounds__(kVecThreads) void k_add_correction(int64_t N, double *x, const TXX *d,
                                                                const SolverState *in) {
    if (in->no_results) return;
    const int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x;
    if (i < N) x[i] = x[i] + (double)d[i];
}

// ---------------------------------------------------------------------------
// Bandwidth references (measurement only): a streaming read and a streaming
// copy of 16-byte words, one pass, each thread kBwUnroll words kBwThreads
// apart with all its loads in flight together (nontemporal, as the matrix
// streams); the read writes one sum per wave so that its loads are kept.
// n4 must be a multiple of kBwThreads * kBwUnroll.
constexpr int kBwThreads = 256, kBwUnroll = 8;
typedef double BwVec __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kBwThreads) void k_bw_read(const double2 *src2, double *out) {
    const BwVec *src = reinterpret_cast<const BwVec *>(src2) + (size_t)blockIdx.x * kBwThreads * kBwUnroll +
                       threadIdx.x;
    BwVec v[kBwUnroll];
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) v[u] = __builtin_nontemporal_load(&src[u * kBwThreads]);
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) acc += v[u].x + v[u].y;
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) out[blockIdx.x * (kBwThreads / 64) + (threadIdx.x >> 6)] = acc;
}
__global__ __launch_bounds__(kBwThreads) void k_bw_copy(const double2 *src2, double2 *dst2) {
    const size_t base = (size_t)blockIdx.x * kBwThreads * kBwUnroll + threadIdx.x;
    const BwVec *src = reinterpret_cast<const BwVec *>(src2) + base;
    BwVec *dst = reinterpret_cast<BwVec *>(dst2) + base;
    BwVec v[kBwUnroll];
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) v[u] = __builtin_nontemporal_load(&src[u * kBwThreads]);
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) __builtin_nontemporal_store(v[u], &dst[u * kBwThreads]);
}

// ---------------------------------------------------------------------------
// Permutations between the caller's ordering and the colour ordering.
__global__ void k_gather_rows(int n, int bs, const int32_t *perm, const double *in, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * bs) return;
    const int64_t k = i / bs, c = i - k * bs;
    out[i] = in[(int64_t)perm[k] * bs + c];
}
// two vectors in one launch (the right-hand side and the start vector)
__global__ void k_gather_rows2(int n, int bs, const int32_t *perm, const double *in1, double *out1,
                               const double *in2, double *out2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * bs) return;
    const int64_t k = i / bs, c = i - k * bs;
    const int64_t j = (int64_t)perm[k] * bs + c;
    const double a = in1[j], b = in2[j];
    out1[i] = a;
    out2[i] = b;
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
    // every loop unrolled and the pivot swap done with selects over the
    // rows, so that m and v stay in registers (a swap that indexes rows by
    // the pivot puts them in scratch memory); the arithmetic and its order
    // are those of the textbook elimination
    double m[BS][BS], v[BS][BS];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            m[r][c] = a[r * BS + c];
            v[r][c] = r == c ? 1.0 : 0.0;
        }
#pragma unroll
    for (int c = 0; c < BS; ++c) {
        int piv = c;
        double best = fabs(m[c][c]);
#pragma unroll
        for (int r = c + 1; r < BS; ++r)
            if (fabs(m[r][c]) > best) { best = fabs(m[r][c]); piv = r; }
        if (!(best > 0.0)) return false;
#pragma unroll
        for (int r = c + 1; r < BS; ++r) {
            const bool sw = piv == r;
#pragma unroll
            for (int q = 0; q < BS; ++q) {
                const double mc = m[c][q], mr = m[r][q], vc = v[c][q], vr = v[r][q];
                m[c][q] = sw ? mr : mc;
                m[r][q] = sw ? mc : mr;
                v[c][q] = sw ? vr : vc;
                v[r][q] = sw ? vc : vr;
            }
        }
        const double p = m[c][c];
#pragma unroll
        for (int q = 0; q < BS; ++q) { m[c][q] = m[c][q] / p; v[c][q] = v[c][q] / p; }
#pragma unroll
        for (int r = 0; r < BS; ++r) {
            if (r == c) continue;
            const double f = m[r][c];
#pragma unroll
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
