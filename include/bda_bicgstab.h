/*
 * bda_bicgstab.h -- C ABI of the MI355X ILU0-preconditioned BiCGSTAB solver.
 *
 * The reference exposes its solver through a C++-linkage static library
 * (linearalgebra/ilu0bicgstab/xilinx/alveo_u280/vitis_20192/OPM-integration/makefile:32,63-66)
 * whose entry points take raw host arrays in a per-colour device format and
 * OpenCL handles (src/sda_app/common/fpga_functions_bicgstab.hpp:30-137).
 * The caller (the Flow FPGA backend of opm-simulators) had to colour,
 * reorder and factorise the matrix itself.  This ABI keeps the reference's
 * lifecycle -- allocate once at the maximum size, update values per linear
 * solve without reallocating, run the whole BiCGSTAB loop on the device,
 * decode a status record -- but takes a plain block-CSR matrix and does
 * colouring, colour-ordered reordering and block ILU0 itself.
 *
 * Conventions (all functions):
 *   - return 0 on success, non-zero on failure (the reference returns 0/1,
 *     fpga_functions_bicgstab.cpp:206-216); the message is available from
 *     bda_last_error();
 *   - no C++ exception crosses this boundary;
 *   - host pointers are owned by the caller and copied; device memory is
 *     owned by the handle; "_dev" entry points take device pointers
 *     (hipMalloc'ed, e.g. a torch tensor's data_ptr) and are synchronous;
 *   - one HIP stream per handle, every call blocks until its work is done;
 *     the stream is a blocking stream, so "_dev" inputs written by work on
 *     the legacy default stream are complete before the handle reads them;
 *     inputs written on other streams must be synchronised by the caller;
 *   - a handle must not be used from two threads at once; handles used from
 *     different threads run concurrently (the library's copies are on the
 *     handle's stream, never on the legacy default stream).
 *
 * Vectors passed to bda_solve*, bda_adjoint_solve* and bda_grad_values are in
 * the caller's ORIGINAL row ordering; the handle permutes internally.
 * bda_spmv_perm / bda_precond_apply_perm / bda_get_factors operate in the
 * colour-ordered (permuted) numbering and exist for component-level checks.
 */
#ifndef BDA_BICGSTAB_H
#define BDA_BICGSTAB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bda_handle_s *bda_handle;

/* Solve status, mirrors the fields decode_debuginfo_bicgstab() extracts from
 * the kernel's status line (src/sda_app/common/bicgstab_utils.cpp:90-280,
 * wrapper bits at bicgstab_solver_wrapper_2r_3r3w_simple-ddr_v1.vhd:75-82),
 * plus the breakdown / NaN flags the reference does not have. */
typedef struct bda_result {
    double iterations;        /* half_iterations/2 + 0.5 (fpga_functions_bicgstab.cpp:922-924) */
    int32_t half_iterations;  /* the kernel's iteration counter in half steps */
    int32_t last_norm_idx;    /* newest entry of norms[1..3] */
    double norms[4];          /* norms[0] = ||r0||, norms[1..3] ring of the last three norms */
    double final_norm;        /* last recurrence residual norm (||s|| or ||r||) */
    double rel_residual;      /* final_norm / norms[0] */
    int32_t converged;        /* stopped because norm <= threshold */
    int32_t no_results;       /* ||r0|| already below threshold: x untouched (wrapper bit 1) */
    int32_t max_iter_reached; /* stopped by the iteration cap */
    int32_t breakdown;        /* rho, v.rt, t.t or omega became exactly zero */
    int32_t nan_detected;     /* a norm or scalar became NaN/Inf */
    int32_t aborted;          /* host watchdog fired (wrapper bit 0, param0) */
    int32_t outer_iterations; /* refinement steps (mixed form), 0 otherwise */
    int32_t inner_half_iterations; /* total inner half iterations (mixed form) */
    double solve_ms;          /* wall time of the device solve (setup excluded) */
} bda_result;

/* Capacity / configuration query: counterpart of fpga_kernel_query()
 * (fpga_functions_bicgstab.cpp:1332-1452).  There are no hard limits beyond
 * device memory; the fields report what this build assumes. */
typedef struct bda_limits {
    int32_t max_block_size;        /* 4 */
    int32_t max_blocks_per_row;    /* longest block row the streaming kernels accept */
    int32_t max_iterations;        /* 32767: 15-bit full-iteration counter */
    int32_t rows_per_tile;         /* block rows per workgroup tile */
    int64_t device_memory_bytes;   /* total HBM of the current device */
    int32_t compute_units;
    int32_t wavefront_size;
} bda_limits;

/* Synthetic reservoir-grid generator (7-point stencil, 3x3/4x4 blocks). */
typedef struct bda_grid_spec {
    int32_t nx, ny, nz;       /* grid */
    int32_t bs;               /* block size, 3 or 4 */
    uint64_t seed;
    double perm_sigma;        /* log-permeability standard deviation */
    double layer_corr;        /* share of log-perm variance common to a z layer (0..1) */
    double inactive_frac;     /* fraction of cells marked inactive and removed */
    double nnc_frac;          /* fraction of active cells given a non-neighbour connection */
    int32_t nnc_dist;         /* Chebyshev radius of such connections */
    double pert_max;          /* off-diagonal block = -T (I + P), P ~ U[0, pert_max] */
    double upwind;            /* extra -upwind*T*I on the higher-index cell's coupling */
    double diag_noise;        /* std of the N(0, s) entries added to the diagonal block */
} bda_grid_spec;

typedef struct bda_grid_s *bda_grid;

/* ---- generator ------------------------------------------------------- */
int bda_grid_create(const bda_grid_spec *spec, bda_grid *out);
int bda_grid_sizes(bda_grid g, int32_t *n, int32_t *nnzb, int32_t *bs);
int bda_grid_copy(bda_grid g, int32_t *rowptr, int32_t *colidx, double *values, double *b);
int bda_grid_destroy(bda_grid g);

/* Host-only colouring of a block pattern, identical to what bda_analyse
 * computes; colour_ptr needs room for n+1 entries.  No device is touched. */
int bda_colour_pattern(int32_t n, const int32_t *rowptr, const int32_t *colidx, int32_t level_scheduling,
                       int32_t *colour, int32_t *perm, int32_t *colour_ptr, int32_t *num_colours);

/* ---- reference-compatible status and limits lines (host only) ---------
 * 16 words = the kernel's status line 0 and its first debug line, laid out
 * as the reference wrapper writes them (bicgstab_solver_wrapper_2r_3r3w_
 * simple-ddr_v1.vhd:51-82) and its host decodes them
 * (decode_debuginfo_bicgstab, bicgstab_utils.cpp:90-280): abort and
 * "no results" bits, cycles in units of the reference's 250 MHz kernel
 * clock, "BDA" signature, half-iteration count, the four norms.
 * Breakdown and NaN have no field there; they are reported in bda_result. */
int bda_encode_status(const bda_result *res, uint64_t line[16]);
/* 8 words = the limits line of a kernel query (fpga_kernel_query,
 * fpga_functions_bicgstab.cpp:1332-1452; decode_debuginfo_bicgstab_query,
 * bicgstab_utils.cpp:282-369), filled with this solver's capacities for a
 * handle of n block rows, nnzb_max blocks and block size bs. */
int bda_encode_query(const bda_limits *limits, int32_t n, int64_t nnzb_max, int32_t bs, uint64_t line[8]);

/* ---- diagnostics --------------------------------------------------------- */
/* Process-wide level of the diagnostic lines the library prints on stderr
 * ("INFO: ..." and "ERROR: ...", flushed at once): 0 none (the default),
 * 1 handle creation, analysis and factor summaries, one line per solve and
 * every error, 2 also per-colour sizes, the preconditioner's launch plan
 * and the norm ring of each solve.  The initial level comes from the
 * environment variable BDA_DEBUG_LEVEL.  Replaces the reference's
 * compile-time BDA_DEBUG_LEVEL printouts (src/sda_app/common/bda_utils.hpp:
 * 46-57; e.g. the per-colour sizes of fpga_functions_bicgstab.cpp:146-167).
 * bda_set_debug_level returns 2 for a negative level. */
int bda_set_debug_level(int32_t level);
int bda_get_debug_level(void);

/* ---- handle lifecycle -------------------------------------------------- */
/* Allocate for a system of n block rows and at most nnzb_max blocks of
 * size bs x bs (bs in {3,4}).  Replaces fpga_setup_host_datamem +
 * fpga_setup_device_datamem (fpga_functions_bicgstab.cpp:84-496,729-798). */
int bda_create(int32_t bs, int32_t n, int32_t nnzb_max, bda_handle *out);
int bda_destroy(bda_handle h);
const char *bda_last_error(bda_handle h);
int bda_query(bda_handle h, bda_limits *limits);

/* Block-CSR upload: rowptr[n+1], colidx[nnzb] (sorted per row, diagonal
 * present), values[nnzb*bs*bs] row-major blocks.  Replaces
 * fpga_copy_host_datamem (fpga_functions_bicgstab.cpp:506-691). */
int bda_upload_bcsr(bda_handle h, const int32_t *rowptr, const int32_t *colidx,
                    const double *values);
/* New values on the uploaded pattern (Newton step); no reallocation. */
int bda_update_values(bda_handle h, const double *values);
int bda_update_values_dev(bda_handle h, const double *d_values);
/* A counter that changes whenever the handle's matrix values or factors
 * change (upload, value update, factorisation), so that a caller holding
 * results of an earlier factorisation (an autograd backward pass) can tell
 * whether the handle still holds that system.  No reference counterpart
 * (the reference's buffers are rewritten by its caller only). */
int bda_values_generation(bda_handle h, uint64_t *gen);

/* The caller's own colour-ordered system with its own factors, as the
 * reference kernel receives them from its caller (fpga_copy_host_datamem,
 * fpga_functions_bicgstab.cpp:606-629; the colour boundaries of
 * rtl/bicgstab_solver_2r-3r3w/sparstition.vhd:673-717): rows already in
 * colour order, colour c holding rows [colour_ptr[c], colour_ptr[c+1])
 * (rows of one colour must not be coupled), A in block CSR (sorted columns,
 * diagonal present), lu_values[nnzb*bs*bs] on A's pattern holding L on the
 * strictly lower blocks (unit diagonal implicit) and U on the strictly upper
 * ones (diagonal blocks ignored), dinv[n*bs*bs] the inverted diagonal blocks
 * of U.  Nothing is coloured, reordered or factorised: the solves, the
 * adjoint and the component access then run on exactly these factors, in the
 * given numbering (the permutation is the identity).  bda_wire_import turns
 * the reference's per-colour device arrays into these arrays. */
int bda_upload_factored(bda_handle h, int32_t ncolours, const int32_t *colour_ptr, const int32_t *rowptr,
                        const int32_t *colidx, const double *values, const double *lu_values, const double *dinv);

/* Greedy colouring + colour-ordered permutation (host).  Counterpart of the
 * colour sizes / P-index arrays the reference consumes
 * (fpga_functions_bicgstab.cpp:606-613).  level_scheduling != 0 selects
 * level sets of the lower-triangular graph instead of colours (the flag of
 * fpga_setup_host_datamem, fpga_functions_bicgstab.cpp:84). */
int bda_analyse(bda_handle h, int32_t level_scheduling);
int bda_num_colours(bda_handle h, int32_t *num_colours);
/* colour[n] per original row, perm[n] (perm[k] = original row at colour-
 * ordered position k), colour_ptr[num_colours+1]. */
int bda_get_colouring(bda_handle h, int32_t *colour, int32_t *perm, int32_t *colour_ptr);

/* Block ILU0 in colour order on the device (FP64; FP32 copies of the
 * factors and of A are made for the mixed form).  No reference counterpart:
 * the reference receives L, U and the inverted diagonal blocks ready-made
 * (fpga_functions_bicgstab.cpp:615-617,629). */
int bda_factor(bda_handle h);
/* Makes the FP32 copies of A, L and U that the precision form runs on, from
 * the current factorisation.  They are otherwise made by the first
 * precision-form call after each bda_factor; this call lets a caller do
 * (and time) that step with the factorisation.  No reference counterpart. */
int bda_prepare_mixed(bda_handle h);

/* Host watchdog in seconds (0 = off), counterpart of param0's abort cycles
 * (bicgstab_solver_wrapper_2r_3r3w_simple-ddr_v1.vhd:605-643); abs_tol != 0
 * selects the absolute compare of config_bits bit 1 (solver.vhd:637-641);
 * poll_iters (1..3) sets how far the host queues ahead of the device, which
 * stops itself and raises a host-visible flag: in the middle of each
 * half-iteration the host waits until at most 2 * poll_iters - 1 earlier
 * half-iterations are unfinished. */
int bda_set_options(bda_handle h, double timeout_s, int32_t abs_tol, int32_t poll_iters);

/* Plain FP64 ILU0-BiCGSTAB, the reference's microprogram
 * (solver_pkg.vhd:120-141): x holds x0 on entry and the solution on exit.
 * maxit follows the kernel's max_iter parameter: maxit full iterations are
 * run, the next one stops after its first half.  Replaces
 * fpga_set_kernel_parameters + fpga_kernel_run + fpga_copy_from_device_debugbuf
 * + fpga_map_results (fpga_functions_bicgstab.cpp:1246-1325,890-934,1073-1190). */
int bda_solve(bda_handle h, const double *b, double *x, double tol, int32_t maxit,
              bda_result *res);
int bda_solve_dev(bda_handle h, const double *d_b, double *d_x, double tol, int32_t maxit,
                  bda_result *res);
/* The last recurrence residual (s or r) in original ordering (the R1/R2
 * result buffer of fpga_map_results). */
int bda_get_residual(bda_handle h, double *r);

/* Precision form: FP32 matrix/factor values, FP32 products, FP64 sums,
 * FP64 diagonal-block inverses, dots and scalars; inner BiCGSTAB to
 * inner_tol (relative), FP64 iterative refinement to tol (relative to the
 * initial FP64 residual).  vec32 != 0 also stores the inner Krylov vectors
 * in FP32.  No reference counterpart (the reference is FP64 only,
 * alveo_u280/rtl/pkg/constants.vhd:98-100). */
int bda_solve_mixed(bda_handle h, const double *b, double *x, double tol, int32_t max_outer,
                    double inner_tol, int32_t inner_maxit, int32_t vec32, bda_result *res);
int bda_solve_mixed_dev(bda_handle h, const double *d_b, double *d_x, double tol,
                        int32_t max_outer, double inner_tol, int32_t inner_maxit,
                        int32_t vec32, bda_result *res);
/* Refinement policy of the precision form.  adaptive != 0 (default): the
 * inner tolerance of refinement step k is max(inner_tol, eta*thr/||r_k||),
 * capped at 0.5, i.e. a step asks only for the reduction the outer target
 * still needs; adaptive = 0: every inner solve runs to inner_tol.
 * eta in (0,1], default 0.9: the last inner run stops close to the outer
 * target, as the FP64 solve does (measured on three grids: 0.9 needs 0.5 to
 * 1.5 fewer inner iterations than 0.5 at the same outer count, with an x
 * error below the FP64 solve's at the same tolerance). */
int bda_set_refinement(bda_handle h, int32_t adaptive, double eta);

/* Adjoint: solve A^T lambda = w with the transposed colour-ordered ILU0
 * (U^T sweep over colours ascending, then L^T descending).  lambda holds the
 * start vector on entry.  No reference counterpart. */
int bda_adjoint_solve(bda_handle h, const double *w, double *lambda, double tol,
                      int32_t maxit, bda_result *res);
int bda_adjoint_solve_dev(bda_handle h, const double *d_w, double *d_lambda, double tol,
                          int32_t maxit, bda_result *res);
/* The adjoint in the precision form: the refinement of bda_solve_mixed
 * (FP64 residual w - A^T lambda, FP32-value inner BiCGSTAB on the transposed
 * ILU0) for A^T lambda = w; the same arguments and refinement policy.  No
 * reference counterpart. */
int bda_adjoint_solve_mixed(bda_handle h, const double *w, double *lambda, double tol,
                            int32_t max_outer, double inner_tol, int32_t inner_maxit,
                            int32_t vec32, bda_result *res);
int bda_adjoint_solve_mixed_dev(bda_handle h, const double *d_w, double *d_lambda, double tol,
                                int32_t max_outer, double inner_tol, int32_t inner_maxit,
                                int32_t vec32, bda_result *res);
/* grad[nnzb*bs*bs] (original block layout) = -lambda_i x_j^T on every block
 * (i,j) of the pattern: dL/dA for L(x(A)) with dL/dx = w. */
int bda_grad_values(bda_handle h, const double *lambda, const double *x, double *grad);
int bda_grad_values_dev(bda_handle h, const double *d_lambda, const double *d_x,
                        double *d_grad);

/* ---- component access (colour-ordered numbering) ---------------------- */
/* y = A x (transpose != 0: y = A^T x). */
int bda_spmv_perm(bda_handle h, const double *x, double *y, int32_t transpose);
/* z = M^{-1} p with M = LU (transpose != 0: z = M^{-T} p). */
int bda_precond_apply_perm(bda_handle h, const double *p, double *z, int32_t transpose);
/* The same in each kernel form the solves run, so that every form's kernels
 * can be checked on their own.  prec selects the form and the vector type:
 *   0  FP64 values, double vectors (bda_solve, the adjoint);
 *   1  FP32 values, float vectors (bda_solve_mixed with vec32 != 0): FP32
 *      products, FP64 sums, D^-1 in FP64, each output rounded to FP32;
 *   2  FP32 values, double vectors (bda_solve_mixed with vec32 = 0): each x
 *      entry rounded to FP32 before its FP32 product, FP64 sums.
 * red selects the SpMV instance with the reductions the solve fuses into it:
 * 0 none; 1 the v = A phat launch, dots[0] = y.w; 3 the t = A shat launch,
 * dots[0] = y.w and dots[1] = y.y (w: a vector of the same type as x; forward
 * only).  The reductions are summed in the solve's fixed order.  The
 * transposed FP32 operators exist for this access only (the adjoint runs in
 * FP64).  Counterparts: the SpMV mode and the ILU0 sweeps of the reference
 * kernel (rtl/bicgstab_solver_2r-3r3w/spmvp.vhd:240-383, ilu0.vhd:385-416)
 * with dot1/dot2 (vector_ops.vhd:301-309). */
int bda_spmv_perm_ex(bda_handle h, int32_t prec, const void *x, void *y, int32_t transpose, const void *w,
                     int32_t red, double dots[2]);
int bda_precond_apply_perm_ex(bda_handle h, int32_t prec, const void *p, void *z, int32_t transpose);
/* Kernel launches one application of the preconditioner makes (transpose
 * != 0: of the adjoint's), after bda_analyse: one per colour and sweep, less
 * the launches the colouring lets the library merge.  No counterpart in the
 * reference, whose sweeps run colour by colour inside one kernel run
 * (rtl/bicgstab_solver_2r-3r3w/int_read_unit.vhd:287-302); for callers that
 * budget launch latency. */
int bda_precond_launches(bda_handle h, int32_t transpose, int32_t *launches);
/* y = A x (transpose != 0: A^T x) on the caller's device vectors (float for
 * prec 1, double otherwise) in the colour-ordered numbering, with the fused
 * reductions of bda_spmv_perm_ex; for callers that drive their own Krylov
 * loop on the handle's matrix.  Counterpart of the reference kernel's SpMV
 * mode (rtl/bicgstab_solver_2r-3r3w/spmvp.vhd:240-383, sparstition.vhd:
 * 655-878) with dot1/dot2 (vector_ops.vhd:301-309).  Its partial sums and
 * scalar state are the handle's; the first FP32 or transposed call makes
 * the handle's FP32 or transposed copies of the matrix. */
int bda_spmv_dev(bda_handle h, int32_t prec, const void *d_x, void *d_y, int32_t transpose, const void *d_w,
                 int32_t red, double dots[2]);
/* bda_spmv_dev without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc, hipFree or allocating routine: the
 * fused reductions keep their partial sums and scalar state in a 1 MiB array
 * of the library's device code (one per GPU, held until the process ends),
 * and the handle's copies of the matrix must already exist and be current
 * (prec 1/2: the FP32 copies, made by a bda_spmv_dev or precision-form call
 * after the last bda_factor; transpose: the transposed copies, made by a
 * transposed bda_spmv_dev or an adjoint solve), else it fails with a
 * message.  A system whose SpMV writes more than 65,504 partial sums per
 * reduction (its launch's tile count) is rejected when red != 0.  All calls
 * on one GPU share its array: calls that overlap in time on one GPU (two
 * threads, two streams) race.  Results equal bda_spmv_dev's bit for bit. */
int bda_spmv_dev_noalloc(bda_handle h, int32_t prec, const void *d_x, void *d_y, int32_t transpose,
                         const void *d_w, int32_t red, double dots[2]);
/* Factors in the colour-ordered numbering: L and U in the strictly-lower /
 * strictly-upper block pattern of the permuted matrix (block counts from
 * bda_factor_sizes), dinv[n*bs*bs] the inverted diagonal blocks. */
int bda_factor_sizes(bda_handle h, int32_t *nnzb_lower, int32_t *nnzb_upper);
int bda_get_factors(bda_handle h, int32_t *l_rowptr, int32_t *l_colidx, double *l_values,
                    int32_t *u_rowptr, int32_t *u_colidx, double *u_values, double *dinv);

/* ---- norm history ------------------------------------------------------
 * Replaces the reference's periodic debug lines, which carry the last four
 * norms every debug_rate cycles (solver.vhd:1074-1148): with a capacity set,
 * every run on the handle's main state records ||r0|| and then the residual
 * norm of every half iteration, in order, on the device.  bda_get_history
 * copies up to `capacity` of them from the last run (for the precision form:
 * the last inner run) and reports how many the run produced (h + 2 for a run
 * that stopped after h half iterations). */
int bda_set_history(bda_handle h, int32_t capacity);
int bda_get_history(bda_handle h, double *norms, int32_t capacity, int32_t *count);

/* ---- the reference's per-colour device format (host only) ---------------
 * The arrays the reference's host library copies into the FPGA's data
 * buffers: vector_pointers[0..20] and vectorSizes[0..17] as
 * fpga_copy_host_datamem reads them (fpga_functions_bicgstab.cpp:506-629),
 * for A (slots 0-4), L (6-10), U (12-16), D^-1 (18), b (19) and x0 (20):
 *   [6k+0] int32[8 + 4*parts]  8 header words, then {row, arow, col, val}
 *   [6k+1] uint32[sizes[6k+3]] gathered-entry (P) indices, 16 per partition
 *   [6k+2] double*[1]          -> double[nnz_val_sizes[k]] stream values
 *   [6k+3] uint16[sizes[6k+1]] slot of each stream entry in the P list
 *   [6k+4] uint8[sizes[6k+4]]  row advance of each entry, 64 per partition
 *   [18]   double[sizes[5]]    D^-1, 4 per scalar row of U's partitions
 *   [19], [20] double[N]       b and x0 in colour order
 * A "colour" of the format is a partition of a colour within the device
 * limits (constants.vhd:80-98: 2048 rows, 8192 gathered entries, 32768
 * entries).  device_fits: the system also fits the U280 design as a whole
 * (N <= 163840, <= 256 partitions and <= 2^21 entries per matrix). */
typedef struct bda_wire_sizes {
    int32_t vector_sizes[18];
    int32_t nnz_val_sizes[3];
    int32_t partitions[3];
    int32_t device_fits;
} bda_wire_sizes;

/* Replaces the reference's consumer-side inputs (fpga_functions_bicgstab.
 * cpp:606-629): builds the arrays from the handle's analysed and factored
 * system; b and x0 in the caller's ordering (either may be NULL: zeros). */
int bda_wire_export_sizes(bda_handle h, bda_wire_sizes *ws);
int bda_wire_export(bda_handle h, const double *b, const double *x0, void *vector_pointers[21]);

/* The same from colour-ordered host arrays (no handle, no GPU): A, L, U as
 * block CSR with sorted columns, dinv[n*bs*bs], b_perm and x0_perm in colour
 * order (may be NULL). */
int bda_wire_build_sizes(int32_t bs, int32_t n, const int32_t *colour_ptr, int32_t ncolours,
                         const int32_t *a_rowptr, const int32_t *a_colidx, const int32_t *l_rowptr,
                         const int32_t *l_colidx, const int32_t *u_rowptr, const int32_t *u_colidx,
                         bda_wire_sizes *ws);
int bda_wire_build(int32_t bs, int32_t n, const int32_t *colour_ptr, int32_t ncolours, const int32_t *a_rowptr,
                   const int32_t *a_colidx, const double *a_values, const int32_t *l_rowptr,
                   const int32_t *l_colidx, const double *l_values, const int32_t *u_rowptr,
                   const int32_t *u_colidx, const double *u_values, const double *dinv, const double *b_perm,
                   const double *x0_perm, void *vector_pointers[21]);

/* The five data buffers and the 5-line setup block of the FPGA kernel:
 * fpga_setup_host_datamem's layout (fpga_functions_bicgstab.cpp:84-496) and
 * fpga_copy_host_datamem's copy (:506-629).  bda_wire_layout gives the
 * buffer sizes in bytes and the 40 setup words; bda_wire_pack fills
 * caller-allocated buffers (the dump_input_data_<b>_seq_<s>.bin files of
 * :649-665 are these buffers); bda_wire_unpack points vector_pointers into
 * packed buffers (value_arrays holds the three double* the double** slots
 * point to). */
int bda_wire_layout(const int32_t vector_sizes[18], const int32_t nnz_val_sizes[3], uint32_t config_bits,
                    uint32_t total_size[5], uint64_t setup[40]);
int bda_wire_pack(const int32_t vector_sizes[18], const int32_t nnz_val_sizes[3], uint32_t config_bits,
                  void *const vector_pointers[21], uint8_t *buffers[5]);
int bda_wire_unpack(uint8_t *const buffers[5], int32_t vector_sizes[18], int32_t nnz_val_sizes[3],
                    void *vector_pointers[21], double *value_arrays[3]);

/* Import: the arrays back into colour-ordered block CSR.  A's pattern is
 * the union of the blocks its stream touches; lu_values (on A's pattern)
 * receives L on the strictly lower and U on the strictly upper blocks;
 * part_rows[parts+1] the first scalar row of each of A's partitions. */
int bda_wire_import_sizes(int32_t bs, const int32_t vector_sizes[18], void *const vector_pointers[21],
                          int32_t *n, int64_t *nnzb, int32_t *num_parts);
int bda_wire_import(int32_t bs, const int32_t vector_sizes[18], void *const vector_pointers[21], int32_t *rowptr,
                    int32_t *colidx, double *values, double *lu_values, double *dinv, int32_t *part_rows,
                    double *b, double *x0);

/* ---- measurement ------------------------------------------------------ */
/* Kernel classes timed with HIP events on the handle's stream while
 * profiling is on: 0 = SpMV, 1 = lower sweep, 2 = upper sweep, 3 = vector
 * updates.  Counts and summed milliseconds since the last reset. */
int bda_set_profiling(bda_handle h, int32_t on);
int bda_get_kernel_stats(bda_handle h, int32_t kind, int64_t *launches, double *total_ms);
/* The individual durations (ms) of the launches counted above, in launch
 * order: up to capacity values go to ms, *count receives how many there
 * are.  Kind 0 brackets the streaming SpMV kernel alone (not the reduction
 * that follows it).  A diagnostic of this build; no counterpart in the
 * reference. */
int bda_get_kernel_times(bda_handle h, int32_t kind, float *ms, int32_t capacity, int32_t *count);
/* reps back-to-back launches of the solve's SpMV kernel (v = A phat with
 * the fused v.rt partial sums; FP64, or FP32 values and vectors with
 * fp32 != 0) on device-resident vectors; average launch time in ms from two
 * HIP events on the handle's stream. */
int bda_time_spmv(bda_handle h, int32_t reps, int32_t fp32, double *avg_ms);
/* reps back-to-back preconditioner applications z = M^-1 p (all launches of
 * the L and U sweeps; FP64, or the precision form's FP32 values and vectors
 * with fp32 != 0) on device-resident vectors; average time of one
 * application in ms from two HIP events on the handle's stream. */
int bda_time_precond(bda_handle h, int32_t reps, int32_t fp32, double *avg_ms);
/* The device's practical streaming bandwidth, for reading measured rates
 * against: a one-pass read of `bytes` (rounded down to 32 KiB; 16-byte
 * nontemporal loads, eight per thread in flight together, one sum per wave
 * written) and a copy of the same size, reps launches each after one
 * warm-up, timed with HIP events on the handle's stream; GB/s = bytes moved
 * / average time.  Uses two scratch buffers of `bytes` for the call.  A
 * measurement aid of this build, no reference counterpart. */
int bda_measure_bandwidth(bda_handle h, int64_t bytes, int32_t reps, double *read_gbs, double *copy_gbs);
/* Algorithmic bytes of one launch of the kernel classes above (FP64 form). */
int bda_algorithmic_bytes(bda_handle h, int32_t kind, int32_t fp32, double *bytes);

#ifdef __cplusplus
}
#endif

#endif /* BDA_BICGSTAB_H */
