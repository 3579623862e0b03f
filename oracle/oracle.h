/*
 * CPU oracle of the ILU0-BiCGSTAB path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference algorithm, used by tests/ and by
 * bench.py's CPU baseline as the checker.  The product never links, loads or
 * calls it.  Each function cites the reference file:line it follows
 * (paths relative to linearalgebra/ilu0bicgstab/xilinx/src/).
 */
#ifndef BDA_ORACLE_H
#define BDA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system *orc_handle;

typedef struct orc_result {
    double iterations;
    int32_t half_iterations;
    int32_t last_norm_idx;
    double norms[4];
    double final_norm;
    double rel_residual;
    int32_t converged, no_results, max_iter_reached, breakdown, nan_detected, aborted;
    int32_t outer_iterations, inner_half_iterations;
    double solve_ms;
} orc_result;

void orc_set_threads(int32_t threads);
int32_t orc_get_threads(void);

/* Greedy colouring (first fit, natural order, symmetrised graph) or level
 * sets; returns the number of colours, -1 on error. */
int32_t orc_colour(int32_t n, const int32_t *rowptr, const int32_t *colidx, int32_t level_scheduling,
                   int32_t *colour);

int orc_create(int32_t bs, int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *values,
               int32_t level_scheduling, orc_handle *out);
int orc_destroy(orc_handle h);
int orc_num_colours(orc_handle h, int32_t *nc);
int orc_get_colouring(orc_handle h, int32_t *colour, int32_t *perm, int32_t *colour_ptr);
int orc_update_values(orc_handle h, const double *values);
int orc_factor(orc_handle h);
int orc_factor_sizes(orc_handle h, int32_t *nl, int32_t *nu);
int orc_get_factors(orc_handle h, int32_t *l_rowptr, int32_t *l_colidx, double *l_values, int32_t *u_rowptr,
                    int32_t *u_colidx, double *u_values, double *dinv);
/* colour-ordered numbering */
int orc_spmv_perm(orc_handle h, const double *x, double *y, int32_t transpose);
int orc_precond_perm(orc_handle h, const double *p, double *z, int32_t transpose);
/* prec 0: FP64 (double vectors); 1: FP32 values, float vectors; 2: FP32
 * values, double vectors -- the three kernel forms of the solves */
int orc_spmv_perm_ex(orc_handle h, int32_t prec, const void *x, void *y, int32_t transpose);
int orc_precond_perm_ex(orc_handle h, int32_t prec, const void *p, void *z, int32_t transpose);
/* original numbering */
int orc_solve(orc_handle h, const double *b, double *x, double tol, int32_t maxit, int32_t abs_tol,
              orc_result *res);
int orc_solve_mixed(orc_handle h, const double *b, double *x, double tol, int32_t max_outer, double inner_tol,
                    int32_t inner_maxit, int32_t vec32, int32_t adaptive, double eta, orc_result *res);
int orc_adjoint_solve(orc_handle h, const double *w, double *lam, double tol, int32_t maxit, orc_result *res);
/* A^T lam = w in the precision form (the refinement of orc_solve_mixed on the transposed operators). */
int orc_adjoint_solve_mixed(orc_handle h, const double *w, double *lam, double tol, int32_t max_outer,
                            double inner_tol, int32_t inner_maxit, int32_t vec32, int32_t adaptive, double eta,
                            orc_result *res);
/* ||r0|| and every half-iteration's residual norm of the last orc_solve / orc_adjoint_solve. */
int orc_get_history(orc_handle h, double *norms, int32_t capacity, int32_t *count);
int orc_grad(orc_handle h, const double *lam, const double *x, double *grad);

#ifdef __cplusplus
}
#endif
#endif
