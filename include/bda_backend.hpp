/*
 * bda_backend.hpp -- a linear-solver backend object over the C ABI of
 * bda_bicgstab.h, in the shape of the solver slot the reference's caller
 * fills (the FPGA backend of opm-simulators, which links the reference's
 * static library, linearalgebra/ilu0bicgstab/xilinx/alveo_u280/vitis_20192/
 * OPM-integration/makefile:32,63-66, and drives the fpga_* sequence of
 * src/sda_app/common/fpga_functions_bicgstab.hpp:30-137 once per linear
 * solve of a Newton iteration).
 *
 * Header-only C++17; links against libfpga_amd.so.  One object per thread,
 * every call blocks, device memory stays resident between solves:
 *
 *   bda_backend::Mi355xBackend be(block_size, maxit, tolerance);
 *   for each Newton step:
 *       be.solve_system(n, nnzb, rowptr, colidx, values, b, stats);   // x0 = 0
 *       be.get_result(x);
 *
 * The first call sets the handle up for the system's size (fpga_setup_host_
 * datamem + fpga_setup_device_datamem), uploads the block CSR matrix and
 * colours it (fpga_copy_host_datamem and the caller's own colouring); later
 * calls with the same sparsity pattern only copy the new values; a changed
 * pattern is analysed again, a larger one reallocates.  Every call
 * factorises (block ILU0 on the GPU) and solves (fpga_set_kernel_parameters
 * + fpga_kernel_run + fpga_map_results).
 *
 * solve_prefactored() is the reference's own boundary: a system the caller
 * has already coloured, reordered and factorised (colour sizes, A, L, U and
 * the inverted diagonal blocks, fpga_functions_bicgstab.cpp:606-629), solved
 * as given.
 */
#ifndef BDA_BACKEND_HPP
#define BDA_BACKEND_HPP

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "bda_bicgstab.h"

namespace bda_backend {

enum class SolverStatus {
    Success = 0,                     // the solve ran (converged or not: see Stats)
    AnalysisFailed = 1,              // upload or colouring refused the system
    CreatePreconditionerFailed = 2,  // block ILU0 failed (singular diagonal block)
    UnknownError = 3,                // the solve itself reported an error
};

struct Stats {
    int iterations = 0;         // full iterations, half steps rounded up (h/2 + 0.5, fpga_functions_bicgstab.cpp:922-924)
    int half_iterations = 0;    // the kernel's counter
    double reduction = 0.0;     // final residual norm / initial residual norm
    bool converged = false;
    double conv_rate = 0.0;     // reduction^(1/iterations)
    double elapsed = 0.0;       // seconds, wall clock around the solve
    bda_result raw{};           // the full status record
};

class Mi355xBackend {
public:
    // precision_form: FP32 inner solves with FP64 refinement to the same
    // tolerance (bda_solve_mixed) instead of the plain FP64 solve.
    Mi355xBackend(int block_size, int maxit, double tolerance, bool precision_form = false)
        : bs_(block_size), maxit_(maxit), tol_(tolerance), mixed_(precision_form) {}
    ~Mi355xBackend() { release(); }
    Mi355xBackend(const Mi355xBackend &) = delete;
    Mi355xBackend &operator=(const Mi355xBackend &) = delete;

    // A (n block rows, nnzb blocks, row-major bs x bs blocks) x = b in the
    // caller's ordering; x starts from zero.
    SolverStatus solve_system(int n, int nnzb, const int32_t *rowptr, const int32_t *colidx, const double *values,
                              const double *b, Stats &stats) {
        if (!same_pattern(n, nnzb, rowptr, colidx)) {
            if (!ensure_handle(n, nnzb)) return SolverStatus::AnalysisFailed;
            if (bda_upload_bcsr(h_, rowptr, colidx, values) != 0 || bda_analyse(h_, 0) != 0)
                return fail(SolverStatus::AnalysisFailed);
            rowptr_.assign(rowptr, rowptr + n + 1);
            colidx_.assign(colidx, colidx + nnzb);
        } else if (bda_update_values(h_, values) != 0) {
            return fail(SolverStatus::AnalysisFailed);
        }
        if (bda_factor(h_) != 0) return fail(SolverStatus::CreatePreconditionerFailed);
        return run(b, stats);
    }

    // The caller's own colour-ordered system and factors: colour_ptr[ncolours+1],
    // A in block CSR, lu_values on A's pattern (L strictly lower, U strictly
    // upper), dinv the inverted diagonal blocks; b and x in the colour order.
    SolverStatus solve_prefactored(int n, int nnzb, int ncolours, const int32_t *colour_ptr, const int32_t *rowptr,
                                   const int32_t *colidx, const double *values, const double *lu_values,
                                   const double *dinv, const double *b, Stats &stats) {
        if (!ensure_handle(n, nnzb)) return SolverStatus::AnalysisFailed;
        if (bda_upload_factored(h_, ncolours, colour_ptr, rowptr, colidx, values, lu_values, dinv) != 0)
            return fail(SolverStatus::AnalysisFailed);
        rowptr_.clear();   // the next solve_system analyses its pattern again
        colidx_.clear();
        return run(b, stats);
    }

    void get_result(double *x) const { std::memcpy(x, x_.data(), x_.size() * sizeof(double)); }
    const std::string &last_error() const { return err_; }
    bda_handle handle() const { return h_; }

private:
    bool same_pattern(int n, int nnzb, const int32_t *rowptr, const int32_t *colidx) const {
        if (!h_ || n != n_ || (size_t)n + 1 != rowptr_.size() || (size_t)nnzb != colidx_.size()) return false;
        return std::memcmp(rowptr, rowptr_.data(), rowptr_.size() * sizeof(int32_t)) == 0 &&
               std::memcmp(colidx, colidx_.data(), colidx_.size() * sizeof(int32_t)) == 0;
    }

    // a handle for n block rows and at least nnzb blocks (reallocated when
    // the system grows or changes its row count)
    bool ensure_handle(int n, int nnzb) {
        if (h_ && n == n_ && nnzb <= nnzb_max_) return true;
        release();
        if (bda_create(bs_, n, nnzb, &h_) != 0) {
            err_ = "bda_create failed";
            h_ = nullptr;
            return false;
        }
        n_ = n;
        nnzb_max_ = nnzb;
        return true;
    }

    SolverStatus run(const double *b, Stats &stats) {
        x_.assign((size_t)n_ * bs_, 0.0);
        bda_result r{};
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = mixed_ ? bda_solve_mixed(h_, b, x_.data(), tol_, 20, 1e-4, maxit_, 1, &r)
                              : bda_solve(h_, b, x_.data(), tol_, maxit_, &r);
        stats.elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc != 0) return fail(SolverStatus::UnknownError);
        stats.raw = r;
        stats.half_iterations = mixed_ ? r.inner_half_iterations : r.half_iterations;
        stats.iterations = (int)(r.iterations + 0.5);
        stats.reduction = r.rel_residual;
        stats.converged = r.converged != 0;
        stats.conv_rate = stats.iterations > 0 ? std::pow(stats.reduction, 1.0 / stats.iterations) : 0.0;
        return SolverStatus::Success;
    }

    SolverStatus fail(SolverStatus s) {
        err_ = h_ ? bda_last_error(h_) : "no handle";
        return s;
    }

    void release() {
        if (h_) bda_destroy(h_);
        h_ = nullptr;
        n_ = 0;
        nnzb_max_ = 0;
        rowptr_.clear();
        colidx_.clear();
    }

    int bs_, maxit_;
    double tol_;
    bool mixed_;
    bda_handle h_ = nullptr;
    int n_ = 0, nnzb_max_ = 0;
    std::vector<int32_t> rowptr_, colidx_;
    std::vector<double> x_;
    std::string err_;
};

}  // namespace bda_backend

#endif  // BDA_BACKEND_HPP
