// LibTorch autograd Function around the ILU0-BiCGSTAB handle.
//
//   x = solve(values, b)        forward: new values on the handle's pattern,
//                               block ILU0, FP64 ILU0-BiCGSTAB
//   dL/db      = lambda         backward: A^T lambda = dL/dx solved with the
//   dL/dvalues = -lambda x^T    transposed colour-ordered ILU0 (U^T sweep over
//                               colours ascending, then L^T descending),
//                               gradient restricted to the block pattern
//
// Tensors are float64 on the handle's GPU, contiguous: values (nnzb, bs, bs)
// in the layout the handle was created with, b (n*bs).  The handle is passed
// as an integer (bda_handle of include/bda_bicgstab.h); every library call is
// synchronous, and torch's device work is synchronised before each call.
#include <torch/cuda.h>
#include <torch/extension.h>

#include <sstream>
#include <stdexcept>

#include "../../include/bda_bicgstab.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

bda_handle as_handle(int64_t h) { return reinterpret_cast<bda_handle>(static_cast<intptr_t>(h)); }

void check(bda_handle h, int rc, const char *what) {
    if (rc != 0) {
        std::ostringstream os;
        os << what << " failed: " << bda_last_error(h);
        throw std::runtime_error(os.str());
    }
}

void check_tensor(const Tensor &t, const char *name) {
    TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor");
    TORCH_CHECK(t.scalar_type() == torch::kFloat64, name, " must be float64");
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

constexpr int32_t kMaxOuter = 20;     // refinement steps of the precision form
constexpr double kInnerTol = 1e-4;    // its inner tolerance

struct BdaSolve : public torch::autograd::Function<BdaSolve> {
    // precision_form != 0: the forward solve and the adjoint run in the
    // precision form (FP64 refinement around FP32-value inner solves, inner
    // tolerance 1e-4, maxit and adj_maxit bound each inner run).
    static Tensor forward(AutogradContext *ctx, Tensor values, Tensor b, int64_t handle, double tol, int64_t maxit,
                          double adj_tol, int64_t adj_maxit, int64_t precision_form) {
        bda_handle h = as_handle(handle);
        values = values.contiguous();
        b = b.contiguous();
        check_tensor(values, "values");
        check_tensor(b, "b");
        torch::cuda::synchronize();
        check(h, bda_update_values_dev(h, values.data_ptr<double>()), "bda_update_values_dev");
        check(h, bda_factor(h), "bda_factor");
        Tensor x = torch::zeros_like(b);
        bda_result res;
        if (precision_form)
            check(h,
                  bda_solve_mixed_dev(h, b.data_ptr<double>(), x.data_ptr<double>(), tol, kMaxOuter, kInnerTol,
                                      (int32_t)maxit, 1, &res),
                  "bda_solve_mixed_dev");
        else
            check(h, bda_solve_dev(h, b.data_ptr<double>(), x.data_ptr<double>(), tol, (int32_t)maxit, &res),
                  "bda_solve_dev");
        TORCH_CHECK(res.converged || res.max_iter_reached, "forward solve failed (breakdown or NaN)");
        uint64_t gen = 0;
        check(h, bda_values_generation(h, &gen), "bda_values_generation");
        ctx->save_for_backward({x, values});
        ctx->saved_data["handle"] = handle;
        ctx->saved_data["generation"] = (int64_t)gen;
        ctx->saved_data["adj_tol"] = adj_tol;
        ctx->saved_data["adj_maxit"] = adj_maxit;
        ctx->saved_data["precision_form"] = precision_form;
        ctx->saved_data["nvalues"] = values.numel();
        ctx->saved_data["vshape"] = values.sizes().vec();
        return x;
    }

    static tensor_list backward(AutogradContext *ctx, tensor_list grad_out) {
        bda_handle h = as_handle(ctx->saved_data["handle"].toInt());
        const auto saved = ctx->get_saved_variables();
        const Tensor x = saved[0], values = saved[1];
        Tensor w = grad_out[0].contiguous().to(torch::kFloat64);
        check_tensor(w, "grad_output");
        Tensor lam = torch::zeros_like(w);
        torch::cuda::synchronize();
        // The handle may hold another system by now (a second forward call,
        // a value update or a factorisation since this forward): the adjoint
        // must use the values this forward solved with, so they are put back.
        uint64_t gen = 0;
        check(h, bda_values_generation(h, &gen), "bda_values_generation");
        if ((int64_t)gen != ctx->saved_data["generation"].toInt()) {
            check(h, bda_update_values_dev(h, values.data_ptr<double>()), "bda_update_values_dev");
            check(h, bda_factor(h), "bda_factor");
        }
        bda_result res;
        const double adj_tol = ctx->saved_data["adj_tol"].toDouble();
        const int32_t adj_maxit = (int32_t)ctx->saved_data["adj_maxit"].toInt();
        if (ctx->saved_data["precision_form"].toInt())
            check(h,
                  bda_adjoint_solve_mixed_dev(h, w.data_ptr<double>(), lam.data_ptr<double>(), adj_tol, kMaxOuter,
                                              kInnerTol, adj_maxit, 1, &res),
                  "bda_adjoint_solve_mixed_dev");
        else
            check(h, bda_adjoint_solve_dev(h, w.data_ptr<double>(), lam.data_ptr<double>(), adj_tol, adj_maxit, &res),
                  "bda_adjoint_solve_dev");
        TORCH_CHECK(res.converged || res.no_results || res.max_iter_reached, "adjoint solve failed");
        Tensor gvals = torch::empty(ctx->saved_data["vshape"].toIntVector(), w.options());
        check(h, bda_grad_values_dev(h, lam.data_ptr<double>(), x.data_ptr<double>(), gvals.data_ptr<double>()),
              "bda_grad_values_dev");
        return {gvals, lam, Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

Tensor solve(Tensor values, Tensor b, int64_t handle, double tol, int64_t maxit, double adj_tol, int64_t adj_maxit,
             int64_t precision_form) {
    return BdaSolve::apply(values, b, handle, tol, maxit, adj_tol, adj_maxit, precision_form);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "ILU0-BiCGSTAB solve with a hand-written adjoint (torch::autograd::Function)";
    m.def("solve", &solve, "x = A(values)^-1 b, differentiable in values and b", py::arg("values"), py::arg("b"),
          py::arg("handle"), py::arg("tol"), py::arg("maxit"), py::arg("adj_tol"), py::arg("adj_maxit"),
          py::arg("precision_form") = 0);
}
