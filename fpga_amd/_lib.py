"""ctypes binding of ``libfpga_amd.so`` (the C ABI of include/bda_bicgstab.h).

The library is the only compute path: if it is missing, importing this
module raises, there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfpga_amd.so"

i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double
P_i32 = C.POINTER(C.c_int32)
P_f64 = C.POINTER(C.c_double)
vp = C.c_void_p


class BdaResult(C.Structure):
    _fields_ = [
        ("iterations", f64),
        ("half_iterations", i32),
        ("last_norm_idx", i32),
        ("norms", f64 * 4),
        ("final_norm", f64),
        ("rel_residual", f64),
        ("converged", i32),
        ("no_results", i32),
        ("max_iter_reached", i32),
        ("breakdown", i32),
        ("nan_detected", i32),
        ("aborted", i32),
        ("outer_iterations", i32),
        ("inner_half_iterations", i32),
        ("solve_ms", f64),
    ]


class BdaLimits(C.Structure):
    _fields_ = [
        ("max_block_size", i32),
        ("max_blocks_per_row", i32),
        ("max_iterations", i32),
        ("rows_per_tile", i32),
        ("device_memory_bytes", i64),
        ("compute_units", i32),
        ("wavefront_size", i32),
    ]


class BdaGridSpec(C.Structure):
    _fields_ = [
        ("nx", i32), ("ny", i32), ("nz", i32), ("bs", i32),
        ("seed", C.c_uint64),
        ("perm_sigma", f64), ("layer_corr", f64), ("inactive_frac", f64),
        ("nnc_frac", f64), ("nnc_dist", i32),
        ("pert_max", f64), ("upwind", f64), ("diag_noise", f64),
    ]


class BdaWireSizes(C.Structure):
    _fields_ = [
        ("vector_sizes", i32 * 18),
        ("nnz_val_sizes", i32 * 3),
        ("partitions", i32 * 3),
        ("device_fits", i32),
    ]


# name -> (restype, argtypes); every symbol include/bda_bicgstab.h declares
SIGNATURES = {
    "bda_grid_create": (i32, [C.POINTER(BdaGridSpec), C.POINTER(vp)]),
    "bda_grid_sizes": (i32, [vp, P_i32, P_i32, P_i32]),
    "bda_grid_copy": (i32, [vp, vp, vp, vp, vp]),
    "bda_grid_destroy": (i32, [vp]),
    "bda_colour_pattern": (i32, [i32, vp, vp, i32, vp, vp, vp, P_i32]),
    "bda_encode_status": (i32, [C.POINTER(BdaResult), vp]),
    "bda_encode_query": (i32, [C.POINTER(BdaLimits), i32, i64, i32, vp]),
    "bda_set_debug_level": (i32, [i32]),
    "bda_get_debug_level": (i32, []),
    "bda_create": (i32, [i32, i32, i32, C.POINTER(vp)]),
    "bda_destroy": (i32, [vp]),
    "bda_last_error": (C.c_char_p, [vp]),
    "bda_query": (i32, [vp, C.POINTER(BdaLimits)]),
    "bda_upload_bcsr": (i32, [vp, vp, vp, vp]),
    "bda_update_values": (i32, [vp, vp]),
    "bda_update_values_dev": (i32, [vp, vp]),
    "bda_values_generation": (i32, [vp, C.POINTER(C.c_uint64)]),
    "bda_upload_factored": (i32, [vp, i32, vp, vp, vp, vp, vp, vp]),
    "bda_analyse": (i32, [vp, i32]),
    "bda_num_colours": (i32, [vp, P_i32]),
    "bda_get_colouring": (i32, [vp, vp, vp, vp]),
    "bda_factor": (i32, [vp]),
    "bda_prepare_mixed": (i32, [vp]),
    "bda_set_options": (i32, [vp, f64, i32, i32]),
    "bda_solve": (i32, [vp, vp, vp, f64, i32, C.POINTER(BdaResult)]),
    "bda_solve_dev": (i32, [vp, vp, vp, f64, i32, C.POINTER(BdaResult)]),
    "bda_get_residual": (i32, [vp, vp]),
    "bda_solve_mixed": (i32, [vp, vp, vp, f64, i32, f64, i32, i32, C.POINTER(BdaResult)]),
    "bda_solve_mixed_dev": (i32, [vp, vp, vp, f64, i32, f64, i32, i32, C.POINTER(BdaResult)]),
    "bda_set_refinement": (i32, [vp, i32, f64]),
    "bda_adjoint_solve": (i32, [vp, vp, vp, f64, i32, C.POINTER(BdaResult)]),
    "bda_adjoint_solve_dev": (i32, [vp, vp, vp, f64, i32, C.POINTER(BdaResult)]),
    "bda_adjoint_solve_mixed": (i32, [vp, vp, vp, f64, i32, f64, i32, i32, C.POINTER(BdaResult)]),
    "bda_adjoint_solve_mixed_dev": (i32, [vp, vp, vp, f64, i32, f64, i32, i32, C.POINTER(BdaResult)]),
    "bda_grad_values": (i32, [vp, vp, vp, vp]),
    "bda_grad_values_dev": (i32, [vp, vp, vp, vp]),
    "bda_spmv_perm": (i32, [vp, vp, vp, i32]),
    "bda_precond_apply_perm": (i32, [vp, vp, vp, i32]),
    "bda_spmv_perm_ex": (i32, [vp, i32, vp, vp, i32, vp, i32, P_f64]),
    "bda_precond_apply_perm_ex": (i32, [vp, i32, vp, vp, i32]),
    "bda_precond_launches": (i32, [vp, i32, P_i32]),
    "bda_spmv_dev": (i32, [vp, i32, vp, vp, i32, vp, i32, P_f64]),
    "bda_spmv_dev_noalloc": (i32, [vp, i32, vp, vp, i32, vp, i32, P_f64]),
    "bda_factor_sizes": (i32, [vp, P_i32, P_i32]),
    "bda_get_factors": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "bda_set_profiling": (i32, [vp, i32]),
    "bda_get_kernel_stats": (i32, [vp, i32, C.POINTER(i64), P_f64]),
    "bda_get_kernel_times": (i32, [vp, i32, vp, i32, P_i32]),
    "bda_time_spmv": (i32, [vp, i32, i32, P_f64]),
    "bda_time_precond": (i32, [vp, i32, i32, P_f64]),
    "bda_measure_bandwidth": (i32, [vp, i64, i32, P_f64, P_f64]),
    "bda_algorithmic_bytes": (i32, [vp, i32, i32, P_f64]),
    "bda_set_history": (i32, [vp, i32]),
    "bda_get_history": (i32, [vp, vp, i32, P_i32]),
    "bda_wire_export_sizes": (i32, [vp, C.POINTER(BdaWireSizes)]),
    "bda_wire_export": (i32, [vp, vp, vp, vp]),
    "bda_wire_build_sizes": (i32, [i32, i32, vp, i32, vp, vp, vp, vp, vp, vp, C.POINTER(BdaWireSizes)]),
    "bda_wire_build": (i32, [i32, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bda_wire_layout": (i32, [vp, vp, C.c_uint32, vp, vp]),
    "bda_wire_pack": (i32, [vp, vp, C.c_uint32, vp, vp]),
    "bda_wire_unpack": (i32, [vp, vp, vp, vp, vp]),
    "bda_wire_import_sizes": (i32, [i32, vp, vp, P_i32, C.POINTER(i64), P_i32]),
    "bda_wire_import": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
}


def _bind_torch_runtime():
    """torch and this library both run on the HIP runtime (soname
    libamdhip64.so.7).  torch loads its own bundled copy under another name,
    so if the library were loaded first the process would hold two runtimes
    and torch would see no GPU.  When torch is installed it is imported
    first; the library's dependency then resolves to torch's copy."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path: Path | str | None = None) -> C.CDLL:
    _bind_torch_runtime()
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is not built: the HIP solver library is required "
            "(build it with `python -m fpga_amd.build`)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()
