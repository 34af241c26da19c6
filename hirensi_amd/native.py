"""ctypes binding of ``libhirensi.so`` (the C ABI of ``include/hirensi.h``).

The library is loaded from the package directory; a missing library is an
error, there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhirensi.so")

HR_DEVICE_PTRS = 1
HR_K, HR_C, HR_G = 0, 1, 2
HR_HISTORY_NONE, HR_HISTORY_STORE, HR_HISTORY_CHECKPOINT, HR_HISTORY_REVOLVE = 0, 1, 2, 3
HR_REG_TV, HR_REG_TN = 0, 1

EXPORTS = (
    "hr_ctx_create", "hr_ctx_destroy", "hr_last_error", "hr_get_sizes", "hr_get_store_eq",
    "hr_set_pml", "hr_set_material", "hr_get_diag_m", "hr_scale", "hr_diagonal_scale",
    "hr_reset_scaling", "hr_apply", "hr_time_function", "hr_forward", "hr_adjoint_gradient",
    "hr_misfit", "hr_recompute_ratio", "hr_set_tiling", "hr_get_tiling", "hr_set_profiling", "hr_get_kernel_stats",
    "hr_forward_shots", "hr_gradient_mass", "hr_regularization", "hr_regularize_gradient",
    "hr_forward_cd", "hr_set_history_budget", "hr_get_history_plan", "hr_revolve_schedule",
    "hr_gradient_accumulate", "hr_set_numbering", "hr_surface_force", "hr_set_energy_interval", "hr_get_energy",
    "hr_autotune", "hr_misfit_noalloc", "hr_set_graph", "hr_get_graph_launches", "hr_synchronize",
)

# include/hirensi_nn.h (signatures set by hirensi_amd.precond.lib)
NN_EXPORTS = (
    "hr_nn_conv3d", "hr_nn_conv3d_wgrad", "hr_nn_transpose_weights", "hr_nn_leaky_relu", "hr_nn_leaky_relu_grad",
    "hr_nn_avgpool2", "hr_nn_avgpool2_grad", "hr_nn_upsample2", "hr_nn_upsample2_grad", "hr_nn_mse", "hr_nn_adam",
    "hr_nn_to_bf16",
)


class MeshDesc(C.Structure):
    _fields_ = [
        ("nex", C.c_int), ("ney", C.c_int), ("nez", C.c_int),
        ("npml", C.c_int * 5),
        ("h", C.c_double * 3),
        ("top_stress_bc", C.c_int),
        ("flags", C.c_int),
    ]


class HrError(RuntimeError):
    pass


_lib = None

_vp = C.c_void_p
_dp = C.c_void_p      # double* passed as raw address (host numpy or device pointer)


def _init_torch_runtime():
    """torch ships its own HIP runtime.  Device-pointer contexts share HBM
    with torch tensors, so the library must bind to the runtime torch uses:
    that runtime is loaded and initialised before the library when a GPU is
    present (loading the library first leaves torch without a device)."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HrError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    _init_torch_runtime()
    L = C.CDLL(LIB_PATH)
    L.hr_ctx_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(_vp)]
    L.hr_ctx_destroy.argtypes = [_vp]
    L.hr_last_error.argtypes = [_vp]
    L.hr_last_error.restype = C.c_char_p
    L.hr_get_sizes.argtypes = [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.hr_get_store_eq.argtypes = [_vp, _dp]
    L.hr_set_pml.argtypes = [_vp, _dp]
    L.hr_set_material.argtypes = [_vp, _dp, _dp, _dp]
    L.hr_get_diag_m.argtypes = [_vp, _dp]
    L.hr_scale.argtypes = [_vp, C.c_int, C.c_double]
    L.hr_diagonal_scale.argtypes = [_vp, C.c_int, _dp, _dp]
    L.hr_reset_scaling.argtypes = [_vp, C.c_int]
    L.hr_apply.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp, _dp]
    L.hr_time_function.argtypes = [C.c_double, C.c_int, C.c_double]
    L.hr_time_function.restype = C.c_double
    L.hr_forward.argtypes = [_vp, C.c_int, C.c_double, _dp, C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int]
    L.hr_forward_shots.argtypes = [_vp, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp,
                                   C.c_int, C.c_int]
    L.hr_adjoint_gradient.argtypes = [_vp, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.hr_misfit.argtypes = [_vp, C.c_int, C.c_double, C.c_int, _dp, _dp]
    L.hr_recompute_ratio.argtypes = [_vp]
    L.hr_recompute_ratio.restype = C.c_double
    L.hr_set_tiling.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.hr_get_tiling.argtypes = [_vp, _vp, _vp, _vp]
    L.hr_set_profiling.argtypes = [_vp, C.c_int]
    L.hr_get_kernel_stats.argtypes = [_vp, _dp, _dp]
    L.hr_forward_cd.argtypes = [_vp, C.c_int, C.c_double, _dp, C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.hr_gradient_mass.argtypes = [_vp, _dp, _dp]
    L.hr_regularization.argtypes = [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.hr_regularize_gradient.argtypes = [_vp, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.hr_set_history_budget.argtypes = [_vp, C.c_double]
    L.hr_get_history_plan.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.hr_revolve_schedule.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.hr_gradient_accumulate.argtypes = [_vp, _dp, _dp, C.c_double, _dp, _dp]
    L.hr_set_numbering.argtypes = [_vp, _dp]
    L.hr_surface_force.argtypes = [_vp, _dp, _dp, _dp, C.POINTER(C.c_longlong)]
    L.hr_set_energy_interval.argtypes = [_vp, C.c_int]
    L.hr_autotune.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.hr_get_energy.argtypes = [_vp, _dp, C.POINTER(C.c_int)]
    L.hr_misfit_noalloc.argtypes = [_vp, C.c_int, C.c_double, C.c_int, _vp, _vp]
    L.hr_set_graph.argtypes = [_vp, C.c_int]
    L.hr_synchronize.argtypes = [_vp]
    L.hr_get_graph_launches.argtypes = [_vp, C.POINTER(C.c_longlong)]
    for name in EXPORTS:
        getattr(L, name).restype = getattr(L, name).restype or C.c_int
    _lib = L
    return L
