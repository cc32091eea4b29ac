"""ctypes binding of libwavebem_bem.so (the C-ABI in include/wavebem_bem.h).

There is no fallback: if the HIP library is missing, importing the device path raises.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WAVEBEM_LIB") or os.path.join(HERE, "libwavebem_bem.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "wavebem_bem.h")

WB_OK, WB_E_ARG, WB_E_OOM, WB_E_NOCONV, WB_E_HIP, WB_E_STATE, WB_E_UNSUPPORTED = range(7)
WB_PREC_FP64, WB_PREC_FP32_IR = 0, 1
WB_PRECOND_NONE, WB_PRECOND_BLOCK_JACOBI = 0, 1

STATUS_NAMES = {0: "WB_OK", 1: "WB_E_ARG", 2: "WB_E_OOM", 3: "WB_E_NOCONV", 4: "WB_E_HIP",
                5: "WB_E_STATE", 6: "WB_E_UNSUPPORTED"}


class WbConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int),
        ("quadrature_order", ctypes.c_int),
        ("singular_quadrature_order", ctypes.c_int),
        ("solver_tolerance", ctypes.c_double),
        ("solver_max_steps", ctypes.c_int),
        ("gmres_max_tmp_vectors", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("preconditioner", ctypes.c_int),
        ("precond_block", ctypes.c_int),
        ("ir_max_sweeps", ctypes.c_int),
        ("ir_inner_reduction", ctypes.c_double),
        ("assembly_tile", ctypes.c_int),
        ("ir_inner_format", ctypes.c_int),
        ("ir_inner16_reduction", ctypes.c_double),
        ("precond_coarse", ctypes.c_int),
        ("assembly_strips", ctypes.c_int),
        ("ir_inner_lr_tol", ctypes.c_double),
    ]


class WbSolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("restarts", ctypes.c_int),
        ("ir_sweeps", ctypes.c_int),
        ("residual", ctypes.c_double),
        ("true_rel_residual", ctypes.c_double),
        ("t_alpha_ms", ctypes.c_double),
        ("t_rhs_ms", ctypes.c_double),
        ("t_precond_ms", ctypes.c_double),
        ("t_gmres_ms", ctypes.c_double),
        ("sweep_iterations", ctypes.c_int * 8),
        ("sweep_residual", ctypes.c_double * 8),
    ]

    def as_dict(self):
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = list(v) if not isinstance(v, (int, float)) else v
        return d


WB_K_NAMES = ["assemble", "singular", "gemv64", "gemv32", "gemvhl", "gemv16", "gemvc",
              "inner_build"]


class WbKernelStats(ctypes.Structure):
    _fields_ = [
        ("launches", ctypes.c_int * len(WB_K_NAMES)),
        ("total_ms", ctypes.c_double * len(WB_K_NAMES)),
        ("rows", ctypes.c_longlong * len(WB_K_NAMES)),
        ("calls", ctypes.c_int * len(WB_K_NAMES)),
    ]

    def as_dict(self):
        return {name: dict(launches=self.launches[k], total_ms=self.total_ms[k],
                           rows=self.rows[k], calls=self.calls[k])
                for k, name in enumerate(WB_K_NAMES)}


_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i8p = ctypes.POINTER(ctypes.c_int8)
_vp = ctypes.c_void_p

_SIGS = {
    "wb_config_default": (None, [ctypes.POINTER(WbConfig)]),
    "wb_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(WbConfig)]),
    "wb_destroy": (None, [_vp]),
    "wb_last_error": (ctypes.c_char_p, [_vp]),
    "wb_version": (ctypes.c_char_p, []),
    "wb_set_mesh": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _i32p, _i8p, _i32p,
                                   _i32p, _i32p]),
    "wb_set_masks": (ctypes.c_int, [_vp, _dp, _dp]),
    "wb_set_support_points": (ctypes.c_int, [_vp, _dp]),
    "wb_assemble": (ctypes.c_int, [_vp]),
    "wb_release_matrices": (ctypes.c_int, [_vp]),
    "wb_get_matrices": (ctypes.c_int, [_vp, _dp, _dp]),
    "wb_get_matrix_rows": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _dp, _dp]),
    "wb_inner_vmult": (ctypes.c_int, [_vp, _dp, _dp]),
    "wb_constrained_vmult_noalloc": (ctypes.c_int, [_vp, _vp, _vp]),
    "wb_inner_operator_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "wb_host_cluster_plan": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _i32p, _i32p, _i32p,
                                            _i32p, _i32p, ctypes.POINTER(ctypes.c_uint8),
                                            _i32p]),
    "wb_host_hm_work_items": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_uint8),
                                             _i32p, _i32p, _i32p, _i32p]),
    "wb_get_coarse_matrix": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int), _dp]),
    "wb_solve_adjoint": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_void_p]),
    "wb_residual_vjp": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int, _dp]),
    "wb_residual_jvp": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int, _dp]),
    "wb_solve_rhs": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_void_p]),
    "wb_compute_alpha": (ctypes.c_int, [_vp, _dp]),
    "wb_vmult": (ctypes.c_int, [_vp, _dp, _dp]),
    "wb_compute_rhs": (ctypes.c_int, [_vp, _dp, _dp]),
    "wb_set_constraints": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, _dp, _dp]),
    "wb_compute_normals": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_int)]),
    "wb_compute_surface_gradients": (ctypes.c_int, [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]),
    "wb_compute_potential_gradients": (ctypes.c_int, [_vp, _dp, _dp, _dp,
                                                      ctypes.POINTER(ctypes.c_int)]),
    "wb_get_masks": (ctypes.c_int, [_vp, _dp, _dp]),
    "wb_prepare_bem_vectors": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, ctypes.c_int, _i32p,
                                              ctypes.c_int, _i32p, _dp, _dp,
                                              ctypes.POINTER(ctypes.c_uint8), _dp, _dp, _dp]),
    "wb_compute_constraints": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]),
    "wb_constrained_vmult": (ctypes.c_int, [_vp, _dp, _dp]),
    "wb_solve_system": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.POINTER(WbSolveStats)]),
    "wb_solve_system_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(WbSolveStats)]),
    "wb_solve": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.POINTER(WbSolveStats)]),
    "wb_residual": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "wb_solve_system_multi": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp,
                                             ctypes.POINTER(WbSolveStats)]),
    "wb_host_quadrature": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "wb_host_mesh_plan": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _i32p, _i32p, _i32p,
                                         _i32p, _i32p, ctypes.c_int, _i32p, _i32p, _i32p]),
    "wb_host_strip_plan": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _i32p, _i32p,
                                          ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p,
                                          _i32p]),
    "wb_host_compute_constraints": (ctypes.c_int, [ctypes.c_int, _i32p, _i32p, _dp, _dp, _dp,
                                                   _dp, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                                   _i32p, _dp, _dp]),
    "wb_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "wb_get_kernel_stats": (ctypes.c_int, [_vp, ctypes.POINTER(WbKernelStats)]),
    "wb_device_bytes": (ctypes.c_size_t, [_vp]),
    "wb_synchronize": (ctypes.c_int, [_vp]),
}


def header_symbols():
    """Function names declared in include/wavebem_bem.h."""
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wb_[a-z_0-9]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m wavebem_amd.build` "
                "(the HIP library is required; there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
