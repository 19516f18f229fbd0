"""ctypes binding of the C ABI declared in ``include/gigalens_amd.h``.

The HIP library is the only compute path: importing this module does not
load it, but every call that needs it raises ``NativeLibraryMissing`` when
``gigalens_amd/_lib/libgigalens_amd.so`` has not been built.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

LIB_PATH = os.environ.get("GIGALENS_AMD_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib",
                                                                 "libgigalens_amd.so")

GL_MAX_COMPONENTS = 8
GL_MAX_COMPONENT_PARAMS = 10
GL_OK, GL_ERR_ARG, GL_ERR_HIP, GL_ERR_UNSUPPORTED = 0, 1, 2, 3
GL_MAX_PARAMS = 128
GL_MAX_EPL_ITER = 100

GL_EPL, GL_SHEAR, GL_SIS, GL_SIE, GL_TNFW = 1, 2, 3, 4, 5
GL_SERSIC, GL_SERSIC_ELLIPSE, GL_SHAPELETS, GL_CORE_SERSIC = 10, 11, 12, 13
GL_ROLE_LENS, GL_ROLE_LENS_LIGHT, GL_ROLE_SOURCE = 0, 1, 2
GL_PRIOR_NORMAL, GL_PRIOR_LOGNORMAL, GL_PRIOR_UNIFORM, GL_PRIOR_TRUNCNORMAL = 1, 2, 3, 4
GL_LIK_FORWARD, GL_LIK_BACKWARD = 0, 1

EXPORTED_SYMBOLS = (
    "gl_last_error",
    "gl_plan_create",
    "gl_plan_destroy",
    "gl_plan_query",
    "gl_simulate",
    "gl_simulate_vjp",
    "gl_log_prob",
    "gl_map_adam_step",
    "gl_lstsq_simulate",
    "gl_log_prob_lenses",
    "gl_map_adam_step_lenses",
    "gl_map_adam_step_noalloc",
    "gl_batch_create",
    "gl_batch_create_ws",
    "gl_batch_create_ex",
    "gl_batch_query",
    "gl_batch_workspace_bytes",
    "gl_batch_num_chunks",
    "gl_batch_destroy",
    "gl_batch_size",
    "gl_batch_log_prob",
    "gl_batch_map_adam_step",
    "gl_svi_diag_sample",
    "gl_svi_diag_update",
    "gl_log_prob_masked",
    "gl_hmc_begin",
    "gl_hmc_position",
    "gl_hmc_momentum",
    "gl_hmc_end",
    "gl_hmc_adapt",
    "gl_noise_poisson_gauss",
    "gl_leapfrog_step",
    "gl_cnn_create",
    "gl_cnn_destroy",
    "gl_cnn_num_params",
    "gl_cnn_set_params",
    "gl_cnn_get_params",
    "gl_cnn_forward",
    "gl_cnn_loss_grad",
    "gl_cnn_adamw_step",
)


class NativeLibraryMissing(RuntimeError):
    pass


class GlError(RuntimeError):
    pass


class ComponentDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("role", ctypes.c_int),
        ("param_col", ctypes.c_int * GL_MAX_COMPONENT_PARAMS),
        ("param_const", ctypes.c_float * GL_MAX_COMPONENT_PARAMS),
        ("niter", ctypes.c_int),
        ("n_max", ctypes.c_int),
        ("interpolate", ctypes.c_int),
        ("use_lstsq", ctypes.c_int),
        ("amp_col0", ctypes.c_int),
    ]


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("num_pix", ctypes.c_int),
        ("supersample", ctypes.c_int),
        ("transform_sub", ctypes.c_float * 4),
        ("origin", ctypes.c_float * 2),
        ("conversion_factor", ctypes.c_float),
        ("kernel_size", ctypes.c_int),
        ("kernel", ctypes.POINTER(ctypes.c_float)),
        ("n_components", ctypes.c_int),
        ("components", ctypes.POINTER(ComponentDesc)),
        ("n_params", ctypes.c_int),
        ("prior_kind", ctypes.POINTER(ctypes.c_int)),
        ("prior_args", ctypes.POINTER(ctypes.c_float)),
        ("observed", ctypes.POINTER(ctypes.c_float)),
        ("background_rms", ctypes.c_float),
        ("exp_time", ctypes.c_float),
        ("likelihood", ctypes.c_int),
        ("fldj_at_x", ctypes.c_int),
        ("epl_batch_series", ctypes.c_int),
        ("psf_rank_tol", ctypes.c_float),
        ("max_batch", ctypes.c_int),
        ("noise_map", ctypes.POINTER(ctypes.c_float)),
        ("n_observed", ctypes.c_int),
        ("staged", ctypes.c_int),
        ("conv_rows", ctypes.c_int),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("num_pix", "supersample", "n_params", "kernel_size", "folded_size",
                                            "sep_rank", "lds_bytes", "block", "adj_rank")]


class BatchOpts(ctypes.Structure):
    _fields_ = [
        ("workspace_bytes", ctypes.c_size_t),
        ("band", ctypes.c_int),
        ("band_rows", ctypes.c_int),
        ("band_lds_bytes", ctypes.c_int),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("total", "lenses_fused", "lenses_staged", "lenses_band", "chunks",
                                            "band_grid", "band_lds_bytes", "band_rows_min", "band_rows_max")] + [
        ("band_scratch_bytes", ctypes.c_size_t), ("workspace_bytes", ctypes.c_size_t)]


class CnnDesc(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int),
        ("width", ctypes.c_int),
        ("channels", ctypes.c_int * 4),
        ("hidden", ctypes.c_int),
        ("n_out", ctypes.c_int),
        ("max_batch", ctypes.c_int),
        ("in_scale", ctypes.c_float),
    ]


_lib: Optional[ctypes.CDLL] = None

_vp = ctypes.c_void_p
_SIGS = {
    "gl_last_error": ([], ctypes.c_char_p),
    "gl_plan_create": ([ctypes.POINTER(PlanDesc), ctypes.POINTER(_vp)], ctypes.c_int),
    "gl_plan_destroy": ([_vp], ctypes.c_int),
    "gl_plan_query": ([_vp, ctypes.POINTER(PlanInfo)], ctypes.c_int),
    "gl_simulate": ([_vp, _vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "gl_simulate_vjp": ([_vp, _vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
    "gl_log_prob": ([_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "gl_lstsq_simulate": ([_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp], ctypes.c_int),
    "gl_map_adam_step": ([_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, ctypes.c_float, _vp, _vp, _vp], ctypes.c_int),
    "gl_log_prob_lenses": ([_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "gl_map_adam_step_lenses": ([_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp], ctypes.c_int),
    "gl_map_adam_step_noalloc": ([_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, _vp, _vp, _vp], ctypes.c_int),
    "gl_batch_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_vp)],
                        ctypes.c_int),
    "gl_batch_create_ws": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_size_t,
                            ctypes.POINTER(_vp)], ctypes.c_int),
    "gl_batch_create_ex": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(BatchOpts), ctypes.POINTER(_vp)], ctypes.c_int),
    "gl_batch_query": ([_vp, ctypes.POINTER(BatchInfo)], ctypes.c_int),
    "gl_batch_workspace_bytes": ([_vp], ctypes.c_size_t),
    "gl_batch_num_chunks": ([_vp], ctypes.c_int),
    "gl_batch_destroy": ([_vp], ctypes.c_int),
    "gl_batch_size": ([_vp], ctypes.c_int),
    "gl_batch_log_prob": ([_vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "gl_batch_map_adam_step": ([_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, _vp, _vp, _vp], ctypes.c_int),
    "gl_svi_diag_sample": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _vp, _vp, _vp,
                            _vp], ctypes.c_int),
    "gl_svi_diag_update": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                            ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, ctypes.c_int, _vp],
                           ctypes.c_int),
    "gl_log_prob_masked": ([_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "gl_hmc_begin": ([ctypes.c_int, ctypes.c_int, _vp, ctypes.c_ulonglong, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                      _vp, _vp, _vp], ctypes.c_int),
    "gl_hmc_position": ([ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
    "gl_hmc_momentum": ([ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
    "gl_hmc_end": ([ctypes.c_int, ctypes.c_int, _vp, ctypes.c_ulonglong, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                    _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "gl_hmc_adapt": ([ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                      ctypes.c_float, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "gl_leapfrog_step": ([_vp, ctypes.c_int, _vp, ctypes.c_float, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "gl_noise_poisson_gauss": ([_vp, ctypes.c_longlong, ctypes.c_float, ctypes.c_float, ctypes.c_ulonglong, _vp, _vp],
                               ctypes.c_int),
    "gl_cnn_create": ([ctypes.POINTER(CnnDesc), ctypes.POINTER(_vp)], ctypes.c_int),
    "gl_cnn_destroy": ([_vp], ctypes.c_int),
    "gl_cnn_num_params": ([_vp], ctypes.c_int),
    "gl_cnn_set_params": ([_vp, _vp, _vp], ctypes.c_int),
    "gl_cnn_get_params": ([_vp, _vp, _vp], ctypes.c_int),
    "gl_cnn_forward": ([_vp, _vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "gl_cnn_loss_grad": ([_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
    "gl_cnn_adamw_step": ([_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                           ctypes.c_float, _vp], ctypes.c_int),
}


def lib() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python -m gigalens_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().gl_last_error().decode(errors="replace")
        raise GlError(f"{what} failed (code {rc}): {msg}")
