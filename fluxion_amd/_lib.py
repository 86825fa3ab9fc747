"""ctypes binding of libfluxion_hip.so (declarations in include/fluxion_hip.h).

Loading fails loudly: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import atexit
import ctypes as C
from pathlib import Path

from . import core

LIB_PATH = Path(__file__).resolve().parent / "libfluxion_hip.so"

c_buf_p = C.c_void_p
c_be_p = C.c_void_p


class FxGrid(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("cols", C.c_uint64), ("dx", C.c_double), ("dy", C.c_double)]


class FxStencilEntry(C.Structure):
    _fields_ = [("dr", C.c_int32), ("dc", C.c_int32), ("weight", C.c_double)]


class FxStencil(C.Structure):
    _fields_ = [("n", C.c_uint32), ("entries", C.POINTER(FxStencilEntry))]


class FxBc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_double), ("b", C.c_double), ("g", C.c_double)]


class FxBoundaries(C.Structure):
    _fields_ = [("top", FxBc), ("bottom", FxBc), ("left", FxBc), ("right", FxBc)]


class FxSolveStats(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("residual", C.c_double), ("reason", C.c_int32),
                ("pad", C.c_int32), ("true_residual", C.c_double)]


class FxOperator(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("coeff", C.c_double), ("sigma", C.c_double),
                ("stencil", C.POINTER(FxStencil)), ("bc", C.POINTER(FxBoundaries)),
                ("kappa", c_buf_p), ("kappa_scalar", C.c_double), ("vx", c_buf_p), ("vy", c_buf_p),
                ("k", c_buf_p)]


class FxMgOptions(C.Structure):
    _fields_ = [("n_pre", C.c_int32), ("n_post", C.c_int32), ("omega", C.c_double),
                ("smoother", C.c_int32), ("transfer", C.c_int32), ("coarsest", C.c_int32),
                ("extra_coarse", C.c_int32), ("precision", C.c_int32), ("fused", C.c_int32),
                ("sigma", C.c_double), ("scale", C.c_double), ("k", c_buf_p),
                ("vx", c_buf_p), ("vy", c_buf_p), ("kappa", C.c_double)]


class FxPrecond(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("handle", C.c_void_p)]


class FxCnOptions(C.Structure):
    _fields_ = [("op", C.c_int32), ("solver", C.c_int32), ("k", c_buf_p), ("mg", C.POINTER(FxMgOptions))]


def grid_c(g: core.Grid) -> FxGrid:
    return FxGrid(g.rows, g.cols, g.dx, g.dy)


class StencilC:
    """Keeps the entry array alive while the FxStencil points at it."""

    def __init__(self, st: core.Stencil):
        n = len(st.entries)
        self.arr = (FxStencilEntry * n)(*[FxStencilEntry(e.dr, e.dc, e.weight) for e in st.entries])
        self.c = FxStencil(n, C.cast(self.arr, C.POINTER(FxStencilEntry)))


def bcs_c(b: core.Boundaries) -> FxBoundaries:
    return FxBoundaries(*[FxBc(int(s.kind), s.a, s.b, s.g) for s in b.sides()])


_P = C.POINTER
_SIGS = {
    "fx_backend_create": (C.c_int, [C.c_int, _P(c_be_p)]),
    "fx_backend_destroy": (None, [c_be_p]),
    "fx_backend_name": (C.c_char_p, [c_be_p]),
    "fx_synchronize": (C.c_int, [c_be_p]),
    "fx_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "fx_backend_stream": (C.c_void_p, [c_be_p]),
    "fx_probe_set": (C.c_int, [c_be_p, C.c_int32]),
    "fx_batch_create": (C.c_int, [c_be_p, C.c_uint32, C.c_void_p, c_buf_p, C.c_void_p, C.c_void_p]),
    "fx_batch_solve": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "fx_batch_destroy": (None, [C.c_void_p]),
    "fx_batch_info": (C.c_int, [C.c_void_p, _P(C.c_int32), _P(C.c_int32), _P(C.c_uint64), _P(C.c_int32)]),
    "fx_batch_level_kernels": (C.c_int, [C.c_void_p, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32)]),
    "fx_probe_read": (C.c_int, [c_be_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "fx_hbm_ceilings": (C.c_int, [c_be_p, C.c_size_t, C.c_int32, C.POINTER(C.c_double)]),
    "fx_allocate": (C.c_int, [c_be_p, C.c_size_t, C.c_int, _P(c_buf_p)]),
    "fx_upload_f64": (C.c_int, [c_be_p, C.c_void_p, C.c_size_t, _P(c_buf_p)]),
    "fx_upload_f32": (C.c_int, [c_be_p, C.c_void_p, C.c_size_t, _P(c_buf_p)]),
    "fx_buffer_free": (None, [c_buf_p]),
    "fx_buffer_len": (C.c_size_t, [c_buf_p]),
    "fx_buffer_dtype": (C.c_int, [c_buf_p]),
    "fx_buffer_device_ptr": (C.c_void_p, [c_buf_p]),
    "fx_copy_to_host_f64": (C.c_int, [c_be_p, c_buf_p, C.c_void_p, C.c_size_t]),
    "fx_copy_to_host_f32": (C.c_int, [c_be_p, c_buf_p, C.c_void_p, C.c_size_t]),
    "fx_write_f64": (C.c_int, [c_be_p, c_buf_p, C.c_void_p, C.c_size_t]),
    "fx_apply_stencil": (C.c_int, [c_be_p, c_buf_p, c_buf_p, _P(FxGrid), _P(FxStencil), _P(FxBoundaries)]),
    "fx_apply_stencil_var": (C.c_int, [c_be_p, c_buf_p, c_buf_p, c_buf_p, _P(FxGrid), _P(FxStencil),
                                       _P(FxBoundaries)]),
    "fx_stencil_axpy": (C.c_int, [c_be_p, C.c_double, c_buf_p, c_buf_p, _P(FxGrid), _P(FxStencil),
                                  _P(FxBoundaries)]),
    "fx_copy": (C.c_int, [c_be_p, c_buf_p, c_buf_p]),
    "fx_fill": (C.c_int, [c_be_p, c_buf_p, C.c_double]),
    "fx_axpy": (C.c_int, [c_be_p, C.c_double, c_buf_p, c_buf_p]),
    "fx_scale": (C.c_int, [c_be_p, c_buf_p, C.c_double]),
    "fx_waxpy": (C.c_int, [c_be_p, C.c_double, c_buf_p, C.c_double, c_buf_p, c_buf_p]),
    "fx_aypx": (C.c_int, [c_be_p, C.c_double, c_buf_p, c_buf_p]),
    "fx_pointwise_mult": (C.c_int, [c_be_p, c_buf_p, c_buf_p, c_buf_p]),
    "fx_pointwise_div": (C.c_int, [c_be_p, c_buf_p, c_buf_p, c_buf_p]),
    "fx_pointwise_max": (C.c_int, [c_be_p, c_buf_p, c_buf_p, c_buf_p]),
    "fx_pointwise_min": (C.c_int, [c_be_p, c_buf_p, c_buf_p, c_buf_p]),
    "fx_reciprocal": (C.c_int, [c_be_p, c_buf_p]),
    "fx_abs_val": (C.c_int, [c_be_p, c_buf_p]),
    "fx_norm_l2": (C.c_int, [c_be_p, c_buf_p, _P(C.c_double)]),
    "fx_reduce_sum": (C.c_int, [c_be_p, c_buf_p, _P(C.c_double)]),
    "fx_reduce_max": (C.c_int, [c_be_p, c_buf_p, _P(C.c_double)]),
    "fx_reduce_min": (C.c_int, [c_be_p, c_buf_p, _P(C.c_double)]),
    "fx_dot": (C.c_int, [c_be_p, c_buf_p, c_buf_p, _P(C.c_double)]),
    "fx_dot_noalloc": (C.c_int, [c_be_p, c_buf_p, c_buf_p, c_buf_p]),
    "fx_dot2": (C.c_int, [c_be_p, c_buf_p, c_buf_p, c_buf_p, c_buf_p, _P(C.c_double)]),
    "fx_restrict": (C.c_int, [c_be_p, c_buf_p, c_buf_p, _P(FxGrid), _P(FxGrid)]),
    "fx_prolong": (C.c_int, [c_be_p, c_buf_p, c_buf_p, _P(FxGrid), _P(FxGrid)]),
    "fx_weighted_jacobi": (C.c_int, [c_be_p, c_buf_p, c_buf_p, C.c_double, _P(FxGrid), _P(FxStencil),
                                     _P(FxBoundaries)]),
    "fx_apply_conv_diff": (C.c_int, [c_be_p, c_buf_p, c_buf_p, c_buf_p, c_buf_p, c_buf_p, _P(FxGrid),
                                     _P(FxBoundaries)]),
    "fx_conv_diff_axpy": (C.c_int, [c_be_p, C.c_double, c_buf_p, c_buf_p, c_buf_p, c_buf_p, c_buf_p,
                                    _P(FxGrid), _P(FxBoundaries)]),
    "fx_zero_boundaries": (C.c_int, [c_be_p, c_buf_p, C.c_uint64, C.c_uint64]),
    "fx_pipelined_cg_fused": (C.c_int, [c_be_p, C.c_double, C.c_double] + [c_buf_p] * 7),
    "fx_multigrid_create": (C.c_int, [c_be_p, _P(FxGrid), _P(FxBoundaries), C.c_uint32, C.c_uint32,
                                      C.c_double, C.c_int32, _P(C.c_void_p)]),
    "fx_multigrid_create_ex": (C.c_int, [c_be_p, _P(FxGrid), _P(FxBoundaries), _P(FxMgOptions),
                                         _P(C.c_void_p)]),
    "fx_multigrid_destroy": (None, [C.c_void_p]),
    "fx_multigrid_depth": (C.c_uint32, [C.c_void_p]),
    "fx_multigrid_vcycle": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p]),
    "fx_multigrid_set_kernels": (C.c_int, [C.c_void_p, C.c_int32]),
    "fx_multigrid_set_wave_rows": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.c_int32]),
    "fx_multigrid_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "fx_solver_create": (C.c_int, [c_be_p, C.c_int32, _P(FxGrid), C.c_uint32, C.c_double, C.c_uint64,
                                   _P(C.c_void_p)]),
    "fx_solver_destroy": (None, [C.c_void_p]),
    "fx_solver_name": (C.c_char_p, [C.c_void_p]),
    "fx_solver_solve": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p, C.c_double, _P(FxStencil), _P(FxBoundaries),
                                  _P(FxSolveStats)]),
    "fx_solver_solve_preconditioned": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p, C.c_double, _P(FxStencil),
                                                 _P(FxBoundaries), _P(FxPrecond), _P(FxSolveStats)]),
    "fx_solver_solve_op": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p, _P(FxOperator), _P(FxPrecond),
                                     _P(FxSolveStats)]),
    "fx_solver_history": (C.c_size_t, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "fx_apply_operator": (C.c_int, [c_be_p, _P(FxOperator), c_buf_p, c_buf_p, _P(FxGrid)]),
    "fx_unet_create": (C.c_int, [c_be_p, C.c_int32, _P(C.c_void_p)]),
    "fx_unet_destroy": (None, [C.c_void_p]),
    "fx_unet_nparams": (C.c_size_t, [C.c_void_p]),
    "fx_unet_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "fx_unet_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "fx_unet_get_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "fx_unet_forward": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p, C.c_int32, C.c_int32, C.c_int32]),
    "fx_unet_loss_grad": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p, C.c_int32, C.c_int32, C.c_int32,
                                    _P(C.c_double)]),
    "fx_unet_loss_grad_energy": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p, C.c_int32, C.c_int32, C.c_int32,
                                    _P(C.c_double)]),
    "fx_unet_adam_step": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float]),
    "fx_learned_create": (C.c_int, [c_be_p, C.c_void_p, C.c_void_p, c_buf_p, _P(FxGrid), C.c_double,
                                    _P(C.c_void_p)]),
    "fx_learned_destroy": (None, [C.c_void_p]),
    "fx_learned_fell_back": (C.c_int32, [C.c_void_p]),
    "fx_learned_set_schedule": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "fx_cn_create": (C.c_int, [c_be_p, _P(FxGrid), C.c_double, C.c_double, _P(FxBoundaries), C.c_double,
                               C.c_uint64, _P(FxCnOptions), _P(C.c_void_p)]),
    "fx_cn_destroy": (None, [C.c_void_p]),
    "fx_cn_step": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p, _P(FxSolveStats)]),
    "fx_cn_step_n": (C.c_int, [C.c_void_p, c_buf_p, C.c_uint64]),
    "fx_cn_dt": (C.c_double, [C.c_void_p]),
    "fx_cn_steps_done": (C.c_uint64, [C.c_void_p]),
    "fx_cn_sim_time": (C.c_double, [C.c_void_p]),
    "fx_cn_solver": (C.c_void_p, [C.c_void_p]),
    "fx_cn_multigrid": (C.c_void_p, [C.c_void_p]),
    "fx_cn_set_preconditioner": (C.c_int, [C.c_void_p, C.POINTER(FxPrecond)]),
    "fx_imex_create": (C.c_int, [c_be_p, _P(FxGrid), C.c_double, C.c_double, C.c_double, C.c_double,
                                 _P(FxBoundaries), C.c_double, C.c_uint64, _P(FxMgOptions), _P(C.c_void_p)]),
    "fx_imex_destroy": (None, [C.c_void_p]),
    "fx_imex_step": (C.c_int, [C.c_void_p, c_buf_p, _P(FxSolveStats)]),
    "fx_imex_step_n": (C.c_int, [C.c_void_p, c_buf_p, C.c_uint64]),
    "fx_imex_dt": (C.c_double, [C.c_void_p]),
    "fx_imex_steps_done": (C.c_uint64, [C.c_void_p]),
    "fx_imex_sim_time": (C.c_double, [C.c_void_p]),
    "fx_bdf_create": (C.c_int, [c_be_p, _P(FxGrid), C.c_double, C.c_double, C.c_uint32, _P(FxBoundaries),
                                C.c_double, C.c_uint64, _P(FxMgOptions), _P(C.c_void_p)]),
    "fx_bdf_destroy": (None, [C.c_void_p]),
    "fx_bdf_step": (C.c_int, [C.c_void_p, c_buf_p, _P(FxSolveStats)]),
    "fx_bdf_step_n": (C.c_int, [C.c_void_p, c_buf_p, C.c_uint64]),
    "fx_bdf_dt": (C.c_double, [C.c_void_p]),
    "fx_bdf_steps_done": (C.c_uint64, [C.c_void_p]),
    "fx_bdf_sim_time": (C.c_double, [C.c_void_p]),
    "fx_bdf_order": (C.c_uint32, [C.c_void_p]),
    "fx_diffusion_create": (C.c_int, [c_be_p, _P(FxGrid), C.c_double, _P(C.c_double), _P(FxBoundaries),
                                      C.c_int32, _P(C.c_void_p)]),
    "fx_diffusion_destroy": (None, [C.c_void_p]),
    "fx_diffusion_enable_graph": (C.c_int, [C.c_void_p]),
    "fx_diffusion_step": (C.c_int, [C.c_void_p, c_buf_p, c_buf_p]),
    "fx_diffusion_step_n": (C.c_int, [C.c_void_p, c_buf_p, C.c_uint64]),
    "fx_diffusion_step_n_with_source": (C.c_int, [C.c_void_p, c_buf_p, C.c_uint64, c_buf_p]),
    "fx_diffusion_dt": (C.c_double, [C.c_void_p]),
    "fx_diffusion_steps_done": (C.c_uint64, [C.c_void_p]),
    "fx_diffusion_sim_time": (C.c_double, [C.c_void_p]),
    "fx_diffusion_scheme": (C.c_int32, [C.c_void_p]),
    "fx_write_raw": (C.c_int, [c_be_p, C.c_char_p, c_buf_p, _P(FxGrid)]),
    "fx_read_raw": (C.c_int, [c_be_p, C.c_char_p, _P(FxGrid), _P(c_buf_p)]),
    "fx_write_raw_host": (C.c_int, [C.c_char_p, C.c_void_p, _P(FxGrid)]),
    "fx_read_raw_host": (C.c_int, [C.c_char_p, _P(FxGrid), C.c_void_p, C.c_size_t]),
}

_lib = None
_exiting = False


def _mark_exiting() -> None:
    global _exiting
    _exiting = True


# Native handles still alive when the interpreter exits are left to the driver: destroy calls made
# during interpreter teardown can run after the HIP runtime has begun its own shutdown.
atexit.register(_mark_exiting)


def live() -> bool:
    """False once the interpreter has begun exiting (destructors then skip native calls)."""
    return not _exiting


def lib() -> C.CDLL:
    """Loads the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m fluxion_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_ERRORS = {
    1: core.PeriodicNotImplemented,
    2: core.DimensionMismatch,
    3: core.InvalidGrid,
    4: core.InvalidStencil,
    10: core.IoError,
}


def last_error() -> str:
    buf = C.create_string_buffer(2048)
    lib().fx_last_error(buf, 2048)
    return buf.value.decode(errors="replace")


def check(status: int, what: str = "") -> None:
    if status != 0:
        exc = _ERRORS.get(status, core.BackendError)
        raise exc(f"{what}: status {status}: {last_error()}")
