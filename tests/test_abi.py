"""The C-ABI library loads and exports every symbol include/fluxion_hip.h declares (CPU only)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "fluxion_hip.h"
LIB = ROOT / "fluxion_amd" / "libfluxion_hip.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fx_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_backend_surface():
    names = declared_functions()
    # One entry point per Backend method (crates/core/src/backend.rs:43-406).
    for op in ["allocate", "upload_f64", "upload_f32", "apply_stencil", "copy", "fill", "axpy", "scale",
               "apply_stencil_var", "stencil_axpy", "norm_l2", "reduce_sum", "reduce_max", "reduce_min", "dot",
               "dot2", "restrict", "prolong", "weighted_jacobi", "pointwise_mult", "pointwise_div", "waxpy",
               "aypx", "reciprocal", "abs_val", "pointwise_max", "pointwise_min", "apply_conv_diff",
               "conv_diff_axpy", "zero_boundaries", "pipelined_cg_fused"]:
        assert f"fx_{op}" in names, op
    assert len(names) > 50


@pytest.mark.skipif(not LIB.exists(), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from fluxion_amd import _lib
    assert set(declared_functions()) <= set(_lib._SIGS), set(declared_functions()) - set(_lib._SIGS)
