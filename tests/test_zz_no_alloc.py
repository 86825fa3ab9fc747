"""fx_dot_noalloc (static device scratch, no allocation) against fx_dot through the C ABI (GPU).

Two calls in a row, the second on a smaller input, so a stale partial left in the shared scratch by
the first call would show; every result must equal fx_dot's bit for bit and the oracle's Kahan dot
within the reductions' 1e-13 relative bar."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from fluxion_amd import _lib as L
from fluxion_amd.core import DimensionMismatch
from fluxion_amd.hip import HipBackend

pytestmark = pytest.mark.gpu


def _dot_pair(be, lib, x, y):
    X, Y = be.upload(x), be.upload(y)
    out = be.allocate(1)
    ref = C.c_double()
    L.check(lib.fx_dot(be.h, X.h, Y.h, C.byref(ref)), "fx_dot")
    L.check(lib.fx_dot_noalloc(be.h, X.h, Y.h, out.h), "fx_dot_noalloc")
    return out.to_numpy()[0], ref.value


def test_dot_noalloc_matches_dot():
    be = HipBackend(0)
    lib = L.lib()
    rng = np.random.default_rng(31)
    for n in (3_000_017, 1 << 20, 4099, 1, 0):
        print(f"fx_dot_noalloc n={n}", flush=True)
        x = rng.standard_normal(n)
        y = rng.standard_normal(n)
        got, ref = _dot_pair(be, lib, x, y)
        assert got == ref, (n, got, ref)
        exact = O.dot(x, y)
        assert abs(got - exact) <= 1e-13 * max(1.0, float(np.abs(x) @ np.abs(y))), (n, got, exact)


def test_dot_noalloc_argument_errors():
    be = HipBackend(0)
    lib = L.lib()
    X, Y = be.upload(np.ones(8)), be.upload(np.ones(9))
    out = be.allocate(1)
    with pytest.raises(DimensionMismatch):
        L.check(lib.fx_dot_noalloc(be.h, X.h, Y.h, out.h), "fx_dot_noalloc")
    empty = be.allocate(0)
    with pytest.raises(DimensionMismatch):
        L.check(lib.fx_dot_noalloc(be.h, X.h, X.h, empty.h), "fx_dot_noalloc")
