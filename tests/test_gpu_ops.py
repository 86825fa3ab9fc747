"""Backend operations on the GPU against the CPU oracle.

Element-wise and stencil results must be bit-identical to the oracle (same
operations, same fused-multiply-add placement); reductions agree to a few ulps
(different summation order, both compensated).
"""
import numpy as np
import pytest

import oracle as O
from fluxion_amd.core import (Boundaries, BoundaryCondition, DimensionMismatch, Grid, InvalidGrid,
                              InvalidStencil, PeriodicNotImplemented, Stencil)

pytestmark = pytest.mark.gpu

ZD = Boundaries.zero_dirichlet()
SHAPES = [(3, 3), (37, 53), (256, 256), (517, 1030)]


def H():
    from fluxion_amd import hip
    return hip


def field(be, g, a):
    return H().Field.from_f64(g, a, be)


def rnd(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


def stencils(g):
    generic = Stencil([(0, 0, -3.5), (-1, 1, 0.25), (1, -1, 0.5), (0, 1, 1.5), (-1, 0, 0.75), (1, 1, -0.125),
                       (0, 0, 0.3)])
    return [Stencil.laplacian_2d_5pt(g.dx, g.dy), Stencil.laplacian_2d_9pt(g.dx, g.dy), generic]


BCS = [ZD,
       Boundaries(BoundaryCondition.dirichlet(1.5), BoundaryCondition.neumann(0.5),
                  BoundaryCondition.robin(2.0, 0.5, 1.0), BoundaryCondition.robin(0.0, 0.0, 3.0))]


@pytest.mark.parametrize("shape", SHAPES)
def test_apply_stencil_bitwise(be, shape):
    rows, cols = shape
    g = Grid(rows, cols, 1.0 / max(cols - 1, 1), 1.0 / max(rows - 1, 1))
    x = rnd(rows * cols, 1)
    fx = field(be, g, x)
    out = H().Field.zeros(g, be=be)
    for st in stencils(g):
        for bc in BCS:
            fx.apply_stencil_into(st, bc, out)
            assert np.array_equal(out.to_vec_f64(), O.apply_stencil(x, g, st, bc)), (shape, len(st), bc)


@pytest.mark.parametrize("shape", SHAPES)
def test_stencil_axpy_and_var_bitwise(be, shape):
    rows, cols = shape
    g = Grid(rows, cols, 0.7 / (cols - 1), 1.3 / (rows - 1))
    x, y, c = rnd(rows * cols, 2), rnd(rows * cols, 3), rnd(rows * cols, 4)
    fx, fc = field(be, g, x), field(be, g, c)
    for st in stencils(g):
        fy = field(be, g, y)
        fy.stencil_axpy(-0.37, fx, st, ZD)
        assert np.array_equal(fy.to_vec_f64(), O.stencil_axpy(-0.37, x, y, g, st, ZD)), (shape, len(st))
        out = H().Field.zeros(g, be=be)
        fx.apply_stencil_var_into(st, BCS[1], fc, out)
        assert np.array_equal(out.to_vec_f64(), O.apply_stencil_var(x, c, g, st, BCS[1]))


def test_stencil_axpy_4096_bitwise(be):
    n = 4096
    g = Grid.square(n, 1.0 / (n - 1))
    x, y = rnd(n * n, 5), rnd(n * n, 6)
    st = Stencil.laplacian_2d_5pt(g.dx, g.dy)
    fy = field(be, g, y)
    fy.stencil_axpy(1e-7, field(be, g, x), st, ZD)
    assert np.array_equal(fy.to_vec_f64(), O.stencil_axpy(1e-7, x, y, g, st, ZD))


@pytest.mark.parametrize("n", [1, 7, 1000, 1 << 20, 3 * (1 << 20) + 5])
def test_blas1_bitwise(be, n):
    g = None
    x, y, z = rnd(n, 7), rnd(n, 8), rnd(n, 9) + 3.0
    hip = H()
    bx, by, bz = be.upload(x), be.upload(y), be.upload(z)

    def F(buf):
        return hip.Field.__new__(hip.Field)

    lib = be._lib
    from fluxion_amd import _lib as L
    L.check(lib.fx_axpy(be.h, 0.3125, bx.h, by.h))
    assert np.array_equal(by.to_numpy(), O.axpy(0.3125, x, y))
    by.write(y)
    L.check(lib.fx_waxpy(be.h, 1.1, bx.h, -2.3, by.h, bz.h))
    assert np.array_equal(bz.to_numpy(), O.waxpy(1.1, x, -2.3, y))
    L.check(lib.fx_aypx(be.h, -0.7, bx.h, by.h))
    assert np.array_equal(by.to_numpy(), O.aypx(-0.7, x, y))
    by.write(y)
    L.check(lib.fx_scale(be.h, by.h, 1.0 / 3.0))
    assert np.array_equal(by.to_numpy(), O.scale(y, 1.0 / 3.0))
    by.write(y)
    for op, fn in enumerate([lib.fx_pointwise_mult, lib.fx_pointwise_div, lib.fx_pointwise_max,
                             lib.fx_pointwise_min]):
        L.check(fn(be.h, bx.h, by.h, bz.h))
        assert np.array_equal(bz.to_numpy(), O.pointwise(op, x, y)), op
    bz.write(z)
    L.check(lib.fx_reciprocal(be.h, bz.h))
    assert np.array_equal(bz.to_numpy(), 1.0 / z)
    L.check(lib.fx_abs_val(be.h, bx.h))
    assert np.array_equal(bx.to_numpy(), np.abs(x))
    L.check(lib.fx_fill(be.h, bx.h, 2.5))
    assert np.all(bx.to_numpy() == 2.5)
    L.check(lib.fx_copy(be.h, by.h, bx.h))
    assert np.array_equal(bx.to_numpy(), y)


@pytest.mark.parametrize("n", [1, 5, 4096, 1 << 20, 16777216])
def test_reductions(be, n):
    import ctypes as C
    from fluxion_amd import _lib as L
    x, y = rnd(n, 10), rnd(n, 11)
    bx, by = be.upload(x), be.upload(y)
    out = C.c_double()
    lib = be._lib
    L.check(lib.fx_dot(be.h, bx.h, by.h, C.byref(out)))
    ref = O.dot(x, y)
    assert abs(out.value - ref) <= 1e-13 * np.dot(np.abs(x), np.abs(y))
    L.check(lib.fx_norm_l2(be.h, bx.h, C.byref(out)))
    assert abs(out.value - O.norm_l2(x)) <= 1e-14 * O.norm_l2(x)
    L.check(lib.fx_reduce_sum(be.h, bx.h, C.byref(out)))
    assert abs(out.value - O.reduce_sum(x)) <= 1e-13 * np.sum(np.abs(x))
    L.check(lib.fx_reduce_max(be.h, bx.h, C.byref(out)))
    assert out.value == O.reduce_max(x)
    L.check(lib.fx_reduce_min(be.h, bx.h, C.byref(out)))
    assert out.value == O.reduce_min(x)
    o2 = (C.c_double * 2)()
    L.check(lib.fx_dot2(be.h, bx.h, by.h, by.h, by.h, o2))
    assert abs(o2[0] - ref) <= 1e-13 * np.dot(np.abs(x), np.abs(y))
    assert abs(o2[1] - O.dot(y, y)) <= 1e-14 * O.dot(y, y)


def test_reduction_closed_forms(be):
    hip = H()
    g = Grid.square(16, 1.0 / 15)
    a, b, c, d = (hip.Field.zeros(g, be=be) for _ in range(4))
    for f, v in ((a, 1.0), (b, 2.0), (c, 3.0), (d, 4.0)):
        f.fill(v)
    d1, d2 = a.dot2(b, c, d)      # solvers.rs:1118-1147
    assert abs(d1 - 512) < 1e-8 and abs(d2 - 3072) < 1e-8
    f16 = be.upload(np.full(16, 3.0))
    import ctypes as C
    out = C.c_double()
    be._lib.fx_norm_l2(be.h, f16.h, C.byref(out))   # ops.rs:1202-1209
    assert abs(out.value - 12.0) < 1e-12
    z = be.upload(np.concatenate([[1e16], np.ones(100000), [-1e16]]))
    be._lib.fx_reduce_sum(be.h, z.h, C.byref(out))
    assert out.value == 100000.0
    nan = be.upload(np.array([1.0, np.nan, 3.0]))
    be._lib.fx_reduce_max(be.h, nan.h, C.byref(out))
    assert out.value == 3.0


def test_stencil_closed_forms(be):
    hip = H()
    n = 32
    h = 1.0 / (n - 1)
    g = Grid.square(n, h)
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    out = hip.Field.zeros(g, be=be)
    field(be, g, (X * X + Y * Y).ravel()).apply_stencil_into(Stencil.laplacian_2d_5pt(h, h), ZD, out)
    assert np.max(np.abs(out.to_numpy()[2:-2, 2:-2] - 4.0)) < 1e-10
    c = 492.4908687776044
    field(be, g, np.full(n * n, c)).apply_stencil_into(Stencil.laplacian_2d_5pt(h, h), ZD, out)
    assert np.max(np.abs(out.to_numpy()[1:-1, 1:-1])) < 1e-10 * (1 + c)


@pytest.mark.parametrize("fine", [(64, 64), (65, 47), (7, 9), (4096, 4096)])
def test_mg_transfer_and_jacobi_bitwise(be, fine):
    hip = H()
    fr, fc = fine
    cr, cc = fr // 2, fc // 2
    fg = Grid(fr, fc, 1.0 / (fc - 1), 1.0 / (fr - 1))
    cg = Grid(cr, cc, 2 * fg.dx, 2 * fg.dy)
    f, c = rnd(fr * fc, 12), rnd(cr * cc, 13)
    Ff, Fc = field(be, fg, f), field(be, cg, c)
    hip.restrict(Ff, Fc)
    assert np.array_equal(Fc.to_vec_f64(), O.restrict(f, fg, cg, 0))
    Fc.write(c)
    hip.prolong(Fc, Ff)
    assert np.array_equal(Ff.to_vec_f64(), O.prolong(c, f, cg, fg, 0))
    st = Stencil.laplacian_2d_5pt(fg.dx, fg.dy)
    b = rnd(fr * fc, 14)
    Ff.write(f)
    Ff.weighted_jacobi(field(be, fg, b), 2.0 / 3.0, st, ZD)
    assert np.array_equal(Ff.to_vec_f64(), O.weighted_jacobi(f, b, 2.0 / 3.0, fg, st, mode=1))


def test_conv_diff_bitwise(be):
    n = 129
    g = Grid.square(n, 1.0 / (n - 1))
    u, k, vx, vy, y = (rnd(n * n, s) for s in range(20, 25))
    k = np.abs(k)
    hip = H()
    Fu, Fk, Fx, Fy = (field(be, g, a) for a in (u, k, vx, vy))
    out = hip.Field.zeros(g, be=be)
    Fu.apply_conv_diff_into(Fk, Fx, Fy, ZD, out)
    assert np.array_equal(out.to_vec_f64(), O.conv_diff(0, 0.0, u, k, vx, vy, g))
    Fo = field(be, g, y)
    Fo.conv_diff_axpy(0.37, Fu, Fk, Fx, Fy, ZD)
    assert np.array_equal(Fo.to_vec_f64(), O.conv_diff(1, 0.37, u, k, vx, vy, g, out=y))


@pytest.mark.parametrize("shape", [(200, 214), (5, 4), (257, 130), (3, 6)])
def test_vark_operator_pair_bitwise(be, shape):
    # Even column counts take the two-column streaming form of the variable-coefficient operator.
    hip = H()
    rows, cols = shape
    g = Grid(rows, cols, 1.0 / (cols - 1), 1.0 / (rows - 1))
    u, k = rnd(rows * cols, 40), np.exp(rnd(rows * cols, 41))
    out = hip.Field.zeros(g, be=be)
    hip.Operator.var_k(field(be, g, k), coeff=0.5, sigma=1.0).apply(field(be, g, u), out)
    assert np.array_equal(out.to_vec_f64(), O.apply_op(O.Op.var_k(k, coeff=0.5, sigma=1.0), u, g))


def test_operators_bitwise(be):
    hip = H()
    n = 200
    g = Grid(n, n + 13, 1.0 / (n + 12), 1.0 / (n - 1))
    u, k, vx, vy = (rnd(n * (n + 13), s) for s in range(30, 34))
    k = np.exp(k)
    Fu, Fk, Fx, Fy = (field(be, g, a) for a in (u, k, vx, vy))
    out = hip.Field.zeros(g, be=be)
    for st in stencils(g):
        hip.Operator.stencil(-0.25, st, ZD).apply(Fu, out)
        assert np.array_equal(out.to_vec_f64(), O.apply_op(O.Op.stencil(-0.25, st), u, g))
    hip.Operator.var_k(Fk, coeff=0.5, sigma=1.0).apply(Fu, out)
    assert np.array_equal(out.to_vec_f64(), O.apply_op(O.Op.var_k(k, coeff=0.5, sigma=1.0), u, g))
    hip.Operator.conv_diff(Fx, Fy, kappa_scalar=-1e-3, coeff=1.0, sigma=0.0).apply(Fu, out)
    assert np.array_equal(out.to_vec_f64(),
                          O.apply_op(O.Op.conv_diff(vx, vy, kappa_scalar=-1e-3, coeff=1.0, sigma=0.0), u, g))
    hip.Operator.conv_diff(Fx, Fy, kappa=Fk, coeff=0.1, sigma=1.0).apply(Fu, out)
    assert np.array_equal(out.to_vec_f64(), O.apply_op(O.Op.conv_diff(vx, vy, kappa=k, coeff=0.1, sigma=1.0), u, g))


def test_zero_boundaries_and_fused_pipelined_update(be):
    hip = H()
    n = 77
    g = Grid.square(n, 1.0 / (n - 1))
    vals = [rnd(n * n, s) for s in range(40, 47)]
    F = [field(be, g, v) for v in vals]
    F[0].zero_boundaries()
    assert np.array_equal(F[0].to_vec_f64(), O.zero_boundaries(vals[0], n, n))
    F[0].write(vals[0])
    a, b = 0.731, -0.219
    hip.pipelined_cg_fused(a, b, *F)
    z, t, p, x, r, w, q = vals
    # pipelined_cg.rs:182-199 in the oracle's primitives.
    z2 = O.axpy(1.0, q, O.scale(z, b))
    t2 = O.axpy(1.0, w, O.scale(t, b))
    p2 = O.axpy(1.0, r, O.scale(p, b))
    expect = [z2, t2, p2, O.axpy(a, p2, x), O.axpy(-a, t2, r), O.axpy(-a, z2, w)]
    for f, e in zip(F[:6], expect):
        assert np.array_equal(f.to_vec_f64(), e)


def test_errors(be):
    hip = H()
    g = Grid.square(8, 1.0 / 7)
    st = Stencil.laplacian_2d_5pt(g.dx, g.dy)
    a, b = hip.Field.zeros(g, be=be), hip.Field.zeros(g, be=be)
    with pytest.raises(PeriodicNotImplemented):
        a.apply_stencil_into(st, Boundaries.uniform(BoundaryCondition.periodic()), b)
    with pytest.raises(InvalidStencil):
        a.stencil_axpy(1.0, b, Stencil([(2, 0, 1.0)]), ZD)
    c = hip.Field.zeros(Grid.square(9, 1.0 / 8), be=be)
    with pytest.raises(DimensionMismatch):
        a.axpy(1.0, c)
    with pytest.raises(DimensionMismatch):
        a.dot(c)
