// fx_mg.hip — geometric multigrid V-cycle (crates/runtime/src/multigrid.rs:12-608).
//
// Levels are FP64 (reference) or FP32 (mixed precision).  The per-point
// arithmetic of every kernel is the same as the oracle's Mg<T> so that FP64
// and FP32 V-cycles are reproducible on the CPU:
//   hierarchy          multigrid.rs:92-146 (coarsest threshold is a parameter)
//   V-cycle order      multigrid.rs:188-230
//   residual           multigrid.rs:476-491 (apply_stencil, scale −1, axpy rhs)
//   2x2 restriction    ops.rs:369-399, injection prolongation ops.rs:401-423
//   Chebyshev          multigrid.rs:317-473
// The smoother is a two-buffer weighted Jacobi sweep with the ops.rs:454-469
// point update; the reference CPU sweep is in place (lexicographic).
#include "fx_mg.hpp"
#include "fx_mgfast.hpp"

#include <cmath>
#include <type_traits>
#include <utility>
#include <vector>

using namespace fxd;

namespace {

template <class T> __device__ __forceinline__ T fmaT(T a, T b, T c);
template <> __device__ __forceinline__ double fmaT<double>(double a, double b, double c) { return fma(a, b, c); }
template <> __device__ __forceinline__ float fmaT<float>(float a, float b, float c) { return fmaf(a, b, c); }

template <class T> struct LW {
	T w[5];
	T inv_diag;
	T omega;
	// Variable-coefficient level operator sigma·u + scale·(−∇·(k∇u)) when k != nullptr.
	const T* k;
	T ix2, iy2, sigma, scale;
	// Convection-diffusion level operator sigma·u + scale·(κΔu + v·∇⁻u) when vx != nullptr:
	// 1/dx, 1/dy and their squares formed as (1/dx)·(1/dx), as ops.rs:686-689 forms them.
	const T *vx, *vy;
	T kap, ix, iy, cix2, ciy2;
};

template <class T> __device__ __forceinline__ T hmT(T a, T b) { return T(2) * a * b / (a + b); }
// sigma·u + scale·(−∇·(k∇u)) at interior point i with harmonic face means; *diag gets the diagonal.
template <class T>
__device__ __forceinline__ T vk_apply(const T* u, const LW<T>& L, int64_t i, int64_t cols, T* diag) {
	const T* k = L.k;
	T kc = k[i], c = u[i];
	T kw = hmT<T>(kc, k[i - 1]), ke = hmT<T>(kc, k[i + 1]);
	T kn = hmT<T>(kc, k[i - cols]), ks = hmT<T>(kc, k[i + cols]);
	T fx = ke * (u[i + 1] - c) - kw * (c - u[i - 1]);
	T fy = ks * (u[i + cols] - c) - kn * (c - u[i - cols]);
	T Lv = -(fx * L.ix2 + fy * L.iy2);
	*diag = fmaT<T>(L.scale, (kw + ke) * L.ix2 + (kn + ks) * L.iy2, L.sigma);
	return fmaT<T>(L.scale, Lv, L.sigma * c);
}

// sigma·u + scale·(κΔu + v·∇⁻u) at interior point i, the arithmetic of k_conv_diff (ops.rs:707-731) in T;
// *diag gets sigma + scale·(|vx|/dx + |vy|/dy − 2κ(1/dx² + 1/dy²)).
template <class T>
__device__ __forceinline__ T cd_apply(const T* u, const LW<T>& L, int64_t i, int64_t cols, T* diag) {
	const T c = u[i];
	const T lap = fmaT<T>(u[i - 1] + u[i + 1], L.cix2, (u[i - cols] + u[i + cols]) * L.ciy2) - T(2) * (L.cix2 + L.ciy2) * c;
	const T diff = L.kap * lap;
	const T vxi = L.vx[i], vyi = L.vy[i];
	const T dudx = vxi >= T(0) ? (c - u[i - 1]) * L.ix : (u[i + 1] - c) * L.ix;
	const T dudy = vyi >= T(0) ? (c - u[i - cols]) * L.iy : (u[i + cols] - c) * L.iy;
	const T conv = fmaT<T>(vxi, dudx, vyi * dudy);
	const T ax = vxi >= T(0) ? vxi : -vxi, ay = vyi >= T(0) ? vyi : -vyi;
	*diag = fmaT<T>(L.scale, fmaT<T>(ax, L.ix, ay * L.iy) - T(2) * L.kap * (L.cix2 + L.ciy2), L.sigma);
	return fmaT<T>(L.scale, diff + conv, L.sigma * c);
}

#define SKIP_RETURN if (skip && *skip) return
#define PT_INDEX                                                           \
	int64_t c = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));  \
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);      \
	if (r >= rows || c >= cols) return;                                    \
	int64_t i = r * cols + c;                                              \
	bool bnd = r == 0 || c == 0 || r == rows - 1 || c == cols - 1;

template <class T>
__global__ __launch_bounds__(BLOCK) void k_cvt_in(const double* b, const double* x, T* rhs, T* u, size_t n, const int* skip) {
	SKIP_RETURN;
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) { rhs[i] = T(b[i]); u[i] = T(x[i]); }
}
template <class T>
__global__ __launch_bounds__(BLOCK) void k_cvt_out(const T* u, double* x, size_t n, const int* skip) {
	SKIP_RETURN;
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) x[i] = double(u[i]);
}
template <class T>
__global__ __launch_bounds__(BLOCK) void k_fillT(T* x, T v, size_t n, const int* skip) {
	SKIP_RETURN;
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) x[i] = v;
}
template <class T>
__global__ __launch_bounds__(BLOCK) void k_axpyT(T a, const T* x, T* y, size_t n, const int* skip) {
	SKIP_RETURN;
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) y[i] = fmaT<T>(a, x[i], y[i]);
}
template <class T>
__global__ __launch_bounds__(BLOCK) void k_scaleT(T* x, T a, size_t n, const int* skip) {
	SKIP_RETURN;
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) x[i] *= a;
}
template <class T>
__global__ __launch_bounds__(BLOCK) void k_copyT(const T* s, T* d, size_t n, const int* skip) {
	SKIP_RETURN;
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) d[i] = s[i];
}
template <class T>
__global__ void k_zero_bndT(T* x, int64_t rows, int64_t cols, const int* skip) {
	SKIP_RETURN;
	int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < cols) { x[t] = T(0); x[(rows - 1) * cols + t] = T(0); }
	else if (t < cols + rows - 2) { int64_t rr = 1 + (t - cols); x[rr * cols] = T(0); x[rr * cols + cols - 1] = T(0); }
}

// Two-buffer damped Jacobi: ax accumulated over entries (0,0),(−1,0),(1,0),(0,−1),(0,1).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_jac(const T* xo, T* xn, const T* b, LW<T> L, int64_t rows, int64_t cols,
                                               const int* skip) {
	SKIP_RETURN;
	PT_INDEX;
	if (bnd) { xn[i] = xo[i]; return; }
	if (L.vx) {
		T d;
		T au = cd_apply<T>(xo, L, i, cols, &d);
		xn[i] = fmaT<T>(L.omega / d, b[i] - au, xo[i]);
		return;
	}
	if (L.k) {
		T d;
		T au = vk_apply<T>(xo, L, i, cols, &d);
		xn[i] = fmaT<T>(L.omega / d, b[i] - au, xo[i]);
		return;
	}
	T ax = T(0);
	ax = fmaT<T>(L.w[0], xo[i], ax);
	ax = fmaT<T>(L.w[1], xo[i - cols], ax);
	ax = fmaT<T>(L.w[2], xo[i + cols], ax);
	ax = fmaT<T>(L.w[3], xo[i - 1], ax);
	ax = fmaT<T>(L.w[4], xo[i + 1], ax);
	T res = b[i] - ax;
	xn[i] = fmaT<T>(L.omega, L.inv_diag * res, xo[i]);
}

__device__ double bcv(fx_bc bc, double self_val, double nb, double spacing) {
	switch (bc.kind) {
	case FX_BC_DIRICHLET: return bc.a;
	case FX_BC_NEUMANN: return fma(spacing, bc.a, nb);
	case FX_BC_ROBIN:
		if (fabs(bc.b) < 1e-30) return fabs(bc.a) < 1e-30 ? 0.0 : bc.g / bc.a;
		return fma(spacing, (bc.g - bc.a * self_val) / bc.b, nb);
	default: return 0.0;
	}
}

// Level stencil (ops.rs:891-938 chain) with bc_value boundaries (ops.rs:799-843).
// mode 0: out = S(u); mode 1: out = rhs − S(u) (multigrid.rs:476-491).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_lstencil(const T* u, const T* rhs, T* out, LW<T> L, fx_boundaries bc,
                                                    double dx, double dy, int mode, int64_t rows, int64_t cols,
                                                    const int* skip) {
	SKIP_RETURN;
	PT_INDEX;
	T s;
	if (bnd) {
		if (r == 0) s = T(bcv(bc.top, double(u[i]), double(u[cols + c]), dy));
		else if (r == rows - 1) s = T(bcv(bc.bottom, double(u[i]), double(u[i - cols]), dy));
		else if (c == 0) s = T(bcv(bc.left, double(u[i]), double(u[i + 1]), dx));
		else s = T(bcv(bc.right, double(u[i]), double(u[i - 1]), dx));
	} else if (L.vx) {
		T d;
		s = cd_apply<T>(u, L, i, cols, &d);
	} else if (L.k) {
		T d;
		s = vk_apply<T>(u, L, i, cols, &d);
	} else {
		s = fmaT<T>(L.w[0], u[i], fmaT<T>(L.w[1], u[i - cols], fmaT<T>(L.w[2], u[i + cols],
		    fmaT<T>(L.w[3], u[i - 1], L.w[4] * u[i + 1]))));
	}
	out[i] = mode == 0 ? s : fmaT<T>(T(1), rhs[i], -s);
}

// Linear transfer between level grids spanning the same domain: fine index i
// sits at coarse coordinate i·(nc−1)/(nf−1) = I0 + w (exact integer split).
struct Lin1 {
	int64_t i0;
	double w;
};
__device__ __forceinline__ Lin1 lin1(int64_t i, int64_t nf, int64_t nc) {
	int64_t num = i * (nc - 1);
	int64_t den = nf - 1;
	return {num / den, double(num % den) * (1.0 / double(den))};
}
// Weight of fine index i toward coarse index I in the linear transfer
// (tabulated split i0 + w): 1 − w, w or 0.
__device__ __forceinline__ double lin_weight(const int32_t* i0, const double* w, int64_t i, int64_t I) {
	const int64_t a = i0[i];
	if (a == I) return 1.0 - w[i];
	if (a + 1 == I) return w[i];
	return 0.0;
}

// Restriction: reference 2x2 average (ops.rs:369-399) or the transpose of the
// linear prolongation scaled by (h/H)² (zero on the coarse boundary).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_restrict(const T* f, T* co, int64_t fr, int64_t fc, int64_t cr, int64_t cc,
                                                    int lin, fx_xtab X, const int* skip) {
	SKIP_RETURN;
	int64_t col = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= cr || col >= cc) return;
	if (!lin) {
		int64_t a = 2 * r, b = 2 * col;
		int64_t a1 = a + 1 < fr - 1 ? a + 1 : fr - 1, b1 = b + 1 < fc - 1 ? b + 1 : fc - 1;
		co[r * cc + col] = T(0.25) * (f[a * fc + b] + f[a1 * fc + b] + f[a * fc + b1] + f[a1 * fc + b1]);
		return;
	}
	if (r == 0 || col == 0 || r == cr - 1 || col == cc - 1) { co[r * cc + col] = T(0); return; }
	const T sy = T(double(cr - 1) / double(fr - 1)), sx = T(double(cc - 1) / double(fc - 1));
	const int64_t rlo = X.lor[r], rhi = X.hir[r], clo = X.loc[col], chi = X.hic[col];
	T acc = T(0);
	for (int64_t i = rlo; i <= rhi; ++i) {
		T wy = T(lin_weight(X.i0r, X.wr, i, r));
		T row = T(0);
		for (int64_t j = clo; j <= chi; ++j) row += T(lin_weight(X.i0c, X.wc, j, col)) * f[i * fc + j];
		acc += wy * row;
	}
	co[r * cc + col] = acc * (sy * sx);
}

// Prolongation (additive): reference injection over the whole grid
// (ops.rs:401-423, coarse index clamped for odd sizes) or linear
// interpolation onto the fine interior.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_prolong(const T* co, T* f, int64_t cr, int64_t cc, int64_t fr, int64_t fc,
                                                   int lin, fx_xtab X, const int* skip) {
	SKIP_RETURN;
	int64_t col = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= fr || col >= fc) return;
	if (!lin) {
		int64_t a = r / 2 < cr - 1 ? r / 2 : cr - 1, b = col / 2 < cc - 1 ? col / 2 : cc - 1;
		f[r * fc + col] += co[a * cc + b];
		return;
	}
	if (r == 0 || col == 0 || r == fr - 1 || col == fc - 1) return;
	T wy = T(X.wr[r]), vy = T(1) - wy, wx = T(X.wc[col]), vx = T(1) - wx;
	int64_t a0 = X.i0r[r], a1 = a0 + 1 < cr ? a0 + 1 : a0;
	int64_t b0 = X.i0c[col], b1 = b0 + 1 < cc ? b0 + 1 : b0;
	T top = vx * co[a0 * cc + b0] + wx * co[a0 * cc + b1];
	T bot = vx * co[a1 * cc + b0] + wx * co[a1 * cc + b1];
	f[r * fc + col] += vy * top + wy * bot;
}

// ── streaming FP64 level kernels (even column counts) ────────────────
// The per-point arithmetic of k_jac / k_lstencil on the two-column traversal of
// fxd::k_pair5: the level's constant 5-point weights or, with VK, its
// coefficient field with harmonic face means.  Results equal the per-point
// kernels' bit for bit.

// Weighted Jacobi sweep: xn = xo + (ω/d)·(b − A·xo) inside, xn = xo on the boundary.
template <bool VK>
struct EpiJacD {
	const double* b;
	double* xn;
	LW<double> L;
	__device__ double pt(const Nb& u, const Nb& k, double bv) const {
		if constexpr (VK) {
			double fs;
			const double lv = vk_lap(u, k, L.ix2, L.iy2, &fs);
			const double d = fma(L.scale, fs, L.sigma);
			const double au = fma(L.scale, lv, L.sigma * u.c);
			return fma(L.omega / d, bv - au, u.c);
		} else {
			double ax = 0.0;
			ax = fma(L.w[0], u.c, ax);
			ax = fma(L.w[1], u.n, ax);
			ax = fma(L.w[2], u.s, ax);
			ax = fma(L.w[3], u.w, ax);
			ax = fma(L.w[4], u.e, ax);
			return fma(L.omega, L.inv_diag * (bv - ax), u.c);
		}
	}
	template <class A>
	__device__ void pair(int64_t i, int64_t, int64_t, bool rin, bool ia, bool ib, const Nb& uA, const Nb& uB,
	                     const Nb& kA, const Nb& kB, A*) const {
		const double2 bv = *(const double2*)(b + i);
		double2 o{uA.c, uB.c};
		if (rin && ia) o.x = pt(uA, kA, bv.x);
		if (rin && ib) o.y = pt(uB, kB, bv.y);
		*(double2*)(xn + i) = o;
	}
};

// Level stencil with bc_value boundaries: out = S(u) (mode 0) or rhs − S(u) (mode 1).
template <bool VK>
struct EpiLStD {
	const double* rhs;
	double* out;
	LW<double> L;
	fx_boundaries bc;
	double dx, dy;
	int64_t rows, cols;
	int mode;
	__device__ double in(const Nb& u, const Nb& k) const {
		if constexpr (VK) {
			double fs;
			return fma(L.scale, vk_lap(u, k, L.ix2, L.iy2, &fs), L.sigma * u.c);
		} else {
			return fma(L.w[0], u.c, fma(L.w[1], u.n, fma(L.w[2], u.s, fma(L.w[3], u.w, L.w[4] * u.e))));
		}
	}
	__device__ double bnd(int64_t r, int64_t c, const Nb& u) const {
		if (r == 0) return bcv(bc.top, u.c, u.s, dy);
		if (r == rows - 1) return bcv(bc.bottom, u.c, u.n, dy);
		if (c == 0) return bcv(bc.left, u.c, u.e, dx);
		return bcv(bc.right, u.c, u.w, dx);
	}
	template <class A>
	__device__ void pair(int64_t i, int64_t r, int64_t c0, bool rin, bool ia, bool ib, const Nb& uA, const Nb& uB,
	                     const Nb& kA, const Nb& kB, A*) const {
		double2 s{rin && ia ? in(uA, kA) : bnd(r, c0, uA), rin && ib ? in(uB, kB) : bnd(r, c0 + 1, uB)};
		if (mode == 1) {
			const double2 rv = *(const double2*)(rhs + i);
			s = double2{fma(1.0, rv.x, -s.x), fma(1.0, rv.y, -s.y)};
		}
		*(double2*)(out + i) = s;
	}
};

// Rows per block of a two-column traversal: as tall as 1024 workgroups allow.
int pair_rpb(int64_t rows, int64_t cols) {
	const int64_t cb = (cols / 2 + BLOCK - 1) / BLOCK;
	int rpb = 64;
	while (rpb > 4 && cb * ((rows + rpb - 1) / rpb) < 1024) rpb /= 2;
	return rpb;
}

// Linear-transfer restriction streamed down the fine rows.  A thread owns one
// coarse column J of its block's 256-column strip; the block walks the fine
// rows of a band of `rb` coarse rows once.  Each fine row's column sum is the
// j-ordered sum of T(lin_weight)·f that k_restrict forms, and it is added, with
// its row weight, to every coarse row whose window [lor, hir] holds that fine
// row: at most three, the lowest open one (Ic) and the two above it, since a
// fine row at coarse coordinate t lies in the windows of the rows in [t−1, t+1].
// A coarse row is written once the fine rows pass its window.  Every coarse
// value therefore gets exactly k_restrict's terms in k_restrict's order.
template <class T, int MAXW>
__global__ __launch_bounds__(BLOCK) void k_restrict_s(const T* __restrict__ f, T* __restrict__ co, int64_t fr,
                                                      int64_t fc, int64_t cr, int64_t cc, int rb, fx_xtab X,
                                                      const int* skip) {
	SKIP_RETURN;
	int bx, by;
	xcd_tile(&bx, &by);
	const int64_t J = (int64_t)bx * BLOCK + threadIdx.x;
	const int64_t I0 = (int64_t)by * rb, I1 = I0 + rb < cr ? I0 + rb : cr;
	if (J >= cc) return;   // no workgroup barrier below
	if (J == 0 || J == cc - 1) {
		for (int64_t I = I0; I < I1; ++I) co[I * cc + J] = T(0);
		return;
	}
	if (I0 == 0) co[J] = T(0);
	if (I1 == cr) co[(cr - 1) * cc + J] = T(0);
	const int64_t Ia = I0 > 1 ? I0 : 1, Ib = I1 < cr - 1 ? I1 : cr - 1;   // interior coarse rows [Ia, Ib)
	if (Ia >= Ib) return;
	const T sy = T(double(cr - 1) / double(fr - 1)), sx = T(double(cc - 1) / double(fc - 1));
	const T sc = sy * sx;
	const int64_t clo = X.loc[J];
	const int nw = (int)(X.hic[J] - clo + 1);
	T cw[MAXW];
#pragma unroll
	for (int q = 0; q < MAXW; ++q) cw[q] = q < nw ? T(lin_weight(X.i0c, X.wc, clo + q, J)) : T(0);
	int64_t Ic = Ia;
	T a0 = T(0), a1 = T(0), a2 = T(0);   // coarse rows Ic, Ic + 1, Ic + 2
	const int64_t ilo = X.lor[Ia], ihi = X.hir[Ib - 1];
	for (int64_t i = ilo; i <= ihi; ++i) {
		while (Ic < Ib && X.hir[Ic] < i) {
			co[Ic * cc + J] = a0 * sc;
			a0 = a1;
			a1 = a2;
			a2 = T(0);
			++Ic;
		}
		const T* fi = f + i * fc + clo;
		T row = T(0);
#pragma unroll
		for (int q = 0; q < MAXW; ++q)
			if (q < nw) row += cw[q] * fi[q];
		if (X.lor[Ic] <= i && i <= X.hir[Ic]) a0 += T(lin_weight(X.i0r, X.wr, i, Ic)) * row;
		if (Ic + 1 < Ib && X.lor[Ic + 1] <= i && i <= X.hir[Ic + 1]) a1 += T(lin_weight(X.i0r, X.wr, i, Ic + 1)) * row;
		if (Ic + 2 < Ib && X.lor[Ic + 2] <= i && i <= X.hir[Ic + 2]) a2 += T(lin_weight(X.i0r, X.wr, i, Ic + 2)) * row;
	}
	while (Ic < Ib) {
		co[Ic * cc + J] = a0 * sc;
		a0 = a1;
		a1 = a2;
		a2 = T(0);
		++Ic;
	}
}
constexpr int RESTRICT_MAXW = 8;

// Linear-transfer restriction, one thread per coarse point with its whole
// window's loads issued together (at most MAXW fine rows and columns): for
// levels too small to give the streaming kernel enough workgroups, whose
// sequential walk would then expose each fine row's latency.  The same terms
// in the same order as k_restrict.
template <class T, int MAXW>
__global__ __launch_bounds__(BLOCK) void k_restrict_p(const T* __restrict__ f, T* __restrict__ co, int64_t fr,
                                                      int64_t fc, int64_t cr, int64_t cc, fx_xtab X, const int* skip) {
	SKIP_RETURN;
	const int64_t col = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	const int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= cr || col >= cc) return;
	if (r == 0 || col == 0 || r == cr - 1 || col == cc - 1) {
		co[r * cc + col] = T(0);
		return;
	}
	const T sy = T(double(cr - 1) / double(fr - 1)), sx = T(double(cc - 1) / double(fc - 1));
	const int64_t rlo = X.lor[r], clo = X.loc[col];
	const int nh = (int)(X.hir[r] - rlo + 1), nw = (int)(X.hic[col] - clo + 1);
	T cw[MAXW], rw[MAXW], v[MAXW][MAXW];
#pragma unroll
	for (int p = 0; p < MAXW; ++p) cw[p] = p < nw ? T(lin_weight(X.i0c, X.wc, clo + p, col)) : T(0);
#pragma unroll
	for (int q = 0; q < MAXW; ++q) rw[q] = q < nh ? T(lin_weight(X.i0r, X.wr, rlo + q, r)) : T(0);
#pragma unroll
	for (int q = 0; q < MAXW; ++q)
#pragma unroll
		for (int p = 0; p < MAXW; ++p) v[q][p] = q < nh && p < nw ? f[(rlo + q) * fc + clo + p] : T(0);
	T acc = T(0);
#pragma unroll
	for (int q = 0; q < MAXW; ++q) {
		if (q < nh) {
			T row = T(0);
#pragma unroll
			for (int p = 0; p < MAXW; ++p)
				if (p < nw) row += cw[p] * v[q][p];
			acc += rw[q] * row;
		}
	}
	co[r * cc + col] = acc * (sy * sx);
}
constexpr int RESTRICT_PMAXW = 5;

// Linear prolongation (additive) streamed down the fine rows: a thread owns one
// fine column; the four coarse values it interpolates between are reloaded only
// when the coarse row pair (i0r[r], i0r[r] + 1) changes.  Per-point arithmetic
// equals k_prolong's.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_prolong_s(const T* __restrict__ co, T* __restrict__ f, int64_t cr,
                                                     int64_t cc, int64_t fr, int64_t fc, int rpb, fx_xtab X,
                                                     const int* skip) {
	SKIP_RETURN;
	int bx, by;
	xcd_tile(&bx, &by);
	const int64_t c = (int64_t)bx * BLOCK + threadIdx.x;
	if (c <= 0 || c >= fc - 1) return;
	const int64_t r0 = (int64_t)by * rpb > 1 ? (int64_t)by * rpb : 1;
	const int64_t r1 = (int64_t)by * rpb + rpb < fr - 1 ? (int64_t)by * rpb + rpb : fr - 1;
	const T wx = T(X.wc[c]), vx = T(1) - wx;
	const int64_t b0 = X.i0c[c], b1 = b0 + 1 < cc ? b0 + 1 : b0;
	int64_t ra = -1;
	T t0 = T(0), t1 = T(0), u0 = T(0), u1 = T(0);   // rows a0 (t) and a1 (u) at columns b0, b1
	for (int64_t r = r0; r < r1; ++r) {
		const int64_t a0 = X.i0r[r], a1 = a0 + 1 < cr ? a0 + 1 : a0;
		if (a0 != ra) {
			t0 = co[a0 * cc + b0];
			t1 = co[a0 * cc + b1];
			u0 = co[a1 * cc + b0];
			u1 = co[a1 * cc + b1];
			ra = a0;
		}
		const T wy = T(X.wr[r]), vy = T(1) - wy;
		const T top = vx * t0 + wx * t1;
		const T bot = vx * u0 + wx * u1;
		f[r * fc + c] += vy * top + wy * bot;
	}
}

dim3 pgrid(int64_t rows, int64_t cols) {
	return dim3((unsigned)((cols + WAVE - 1) / WAVE), (unsigned)((rows + NWAVES - 1) / NWAVES));
}

template <class T>
__global__ void k_cvt_k(const double* k, T* out, size_t n) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] = T(k[i]);
}
// Coarse-level coefficient: injection for halving levels, linear interpolation
// at the coarse node positions for same-domain levels.
template <class T>
__global__ void k_coarsen_k(const T* kf, T* kc, int64_t fr, int64_t fc, int64_t cr, int64_t cc, int lin) {
	int64_t C = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t R = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (R >= cr || C >= cc) return;
	if (!lin) {
		int64_t a = 2 * R < fr - 1 ? 2 * R : fr - 1, b = 2 * C < fc - 1 ? 2 * C : fc - 1;
		kc[R * cc + C] = kf[a * fc + b];
		return;
	}
	Lin1 ly = lin1(R, cr, fr), lx = lin1(C, cc, fc);
	int64_t a0 = ly.i0, a1 = ly.i0 + 1 < fr - 1 ? ly.i0 + 1 : fr - 1;
	int64_t b0 = lx.i0, b1 = lx.i0 + 1 < fc - 1 ? lx.i0 + 1 : fc - 1;
	T wy = T(ly.w), vy = T(1) - wy, wx = T(lx.w), vx = T(1) - wx;
	T top = vx * kf[a0 * fc + b0] + wx * kf[a0 * fc + b1];
	T bot = vx * kf[a1 * fc + b0] + wx * kf[a1 * fc + b1];
	kc[R * cc + C] = vy * top + wy * bot;
}

// Face coefficients of a level: kx[i] = hm(k[i], k[i+1]), ky[i] = hm(k[i], k[i+cols])
// (the harmonic mean is symmetric bit for bit: 2·a·b and a + b are exact in either order).
__global__ void k_faces(const float* k, float* kx, float* ky, int64_t rows, int64_t cols) {
	int64_t c = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= rows || c >= cols) return;
	int64_t i = r * cols + c;
	kx[i] = c + 1 < cols ? fmg::hmf(k[i], k[i + 1]) : 0.0f;
	ky[i] = r + 1 < rows ? fmg::hmf(k[i], k[i + cols]) : 0.0f;
}

// {kx, ky, ω/d} of each point side by side, so that a wave reads a row's three
// coefficients with one 12-byte load per lane instead of three 4-byte loads.
__global__ void k_pack3(const float* kx, const float* ky, const float* wd, float* kp, int64_t n) {
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
		kp[3 * i] = kx[i];
		kp[3 * i + 1] = ky[i];
		kp[3 * i + 2] = wd[i];
	}
}

// ω/diagonal of the variable-coefficient level operator on interior points, from the
// face coefficients (the d of fmg::vkff, so wd equals the ω/d of fmg::jackf).
__global__ void k_wdiag(const float* kx, const float* ky, float* wd, fmg::LWf L, int64_t rows, int64_t cols) {
	int64_t c = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= rows || c >= cols) return;
	int64_t i = r * cols + c;
	if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) {
		wd[i] = 0.0f;
		return;
	}
	float d;
	(void)fmg::vkff(L, kx[i - 1], kx[i], ky[i - cols], ky[i], 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, &d);
	wd[i] = L.omega / d;
}

template <class T>
struct Driver {
	fx_multigrid* mg;
	const int* skip;
	hipStream_t st;

	LW<T> lw(const fx_multigrid::Level& L) const {
		LW<T> o;
		for (int e = 0; e < 5; ++e) o.w[e] = T(L.w[e]);
		o.inv_diag = T(L.inv_diag);
		o.omega = T(mg->opt.omega);
		o.k = (const T*)L.k;
		o.ix2 = T(1.0 / (L.g.dx * L.g.dx));
		o.iy2 = T(1.0 / (L.g.dy * L.g.dy));
		o.sigma = T(mg->opt.sigma);
		o.scale = T(mg->opt.scale);
		o.vx = (const T*)L.vx;
		o.vy = (const T*)L.vy;
		o.kap = T(mg->opt.kappa);
		o.ix = T(1.0 / L.g.dx);
		o.iy = T(1.0 / L.g.dy);
		o.cix2 = T((1.0 / L.g.dx) * (1.0 / L.g.dx));
		o.ciy2 = T((1.0 / L.g.dy) * (1.0 / L.g.dy));
		return o;
	}
	size_t n(const fx_multigrid::Level& L) const { return (size_t)L.g.rows * L.g.cols; }

	fx_status smooth(int l, int count) {
		if (count <= 0) return FX_OK;
		fx_multigrid::Level& L = mg->lv[l];
		if (mg->opt.smoother == FX_SMOOTH_CHEBYSHEV) return cheby(l, count);
		LW<T> w = lw(L);
		for (int it = 0; it < count; ++it) {
			if constexpr (std::is_same_v<T, double>) {
				if (pair_ok(L)) {
					const int64_t rows = (int64_t)L.g.rows, cols = (int64_t)L.g.cols;
					const int rpb = pair_rpb(rows, cols);
					const dim3 grid = fx::stream_grid(rows, cols / 2, rpb);
					const PLoad v{(const double*)L.u};
					if (L.k)
						hipLaunchKernelGGL((k_pair5<0, true, PLoad, EpiJacD<true>>), grid, dim3(BLOCK), 0, st, v,
						                   (const double*)L.k, EpiJacD<true>{(const double*)L.rhs, (double*)L.tmp, w},
						                   rows, cols, rpb, RedScratch{}, skip);
					else
						hipLaunchKernelGGL((k_pair5<0, false, PLoad, EpiJacD<false>>), grid, dim3(BLOCK), 0, st, v,
						                   (const double*)nullptr,
						                   EpiJacD<false>{(const double*)L.rhs, (double*)L.tmp, w}, rows, cols, rpb,
						                   RedScratch{}, skip);
					std::swap(L.u, L.tmp);
					continue;
				}
			}
			hipLaunchKernelGGL(k_jac<T>, pgrid(L.g.rows, L.g.cols), dim3(BLOCK), 0, st, (const T*)L.u, (T*)L.tmp,
			                   (const T*)L.rhs, w, (int64_t)L.g.rows, (int64_t)L.g.cols, skip);
			std::swap(L.u, L.tmp);
		}
		return fx::check_launch("mg jacobi");
	}

	// FP64 levels without a velocity field and with an even column count take the
	// two-column streaming kernels.
	static bool pair_ok(const fx_multigrid::Level& L) { return !L.vx && L.g.cols % 2 == 0 && L.g.cols >= 2; }

	fx_status lstencil(int l, const T* u, T* out, int mode) {
		fx_multigrid::Level& L = mg->lv[l];
		if constexpr (std::is_same_v<T, double>) {
			if (pair_ok(L)) {
				const int64_t rows = (int64_t)L.g.rows, cols = (int64_t)L.g.cols;
				const int rpb = pair_rpb(rows, cols);
				const dim3 grid = fx::stream_grid(rows, cols / 2, rpb);
				if (L.k)
					hipLaunchKernelGGL((k_pair5<0, true, PLoad, EpiLStD<true>>), grid, dim3(BLOCK), 0, st, PLoad{u},
					                   (const double*)L.k,
					                   EpiLStD<true>{(const double*)L.rhs, out, lw(L), mg->bc, L.g.dx, L.g.dy, rows, cols,
					                                 mode},
					                   rows, cols, rpb, RedScratch{}, skip);
				else
					hipLaunchKernelGGL((k_pair5<0, false, PLoad, EpiLStD<false>>), grid, dim3(BLOCK), 0, st, PLoad{u},
					                   (const double*)nullptr,
					                   EpiLStD<false>{(const double*)L.rhs, out, lw(L), mg->bc, L.g.dx, L.g.dy, rows,
					                                  cols, mode},
					                   rows, cols, rpb, RedScratch{}, skip);
				return fx::check_launch("mg level stencil");
			}
		}
		hipLaunchKernelGGL(k_lstencil<T>, pgrid(L.g.rows, L.g.cols), dim3(BLOCK), 0, st, u, (const T*)L.rhs, out, lw(L),
		                   mg->bc, L.g.dx, L.g.dy, mode, (int64_t)L.g.rows, (int64_t)L.g.cols, skip);
		return fx::check_launch("mg level stencil");
	}

	unsigned eb(size_t nn) const { return fx::elem_blocks(nn); }

	// multigrid.rs:317-473.
	fx_status cheby(int l, int count) {
		fx_multigrid::Level& L = mg->lv[l];
		const size_t N = n(L);
		const int64_t rows = L.g.rows, cols = L.g.cols;
		const double emax = 2.0, emin = 0.1 * emax;
		const double scl = 2.0 / (emax + emin), alpha = 1.0 - scl * emin, mu = 1.0 / alpha, omegaprod = 2.0 / alpha;
		const double ix2 = 1.0 / (L.g.dx * L.g.dx), iy2 = 1.0 / (L.g.dy * L.g.dy);
		const double inv_diag = -1.0 / (2.0 * ix2 + 2.0 * iy2);
		unsigned zb = (unsigned)((cols + rows - 2 + 255) / 256);
		auto axpy = [&](T a, const void* x, void* y) {
			hipLaunchKernelGGL(k_axpyT<T>, dim3(eb(N)), dim3(BLOCK), 0, st, a, (const T*)x, (T*)y, N, skip);
		};
		auto scale = [&](void* x, T a) { hipLaunchKernelGGL(k_scaleT<T>, dim3(eb(N)), dim3(BLOCK), 0, st, (T*)x, a, N, skip); };
		auto copy = [&](const void* s, void* d) {
			hipLaunchKernelGGL(k_copyT<T>, dim3(eb(N)), dim3(BLOCK), 0, st, (const T*)s, (T*)d, N, skip);
		};
		auto zero_bnd = [&](void* x) { hipLaunchKernelGGL(k_zero_bndT<T>, dim3(zb), dim3(256), 0, st, (T*)x, rows, cols, skip); };
		double c_km1 = 1.0, c_k = mu;
		FX_CHECK(lstencil(l, (const T*)L.u, (T*)L.tmp, 0));
		copy(L.rhs, L.w0);
		axpy(T(-1), L.tmp, L.w0);
		scale(L.w0, T(scl * inv_diag));
		axpy(T(1), L.u, L.w0);
		zero_bnd(L.w0);
		if (count == 1) { copy(L.w0, L.u); return fx::check_launch("mg chebyshev"); }
		void** B[3] = {&L.u, &L.w0, &L.w1};
		int km1 = 0, k = 1, kp1 = 2;
		for (int i = 1; i < count; ++i) {
			FX_CHECK(lstencil(l, (const T*)*B[k], (T*)L.tmp, 0));
			void* p = *B[kp1];
			copy(L.rhs, p);
			axpy(T(-1), L.tmp, p);
			scale(p, T(inv_diag));
			double c_kp1 = 2.0 * mu * c_k - c_km1;
			double om = omegaprod * c_k / c_kp1;
			scale(p, T(om * scl));
			axpy(T(om), *B[k], p);
			axpy(T(1.0 - om), *B[km1], p);
			zero_bnd(p);
			c_km1 = c_k;
			c_k = c_kp1;
			int t = km1; km1 = k; k = kp1; kp1 = t;
		}
		if (k != 0) copy(*B[k], L.u);
		return fx::check_launch("mg chebyshev");
	}

	fx_status restrict_level(int l) {
		auto& lv = mg->lv;
		const fx_grid& fg = lv[l].g;
		const fx_grid& cg = lv[l + 1].g;
		const int fw = mg->opt.transfer == FX_TRANSFER_FW_BILINEAR;
		const int rb = 8;
		// The streaming kernel where its bands give at least 1024 workgroups (measured at
		// 4096² -> 2048² FP64: 61 µs against 83 µs per point; the per-point kernel is the
		// faster one on every smaller level).
		const bool stream = (int64_t)((cg.cols + BLOCK - 1) / BLOCK) * (int64_t)((cg.rows + rb - 1) / rb) >= 1024;
		if (fw && !stream && lv[l].xt.gmax <= RESTRICT_PMAXW && lv[l].xt.gmaxr <= RESTRICT_PMAXW) {
			hipLaunchKernelGGL((k_restrict_p<T, RESTRICT_PMAXW>), pgrid(cg.rows, cg.cols), dim3(BLOCK), 0, st,
			                   (const T*)lv[l].tmp, (T*)lv[l + 1].rhs, (int64_t)fg.rows, (int64_t)fg.cols,
			                   (int64_t)cg.rows, (int64_t)cg.cols, lv[l].xt, skip);
		} else if (fw && lv[l].xt.gmax <= RESTRICT_MAXW) {
			hipLaunchKernelGGL((k_restrict_s<T, RESTRICT_MAXW>),
			                   dim3((unsigned)((cg.cols + BLOCK - 1) / BLOCK), (unsigned)((cg.rows + rb - 1) / rb)),
			                   dim3(BLOCK), 0, st, (const T*)lv[l].tmp, (T*)lv[l + 1].rhs, (int64_t)fg.rows,
			                   (int64_t)fg.cols, (int64_t)cg.rows, (int64_t)cg.cols, rb, lv[l].xt, skip);
		} else {
			hipLaunchKernelGGL(k_restrict<T>, pgrid(cg.rows, cg.cols), dim3(BLOCK), 0, st, (const T*)lv[l].tmp,
			                   (T*)lv[l + 1].rhs, (int64_t)fg.rows, (int64_t)fg.cols, (int64_t)cg.rows, (int64_t)cg.cols,
			                   fw, lv[l].xt, skip);
		}
		return FX_OK;
	}

	fx_status prolong_level(int l) {
		auto& lv = mg->lv;
		const fx_grid& fg = lv[l].g;
		const fx_grid& cg = lv[l + 1].g;
		const int fw = mg->opt.transfer == FX_TRANSFER_FW_BILINEAR;
		if (fw) {
			// Rows per block: the walk is sequential, so small levels take short walks
			// and more workgroups.
			const int64_t cb = (int64_t)((fg.cols + BLOCK - 1) / BLOCK);
			int rpb = 16;
			while (rpb > 1 && cb * (((int64_t)fg.rows + rpb - 1) / rpb) < 2048) rpb /= 2;
			hipLaunchKernelGGL(k_prolong_s<T>,
			                   dim3((unsigned)((fg.cols + BLOCK - 1) / BLOCK), (unsigned)((fg.rows + rpb - 1) / rpb)),
			                   dim3(BLOCK), 0, st, (const T*)lv[l + 1].u, (T*)lv[l].u, (int64_t)cg.rows, (int64_t)cg.cols,
			                   (int64_t)fg.rows, (int64_t)fg.cols, rpb, lv[l].xt, skip);
		} else {
			hipLaunchKernelGGL(k_prolong<T>, pgrid(fg.rows, fg.cols), dim3(BLOCK), 0, st, (const T*)lv[l + 1].u,
			                   (T*)lv[l].u, (int64_t)cg.rows, (int64_t)cg.cols, (int64_t)fg.rows, (int64_t)fg.cols, fw,
			                   lv[l].xt, skip);
		}
		return FX_OK;
	}

	fx_status run(double* z, const double* r) {
		auto& lv = mg->lv;
		const int nl = (int)lv.size();
		const auto& o = mg->opt;
		size_t n0 = n(lv[0]);
		// FP64 levels work on the caller's vectors in place of the copies
		// rhs ← r, u ← z, z ← u (multigrid.rs:194-195, 228): level 0 reads r as its
		// right-hand side and iterates in z, the level's own u buffer standing aside.
		// The values are those of the copies.  Not when z and r share memory.
		fx_multigrid::Level& L0 = lv[0];
		void* const own_u = L0.u;
		void* const own_rhs = L0.rhs;
		void* const own_tmp = L0.tmp;
		const bool alias = std::is_same_v<T, double> && (const void*)z != (const void*)r;
		if (alias) {
			L0.u = z;
			L0.rhs = const_cast<double*>(r);
		} else {
			hipLaunchKernelGGL(k_cvt_in<T>, dim3(eb(n0)), dim3(BLOCK), 0, st, r, (const double*)z, (T*)lv[0].rhs,
			                   (T*)lv[0].u, n0, skip);
		}
		fx_status s = FX_OK;
		for (int l = 0; l < nl - 1 && s == FX_OK; ++l) {
			s = smooth(l, o.n_pre);
			if (s == FX_OK) s = lstencil(l, (const T*)lv[l].u, (T*)lv[l].tmp, 1);
			if (s == FX_OK) s = restrict_level(l);
			size_t nc = n(lv[l + 1]);
			hipLaunchKernelGGL(k_fillT<T>, dim3(eb(nc)), dim3(BLOCK), 0, st, (T*)lv[l + 1].u, T(0), nc, skip);
		}
		if (s == FX_OK) s = smooth(nl - 1, o.n_pre + o.n_post + o.extra_coarse);
		for (int l = nl - 2; l >= 0 && s == FX_OK; --l) {
			s = prolong_level(l);
			if (s == FX_OK) s = smooth(l, o.n_post);
		}
		if (s == FX_OK) {
			if (alias) {
				if (L0.u != (void*)z)
					hipLaunchKernelGGL(k_copyT<T>, dim3(eb(n0)), dim3(BLOCK), 0, st, (const T*)L0.u, (T*)z, n0, skip);
			} else {
				hipLaunchKernelGGL(k_cvt_out<T>, dim3(eb(n0)), dim3(BLOCK), 0, st, (const T*)lv[0].u, z, n0, skip);
			}
		}
		if (alias) {
			// The smoother swapped u and tmp between z and the level's tmp buffer.
			L0.u = own_u;
			L0.tmp = own_tmp;
			L0.rhs = own_rhs;
		}
		FX_CHECK(s);
		return fx::check_launch("mg vcycle");
	}
};

void free_levels(fx_multigrid* mg) {
	for (auto& L : mg->lv) {
		fx::dev_free(L.u); fx::dev_free(L.rhs); fx::dev_free(L.tmp); fx::dev_free(L.w0); fx::dev_free(L.w1);
		fx::dev_free(L.k);
		fx::dev_free(L.vx);
		fx::dev_free(L.vy);
		fx::dev_free(L.kx);
		fx::dev_free(L.ky);
		fx::dev_free(L.wd);
		fx::dev_free(L.kp);
		fx::dev_free(L.xbuf);
	}
	mg->lv.clear();
}

// Tabulates the linear transfer from level grid f to the next coarser grid c
// (fine index i → coarse coordinate i·(nc−1)/(nf−1) = i0 + w, split exactly in
// integers; coarse index I gathers fine interior indices lo..hi) and the LDS
// region of the fused restriction's widest tile.
fx_status build_xtab(fx_backend* be, fx_multigrid::Level& L, const fx_grid& c) {
	const int64_t fr = (int64_t)L.g.rows, fc = (int64_t)L.g.cols, cr = (int64_t)c.rows, cc = (int64_t)c.cols;
	FX_REQUIRE(fr < (int64_t)1 << 30 && fc < (int64_t)1 << 30, FX_ERR_GRID, "grid too large for transfer tables");
	const size_t nd = (size_t)(fr + fc), ni = (size_t)(fr + fc + 2 * cr + 2 * cc);
	std::vector<double> dv(nd);
	std::vector<int32_t> iv(ni);
	double* wr = dv.data();
	double* wc = wr + fr;
	int32_t* i0r = iv.data();
	int32_t* i0c = i0r + fr;
	int32_t* lor = i0c + fc;
	int32_t* hir = lor + cr;
	int32_t* loc = hir + cr;
	int32_t* hic = loc + cc;
	auto axis = [](int64_t nf, int64_t nc, int32_t* i0, double* w, int32_t* lo, int32_t* hi) {
		const int64_t den = nf - 1, dc = nc - 1;
		for (int64_t i = 0; i < nf; ++i) {
			const int64_t num = i * (nc - 1);
			i0[i] = (int32_t)(num / den);
			w[i] = double(num % den) * (1.0 / double(den));
		}
		for (int64_t I = 0; I < nc; ++I) {
			const int64_t a = (I - 1) * (nf - 1);
			int64_t l = a <= 0 ? 1 : (a + dc - 1) / dc;
			int64_t h = ((I + 1) * (nf - 1)) / dc;
			if (l < 1) l = 1;
			if (h > nf - 2) h = nf - 2;
			lo[I] = (int32_t)l;
			hi[I] = (int32_t)h;
		}
	};
	axis(fr, cr, i0r, wr, lor, hir);
	axis(fc, cc, i0c, wc, loc, hic);
	auto span = [](int64_t nc, int tile, const int32_t* lo, const int32_t* hi) {
		int64_t m = 1;
		for (int64_t I0 = 0; I0 < nc; I0 += tile) {
			const int64_t a = I0 > 1 ? I0 : 1, b = I0 + tile - 1 < nc - 2 ? I0 + tile - 1 : nc - 2;
			if (a <= b && hi[b] >= lo[a] && hi[b] - lo[a] + 1 > m) m = hi[b] - lo[a] + 1;
		}
		return m;
	};
	void* p = nullptr;
	FX_CHECK(fx::dev_alloc(be, nd * sizeof(double) + ni * sizeof(int32_t), &p));
	L.xbuf = p;
	FX_HIP(hipMemcpyAsync(p, dv.data(), nd * sizeof(double), hipMemcpyHostToDevice, be->stream));
	FX_HIP(hipMemcpyAsync((char*)p + nd * sizeof(double), iv.data(), ni * sizeof(int32_t), hipMemcpyHostToDevice,
	                      be->stream));
	FX_HIP(hipStreamSynchronize(be->stream));
	const double* dw = (const double*)p;
	const int32_t* di = (const int32_t*)((char*)p + nd * sizeof(double));
	fx_xtab& X = L.xt;
	X.wr = dw;
	X.wc = dw + fr;
	X.i0r = di;
	X.i0c = di + fr;
	X.lor = X.i0c + fc;
	X.hir = X.lor + cr;
	X.loc = X.hir + cr;
	X.hic = X.loc + cc;
	X.sy = float(double(cr - 1) / double(fr - 1));
	X.sx = float(double(cc - 1) / double(fc - 1));
	X.ry = (int)span(cr, fmg::RCY, lor, hir) + 4;
	X.rx = (int)span(cc, fmg::RCX, loc, hic) + 4;
	X.gmax = 0;
	for (int64_t J = 1; J + 1 < cc; ++J) X.gmax = std::max<int>(X.gmax, hic[J] - loc[J] + 1);
	X.gmaxr = 0;
	for (int64_t I = 1; I + 1 < cr; ++I) X.gmaxr = std::max<int>(X.gmaxr, hir[I] - lor[I] + 1);
	// Fine window (two columns of halo each side) of the wave restriction's strips of RCX interior
	// coarse columns.
	int64_t wspan = 1;
	for (int64_t Ja = 1; Ja + 1 < cc; Ja += fmg::RCX) {
		const int64_t Jb = std::min<int64_t>(Ja + fmg::RCX - 1, cc - 2);
		if (hic[Jb] >= loc[Ja]) wspan = std::max<int64_t>(wspan, hic[Jb] - loc[Ja] + 1);
	}
	X.wx = (int)wspan + 4;
	X.adv = 0;
	for (int64_t Ja = 1; Ja + fmg::RCX + 1 < cc; Ja += fmg::RCX)
		X.adv = std::max<int>(X.adv, loc[Ja + fmg::RCX] - loc[Ja]);
	return FX_OK;
}

__global__ __launch_bounds__(BLOCK) void k_to_f32(const double* r, float* b, size_t n, const int* skip) {
	SKIP_RETURN;
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) b[i] = float(r[i]);
}

}  // namespace

namespace fx {
fx_status mg_apply(fx_multigrid* mg, double* z, const double* r, const int* skip) {
	if (mg->opt.fused) {
		// Fused cycle: the initial guess is taken as zero.
		size_t n0 = (size_t)mg->lv[0].g.rows * mg->lv[0].g.cols;
		hipStream_t st = mg->be->stream;
		hipLaunchKernelGGL(k_to_f32, dim3(fx::elem_blocks(n0)), dim3(BLOCK), 0, st, r, (float*)mg->lv[0].rhs, n0, skip);
		FX_CHECK(fmg::fast_cycle(mg, skip, fmg::NoEpi{}));
		hipLaunchKernelGGL(k_cvt_out<float>, dim3(fx::elem_blocks(n0)), dim3(BLOCK), 0, st, (const float*)mg->lv[0].tmp, z,
		                   n0, skip);
		return fx::check_launch("mg fused");
	}
	if (mg->f32) return Driver<float>{mg, skip, mg->be->stream}.run(z, r);
	return Driver<double>{mg, skip, mg->be->stream}.run(z, r);
}
}  // namespace fx

using namespace fx;

extern "C" {

fx_status fx_multigrid_create_ex(fx_backend* be, const fx_grid* grid, const fx_boundaries* bc, const fx_mg_options* opt,
                                 fx_multigrid** out) {
	FX_REQUIRE(be && out && opt, FX_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	FX_CHECK(check_grid(grid));
	FX_CHECK(check_bc(bc));
	FX_REQUIRE(opt->precision == FX_F32 || opt->precision == FX_F64, FX_ERR_ARGUMENT, "bad precision");
	FX_REQUIRE(opt->smoother == FX_SMOOTH_JACOBI || opt->smoother == FX_SMOOTH_CHEBYSHEV, FX_ERR_ARGUMENT, "bad smoother");
	FX_REQUIRE(opt->transfer == FX_TRANSFER_REFERENCE || opt->transfer == FX_TRANSFER_FW_BILINEAR, FX_ERR_ARGUMENT,
	           "bad transfer");
	if (opt->k) {
		FX_CHECK(check_buf(opt->k, grid->rows * grid->cols, FX_F64, "k"));
		FX_REQUIRE(opt->smoother == FX_SMOOTH_JACOBI, FX_ERR_UNSUPPORTED,
		           "variable-coefficient levels use the Jacobi smoother");
	}
	FX_REQUIRE(!opt->vx == !opt->vy, FX_ERR_ARGUMENT, "give both velocity components or neither");
	if (opt->vx) {
		FX_CHECK(check_buf(opt->vx, grid->rows * grid->cols, FX_F64, "vx"));
		FX_CHECK(check_buf(opt->vy, grid->rows * grid->cols, FX_F64, "vy"));
		FX_REQUIRE(!opt->k && !opt->fused && opt->smoother == FX_SMOOTH_JACOBI, FX_ERR_UNSUPPORTED,
		           "convection-diffusion levels use the unfused Jacobi cycle without a coefficient field");
		FX_REQUIRE(std::isfinite(opt->kappa), FX_ERR_ARGUMENT, "kappa must be finite");
	}
	fx_multigrid* mg = new fx_multigrid();
	mg->be = be;
	mg->opt = *opt;
	mg->bc = *bc;
	mg->f32 = opt->precision == FX_F32;
	const size_t es = mg->f32 ? 4 : 8;
	const uint64_t lim = opt->coarsest > 0 ? (uint64_t)opt->coarsest : 4;
	fx_grid g = *grid;
	for (;;) {
		fx_multigrid::Level L;
		L.g = g;
		double ix2 = 1.0 / (g.dx * g.dx), iy2 = 1.0 / (g.dy * g.dy);
		L.w[0] = opt->sigma + opt->scale * (-2.0 * (ix2 + iy2));
		L.w[1] = opt->scale * iy2;
		L.w[2] = opt->scale * iy2;
		L.w[3] = opt->scale * ix2;
		L.w[4] = opt->scale * ix2;
		L.inv_diag = 1.0 / L.w[0];
		size_t bytes = (size_t)g.rows * g.cols * es;
		fx_status s = FX_OK;
		void** bufs[5] = {&L.u, &L.rhs, &L.tmp, &L.w0, &L.w1};
		int nb = opt->smoother == FX_SMOOTH_CHEBYSHEV ? 5 : 3;
		for (int i = 0; i < nb && s == FX_OK; ++i) s = dev_alloc(be, bytes, bufs[i]);
		if (s == FX_OK && opt->k) {
			s = dev_alloc(be, bytes, &L.k);
			if (s == FX_OK) {
				size_t n = (size_t)g.rows * g.cols;
				if (mg->lv.empty()) {
					if (mg->f32) hipLaunchKernelGGL(k_cvt_k<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
					                                be->stream, (const double*)opt->k->ptr, (float*)L.k, n);
					else hipLaunchKernelGGL(k_cvt_k<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
					                        be->stream, (const double*)opt->k->ptr, (double*)L.k, n);
				} else {
					const fx_multigrid::Level& F = mg->lv.back();
					int lin = opt->transfer == FX_TRANSFER_FW_BILINEAR;
					if (mg->f32) hipLaunchKernelGGL(k_coarsen_k<float>, pgrid(g.rows, g.cols), dim3(BLOCK), 0, be->stream,
					                                (const float*)F.k, (float*)L.k, (int64_t)F.g.rows, (int64_t)F.g.cols,
					                                (int64_t)g.rows, (int64_t)g.cols, lin);
					else hipLaunchKernelGGL(k_coarsen_k<double>, pgrid(g.rows, g.cols), dim3(BLOCK), 0, be->stream,
					                        (const double*)F.k, (double*)L.k, (int64_t)F.g.rows, (int64_t)F.g.cols,
					                        (int64_t)g.rows, (int64_t)g.cols, lin);
				}
				s = check_launch("multigrid coefficient");
			}
		}
		if (s == FX_OK && opt->vx) {
			const size_t n = (size_t)g.rows * g.cols;
			void** vb[2] = {&L.vx, &L.vy};
			const fx_buf* vsrc[2] = {opt->vx, opt->vy};
			for (int q = 0; q < 2 && s == FX_OK; ++q) {
				s = dev_alloc(be, bytes, vb[q]);
				if (s != FX_OK) break;
				if (mg->lv.empty()) {
					if (mg->f32) hipLaunchKernelGGL(k_cvt_k<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
					                                be->stream, (const double*)vsrc[q]->ptr, (float*)*vb[q], n);
					else hipLaunchKernelGGL(k_cvt_k<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
					                        be->stream, (const double*)vsrc[q]->ptr, (double*)*vb[q], n);
				} else {
					const fx_multigrid::Level& F = mg->lv.back();
					const void* fv = q == 0 ? F.vx : F.vy;
					int lin = opt->transfer == FX_TRANSFER_FW_BILINEAR;
					if (mg->f32) hipLaunchKernelGGL(k_coarsen_k<float>, pgrid(g.rows, g.cols), dim3(BLOCK), 0, be->stream,
					                                (const float*)fv, (float*)*vb[q], (int64_t)F.g.rows, (int64_t)F.g.cols,
					                                (int64_t)g.rows, (int64_t)g.cols, lin);
					else hipLaunchKernelGGL(k_coarsen_k<double>, pgrid(g.rows, g.cols), dim3(BLOCK), 0, be->stream,
					                        (const double*)fv, (double*)*vb[q], (int64_t)F.g.rows, (int64_t)F.g.cols,
					                        (int64_t)g.rows, (int64_t)g.cols, lin);
				}
				s = check_launch("multigrid velocity");
			}
		}
		mg->lv.push_back(L);
		if (s != FX_OK) { free_levels(mg); delete mg; return s; }
		if (g.rows <= lim || g.cols <= lim) break;
		if (opt->transfer == FX_TRANSFER_REFERENCE) {
			// multigrid.rs:121-135.
			uint64_t cr = g.rows / 2, cc = g.cols / 2;
			if (cr < 3 || cc < 3) break;
			g = fx_grid{cr, cc, g.dx * 2.0, g.dy * 2.0};
		} else {
			// Every level spans the fine domain: odd sizes nest ((n+1)/2), even sizes halve.
			uint64_t cr = g.rows % 2 ? (g.rows + 1) / 2 : g.rows / 2;
			uint64_t cc = g.cols % 2 ? (g.cols + 1) / 2 : g.cols / 2;
			if (cr < 3 || cc < 3) break;
			g = fx_grid{cr, cc, g.dx * double(g.cols - 1) / double(cc - 1), g.dy * double(g.rows - 1) / double(cr - 1)};
		}
	}
	if (opt->transfer == FX_TRANSFER_FW_BILINEAR) {
		for (size_t l = 0; l + 1 < mg->lv.size(); ++l) {
			fx_status s = build_xtab(be, mg->lv[l], mg->lv[l + 1].g);
			if (s != FX_OK) { free_levels(mg); delete mg; return s; }
		}
	}
	if (opt->fused && opt->k && mg->f32) {
		for (auto& L : mg->lv) {
			const size_t bytes = (size_t)L.g.rows * L.g.cols * sizeof(float);
			fx_status s = dev_alloc(be, bytes, &L.kx);
			if (s == FX_OK) s = dev_alloc(be, bytes, &L.ky);
			if (s == FX_OK) s = dev_alloc(be, bytes, &L.wd);
			if (s == FX_OK) {
				hipLaunchKernelGGL(k_faces, pgrid(L.g.rows, L.g.cols), dim3(BLOCK), 0, be->stream, (const float*)L.k,
				                   (float*)L.kx, (float*)L.ky, (int64_t)L.g.rows, (int64_t)L.g.cols);
				s = check_launch("multigrid face coefficients");
			}
			if (s == FX_OK) {
				hipLaunchKernelGGL(k_wdiag, pgrid(L.g.rows, L.g.cols), dim3(BLOCK), 0, be->stream, (const float*)L.kx,
				                   (const float*)L.ky, (float*)L.wd, fmg::lwf(mg, L), (int64_t)L.g.rows,
				                   (int64_t)L.g.cols);
				s = check_launch("multigrid omega/diagonal");
			}
			if (s == FX_OK) s = dev_alloc(be, 3 * bytes, &L.kp);
			if (s == FX_OK) {
				const int64_t n = (int64_t)L.g.rows * (int64_t)L.g.cols;
				hipLaunchKernelGGL(k_pack3, dim3((unsigned)std::min<int64_t>((n + BLOCK - 1) / BLOCK, 65536)),
				                   dim3(BLOCK), 0, be->stream, (const float*)L.kx, (const float*)L.ky,
				                   (const float*)L.wd, (float*)L.kp, n);
				s = check_launch("multigrid packed coefficients");
			}
			if (s != FX_OK) { free_levels(mg); delete mg; return s; }
		}
	}
	if (opt->fused && !fmg::fused_ok(mg)) {
		free_levels(mg);
		delete mg;
		set_error("the fused V-cycle needs FP32 levels, Jacobi smoothing, linear transfers, 1-2 pre/post sweeps "
		          "and a coarsest level of at most 4096 points");
		return FX_ERR_UNSUPPORTED;
	}
	*out = mg;
	return FX_OK;
}

fx_status fx_multigrid_create(fx_backend* be, const fx_grid* grid, const fx_boundaries* bc, uint32_t n_pre,
                              uint32_t n_post, double omega, int32_t smoother, fx_multigrid** out) {
	fx_mg_options o{};
	o.n_pre = (int32_t)n_pre;
	o.n_post = (int32_t)n_post;
	o.omega = omega;
	o.smoother = smoother;
	o.transfer = FX_TRANSFER_REFERENCE;
	o.coarsest = 4;
	o.extra_coarse = 10;
	o.precision = FX_F64;
	o.fused = 0;
	o.sigma = 0.0;
	o.scale = 1.0;
	o.k = nullptr;
	return fx_multigrid_create_ex(be, grid, bc, &o, out);
}

void fx_multigrid_destroy(fx_multigrid* mg) {
	if (!mg) return;
	if (mg->be) (void)hipStreamSynchronize(mg->be->stream);
	free_levels(mg);
	delete mg;
}

uint32_t fx_multigrid_depth(const fx_multigrid* mg) { return mg ? (uint32_t)mg->lv.size() : 0; }

fx_status fx_multigrid_vcycle(fx_multigrid* mg, fx_buf* x, const fx_buf* b) {
	FX_REQUIRE(mg, FX_ERR_ARGUMENT, "multigrid is null");
	size_t n = mg->lv[0].g.rows * mg->lv[0].g.cols;
	FX_CHECK(check_buf(x, n, FX_F64, "x"));
	FX_CHECK(check_buf(b, n, FX_F64, "b"));
	return mg_apply(mg, (double*)x->ptr, (const double*)b->ptr, nullptr);
}

fx_status fx_multigrid_set_kernels(fx_multigrid* mg, int32_t kernels) {
	FX_REQUIRE(mg, FX_ERR_ARGUMENT, "multigrid is null");
	FX_REQUIRE(kernels == FX_MGK_TILED || kernels == FX_MGK_RING || kernels == FX_MGK_WAVE, FX_ERR_ARGUMENT,
	           "unknown level-kernel variant %d", kernels);
	mg->kernels = kernels;
	return FX_OK;
}

fx_status fx_multigrid_autotune(fx_multigrid* mg, int32_t reps, double* ms_out) {
	FX_REQUIRE(mg, FX_ERR_ARGUMENT, "multigrid is null");
	FX_REQUIRE(reps >= 1 && reps <= 1000, FX_ERR_ARGUMENT, "reps must lie in [1, 1000]");
	FX_REQUIRE(fmg::fused_ok(mg), FX_ERR_UNSUPPORTED, "autotuning needs a fused FP32 hierarchy");
	return fmg::autotune(mg, reps, ms_out);
}

fx_status fx_multigrid_set_wave_rows(fx_multigrid* mg, uint32_t level, int32_t post_rows, int32_t restrict_rows) {
	FX_REQUIRE(mg, FX_ERR_ARGUMENT, "multigrid is null");
	FX_REQUIRE(level + 1 < mg->lv.size(), FX_ERR_ARGUMENT, "level %u has no coarser level (depth %zu)", level,
	           mg->lv.size());
	const int lo = level == 0 ? 0 : -1;
	FX_REQUIRE(post_rows >= lo && post_rows <= 4096 && restrict_rows >= lo && restrict_rows <= 4096,
	           FX_ERR_ARGUMENT, "segment heights must lie in [%d, 4096] on level %u", lo, level);
	mg->wrows_post.resize(mg->lv.size(), 0);
	mg->wrows_restrict.resize(mg->lv.size(), 0);
	mg->wrows_post[level] = post_rows;
	mg->wrows_restrict[level] = restrict_rows;
	if (level == 0) mg->wrows_restrict_upd = restrict_rows, mg->wrows_post_epi = post_rows;
	return FX_OK;
}

}  // extern "C"
