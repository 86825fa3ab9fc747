// oracle.cpp — CPU FP64 restatement of fluxion's matrix-free solve path.
//
// TEST INFRASTRUCTURE ONLY. This file is the parity checker for the HIP
// library in fluxion_amd/ and the timed CPU baseline ("port") in bench.py.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it. Nothing in the product path calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the fluxion workspace root).  Build flags matter for parity: g++ -O3
// -ffp-contract=off -fopenmp, so a fused multiply-add happens exactly where
// the reference writes `mul_add` (std::fma here) and nowhere else.
//
// Parallel structure mirrors backend-cpu: rows in parallel for stencils
// (crates/backend-cpu/src/ops.rs:114-124, 256-266), 1024-element chunks for
// BLAS-1 (ops.rs:198-205) and 4096-element Kahan chunks for reductions
// (ops.rs:272-367).  Chunk partials are merged with kahan_merge
// (ops.rs:1129-1137) in chunk order; rayon's reduce tree shape is not fixed,
// so this left fold is one of the orders the reference may take.
//
// Extensions with no reference counterpart (two-buffer Jacobi, full-weighting
// and bilinear transfer, FP32 V-cycle, conv-diff and variable-k operator modes,
// flexible CG, preconditioned pipelined CG) are marked "EXTENSION".

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>

namespace {

constexpr int OK = 0;
constexpr int ERR_PERIODIC = 1;   // ops.rs:11-23 reject_periodic
constexpr int ERR_DIM = 2;        // CoreError::DimensionMismatch
constexpr int ERR_GRID = 3;       // grid.rs:15-36
constexpr int ERR_STENCIL = 4;    // stencil.rs:23-30 (empty) / radius > 1
constexpr int ERR_ARG = 5;

enum { BC_DIRICHLET = 0, BC_NEUMANN = 1, BC_ROBIN = 2, BC_PERIODIC = 3 };

struct Entry { int dr, dc; double w; };
struct Bc { int kind; double a, b, g; };
struct Bcs { Bc top, bottom, left, right; };

struct Grid {
	int64_t rows, cols;
	double dx, dy;
	int64_t len() const { return rows * cols; }
};

// ── bc_value — ops.rs:1141-1180 ───────────────────────────────────────
double bc_value(const Bc& bc, double self_val, double nb, double spacing) {
	switch (bc.kind) {
	case BC_DIRICHLET: return bc.a;
	case BC_NEUMANN: return std::fma(spacing, bc.a, nb);
	case BC_ROBIN: {
		double alpha = bc.a, beta = bc.b, g = bc.g;
		if (std::fabs(beta) < 1e-30) {
			if (std::fabs(alpha) < 1e-30) return 0.0;
			return g / alpha;
		}
		double du_dn = (g - alpha * self_val) / beta;
		return std::fma(spacing, du_dn, nb);
	}
	default: return 0.0;
	}
}

bool has_periodic(const Bcs& b) {
	return b.top.kind == BC_PERIODIC || b.bottom.kind == BC_PERIODIC ||
	       b.left.kind == BC_PERIODIC || b.right.kind == BC_PERIODIC;
}

// ── apply_boundaries — ops.rs:799-843 ─────────────────────────────────
void apply_boundaries(const double* src, double* dst, int64_t rows, int64_t cols,
                      double dx, double dy, const Bcs& bc) {
	int64_t lr = rows - 1, lc = cols - 1;
	for (int64_t c = 0; c < cols; ++c) {
		dst[c] = bc_value(bc.top, src[c], src[cols + c], dy);
		dst[lr * cols + c] = bc_value(bc.bottom, src[lr * cols + c], src[(lr - 1) * cols + c], dy);
	}
	for (int64_t r = 1; r < lr; ++r) {
		int64_t i = r * cols;
		dst[i] = bc_value(bc.left, src[i], src[i + 1], dx);
		dst[i + lc] = bc_value(bc.right, src[i + lc], src[i + lc - 1], dx);
	}
}

// Weights by offset, last matching entry wins (ops.rs:908-920 match loop).
struct W9 { double c = 0, n = 0, s = 0, w = 0, e = 0, nw = 0, ne = 0, sw = 0, se = 0; };
W9 extract(const Entry* es, int n, bool diag) {
	W9 k;
	for (int i = 0; i < n; ++i) {
		const Entry& e = es[i];
		if (e.dr == 0 && e.dc == 0) k.c = e.w;
		else if (e.dr == -1 && e.dc == 0) k.n = e.w;
		else if (e.dr == 1 && e.dc == 0) k.s = e.w;
		else if (e.dr == 0 && e.dc == -1) k.w = e.w;
		else if (e.dr == 0 && e.dc == 1) k.e = e.w;
		else if (diag && e.dr == -1 && e.dc == -1) k.nw = e.w;
		else if (diag && e.dr == -1 && e.dc == 1) k.ne = e.w;
		else if (diag && e.dr == 1 && e.dc == -1) k.sw = e.w;
		else if (diag && e.dr == 1 && e.dc == 1) k.se = e.w;
	}
	return k;
}

inline double chain5(const W9& k, const double* s, int64_t i, int64_t cols) {
	// ops.rs:924-935 / 1093-1106: c, n, s, w, e*.
	return std::fma(k.c, s[i], std::fma(k.n, s[i - cols], std::fma(k.s, s[i + cols],
	       std::fma(k.w, s[i - 1], k.e * s[i + 1]))));
}
inline double chain9(const W9& k, const double* s, int64_t i, int64_t cols) {
	// ops.rs:984-1007 / 1058-1083: c, n, s, w, e, nw, ne, sw, se*.
	return std::fma(k.c, s[i], std::fma(k.n, s[i - cols], std::fma(k.s, s[i + cols],
	       std::fma(k.w, s[i - 1], std::fma(k.e, s[i + 1], std::fma(k.nw, s[i - cols - 1],
	       std::fma(k.ne, s[i - cols + 1], std::fma(k.sw, s[i + cols - 1], k.se * s[i + cols + 1]))))))));
}
inline double chain_generic(const Entry* es, int n, const double* s, int64_t r, int64_t c, int64_t cols) {
	// ops.rs:872-883: acc = w.mul_add(x, acc) in entry order.
	double acc = 0.0;
	for (int j = 0; j < n; ++j)
		acc = std::fma(es[j].w, s[(r + es[j].dr) * cols + (c + es[j].dc)], acc);
	return acc;
}

int check_stencil(const Entry* es, int n) {
	if (n <= 0) return ERR_STENCIL;
	for (int i = 0; i < n; ++i)
		if (es[i].dr < -1 || es[i].dr > 1 || es[i].dc < -1 || es[i].dc > 1) return ERR_STENCIL;
	return OK;
}

// ── apply_stencil — ops.rs:77-127, apply_interior_row 850-885 ─────────
int apply_stencil(const double* in, double* out, const Grid& g, const Entry* es, int n, const Bcs& bc) {
	if (has_periodic(bc)) return ERR_PERIODIC;
	if (int e = check_stencil(es, n)) return e;
	int64_t rows = g.rows, cols = g.cols;
	apply_boundaries(in, out, rows, cols, g.dx, g.dy, bc);
	W9 k = extract(es, n, n == 9);
#pragma omp parallel for schedule(static)
	for (int64_t r = 1; r < rows - 1; ++r) {
		for (int64_t c = 1; c < cols - 1; ++c) {
			int64_t i = r * cols + c;
			double v;
			if (n == 5) v = chain5(k, in, i, cols);
			else if (n == 9) v = chain9(k, in, i, cols);
			else v = chain_generic(es, n, in, r, c, cols);
			out[i] = v;
		}
	}
	return OK;
}

// ── stencil_axpy — ops.rs:226-269, fused_5pt_axpy_row 1016-1125 ───────
int stencil_axpy(double alpha, const double* x, double* y, const Grid& g, const Entry* es, int n, const Bcs& bc) {
	if (has_periodic(bc)) return ERR_PERIODIC;
	if (int e = check_stencil(es, n)) return e;
	int64_t rows = g.rows, cols = g.cols;
	W9 k = extract(es, n, n == 9);
#pragma omp parallel for schedule(static)
	for (int64_t r = 1; r < rows - 1; ++r) {
		for (int64_t c = 1; c < cols - 1; ++c) {
			int64_t i = r * cols + c;
			double lap;
			if (n == 9) lap = chain9(k, x, i, cols);
			else if (n == 5) lap = chain5(k, x, i, cols);
			else lap = chain_generic(es, n, x, r, c, cols);
			y[i] = std::fma(alpha, lap, y[i]);
		}
	}
	return OK;
}

// ── BLAS-1 — ops.rs:157-223, 476-672 ──────────────────────────────────
constexpr int64_t CH = 1024;
void axpy(double alpha, const double* x, double* y, int64_t n) {
#pragma omp parallel for schedule(static)
	for (int64_t b = 0; b < n; b += CH)
		for (int64_t i = b; i < std::min(n, b + CH); ++i) y[i] = std::fma(alpha, x[i], y[i]);
}
void scale(double* x, double alpha, int64_t n) {
#pragma omp parallel for schedule(static)
	for (int64_t b = 0; b < n; b += CH)
		for (int64_t i = b; i < std::min(n, b + CH); ++i) x[i] *= alpha;
}
void copy(const double* s, double* d, int64_t n) { std::memcpy(d, s, sizeof(double) * n); }
void fill(double* d, double v, int64_t n) { std::fill(d, d + n, v); }

// ── Reductions — ops.rs:272-367, kahan_merge 1129-1137 ────────────────
constexpr int64_t RCH = 4096;
struct KS { double s, c; };
inline KS kahan_merge(KS a, KS b) {
	double y = b.s - (a.c + b.c);
	double t = a.s + y;
	double c = (t - a.s) - y;
	return {t, c};
}
template <class F>
double kahan_reduce(int64_t n, F term_chunk) {
	int64_t nch = (n + RCH - 1) / RCH;
	std::vector<KS> part(nch);
#pragma omp parallel for schedule(static)
	for (int64_t ch = 0; ch < nch; ++ch) part[ch] = term_chunk(ch * RCH, std::min(n, (ch + 1) * RCH));
	KS acc{0.0, 0.0};
	for (int64_t ch = 0; ch < nch; ++ch) acc = kahan_merge(acc, part[ch]);
	return acc.s;
}
double dot(const double* x, const double* y, int64_t n) {
	// ops.rs:334-367: prod = x*y (no fma), Kahan on prod.
	return kahan_reduce(n, [&](int64_t b, int64_t e) {
		double acc = 0.0, comp = 0.0;
		for (int64_t i = b; i < e; ++i) {
			double prod = x[i] * y[i];
			double ty = prod - comp;
			double t = acc + ty;
			comp = (t - acc) - ty;
			acc = t;
		}
		return KS{acc, comp};
	});
}
double norm_l2(const double* x, int64_t n) {
	// ops.rs:272-292: y = v.mul_add(v, -comp).
	double s = kahan_reduce(n, [&](int64_t b, int64_t e) {
		double acc = 0.0, comp = 0.0;
		for (int64_t i = b; i < e; ++i) {
			double y = std::fma(x[i], x[i], -comp);
			double t = acc + y;
			comp = (t - acc) - y;
			acc = t;
		}
		return KS{acc, comp};
	});
	return std::sqrt(s);
}
double reduce_sum(const double* x, int64_t n) {
	// ops.rs:295-316.
	return kahan_reduce(n, [&](int64_t b, int64_t e) {
		double acc = 0.0, comp = 0.0;
		for (int64_t i = b; i < e; ++i) {
			double y = x[i] - comp;
			double t = acc + y;
			comp = (t - acc) - y;
			acc = t;
		}
		return KS{acc, comp};
	});
}

// ── Multigrid transfer / smoothing — ops.rs:369-473 ───────────────────
template <class T>
void restrict_ref(const T* f, T* c, const Grid& fg, const Grid& cg) {
	// ops.rs:369-399: 2x2 average, clamped second index.
	int64_t fc = fg.cols, cc = cg.cols, fr_max = fg.rows - 1, fc_max = fc - 1;
#pragma omp parallel for schedule(static)
	for (int64_t r = 0; r < cg.rows; ++r)
		for (int64_t col = 0; col < cc; ++col) {
			int64_t fr = r * 2, fcc = col * 2;
			int64_t fr1 = std::min(fr + 1, fr_max), fc1 = std::min(fcc + 1, fc_max);
			c[r * cc + col] = T(0.25) * (f[fr * fc + fcc] + f[fr1 * fc + fcc] + f[fr * fc + fc1] + f[fr1 * fc + fc1]);
		}
}
template <class T>
void prolong_ref(const T* c, T* f, const Grid& cg, const Grid& fg) {
	// ops.rs:401-423: fine[r][col] += coarse[r/2][col/2].  The reference
	// indexes past the coarse array when the fine size is odd (7 -> 3);
	// the coarse index is clamped here, which is identical for even sizes.
	int64_t fc = fg.cols, cc = cg.cols;
#pragma omp parallel for schedule(static)
	for (int64_t r = 0; r < fg.rows; ++r) {
		int64_t cr = std::min(r / 2, cg.rows - 1);
		for (int64_t col = 0; col < fc; ++col) f[r * fc + col] += c[cr * cc + std::min(col / 2, cc - 1)];
	}
}

// EXTENSION: linear transfer between level grids that span the same domain.
// Level l has n_l nodes at spacing L/(n_l − 1); coarse sizes are (n+1)/2 for odd
// n (nested: the transfers are textbook full weighting and bilinear
// interpolation) and n/2 for even n (non-nested).  Prolongation interpolates
// linearly in each direction; restriction is its transpose scaled by
// (h/H)² per level so that it averages.
// Fine index i sits at coarse coordinate i·(nc−1)/(nf−1) = I0 + w.
struct Lin1 { int64_t i0; double w; };
inline Lin1 lin1(int64_t i, int64_t nf, int64_t nc) {
	int64_t num = i * (nc - 1);
	int64_t den = nf - 1;
	return {num / den, double(num % den) * (1.0 / double(den))};
}
template <class T>
void prolong_lin(const T* c, T* f, const Grid& cg, const Grid& fg) {
	int64_t fc = fg.cols, cc = cg.cols;
#pragma omp parallel for schedule(static)
	for (int64_t r = 1; r < fg.rows - 1; ++r) {
		Lin1 ly = lin1(r, fg.rows, cg.rows);
		T wy = T(ly.w), vy = T(1) - wy;
		int64_t a0 = ly.i0, a1 = ly.i0 + 1 < cg.rows ? ly.i0 + 1 : ly.i0;
		for (int64_t col = 1; col < fc - 1; ++col) {
			Lin1 lx = lin1(col, fg.cols, cg.cols);
			T wx = T(lx.w), vx = T(1) - wx;
			int64_t b0 = lx.i0, b1 = lx.i0 + 1 < cc ? lx.i0 + 1 : lx.i0;
			T top = vx * c[a0 * cc + b0] + wx * c[a0 * cc + b1];
			T bot = vx * c[a1 * cc + b0] + wx * c[a1 * cc + b1];
			f[r * fc + col] += vy * top + wy * bot;
		}
	}
}
// Weight of coarse node I in the interpolation of fine node i.
inline double lin_weight(int64_t i, int64_t I, int64_t nf, int64_t nc) {
	Lin1 l = lin1(i, nf, nc);
	if (l.i0 == I) return 1.0 - l.w;
	if (l.i0 + 1 == I) return l.w;
	return 0.0;
}
inline void lin_range(int64_t I, int64_t nf, int64_t nc, int64_t* lo, int64_t* hi) {
	int64_t den = nc - 1;
	int64_t a = (I - 1) * (nf - 1);
	*lo = a <= 0 ? 1 : (a + den - 1) / den;
	*hi = ((I + 1) * (nf - 1)) / den;
	if (*lo < 1) *lo = 1;
	if (*hi > nf - 2) *hi = nf - 2;
}
template <class T>
void restrict_lin(const T* f, T* c, const Grid& fg, const Grid& cg) {
	int64_t fc = fg.cols, cc = cg.cols;
	T sy = T(double(cg.rows - 1) / double(fg.rows - 1)), sx = T(double(cg.cols - 1) / double(fg.cols - 1));
#pragma omp parallel for schedule(static)
	for (int64_t R = 0; R < cg.rows; ++R)
		for (int64_t Cc = 0; Cc < cc; ++Cc) {
			if (R == 0 || Cc == 0 || R == cg.rows - 1 || Cc == cc - 1) { c[R * cc + Cc] = T(0); continue; }
			int64_t rlo, rhi, clo, chi;
			lin_range(R, fg.rows, cg.rows, &rlo, &rhi);
			lin_range(Cc, fg.cols, cg.cols, &clo, &chi);
			T acc = T(0);
			for (int64_t r = rlo; r <= rhi; ++r) {
				T wy = T(lin_weight(r, R, fg.rows, cg.rows));
				T row = T(0);
				for (int64_t col = clo; col <= chi; ++col) row += T(lin_weight(col, Cc, fg.cols, cg.cols)) * f[r * fc + col];
				acc += wy * row;
			}
			c[R * cc + Cc] = acc * (sy * sx);
		}
}

template <class T> inline T fmaT(T a, T b, T c);
template <> inline double fmaT<double>(double a, double b, double c) { return std::fma(a, b, c); }
template <> inline float fmaT<float>(float a, float b, float c) { return std::fmaf(a, b, c); }

// Level stencil entries in laplacian_2d_5pt order (stencil.rs:56-72):
// (0,0), (-1,0), (1,0), (0,-1), (0,1).
template <class T> struct Lvl5 { T w[5]; T inv_diag; };

// ops.rs:425-473: in-place lexicographic sweep (damped Gauss-Seidel).
template <class T>
void smooth_ref_gs(T* x, const T* b, T omega, const Grid& g, const Lvl5<T>& L) {
	int64_t rows = g.rows, cols = g.cols;
	const int dr[5] = {0, -1, 1, 0, 0}, dc[5] = {0, 0, 0, -1, 1};
	for (int64_t r = 1; r < rows - 1; ++r)
		for (int64_t c = 1; c < cols - 1; ++c) {
			int64_t i = r * cols + c;
			T ax = T(0);
			for (int e = 0; e < 5; ++e) ax = fmaT<T>(L.w[e], x[(r + dr[e]) * cols + (c + dc[e])], ax);
			T res = b[i] - ax;
			x[i] = fmaT<T>(omega, L.inv_diag * res, x[i]);
		}
}
// EXTENSION: two-buffer weighted Jacobi, same per-point arithmetic as above
// but reading only the previous iterate.
template <class T>
void smooth_jacobi(const T* xo, T* xn, const T* b, T omega, const Grid& g, const Lvl5<T>& L) {
	int64_t rows = g.rows, cols = g.cols;
	const int dr[5] = {0, -1, 1, 0, 0}, dc[5] = {0, 0, 0, -1, 1};
#pragma omp parallel for schedule(static)
	for (int64_t r = 0; r < rows; ++r)
		for (int64_t c = 0; c < cols; ++c) {
			int64_t i = r * cols + c;
			if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) { xn[i] = xo[i]; continue; }
			T ax = T(0);
			for (int e = 0; e < 5; ++e) ax = fmaT<T>(L.w[e], xo[(r + dr[e]) * cols + (c + dc[e])], ax);
			T res = b[i] - ax;
			xn[i] = fmaT<T>(omega, L.inv_diag * res, xo[i]);
		}
}

// ── conv-diff — ops.rs:674-791 ────────────────────────────────────────
// kappa == nullptr selects the scalar diffusivity `ks` (EXTENSION).
void conv_diff_core(double alpha, const double* u, double* out, const double* kappa, double ks,
                    const double* vx, const double* vy, const Grid& g, int mode, double sigma) {
	// mode 0: apply (boundary 0, out = diff+conv); 1: axpy (out = fma(alpha, diff+conv, out));
	// 2: operator (out = fma(alpha, diff+conv, sigma*u), boundary out = u).
	int64_t rows = g.rows, cols = g.cols;
	double inv_dx = 1.0 / g.dx, inv_dy = 1.0 / g.dy;
	double inv_dx2 = inv_dx * inv_dx, inv_dy2 = inv_dy * inv_dy;
	if (mode == 0 || mode == 2) {
		for (int64_t c = 0; c < cols; ++c) {
			out[c] = mode == 0 ? 0.0 : u[c];
			out[(rows - 1) * cols + c] = mode == 0 ? 0.0 : u[(rows - 1) * cols + c];
		}
		for (int64_t r = 1; r < rows - 1; ++r) {
			out[r * cols] = mode == 0 ? 0.0 : u[r * cols];
			out[r * cols + cols - 1] = mode == 0 ? 0.0 : u[r * cols + cols - 1];
		}
	}
#pragma omp parallel for schedule(static) if (mode == 2)
	for (int64_t r = 1; r < rows - 1; ++r)
		for (int64_t col = 1; col < cols - 1; ++col) {
			int64_t i = r * cols + col;
			double c = u[i];
			double lap = std::fma(u[i - 1] + u[i + 1], inv_dx2, (u[i - cols] + u[i + cols]) * inv_dy2) -
			             2.0 * (inv_dx2 + inv_dy2) * c;
			double diff = (kappa ? kappa[i] : ks) * lap;
			double vxi = vx[i], vyi = vy[i];
			double dudx = vxi >= 0.0 ? (c - u[i - 1]) * inv_dx : (u[i + 1] - c) * inv_dx;
			double dudy = vyi >= 0.0 ? (c - u[i - cols]) * inv_dy : (u[i + cols] - c) * inv_dy;
			double conv = std::fma(vxi, dudx, vyi * dudy);
			if (mode == 0) out[i] = diff + conv;
			else if (mode == 1) out[i] = std::fma(alpha, diff + conv, out[i]);
			else out[i] = std::fma(alpha, diff + conv, sigma * c);
		}
}

// EXTENSION: variable-coefficient operator sigma*u + alpha*(-div(k grad u)),
// face coefficient = harmonic mean of the two node values.
inline double hmean(double a, double b) { return 2.0 * a * b / (a + b); }
void vark_apply(double sigma, double alpha, const double* k, const double* u, double* out, const Grid& g) {
	int64_t rows = g.rows, cols = g.cols;
	double inv_dx2 = 1.0 / (g.dx * g.dx), inv_dy2 = 1.0 / (g.dy * g.dy);
#pragma omp parallel for schedule(static)
	for (int64_t r = 0; r < rows; ++r)
		for (int64_t col = 0; col < cols; ++col) {
			int64_t i = r * cols + col;
			if (r == 0 || col == 0 || r == rows - 1 || col == cols - 1) { out[i] = u[i]; continue; }
			double kc = k[i], c = u[i];
			double kw = hmean(kc, k[i - 1]), ke = hmean(kc, k[i + 1]);
			double kn = hmean(kc, k[i - cols]), ks = hmean(kc, k[i + cols]);
			double fx = ke * (u[i + 1] - c) - kw * (c - u[i - 1]);
			double fy = ks * (u[i + cols] - c) - kn * (c - u[i - cols]);
			double L = -(fx * inv_dx2 + fy * inv_dy2);
			out[i] = std::fma(alpha, L, sigma * c);
		}
}

}  // namespace

// ════════════════════════════════════════════════════════════════════════
// C ABI for ctypes (tests / bench cpu_baseline)
// ════════════════════════════════════════════════════════════════════════

extern "C" {

struct orc_stencil { int32_t n; int32_t dr[32]; int32_t dc[32]; double w[32]; };
struct orc_bcs { int32_t kind[4]; double param[12]; };   // top, bottom, left, right; (a, b, g)

// Operator description shared by every solver.
//  kind 0: reference  A·v = v + coeff·S(v)      (copy + stencil_axpy: cg.rs:104-111)
//  kind 1: EXTENSION  conv-diff  sigma·v + coeff·(κΔv + w·∇⁻v) on the interior, identity rows on the boundary
//  kind 2: EXTENSION  variable-k sigma·v + coeff·(−∇·(k∇v)) on the interior, identity rows on the boundary
struct orc_op {
	int32_t kind;
	int32_t pad;
	double coeff;
	double sigma;
	orc_stencil st;
	orc_bcs bc;
	const double* kappa;
	double kappa_scalar;
	const double* vx;
	const double* vy;
	const double* k;
};

struct orc_mg_cfg {
	int32_t n_pre, n_post;
	double omega;
	int32_t smoother;      // 0 reference in-place sweep, 1 two-buffer Jacobi, 2 Chebyshev
	int32_t transfer;      // 0 reference (2x2 average + injection), 1 full weighting + bilinear
	int32_t coarsest;      // stop coarsening once rows or cols <= coarsest (reference: 4)
	int32_t extra_coarse;  // sweeps added on the coarsest level (reference: 10)
	int32_t precision;     // 0 fp64, 1 fp32 V-cycle
	int32_t pad;
	double sigma, scale;   // level operator sigma·I + scale·Δ_h (reference: 0, 1)
	orc_bcs bc;
	const double* k;       // EXTENSION: nodal coefficient -> level operator sigma·I + scale·(−∇·k∇)
	// EXTENSION: nodal velocity -> level operator sigma·I + scale·(κΔ + v·∇⁻), the upwind
	// convection-diffusion of ops.rs:707-731 rediscretised on every level (velocity interpolated
	// like k); kappa is the scalar κ.
	const double* vx;
	const double* vy;
	double kappa;
};

struct orc_stats { int64_t iterations; double residual; int32_t reason; int32_t pad; };
enum { CONVERGED = 0, MAX_ITERATIONS = 1, BREAKDOWN = 2 };

}  // extern "C"

namespace {

Grid mkgrid(int64_t rows, int64_t cols, double dx, double dy) { return Grid{rows, cols, dx, dy}; }
int valid_grid(const Grid& g) {
	if (g.rows < 3 || g.cols < 3) return ERR_GRID;
	if (!std::isfinite(g.dx) || !std::isfinite(g.dy) || g.dx <= 0 || g.dy <= 0) return ERR_GRID;
	return OK;
}
std::vector<Entry> entries(const orc_stencil& s) {
	std::vector<Entry> e(std::max(0, std::min(32, (int)s.n)));
	for (size_t i = 0; i < e.size(); ++i) e[i] = {s.dr[i], s.dc[i], s.w[i]};
	return e;
}
Bcs bcs(const orc_bcs& b) {
	Bc s[4];
	for (int i = 0; i < 4; ++i) s[i] = {b.kind[i], b.param[3 * i], b.param[3 * i + 1], b.param[3 * i + 2]};
	return {s[0], s[1], s[2], s[3]};
}

// apply_op — cg.rs:104-111 (copy + stencil_axpy) and the extension kinds.
int apply_op(const orc_op& op, const double* x, double* y, const Grid& g) {
	if (op.kind == 0) {
		auto es = entries(op.st);
		copy(x, y, g.len());
		return stencil_axpy(op.coeff, x, y, g, es.data(), (int)es.size(), bcs(op.bc));
	}
	if (op.kind == 1) {
		conv_diff_core(op.coeff, x, y, op.kappa, op.kappa_scalar, op.vx, op.vy, g, 2, op.sigma);
		return OK;
	}
	if (op.kind == 2) { vark_apply(op.sigma, op.coeff, op.k, x, y, g); return OK; }
	return ERR_ARG;
}

// ── Multigrid — multigrid.rs:12-608 ───────────────────────────────────
template <class T>
struct MgLevel {
	Grid g;
	Lvl5<T> L;
	std::vector<T> u, rhs, tmp, w0, w1;
	std::vector<T> k;   // variable-coefficient levels only
	std::vector<T> vx, vy;   // convection-diffusion levels only
};

// EXTENSION: variable-coefficient level operator sigma·u + scale·(−∇·(k∇u)) at an interior
// point, harmonic face means; `diag` receives its diagonal.
template <class T> inline T hmT(T a, T b) { return T(2) * a * b / (a + b); }
template <class T>
inline T vk_apply(const T* u, const T* k, int64_t i, int64_t cols, T ix2, T iy2, T sigma, T scale, T* diag) {
	T kc = k[i], c = u[i];
	T kw = hmT<T>(kc, k[i - 1]), ke = hmT<T>(kc, k[i + 1]);
	T kn = hmT<T>(kc, k[i - cols]), ks = hmT<T>(kc, k[i + cols]);
	T fx = ke * (u[i + 1] - c) - kw * (c - u[i - 1]);
	T fy = ks * (u[i + cols] - c) - kn * (c - u[i - cols]);
	T L = -(fx * ix2 + fy * iy2);
	*diag = fmaT<T>(scale, (kw + ke) * ix2 + (kn + ks) * iy2, sigma);
	return fmaT<T>(scale, L, sigma * c);
}

// EXTENSION: convection-diffusion level operator sigma·u + scale·(κΔu + v·∇⁻u) at an interior point,
// the arithmetic of conv_diff_core (ops.rs:707-731) in T; `diag` receives its diagonal
// sigma + scale·(|vx|/dx + |vy|/dy − 2κ(1/dx² + 1/dy²)).
template <class T>
inline T cd_apply(const T* u, const T* vx, const T* vy, int64_t i, int64_t cols, T ix, T iy, T ix2, T iy2, T kap,
                  T sigma, T scale, T* diag) {
	T c = u[i];
	T lap = fmaT<T>(u[i - 1] + u[i + 1], ix2, (u[i - cols] + u[i + cols]) * iy2) - T(2) * (ix2 + iy2) * c;
	T diff = kap * lap;
	T vxi = vx[i], vyi = vy[i];
	T dudx = vxi >= T(0) ? (c - u[i - 1]) * ix : (u[i + 1] - c) * ix;
	T dudy = vyi >= T(0) ? (c - u[i - cols]) * iy : (u[i + cols] - c) * iy;
	T conv = fmaT<T>(vxi, dudx, vyi * dudy);
	T ax = vxi >= T(0) ? vxi : -vxi, ay = vyi >= T(0) ? vyi : -vyi;
	*diag = fmaT<T>(scale, fmaT<T>(ax, ix, ay * iy) - T(2) * kap * (ix2 + iy2), sigma);
	return fmaT<T>(scale, diff + conv, sigma * c);
}

// Coefficient on a coarse level: injection for halving levels, linear
// interpolation at the coarse node positions for same-domain levels.
template <class T>
void coarsen_k(const T* kf, const Grid& fg, T* kc, const Grid& cg, int transfer) {
	for (int64_t R = 0; R < cg.rows; ++R)
		for (int64_t Cc = 0; Cc < cg.cols; ++Cc) {
			if (transfer == 0) {
				kc[R * cg.cols + Cc] = kf[std::min(2 * R, fg.rows - 1) * fg.cols + std::min(2 * Cc, fg.cols - 1)];
				continue;
			}
			Lin1 ly = lin1(R, cg.rows, fg.rows), lx = lin1(Cc, cg.cols, fg.cols);
			int64_t a0 = ly.i0, a1 = std::min(ly.i0 + 1, fg.rows - 1);
			int64_t b0 = lx.i0, b1 = std::min(lx.i0 + 1, fg.cols - 1);
			T wy = T(ly.w), vy = T(1) - wy, wx = T(lx.w), vx = T(1) - wx;
			T top = vx * kf[a0 * fg.cols + b0] + wx * kf[a0 * fg.cols + b1];
			T bot = vx * kf[a1 * fg.cols + b0] + wx * kf[a1 * fg.cols + b1];
			kc[R * cg.cols + Cc] = vy * top + wy * bot;
		}
}

template <class T>
struct Mg {
	orc_mg_cfg cfg;
	Bcs bc;
	std::vector<MgLevel<T>> lv;

	// multigrid.rs:92-146 (hierarchy; coarsest threshold is a parameter here).
	explicit Mg(const Grid& fine, const orc_mg_cfg& c) : cfg(c), bc(bcs(c.bc)) {
		Grid g = fine;
		int64_t lim = c.coarsest > 0 ? c.coarsest : 4;
		for (;;) {
			MgLevel<T> L;
			L.g = g;
			double ix2 = 1.0 / (g.dx * g.dx), iy2 = 1.0 / (g.dy * g.dy);
			double w[5] = {c.sigma + c.scale * (-2.0 * (ix2 + iy2)), c.scale * iy2, c.scale * iy2, c.scale * ix2,
			               c.scale * ix2};
			for (int e = 0; e < 5; ++e) L.L.w[e] = T(w[e]);
			L.L.inv_diag = T(1.0 / w[0]);
			size_t n = (size_t)g.len();
			L.u.assign(n, T(0)); L.rhs.assign(n, T(0)); L.tmp.assign(n, T(0));
			L.w0.assign(n, T(0)); L.w1.assign(n, T(0));
			if (c.k) {
				L.k.assign(n, T(0));
				if (lv.empty()) {
					for (size_t i = 0; i < n; ++i) L.k[i] = T(c.k[i]);
				} else {
					const MgLevel<T>& F = lv.back();
					coarsen_k(F.k.data(), F.g, L.k.data(), g, c.transfer);
				}
			}
			if (c.vx && c.vy) {
				L.vx.assign(n, T(0));
				L.vy.assign(n, T(0));
				if (lv.empty()) {
					for (size_t i = 0; i < n; ++i) { L.vx[i] = T(c.vx[i]); L.vy[i] = T(c.vy[i]); }
				} else {
					const MgLevel<T>& F = lv.back();
					coarsen_k(F.vx.data(), F.g, L.vx.data(), g, c.transfer);
					coarsen_k(F.vy.data(), F.g, L.vy.data(), g, c.transfer);
				}
			}
			lv.push_back(std::move(L));
			if (g.rows <= lim || g.cols <= lim) break;
			if (c.transfer == 0) {
				int64_t cr = g.rows / 2, cc = g.cols / 2;
				if (cr < 3 || cc < 3) break;
				g = mkgrid(cr, cc, g.dx * 2.0, g.dy * 2.0);
			} else {
				// Same domain on every level: odd sizes nest ((n+1)/2), even sizes halve.
				int64_t cr = g.rows % 2 ? (g.rows + 1) / 2 : g.rows / 2;
				int64_t cc = g.cols % 2 ? (g.cols + 1) / 2 : g.cols / 2;
				if (cr < 3 || cc < 3) break;
				g = mkgrid(cr, cc, g.dx * double(g.cols - 1) / double(cc - 1), g.dy * double(g.rows - 1) / double(cr - 1));
			}
		}
	}

	void smooth(int l, int n) {
		MgLevel<T>& L = lv[l];
		T om = T(cfg.omega);
		if (cfg.smoother == 2) { smooth_cheby(l, n); return; }
		for (int it = 0; it < n; ++it) {
			if (!L.k.empty() || !L.vx.empty()) {
				smooth_jacobi_k(l, L.u.data(), L.tmp.data());
				std::swap(L.u, L.tmp);
			} else if (cfg.smoother == 0) smooth_ref_gs<T>(L.u.data(), L.rhs.data(), om, L.g, L.L);
			else {
				smooth_jacobi<T>(L.u.data(), L.tmp.data(), L.rhs.data(), om, L.g, L.L);
				std::swap(L.u, L.tmp);
			}
		}
	}

	// EXTENSION: two-buffer Jacobi with the variable-coefficient or convection-diffusion level
	// operator: x_new = fma(ω/diag, b − A·x, x).
	void smooth_jacobi_k(int l, const T* xo, T* xn) {
		MgLevel<T>& L = lv[l];
		const int64_t rows = L.g.rows, cols = L.g.cols;
		const T ix2 = T(1.0 / (L.g.dx * L.g.dx)), iy2 = T(1.0 / (L.g.dy * L.g.dy));
		const T sg = T(cfg.sigma), sc = T(cfg.scale), om = T(cfg.omega);
		const T ix = T(1.0 / L.g.dx), iy = T(1.0 / L.g.dy), kap = T(cfg.kappa);
		const T cix2 = T((1.0 / L.g.dx) * (1.0 / L.g.dx)), ciy2 = T((1.0 / L.g.dy) * (1.0 / L.g.dy));
#pragma omp parallel for schedule(static)
		for (int64_t r = 0; r < rows; ++r)
			for (int64_t c = 0; c < cols; ++c) {
				int64_t i = r * cols + c;
				if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) { xn[i] = xo[i]; continue; }
				T d;
				T au = L.vx.empty() ? vk_apply<T>(xo, L.k.data(), i, cols, ix2, iy2, sg, sc, &d)
				                    : cd_apply<T>(xo, L.vx.data(), L.vy.data(), i, cols, ix, iy, cix2, ciy2, kap, sg, sc, &d);
				xn[i] = fmaT<T>(om / d, L.rhs[i] - au, xo[i]);
			}
	}

	// apply_stencil on a level (5-pt path, ops.rs:891-938), boundary via bc_value.
	void level_stencil(int l, const T* in, T* out) {
		MgLevel<T>& L = lv[l];
		int64_t rows = L.g.rows, cols = L.g.cols;
		// bc_value in T precision (double for the reference path).
		auto bcv = [&](const Bc& b, T sv, T nb, double sp) -> T {
			return T(bc_value(b, double(sv), double(nb), sp));
		};
		int64_t lr = rows - 1, lc = cols - 1;
		for (int64_t c = 0; c < cols; ++c) {
			out[c] = bcv(bc.top, in[c], in[cols + c], L.g.dy);
			out[lr * cols + c] = bcv(bc.bottom, in[lr * cols + c], in[(lr - 1) * cols + c], L.g.dy);
		}
		for (int64_t r = 1; r < lr; ++r) {
			out[r * cols] = bcv(bc.left, in[r * cols], in[r * cols + 1], L.g.dx);
			out[r * cols + lc] = bcv(bc.right, in[r * cols + lc], in[r * cols + lc - 1], L.g.dx);
		}
		const T* w = L.L.w;
		const T ix2 = T(1.0 / (L.g.dx * L.g.dx)), iy2 = T(1.0 / (L.g.dy * L.g.dy));
		const T sg = T(cfg.sigma), sc = T(cfg.scale);
		const T ix = T(1.0 / L.g.dx), iy = T(1.0 / L.g.dy), kap = T(cfg.kappa);
		const T cix2 = T((1.0 / L.g.dx) * (1.0 / L.g.dx)), ciy2 = T((1.0 / L.g.dy) * (1.0 / L.g.dy));
#pragma omp parallel for schedule(static)
		for (int64_t r = 1; r < lr; ++r)
			for (int64_t c = 1; c < lc; ++c) {
				int64_t i = r * cols + c;
				if (!L.vx.empty()) {
					T d;
					out[i] = cd_apply<T>(in, L.vx.data(), L.vy.data(), i, cols, ix, iy, cix2, ciy2, kap, sg, sc, &d);
					continue;
				}
				if (!L.k.empty()) {
					T d;
					out[i] = vk_apply<T>(in, L.k.data(), i, cols, ix2, iy2, sg, sc, &d);
					continue;
				}
				out[i] = fmaT<T>(w[0], in[i], fmaT<T>(w[1], in[i - cols], fmaT<T>(w[2], in[i + cols],
				         fmaT<T>(w[3], in[i - 1], w[4] * in[i + 1]))));
			}
	}

	// multigrid.rs:476-491: tmp = S(u); tmp *= -1; tmp += rhs.
	void residual(int l) {
		MgLevel<T>& L = lv[l];
		level_stencil(l, L.u.data(), L.tmp.data());
		int64_t n = L.g.len();
		for (int64_t i = 0; i < n; ++i) L.tmp[i] = fmaT<T>(T(1), L.rhs[i], -L.tmp[i]);
	}

	// multigrid.rs:317-473: first-kind Chebyshev with D^-1, emax=2, emin=0.2.
	void smooth_cheby(int l, int n) {
		if (n == 0) return;
		MgLevel<T>& L = lv[l];
		const double emax = 2.0, emin = 0.1 * emax;
		const double scl = 2.0 / (emax + emin), alpha = 1.0 - scl * emin, mu = 1.0 / alpha, omegaprod = 2.0 / alpha;
		const double ix2 = 1.0 / (L.g.dx * L.g.dx), iy2 = 1.0 / (L.g.dy * L.g.dy);
		const double inv_diag = -1.0 / (2.0 * ix2 + 2.0 * iy2);
		int64_t N = L.g.len();
		auto zb = [&](std::vector<T>& v) {   // field.rs:386-431
			int64_t rows = L.g.rows, cols = L.g.cols;
			for (int64_t c = 0; c < cols; ++c) { v[c] = T(0); v[(rows - 1) * cols + c] = T(0); }
			for (int64_t r = 1; r < rows - 1; ++r) { v[r * cols] = T(0); v[r * cols + cols - 1] = T(0); }
		};
		auto ax = [&](std::vector<T>& y, T a, const std::vector<T>& x) {
			for (int64_t i = 0; i < N; ++i) y[i] = fmaT<T>(a, x[i], y[i]);
		};
		auto sc = [&](std::vector<T>& y, T a) { for (int64_t i = 0; i < N; ++i) y[i] *= a; };
		double c_km1 = 1.0, c_k = mu;
		level_stencil(l, L.u.data(), L.tmp.data());
		L.w0 = L.rhs;
		ax(L.w0, T(-1), L.tmp);
		sc(L.w0, T(scl * inv_diag));
		ax(L.w0, T(1), L.u);
		zb(L.w0);
		if (n == 1) { L.u = L.w0; return; }
		std::vector<T>* B[3] = {&L.u, &L.w0, &L.w1};
		int km1 = 0, k = 1, kp1 = 2;
		for (int i = 1; i < n; ++i) {
			level_stencil(l, B[k]->data(), L.tmp.data());
			std::vector<T>& p = *B[kp1];
			p = L.rhs;
			ax(p, T(-1), L.tmp);
			sc(p, T(inv_diag));
			double c_kp1 = 2.0 * mu * c_k - c_km1;
			double om = omegaprod * c_k / c_kp1;
			sc(p, T(om * scl));
			ax(p, T(om), *B[k]);
			ax(p, T(1.0 - om), *B[km1]);
			zb(p);
			c_km1 = c_k; c_k = c_kp1;
			int t = km1; km1 = k; k = kp1; kp1 = t;
		}
		if (k != 0) L.u = *B[k];
	}

	// multigrid.rs:188-230.
	void vcycle(double* x, const double* b) {
		MgLevel<T>& L0 = lv[0];
		int64_t n0 = L0.g.len();
		for (int64_t i = 0; i < n0; ++i) { L0.rhs[i] = T(b[i]); L0.u[i] = T(x[i]); }
		int nl = (int)lv.size();
		for (int l = 0; l < nl - 1; ++l) {
			smooth(l, cfg.n_pre);
			residual(l);
			if (cfg.transfer == 0) restrict_ref<T>(lv[l].tmp.data(), lv[l + 1].rhs.data(), lv[l].g, lv[l + 1].g);
			else restrict_lin<T>(lv[l].tmp.data(), lv[l + 1].rhs.data(), lv[l].g, lv[l + 1].g);
			std::fill(lv[l + 1].u.begin(), lv[l + 1].u.end(), T(0));
		}
		smooth(nl - 1, cfg.n_pre + cfg.n_post + cfg.extra_coarse);
		for (int l = nl - 2; l >= 0; --l) {
			if (cfg.transfer == 0) prolong_ref<T>(lv[l + 1].u.data(), lv[l].u.data(), lv[l + 1].g, lv[l].g);
			else prolong_lin<T>(lv[l + 1].u.data(), lv[l].u.data(), lv[l + 1].g, lv[l].g);
			smooth(l, cfg.n_post);
		}
		for (int64_t i = 0; i < n0; ++i) x[i] = double(L0.u[i]);
	}
};

// Preconditioner: identity (preconditioner.rs:29-37) or multigrid.
struct Pc {
	int kind = 0;  // 0 identity, 1 mg64, 2 mg32
	Mg<double>* m64 = nullptr;
	Mg<float>* m32 = nullptr;
	// kind 1 with a config builds a multigrid; any other kind is the identity.
	Pc(int k, const Grid& g, const orc_mg_cfg* c) : kind(0) {
		if (k == 1 && c) {
			if (c->precision == 1) { m32 = new Mg<float>(g, *c); kind = 2; }
			else { m64 = new Mg<double>(g, *c); kind = 1; }
		}
	}
	~Pc() { delete m64; delete m32; }
	// `z` is zeroed by the caller first (cg.rs:193-194).
	void apply(double* z, const double* r, int64_t n) {
		if (kind == 0) copy(r, z, n);
		else if (m64) m64->vcycle(z, r);
		else m32->vcycle(z, r);
	}
};

void set_stats(orc_stats* s, int64_t it, double res, int reason) {
	if (s) { s->iterations = it; s->residual = res; s->reason = reason; }
}

}  // namespace

extern "C" {

// ── Backend ops ───────────────────────────────────────────────────────
int orc_apply_stencil(const double* in, double* out, int64_t rows, int64_t cols, double dx, double dy,
                      const orc_stencil* st, const orc_bcs* bc) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	auto es = entries(*st);
	return apply_stencil(in, out, g, es.data(), (int)es.size(), bcs(*bc));
}
int orc_stencil_axpy(double alpha, const double* x, double* y, int64_t rows, int64_t cols, double dx, double dy,
                     const orc_stencil* st, const orc_bcs* bc) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	auto es = entries(*st);
	return stencil_axpy(alpha, x, y, g, es.data(), (int)es.size(), bcs(*bc));
}
// ops.rs:129-154: apply_stencil then out *= coeff.
int orc_apply_stencil_var(const double* in, double* out, const double* coeff, int64_t rows, int64_t cols, double dx,
                          double dy, const orc_stencil* st, const orc_bcs* bc) {
	int e = orc_apply_stencil(in, out, rows, cols, dx, dy, st, bc);
	if (e) return e;
	int64_t n = rows * cols;
	for (int64_t i = 0; i < n; ++i) out[i] *= coeff[i];
	return OK;
}
void orc_axpy(double a, const double* x, double* y, int64_t n) { axpy(a, x, y, n); }
void orc_scale(double* x, double a, int64_t n) { scale(x, a, n); }
void orc_fill(double* x, double v, int64_t n) { fill(x, v, n); }
void orc_copy(const double* s, double* d, int64_t n) { copy(s, d, n); }
double orc_dot(const double* x, const double* y, int64_t n) { return dot(x, y, n); }
double orc_norm_l2(const double* x, int64_t n) { return norm_l2(x, n); }
double orc_reduce_sum(const double* x, int64_t n) { return reduce_sum(x, n); }
double orc_reduce_max(const double* x, int64_t n) {
	// ops.rs:318-331: f64::max ignores NaN; empty -> -inf.
	double m = -std::numeric_limits<double>::infinity();
	for (int64_t i = 0; i < n; ++i) m = std::fmax(m, x[i]);
	return m;
}
double orc_reduce_min(const double* x, int64_t n) {
	double m = std::numeric_limits<double>::infinity();
	for (int64_t i = 0; i < n; ++i) m = std::fmin(m, x[i]);
	return m;
}
// core/src/backend.rs:219-230: default dot2 = two dots.
void orc_dot2(const double* a, const double* b, const double* c, const double* d, int64_t n, double* out) {
	out[0] = dot(a, b, n);
	out[1] = dot(c, d, n);
}
void orc_waxpy(double alpha, const double* x, double beta, const double* y, double* w, int64_t n) {
	// ops.rs:530-559: w = alpha.mul_add(x, beta*y).
	for (int64_t i = 0; i < n; ++i) w[i] = std::fma(alpha, x[i], beta * y[i]);
}
void orc_aypx(double alpha, const double* x, double* y, int64_t n) {
	// ops.rs:562-585: y = alpha.mul_add(y, x).
	for (int64_t i = 0; i < n; ++i) y[i] = std::fma(alpha, y[i], x[i]);
}
void orc_pointwise(int op, const double* x, const double* y, double* z, int64_t n) {
	// ops.rs:476-528, 650-672: 0 mult, 1 div, 2 max, 3 min.
	for (int64_t i = 0; i < n; ++i) {
		double a = x[i], b = y[i];
		z[i] = op == 0 ? a * b : op == 1 ? a / b : op == 2 ? std::fmax(a, b) : std::fmin(a, b);
	}
}
void orc_reciprocal(double* x, int64_t n) { for (int64_t i = 0; i < n; ++i) x[i] = 1.0 / x[i]; }
void orc_abs(double* x, int64_t n) { for (int64_t i = 0; i < n; ++i) x[i] = std::fabs(x[i]); }
void orc_zero_boundaries(double* x, int64_t rows, int64_t cols) {
	for (int64_t c = 0; c < cols; ++c) { x[c] = 0.0; x[(rows - 1) * cols + c] = 0.0; }
	for (int64_t r = 1; r < rows - 1; ++r) { x[r * cols] = 0.0; x[r * cols + cols - 1] = 0.0; }
}
void orc_restrict(const double* f, double* c, int64_t fr, int64_t fc, int64_t cr, int64_t cc, int mode) {
	Grid fg = mkgrid(fr, fc, 1, 1), cg = mkgrid(cr, cc, 2, 2);
	if (mode == 0) restrict_ref<double>(f, c, fg, cg); else restrict_lin<double>(f, c, fg, cg);
}
void orc_prolong(const double* c, double* f, int64_t cr, int64_t cc, int64_t fr, int64_t fc, int mode) {
	Grid fg = mkgrid(fr, fc, 1, 1), cg = mkgrid(cr, cc, 2, 2);
	if (mode == 0) prolong_ref<double>(c, f, cg, fg); else prolong_lin<double>(c, f, cg, fg);
}
// ops.rs:425-473 (mode 0, in place) or EXTENSION two-buffer Jacobi (mode 1).
int orc_weighted_jacobi(double* x, const double* b, double omega, int64_t rows, int64_t cols, double dx, double dy,
                        const orc_stencil* st, int mode) {
	Grid g = mkgrid(rows, cols, dx, dy);
	auto es = entries(*st);
	double diag = -4.0;
	for (auto& e : es) if (e.dr == 0 && e.dc == 0) { diag = e.w; break; }
	double inv_diag = 1.0 / diag;
	std::vector<double> xo(x, x + g.len());
	for (int64_t r = 1; r < rows - 1; ++r)
		for (int64_t c = 1; c < cols - 1; ++c) {
			int64_t i = r * cols + c;
			const double* src = mode == 0 ? x : xo.data();
			double ax = 0.0;
			for (auto& e : es) ax = std::fma(e.w, src[(r + e.dr) * cols + (c + e.dc)], ax);
			double res = b[i] - ax;
			x[i] = std::fma(omega, inv_diag * res, src[i]);
		}
	return OK;
}
void orc_conv_diff(int mode, double alpha, const double* u, double* out, const double* kappa, double ks,
                   const double* vx, const double* vy, int64_t rows, int64_t cols, double dx, double dy, double sigma) {
	conv_diff_core(alpha, u, out, kappa, ks, vx, vy, mkgrid(rows, cols, dx, dy), mode, sigma);
}
void orc_vark_apply(double sigma, double alpha, const double* k, const double* u, double* out, int64_t rows,
                    int64_t cols, double dx, double dy) {
	vark_apply(sigma, alpha, k, u, out, mkgrid(rows, cols, dx, dy));
}
int orc_apply_op(const orc_op* op, const double* x, double* y, int64_t rows, int64_t cols, double dx, double dy) {
	return apply_op(*op, x, y, mkgrid(rows, cols, dx, dy));
}

// ── Multigrid ─────────────────────────────────────────────────────────
int orc_mg_depth(int64_t rows, int64_t cols, double dx, double dy, const orc_mg_cfg* cfg) {
	Mg<double> m(mkgrid(rows, cols, dx, dy), *cfg);
	return (int)m.lv.size();
}
void orc_vcycle(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_mg_cfg* cfg,
                int n_cycles) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (cfg->precision == 1) { Mg<float> m(g, *cfg); for (int i = 0; i < n_cycles; ++i) m.vcycle(x, b); }
	else { Mg<double> m(g, *cfg); for (int i = 0; i < n_cycles; ++i) m.vcycle(x, b); }
}

// ── CgSolver::solve — cg.rs:96-164 ────────────────────────────────────
int orc_cg(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op, double tol,
           int64_t max_iter, orc_stats* st, double* hist) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	int64_t n = g.len();
	std::vector<double> r(n), p(n), ap(n);
	if (int e = apply_op(*op, x, ap.data(), g)) return e;
	copy(b, r.data(), n);
	axpy(-1.0, ap.data(), r.data(), n);
	copy(r.data(), p.data(), n);
	double rs_old = dot(r.data(), r.data(), n);
	for (int64_t i = 0; i < max_iter; ++i) {
		if (hist) hist[i] = std::sqrt(rs_old);
		if (std::sqrt(rs_old) < tol) { set_stats(st, i, std::sqrt(rs_old), CONVERGED); return OK; }
		apply_op(*op, p.data(), ap.data(), g);
		double p_ap = dot(p.data(), ap.data(), n);
		if (std::fabs(p_ap) < 1e-30) { set_stats(st, i, std::sqrt(rs_old), BREAKDOWN); return OK; }
		double alpha = rs_old / p_ap;
		axpy(alpha, p.data(), x, n);
		axpy(-alpha, ap.data(), r.data(), n);
		double rs_new = dot(r.data(), r.data(), n);
		double beta = rs_new / rs_old;
		scale(p.data(), beta, n);
		axpy(1.0, r.data(), p.data(), n);
		rs_old = rs_new;
	}
	set_stats(st, max_iter, std::sqrt(rs_old), MAX_ITERATIONS);
	return OK;
}

// ── CgSolver::solve_preconditioned — cg.rs:178-256 ────────────────────
int orc_pcg(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op, int pc_kind,
            const orc_mg_cfg* cfg, double tol, int64_t max_iter, orc_stats* st, double* hist) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	int64_t n = g.len();
	Pc pc(pc_kind, g, cfg);
	std::vector<double> r(n), z(n), p(n), ap(n);
	if (int e = apply_op(*op, x, ap.data(), g)) return e;
	copy(b, r.data(), n);
	axpy(-1.0, ap.data(), r.data(), n);
	fill(z.data(), 0.0, n);
	pc.apply(z.data(), r.data(), n);
	copy(z.data(), p.data(), n);
	double rz_old = dot(r.data(), z.data(), n);
	for (int64_t i = 0; i < max_iter; ++i) {
		double r_norm = std::sqrt(dot(r.data(), r.data(), n));
		if (hist) hist[i] = r_norm;
		if (r_norm < tol) { set_stats(st, i, r_norm, CONVERGED); return OK; }
		apply_op(*op, p.data(), ap.data(), g);
		double p_ap = dot(p.data(), ap.data(), n);
		if (std::fabs(p_ap) < 1e-30) { set_stats(st, i, r_norm, BREAKDOWN); return OK; }
		double alpha = rz_old / p_ap;
		axpy(alpha, p.data(), x, n);
		axpy(-alpha, ap.data(), r.data(), n);
		fill(z.data(), 0.0, n);
		pc.apply(z.data(), r.data(), n);
		double rz_new = dot(r.data(), z.data(), n);
		double beta = rz_new / rz_old;
		scale(p.data(), beta, n);
		axpy(1.0, z.data(), p.data(), n);
		rz_old = rz_new;
	}
	set_stats(st, max_iter, std::sqrt(dot(r.data(), r.data(), n)), MAX_ITERATIONS);
	return OK;
}

// ── EXTENSION: flexible PCG (Notay FCG(1)) with an FP64 true-residual
// judge: β = −(z·Ap_old)/(p_old·Ap_old), restart from the true residual
// when the recurrence says converged but the true residual does not.
int orc_fcg(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op, int pc_kind,
            const orc_mg_cfg* cfg, double tol, int64_t max_iter, orc_stats* st, double* hist) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	int64_t n = g.len();
	Pc pc(pc_kind, g, cfg);
	std::vector<double> r(n), z(n), p(n), ap(n);
	int64_t it = 0;
	double r_norm = 0.0, prev_tr = INFINITY;
	bool at_floor = false;
	for (int restart = 0; restart < 8; ++restart) {
		apply_op(*op, x, ap.data(), g);
		copy(b, r.data(), n);
		axpy(-1.0, ap.data(), r.data(), n);
		r_norm = std::sqrt(dot(r.data(), r.data(), n));
		if (r_norm < tol) { set_stats(st, it, r_norm, CONVERGED); return OK; }
		// A restart that did not halve the true residual has reached the FP64
		// attainable accuracy; further passes only repeat it.  The recurrence of the
		// pass before met the tolerance, so the solve counts as converged.
		if (restart > 0 && !(r_norm < 0.5 * prev_tr)) { at_floor = true; break; }
		prev_tr = r_norm;
		fill(z.data(), 0.0, n);
		pc.apply(z.data(), r.data(), n);
		copy(z.data(), p.data(), n);
		double rz = dot(r.data(), z.data(), n);
		bool recur_conv = false;
		while (it < max_iter) {
			if (hist) hist[it] = r_norm;
			apply_op(*op, p.data(), ap.data(), g);
			double p_ap = dot(p.data(), ap.data(), n);
			if (std::fabs(p_ap) < 1e-30) { set_stats(st, it, r_norm, BREAKDOWN); return OK; }
			double alpha = rz / p_ap;
			axpy(alpha, p.data(), x, n);
			axpy(-alpha, ap.data(), r.data(), n);
			++it;
			r_norm = std::sqrt(dot(r.data(), r.data(), n));
			if (r_norm < tol) { recur_conv = true; break; }
			fill(z.data(), 0.0, n);
			pc.apply(z.data(), r.data(), n);
			rz = dot(r.data(), z.data(), n);
			double zap = dot(z.data(), ap.data(), n);
			double beta = -zap / p_ap;
			scale(p.data(), beta, n);
			axpy(1.0, z.data(), p.data(), n);
		}
		if (!recur_conv) break;
	}
	apply_op(*op, x, ap.data(), g);
	copy(b, r.data(), n);
	axpy(-1.0, ap.data(), r.data(), n);
	r_norm = std::sqrt(dot(r.data(), r.data(), n));
	set_stats(st, it, r_norm, r_norm < tol || at_floor ? CONVERGED : MAX_ITERATIONS);
	return OK;
}

// ── PipelinedCgSolver::solve — pipelined_cg.rs:96-200 ─────────────────
int orc_pipelined_cg(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op,
                     double tol, int64_t max_iter, orc_stats* st, double* hist) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	int64_t n = g.len();
	std::vector<double> r(n), p(n), w(n), t(n), z(n), q(n);
	if (int e = apply_op(*op, x, q.data(), g)) return e;
	copy(b, r.data(), n);
	axpy(-1.0, q.data(), r.data(), n);
	apply_op(*op, r.data(), w.data(), g);
	double gamma_prev = INFINITY, alpha_prev = INFINITY;
	for (int64_t k = 0; k < max_iter; ++k) {
		double gamma = dot(r.data(), r.data(), n), delta = dot(r.data(), w.data(), n);
		if (hist) hist[k] = std::sqrt(gamma);
		if (std::sqrt(gamma) < tol) { set_stats(st, k, std::sqrt(gamma), CONVERGED); return OK; }
		apply_op(*op, w.data(), q.data(), g);
		double beta = gamma / gamma_prev;
		double alpha = gamma / (delta - beta * gamma / alpha_prev);
		// pipelined_cg.rs:182-199 (CPU order).
		scale(z.data(), beta, n); axpy(1.0, q.data(), z.data(), n);
		scale(t.data(), beta, n); axpy(1.0, w.data(), t.data(), n);
		scale(p.data(), beta, n); axpy(1.0, r.data(), p.data(), n);
		axpy(alpha, p.data(), x, n);
		axpy(-alpha, t.data(), r.data(), n);
		axpy(-alpha, z.data(), w.data(), n);
		gamma_prev = gamma;
		alpha_prev = alpha;
	}
	set_stats(st, max_iter, std::sqrt(dot(r.data(), r.data(), n)), MAX_ITERATIONS);
	return OK;
}

// ── EXTENSION: preconditioned pipelined CG (Ghysels & Vanroose 2014,
// Alg. 3).  u = M r, w = A u; per iteration one fused (γ,δ,‖r‖²) reduction,
// m = M w, n = A m.
}  // extern "C"

namespace {
// EXTENSION: pipelined flexible CG for an FP32 multigrid, inside FP64 true-residual refinement passes.
// One reduction per iteration, the pipelined loop of pipelined_cg.rs:130-159 with the preconditioner
// applied to the current residual (u = M(r), w = A·u): from (r·u, w·u, u·s_old, r·r) the flexible
// β = −(u·s_old)/(p_old·s_old), p·s = w·u + β·(u·s_old) and α = (r·u)/(p·s); then p = u + β·p,
// s = w + β·s (= A·p), x += α·p, r −= α·s.  A pass restarts from x while the true residual is above
// tol and keeps halving (the refinement rule of the flexible CG).
int pfcg(double* x, const double* b, const Grid& g, const orc_op* op, Pc& pc, double tol, int64_t max_iter,
         orc_stats* st, double* hist) {
	const int64_t N = g.len();
	std::vector<double> r(N), u(N), w(N), p(N), s(N);
	int64_t used = 0;
	double prev_tr = INFINITY, rn = 0.0, tr = 0.0;
	bool at_floor = false;
	int reason = CONVERGED;
	for (int pass = 0; pass < 8; ++pass) {
		apply_op(*op, x, w.data(), g);
		copy(b, r.data(), N);
		axpy(-1.0, w.data(), r.data(), N);
		fill(p.data(), 0.0, N);
		fill(s.data(), 0.0, N);
		double pap = 0.0;
		int64_t k = 0;
		reason = CONVERGED;
		for (;; ++k) {
			fill(u.data(), 0.0, N);
			pc.apply(u.data(), r.data(), N);
			apply_op(*op, u.data(), w.data(), g);
			const double gamma = dot(r.data(), u.data(), N), delta = dot(w.data(), u.data(), N);
			const double eps = dot(u.data(), s.data(), N);
			rn = std::sqrt(dot(r.data(), r.data(), N));
			if (hist) hist[used + k] = rn;
			if (used + k >= max_iter) { reason = MAX_ITERATIONS; break; }
			if (rn < tol) break;
			const double beta = k == 0 ? 0.0 : -eps / pap;
			const double eta = k == 0 ? delta : delta + beta * eps;
			if (!(std::fabs(eta) >= 1e-30)) { reason = BREAKDOWN; break; }
			const double alpha = gamma / eta;
			for (int64_t j = 0; j < N; ++j) {
				const double pj = u[j] + beta * p[j];
				const double sj = w[j] + beta * s[j];
				x[j] = std::fma(alpha, pj, x[j]);
				r[j] = std::fma(-alpha, sj, r[j]);
				p[j] = pj;
				s[j] = sj;
			}
			pap = eta;
		}
		used += k;
		apply_op(*op, x, w.data(), g);
		std::vector<double> tres(b, b + N);
		axpy(-1.0, w.data(), tres.data(), N);
		tr = std::sqrt(dot(tres.data(), tres.data(), N));
		if (reason != CONVERGED || tr < tol || used >= max_iter) break;
		if (!(tr < 0.5 * prev_tr)) { at_floor = true; break; }
		prev_tr = tr;
	}
	if (reason == CONVERGED && !(tr < tol) && !at_floor) reason = MAX_ITERATIONS;
	set_stats(st, used, rn, reason);
	return OK;
}
}  // namespace

extern "C" {

// Preconditioned pipelined CG (EXTENSION; PipelinedCgSolver has no preconditioned form): Ghysels &
// Vanroose 2014 Alg. 3 for an exact preconditioner; an FP32 multigrid gets pfcg above.
int orc_ppcg(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op, int pc_kind,
             const orc_mg_cfg* cfg, double tol, int64_t max_iter, orc_stats* st, double* hist) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	int64_t N = g.len();
	Pc pc(pc_kind, g, cfg);
	if (pc.m32) return pfcg(x, b, g, op, pc, tol, max_iter, st, hist);
	std::vector<double> r(N), u(N), w(N), m(N), nn(N), z(N), q(N), s(N), p(N);
	apply_op(*op, x, s.data(), g);
	copy(b, r.data(), N);
	axpy(-1.0, s.data(), r.data(), N);
	fill(u.data(), 0.0, N);
	pc.apply(u.data(), r.data(), N);
	apply_op(*op, u.data(), w.data(), g);
	fill(z.data(), 0, N); fill(q.data(), 0, N); fill(s.data(), 0, N); fill(p.data(), 0, N);
	double gamma_prev = 0, alpha = 0;
	for (int64_t i = 0; i < max_iter; ++i) {
		double gamma = dot(r.data(), u.data(), N), delta = dot(w.data(), u.data(), N);
		double rn = std::sqrt(dot(r.data(), r.data(), N));
		if (hist) hist[i] = rn;
		if (rn < tol) { set_stats(st, i, rn, CONVERGED); return OK; }
		fill(m.data(), 0.0, N);
		pc.apply(m.data(), w.data(), N);
		apply_op(*op, m.data(), nn.data(), g);
		double beta;
		if (i > 0) { beta = gamma / gamma_prev; alpha = gamma / (delta - beta * gamma / alpha); }
		else { beta = 0.0; alpha = gamma / delta; }
		if (!std::isfinite(alpha)) { set_stats(st, i, rn, BREAKDOWN); return OK; }
		for (int64_t j = 0; j < N; ++j) {
			z[j] = nn[j] + beta * z[j];
			q[j] = m[j] + beta * q[j];
			s[j] = w[j] + beta * s[j];
			p[j] = u[j] + beta * p[j];
			x[j] = std::fma(alpha, p[j], x[j]);
			r[j] = std::fma(-alpha, s[j], r[j]);
			u[j] = std::fma(-alpha, q[j], u[j]);
			w[j] = std::fma(-alpha, z[j], w[j]);
		}
		gamma_prev = gamma;
	}
	set_stats(st, max_iter, std::sqrt(dot(r.data(), r.data(), N)), MAX_ITERATIONS);
	return OK;
}

// ── BiCgStabSolver — bicgstab.rs:90-346 (pc_kind 0 = solve(), else
// solve_preconditioned with right preconditioning) ─────────────────────
int orc_bicgstab(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op,
                 int pc_kind, const orc_mg_cfg* cfg, double tol, int64_t max_iter, orc_stats* st, double* hist) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	int64_t n = g.len();
	bool pre = pc_kind != 0;
	Pc pc(pc_kind, g, cfg);
	std::vector<double> r(n), rh(n), p(n), v(n), s(n), t(n), ph(n), sh(n);
	if (int e = apply_op(*op, x, v.data(), g)) return e;
	copy(b, r.data(), n);
	axpy(-1.0, v.data(), r.data(), n);
	copy(r.data(), rh.data(), n);
	double rho = 1.0, alpha = 1.0, omega = 1.0;
	fill(v.data(), 0.0, n);
	fill(p.data(), 0.0, n);
	for (int64_t i = 0; i < max_iter; ++i) {
		double r_norm = std::sqrt(dot(r.data(), r.data(), n));
		if (hist) hist[i] = r_norm;
		if (r_norm < tol) { set_stats(st, i, r_norm, CONVERGED); return OK; }
		double rho_new = dot(rh.data(), r.data(), n);
		if (std::fabs(rho_new) < 1e-30) { set_stats(st, i, r_norm, BREAKDOWN); return OK; }
		double beta = (rho_new / rho) * (alpha / omega);
		axpy(-omega, v.data(), p.data(), n);
		scale(p.data(), beta, n);
		axpy(1.0, r.data(), p.data(), n);
		const double* pa = p.data();
		if (pre) { fill(ph.data(), 0.0, n); pc.apply(ph.data(), p.data(), n); pa = ph.data(); }
		apply_op(*op, pa, v.data(), g);
		double rv = dot(rh.data(), v.data(), n);
		if (std::fabs(rv) < 1e-30) { set_stats(st, i, r_norm, BREAKDOWN); return OK; }
		alpha = rho_new / rv;
		copy(r.data(), s.data(), n);
		axpy(-alpha, v.data(), s.data(), n);
		double s_norm = std::sqrt(dot(s.data(), s.data(), n));
		if (s_norm < tol) { axpy(alpha, pa, x, n); set_stats(st, i, s_norm, CONVERGED); return OK; }
		const double* sa = s.data();
		if (pre) { fill(sh.data(), 0.0, n); pc.apply(sh.data(), s.data(), n); sa = sh.data(); }
		apply_op(*op, sa, t.data(), g);
		double ts = dot(t.data(), s.data(), n);
		double tt = dot(t.data(), t.data(), n);
		if (std::fabs(tt) < 1e-30) { axpy(alpha, pa, x, n); set_stats(st, i, s_norm, BREAKDOWN); return OK; }
		omega = ts / tt;
		axpy(alpha, pa, x, n);
		axpy(omega, sa, x, n);
		copy(s.data(), r.data(), n);
		axpy(-omega, t.data(), r.data(), n);
		rho = rho_new;
	}
	set_stats(st, max_iter, std::sqrt(dot(r.data(), r.data(), n)), MAX_ITERATIONS);
	return OK;
}

// ── GmresSolver — gmres.rs:116-490 (left preconditioning when pc_kind != 0)
}  // extern "C"

namespace {
// GmresSolver::solve / solve_preconditioned (gmres.rs:116-490; left preconditioning), or with
// `flex` the EXTENSION flexible right-preconditioned GMRES(m) (Saad 1993): the cycle starts from the
// unpreconditioned residual, every Arnoldi step applies z_k = M⁻¹v_k (stored) and then A, and the
// correction is x += Σ y_i·z_i, so |g_{k+1}| estimates ‖b − A·x_k‖ whatever M does.
int gmres_impl(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op,
               int64_t restart, int pc_kind, const orc_mg_cfg* cfg, double tol, int64_t max_iter, orc_stats* st,
               double* hist, bool flex) {
	Grid g = mkgrid(rows, cols, dx, dy);
	if (int e = valid_grid(g)) return e;
	int64_t n = g.len(), m = restart;
	bool pre = pc_kind != 0 && !flex;
	Pc pc(pc_kind, g, cfg);
	std::vector<std::vector<double>> vv(m + 1, std::vector<double>(n));
	std::vector<std::vector<double>> zz(flex ? m : 0, std::vector<double>(n));
	std::vector<double> w(n), hh((m + 1) * m), cc(m), ss(m), grs(m + 1);
	auto H = [&](int64_t i, int64_t j) -> double& { return hh[j * (m + 1) + i]; };
	// gmres.rs:461-490.
	auto build = [&](int64_t k) {
		if (k == 0) return;
		std::vector<double> y(k, 0.0);
		for (int64_t i = k - 1; i >= 0; --i) {
			double s = grs[i];
			for (int64_t j = i + 1; j < k; ++j) s -= hh[j * (m + 1) + i] * y[j];
			double diag = hh[i * (m + 1) + i];
			if (std::fabs(diag) < 1e-30) break;
			y[i] = s / diag;
		}
		for (int64_t i = 0; i < k; ++i) axpy(y[i], flex ? zz[i].data() : vv[i].data(), x, n);
	};
	int64_t total = 0;
	for (;;) {
		double beta;
		if (!pre) {
			if (int e = apply_op(*op, x, vv[0].data(), g)) return e;
			scale(vv[0].data(), -1.0, n);
			axpy(1.0, b, vv[0].data(), n);
		} else {
			if (int e = apply_op(*op, x, w.data(), g)) return e;
			scale(w.data(), -1.0, n);
			axpy(1.0, b, w.data(), n);
			fill(vv[0].data(), 0.0, n);
			pc.apply(vv[0].data(), w.data(), n);
		}
		beta = std::sqrt(dot(vv[0].data(), vv[0].data(), n));
		if (beta < tol) { set_stats(st, total, beta, CONVERGED); return OK; }
		scale(vv[0].data(), 1.0 / beta, n);
		std::fill(grs.begin(), grs.end(), 0.0);
		grs[0] = beta;
		std::fill(hh.begin(), hh.end(), 0.0);
		double cur = beta;
		int64_t k = 0;
		while (k < m) {
			if (total >= max_iter) { build(k); set_stats(st, total, cur, MAX_ITERATIONS); return OK; }
			if (hist) hist[total] = cur;
			if (flex) {
				fill(zz[k].data(), 0.0, n);
				pc.apply(zz[k].data(), vv[k].data(), n);
				apply_op(*op, zz[k].data(), w.data(), g);
			} else {
				apply_op(*op, vv[k].data(), w.data(), g);
			}
			if (pre) {
				fill(vv[k + 1].data(), 0.0, n);
				pc.apply(vv[k + 1].data(), w.data(), n);
				copy(vv[k + 1].data(), w.data(), n);
			}
			for (int64_t j = 0; j <= k; ++j) {
				double h = dot(w.data(), vv[j].data(), n);
				H(j, k) = h;
				axpy(-h, vv[j].data(), w.data(), n);
			}
			double hk = std::sqrt(dot(w.data(), w.data(), n));
			H(k + 1, k) = hk;
			if (hk > 1e-30) { copy(w.data(), vv[k + 1].data(), n); scale(vv[k + 1].data(), 1.0 / hk, n); }
			for (int64_t j = 0; j < k; ++j) {
				double hj = H(j, k), hj1 = H(j + 1, k), cj = cc[j], sj = ss[j];
				H(j, k) = std::fma(cj, hj, sj * hj1);
				H(j + 1, k) = std::fma(cj, hj1, -sj * hj);
			}
			double a = H(k, k), bv = H(k + 1, k);
			double rr = std::hypot(a, bv);
			if (std::fabs(rr) < 1e-30) { build(k); set_stats(st, total, cur, BREAKDOWN); return OK; }
			cc[k] = a / rr;
			ss[k] = bv / rr;
			H(k, k) = rr;
			H(k + 1, k) = 0.0;
			double gk = grs[k];
			grs[k] = cc[k] * gk;
			grs[k + 1] = -ss[k] * gk;
			cur = std::fabs(grs[k + 1]);
			++total;
			++k;
			if (cur < tol) break;
			if (hk <= 1e-30) break;
		}
		build(k);
		if (cur < tol) { set_stats(st, total, cur, CONVERGED); return OK; }
		if (total >= max_iter) { set_stats(st, total, cur, MAX_ITERATIONS); return OK; }
	}
}
}  // namespace

extern "C" {

int orc_gmres(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op,
              int64_t restart, int pc_kind, const orc_mg_cfg* cfg, double tol, int64_t max_iter, orc_stats* st,
              double* hist) {
	return gmres_impl(x, b, rows, cols, dx, dy, op, restart, pc_kind, cfg, tol, max_iter, st, hist, false);
}

int orc_fgmres(double* x, const double* b, int64_t rows, int64_t cols, double dx, double dy, const orc_op* op,
               int64_t restart, int pc_kind, const orc_mg_cfg* cfg, double tol, int64_t max_iter, orc_stats* st,
               double* hist) {
	return gmres_impl(x, b, rows, cols, dx, dy, op, restart, pc_kind, cfg, tol, max_iter, st, hist, true);
}

int orc_num_threads() { return omp_get_max_threads(); }

}  // extern "C"
