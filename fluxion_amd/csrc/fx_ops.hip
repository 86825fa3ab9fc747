// fx_ops.hip — the Backend operation set (crates/core/src/backend.rs:43-406)
// as gfx950 kernels behind the C ABI.
//
// Arithmetic follows the reference CPU backend (crates/backend-cpu/src/ops.rs)
// operation by operation, so element-wise and stencil results are bit-identical
// to it; reductions use a deterministic compensated tree instead of rayon's
// Kahan chunks.
#include "fx_internal.hpp"

#include <cmath>
#include <mutex>

using namespace fxd;

namespace {

// ── element-wise ─────────────────────────────────────────────────────
template <class F>
__global__ __launch_bounds__(BLOCK) void k_elem(F f, size_t n) {
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) f(i);
}

template <class F>
fx_status launch_elem(fx_backend* be, F f, size_t n, const char* what) {
	if (n == 0) return FX_OK;
	hipLaunchKernelGGL(k_elem<F>, dim3(fx::elem_blocks(n)), dim3(BLOCK), 0, be->stream, f, n);
	return fx::check_launch(what);
}

struct CopyF { const double* s; double* d; __device__ void operator()(size_t i) const { d[i] = s[i]; } };
struct FillF { double* d; double v; __device__ void operator()(size_t i) const { d[i] = v; } };
struct FillF32 { float* d; float v; __device__ void operator()(size_t i) const { d[i] = v; } };
// ops.rs:203: y = alpha.mul_add(x, y).
struct AxpyF { double a; const double* x; double* y; __device__ void operator()(size_t i) const { y[i] = fma(a, x[i], y[i]); } };
// ops.rs:215-221.
struct ScaleF { double* x; double a; __device__ void operator()(size_t i) const { x[i] *= a; } };
// ops.rs:551-555: w = alpha.mul_add(x, beta*y).
struct WaxpyF { double a; const double* x; double b; const double* y; double* w;
	__device__ void operator()(size_t i) const { w[i] = fma(a, x[i], b * y[i]); } };
// ops.rs:578-582: y = alpha.mul_add(y, x).
struct AypxF { double a; const double* x; double* y; __device__ void operator()(size_t i) const { y[i] = fma(a, y[i], x[i]); } };
// ops.rs:476-528, 650-672 (f64::max / min ignore NaN like fmax / fmin).
struct PwF { int op; const double* x; const double* y; double* z;
	__device__ void operator()(size_t i) const {
		double a = x[i], b = y[i];
		z[i] = op == 0 ? a * b : op == 1 ? a / b : op == 2 ? fmax(a, b) : fmin(a, b);
	} };
struct RecipF { double* x; __device__ void operator()(size_t i) const { x[i] = 1.0 / x[i]; } };
struct AbsF { double* x; __device__ void operator()(size_t i) const { x[i] = fabs(x[i]); } };
struct MulF { double* x; const double* c; __device__ void operator()(size_t i) const { x[i] *= c[i]; } };
// pipelined_cg.rs:182-199 (CPU order: scale then axpy(1, ·), updated vectors).
struct PcgFusedF {
	double a, b;
	double *z, *t, *p, *x, *r, *w;
	const double* q;
	__device__ void operator()(size_t i) const {
		double zi = q[i] + b * z[i];
		double ti = w[i] + b * t[i];
		double pi = r[i] + b * p[i];
		x[i] = fma(a, pi, x[i]);
		r[i] = fma(-a, ti, r[i]);
		w[i] = fma(-a, zi, w[i]);
		z[i] = zi;
		t[i] = ti;
		p[i] = pi;
	}
};

// ── reductions ───────────────────────────────────────────────────────
template <int NS, class T>
__global__ __launch_bounds__(BLOCK) void k_reduce(T term, size_t n, RedScratch rs) {
	Acc acc[NS];
#pragma unroll
	for (int k = 0; k < NS; ++k) acc[k] = Acc{0.0, 0.0};
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
		double v[NS];
		term(i, v);
#pragma unroll
		for (int k = 0; k < NS; ++k) acc_add(acc[k], v[k]);
	}
	block_partial<NS>(acc, rs);
}
// Writes the NS global sums to out[0..NS).
template <int NS>
struct FinStore {
	double* out;
	__device__ void operator()(Acc* a) const {
#pragma unroll
		for (int k = 0; k < NS; ++k) out[k] = acc_value(a[k]);
	}
};
// ops.rs:341-356: prod = x*y (no fma).
struct DotT { const double* x; const double* y; __device__ void operator()(size_t i, double* v) const { v[0] = x[i] * y[i]; } };
struct Dot2T { const double *a, *b, *c, *d;
	__device__ void operator()(size_t i, double* v) const { v[0] = a[i] * b[i]; v[1] = c[i] * d[i]; } };
struct SumT { const double* x; __device__ void operator()(size_t i, double* v) const { v[0] = x[i]; } };
struct SqT { const double* x; __device__ void operator()(size_t i, double* v) const { v[0] = x[i] * x[i]; } };

// max / min (ops.rs:318-331, 587-601): NaN-ignoring, identities ∓inf.
// Block results go to partials[b].hi; k_minmax_finish combines them.
template <int MAX>
__device__ __forceinline__ double mm_block(double m) {
	for (int off = WAVE / 2; off > 0; off >>= 1) {
		double o = __shfl_xor(m, off);
		m = MAX ? fmax(m, o) : fmin(m, o);
	}
	__shared__ double sh[NWAVES];
	if ((threadIdx.x & (WAVE - 1)) == 0) sh[threadIdx.x / WAVE] = m;
	__syncthreads();
	if (threadIdx.x == 0)
		for (int w = 1; w < NWAVES; ++w) m = MAX ? fmax(m, sh[w]) : fmin(m, sh[w]);
	return m;
}
template <int MAX>
__global__ __launch_bounds__(BLOCK) void k_minmax(const double* x, size_t n, RedScratch rs) {
	double m = MAX ? -INFINITY : INFINITY;
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) m = MAX ? fmax(m, x[i]) : fmin(m, x[i]);
	m = mm_block<MAX>(m);
	if (threadIdx.x == 0) rs.partials[blockIdx.x].hi = m;
}
template <int MAX>
__global__ __launch_bounds__(BLOCK) void k_minmax_finish(RedScratch rs, unsigned nb, double* out) {
	double m = MAX ? -INFINITY : INFINITY;
	for (unsigned b = threadIdx.x; b < nb; b += BLOCK) {
		double p = rs.partials[b].hi;
		m = MAX ? fmax(m, p) : fmin(m, p);
	}
	m = mm_block<MAX>(m);
	if (threadIdx.x == 0) *out = m;
}

template <int NS, class T>
fx_status reduce_host(fx_backend* be, T term, size_t n, double* out, const char* what) {
	unsigned nb = fx::red_blocks(n);
	FX_CHECK(fx::ensure_partials(be, (size_t)nb * NS));
	fx::probe_begin(be, FX_PROBE_REDUCE);
	hipLaunchKernelGGL((k_reduce<NS, T>), dim3(nb), dim3(BLOCK), 0, be->stream, term, n, fx::red_scratch(be));
	fx::probe_end(be, FX_PROBE_REDUCE);
	hipLaunchKernelGGL((k_finish<NS, FinStore<NS>>), dim3(1), dim3(FIN_BLOCK), 0, be->stream, fx::red_scratch(be), nb,
	                   FinStore<NS>{be->dres}, (const int*)nullptr);
	FX_CHECK(fx::check_launch(what));
	FX_HIP(hipMemcpyAsync(be->hres, be->dres, sizeof(double) * NS, hipMemcpyDeviceToHost, be->stream));
	FX_HIP(hipStreamSynchronize(be->stream));
	for (int k = 0; k < NS; ++k) out[k] = be->hres[k];
	return FX_OK;
}

// ── stencils ─────────────────────────────────────────────────────────
struct LoadX { const double* x; __device__ double load(int64_t i) const { return x[i]; } };

// Epilogues take one point (interior / boundary, from k_stream5 and k_point3)
// or a pair of adjacent points i, i + 1 with i even (pair, from k_stream5d:
// ia / ib say which of the two are interior, la / lb are their stencil sums and
// c their values), which reads and writes its vectors 16 bytes at a time.
// PAIR says whether the pair form is worth it: it is for epilogues that write a
// second buffer; stencil_axpy's read-modify-write of y measured no faster with it
// (62 µs either way at 4096², the rate of a copy), so it keeps the per-point form.
//
// y += alpha·lap on the interior (ops.rs:1104-1106); boundary untouched.
struct EpiAxpy {
	static constexpr bool PAIR = false;
	double alpha; double* y;
	template <class A> __device__ void interior(int64_t i, double lap, double, A*) const { y[i] = fma(alpha, lap, y[i]); }
	template <class A> __device__ void boundary(int64_t, double, A*) const {}
};
// out = lap on the interior (boundary handled by k_bc_perimeter, and not written here).
struct EpiStore {
	static constexpr bool PAIR = true;
	double* y;
	template <class A> __device__ void interior(int64_t i, double lap, double, A*) const { y[i] = lap; }
	template <class A> __device__ void boundary(int64_t, double, A*) const {}
	__device__ void pair(int64_t i, bool ia, double la, bool ib, double lb, double2) const {
		if (ia && ib) {
			*(double2*)(y + i) = double2{la, lb};
		} else {
			if (ia) y[i] = la;
			if (ib) y[i + 1] = lb;
		}
	}
};
// A·x = copy + stencil_axpy (cg.rs:104-111) in one pass: fma(coeff, lap, x); x on the boundary.
struct EpiOp {
	static constexpr bool PAIR = true;
	double coeff; double* y;
	template <class A> __device__ void interior(int64_t i, double lap, double c, A*) const { y[i] = fma(coeff, lap, c); }
	template <class A> __device__ void boundary(int64_t i, double c, A*) const { y[i] = c; }
	__device__ void pair(int64_t i, bool ia, double la, bool ib, double lb, double2 c) const {
		*(double2*)(y + i) = double2{ia ? fma(coeff, la, c.x) : c.x, ib ? fma(coeff, lb, c.y) : c.y};
	}
};

// One explicit diffusion stage into a second buffer: t = x + coeff·lap inside
// (stencil_axpy's fma on a copy of x), x on the boundary; then t += dt·src
// (axpy) and, for the second stage of Heun's method, t = (t + avg)·0.5 (axpy
// with 1 then scale by 0.5).  Each step is the reference's own operation in its
// order (runtime/src/solver.rs:344-418), so the values are identical; only the
// copies between them are gone.
struct EpiStage {
	static constexpr bool PAIR = true;
	double coeff; double* y; const double* src; double dt; const double* avg;
	__device__ void put(int64_t i, double t) const {
		if (src) t = fma(dt, src[i], t);
		if (avg) t = fma(1.0, avg[i], t) * 0.5;
		y[i] = t;
	}
	template <class A> __device__ void interior(int64_t i, double lap, double c, A*) const { put(i, fma(coeff, lap, c)); }
	template <class A> __device__ void boundary(int64_t i, double c, A*) const { put(i, c); }
	__device__ void pair(int64_t i, bool ia, double la, bool ib, double lb, double2 c) const {
		double2 t{ia ? fma(coeff, la, c.x) : c.x, ib ? fma(coeff, lb, c.y) : c.y};
		if (src) {
			const double2 s2 = *(const double2*)(src + i);
			t.x = fma(dt, s2.x, t.x);
			t.y = fma(dt, s2.y, t.y);
		}
		if (avg) {
			const double2 a2 = *(const double2*)(avg + i);
			t.x = fma(1.0, a2.x, t.x) * 0.5;
			t.y = fma(1.0, a2.y, t.y) * 0.5;
		}
		*(double2*)(y + i) = t;
	}
};

// Classical RK4 (solver.rs:421-494) in four stencil passes.  The reference
// keeps u_old, forms each stage in u and copies u_old back; here u stays u_old
// until the last pass and the stages go to two scratch buffers.  Per point the
// operations are the reference's, in its order: k = S(stage) inside (the
// boundary values of k come from the BC perimeter kernel, as apply_stencil's);
// stage 2 = u_old + ½dtα·S(u_old) inside and u_old on the boundary (copy +
// stencil_axpy); stages 3, 4 = u_old + coeff·k inside and 0 on the boundary
// (copy, axpy, zero_boundaries); u = u_old + c·k1 + 2c·k2 + 2c·k3 + c·k4
// (+ dt·source), one fma per axpy.
struct EpiRk1 {
	static constexpr bool PAIR = true;
	double h; double* k; double* st;
	template <class A> __device__ void interior(int64_t i, double lap, double c, A*) const {
		k[i] = lap;
		st[i] = fma(h, lap, c);
	}
	template <class A> __device__ void boundary(int64_t i, double c, A*) const { st[i] = c; }
	__device__ void pair(int64_t i, bool ia, double la, bool ib, double lb, double2 c) const {
		*(double2*)(st + i) = double2{ia ? fma(h, la, c.x) : c.x, ib ? fma(h, lb, c.y) : c.y};
		EpiStore{k}.pair(i, ia, la, ib, lb, c);
	}
};
struct EpiRkMid {
	static constexpr bool PAIR = true;
	double h; double* k; double* st; const double* uold;
	template <class A> __device__ void interior(int64_t i, double lap, double, A*) const {
		k[i] = lap;
		st[i] = fma(h, lap, uold[i]);
	}
	template <class A> __device__ void boundary(int64_t i, double, A*) const { st[i] = 0.0; }
	__device__ void pair(int64_t i, bool ia, double la, bool ib, double lb, double2 c) const {
		const double2 u = *(const double2*)(uold + i);
		*(double2*)(st + i) = double2{ia ? fma(h, la, u.x) : 0.0, ib ? fma(h, lb, u.y) : 0.0};
		EpiStore{k}.pair(i, ia, la, ib, lb, c);
	}
};
__device__ __forceinline__ double rk4_comb(double c, double u, double k1, double k2, double k3, double k4) {
	return fma(c, k4, fma(2.0 * c, k3, fma(2.0 * c, k2, fma(c, k1, u))));
}
struct EpiRkLast {
	static constexpr bool PAIR = true;
	double c; double* u; const double* k1; const double* k2; const double* k3; const double* src; double dt;
	template <class A> __device__ void interior(int64_t i, double lap, double, A*) const {
		double t = rk4_comb(c, u[i], k1[i], k2[i], k3[i], lap);
		if (src) t = fma(dt, src[i], t);
		u[i] = t;
	}
	template <class A> __device__ void boundary(int64_t, double, A*) const {}
	__device__ void pair(int64_t i, bool ia, double la, bool ib, double lb, double2) const {
		if (!ia && !ib) return;
		double2 t = *(const double2*)(u + i);
		const double2 a = *(const double2*)(k1 + i), b = *(const double2*)(k2 + i), d = *(const double2*)(k3 + i);
		if (ia) t.x = rk4_comb(c, t.x, a.x, b.x, d.x, la);
		if (ib) t.y = rk4_comb(c, t.y, a.y, b.y, d.y, lb);
		if (src) {
			const double2 s2 = *(const double2*)(src + i);
			if (ia) t.x = fma(dt, s2.x, t.x);
			if (ib) t.y = fma(dt, s2.y, t.y);
		}
		*(double2*)(u + i) = t;
	}
};
// The last pass on the boundary: u = u_old + c·k1 + … + c·k4 (+ dt·source) with the k's perimeter values.
__global__ void k_rk4_edge(double* u, const double* k1, const double* k2, const double* k3, const double* k4,
                           const double* src, double c, double dt, int64_t rows, int64_t cols) {
	const int64_t per = 2 * cols + 2 * (rows - 2);
	for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < per; q += (int64_t)gridDim.x * blockDim.x) {
		int64_t i;
		if (q < cols) i = q;                                         // top row
		else if (q < 2 * cols) i = (rows - 1) * cols + (q - cols);   // bottom row
		else {
			const int64_t j = q - 2 * cols, r = 1 + j / 2;           // left and right columns, rows 1 … rows − 2
			i = r * cols + (j % 2 ? cols - 1 : 0);
		}
		double t = rk4_comb(c, u[i], k1[i], k2[i], k3[i], k4[i]);
		if (src) t = fma(dt, src[i], t);
		u[i] = t;
	}
}

// Per-point radius-1 kernel for 9-entry and generic stencils.
template <class E>
__global__ __launch_bounds__(BLOCK) void k_point3(const double* x, E e, W9 w, Ent ent, int mode, int64_t rows, int64_t cols,
                                                  const int* skip) {
	if (skip && *skip) return;
	int64_t c = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= rows || c >= cols) return;
	int64_t i = r * cols + c;
	if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) {
		e.boundary(i, x[i], (Acc*)nullptr);
		return;
	}
	double lap = mode == 9 ? chain9(w, x, i, cols) : chain_generic(ent, x, i, cols);
	e.interior(i, lap, x[i], (Acc*)nullptr);
}

template <class E>
fx_status launch_stencil(fx_backend* be, const double* x, E e, const fx_stencil* st, int64_t rows, int64_t cols,
                         const char* what, const int* skip = nullptr) {
	W9 w;
	Ent ent;
	FX_CHECK(fx::stencil_args(st, &w, &ent));
	if constexpr (E::PAIR) {
		if (st->n == 5 && cols % 2 == 0) {
			// Two columns per thread (16-byte rows; the row starts stay 16-byte aligned).
			const int rpb = stream_rpb(rows, cols / 2);
			hipLaunchKernelGGL(k_stream5d<E>, fx::stream_grid(rows, cols / 2, rpb), dim3(BLOCK), 0, be->stream, x, e,
			                   w, rows, cols, rpb, skip);
			return fx::check_launch(what);
		}
	}
	if (st->n == 5) {
		int rpb = stream_rpb(rows, cols);
		hipLaunchKernelGGL((k_stream5<0, LoadX, E, NoFinal>), fx::stream_grid(rows, cols, rpb), dim3(BLOCK), 0,
		                   be->stream, LoadX{x}, e, NoFinal{}, w, rows, cols, rpb, RedScratch{nullptr}, skip);
	} else {
		dim3 grid((unsigned)((cols + WAVE - 1) / WAVE), (unsigned)((rows + NWAVES - 1) / NWAVES));
		hipLaunchKernelGGL(k_point3<E>, grid, dim3(BLOCK), 0, be->stream, x, e, w, ent, st->n == 9 ? 9 : 0, rows, cols,
		                   skip);
	}
	return fx::check_launch(what);
}

// bc_value (ops.rs:1141-1180).
__device__ double bc_value(fx_bc bc, double self_val, double nb, double spacing) {
	switch (bc.kind) {
	case FX_BC_DIRICHLET: return bc.a;
	case FX_BC_NEUMANN: return fma(spacing, bc.a, nb);
	case FX_BC_ROBIN: {
		if (fabs(bc.b) < 1e-30) {
			if (fabs(bc.a) < 1e-30) return 0.0;
			return bc.g / bc.a;
		}
		double du_dn = (bc.g - bc.a * self_val) / bc.b;
		return fma(spacing, du_dn, nb);
	}
	default: return 0.0;
	}
}

// apply_boundaries (ops.rs:799-843): top/bottom rows own the corners.
// If `coeff` is set the value is multiplied by it (apply_stencil_var, ops.rs:140-152).
__global__ void k_bc_perimeter(const double* src, double* dst, const double* coeff, fx_boundaries bc, int64_t rows,
                               int64_t cols, double dx, double dy) {
	int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	int64_t lr = rows - 1, lc = cols - 1;
	int64_t i;
	double v;
	if (t < cols) {
		i = t;
		v = bc_value(bc.top, src[t], src[cols + t], dy);
	} else if (t < 2 * cols) {
		int64_t c = t - cols;
		i = lr * cols + c;
		v = bc_value(bc.bottom, src[i], src[(lr - 1) * cols + c], dy);
	} else if (t < 2 * cols + (rows - 2)) {
		int64_t r = 1 + (t - 2 * cols);
		i = r * cols;
		v = bc_value(bc.left, src[i], src[i + 1], dx);
	} else if (t < 2 * cols + 2 * (rows - 2)) {
		int64_t r = 1 + (t - 2 * cols - (rows - 2));
		i = r * cols + lc;
		v = bc_value(bc.right, src[i], src[i - 1], dx);
	} else {
		return;
	}
	dst[i] = coeff ? v * coeff[i] : v;
}

// conv-diff (ops.rs:674-791).  mode 0: apply (boundary 0); 1: axpy; 2: operator
// fma(alpha, diff+conv, sigma·u) with u on the boundary.
__global__ __launch_bounds__(BLOCK) void k_conv_diff(int mode, double alpha, double sigma, const double* u, double* out,
                                                     const double* kappa, double ks, const double* vx, const double* vy,
                                                     int64_t rows, int64_t cols, double inv_dx, double inv_dy,
                                                     const int* skip) {
	if (skip && *skip) return;
	int64_t c = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= rows || c >= cols) return;
	int64_t i = r * cols + c;
	if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) {
		if (mode == 0) out[i] = 0.0;
		else if (mode == 2) out[i] = u[i];
		return;
	}
	const double inv_dx2 = inv_dx * inv_dx, inv_dy2 = inv_dy * inv_dy;
	double cv = u[i];
	double lap = fma(u[i - 1] + u[i + 1], inv_dx2, (u[i - cols] + u[i + cols]) * inv_dy2) - 2.0 * (inv_dx2 + inv_dy2) * cv;
	double diff = (kappa ? kappa[i] : ks) * lap;
	double vxi = vx[i], vyi = vy[i];
	double dudx = vxi >= 0.0 ? (cv - u[i - 1]) * inv_dx : (u[i + 1] - cv) * inv_dx;
	double dudy = vyi >= 0.0 ? (cv - u[i - cols]) * inv_dy : (u[i + cols] - cv) * inv_dy;
	double conv = fma(vxi, dudx, vyi * dudy);
	if (mode == 0) out[i] = diff + conv;
	else if (mode == 1) out[i] = fma(alpha, diff + conv, out[i]);
	else out[i] = fma(alpha, diff + conv, sigma * cv);
}

// Variable-coefficient operator sigma·u + alpha·(−∇·(k∇u)), harmonic face means.
__device__ __forceinline__ double hmean(double a, double b) { return 2.0 * a * b / (a + b); }
__global__ __launch_bounds__(BLOCK) void k_vark(double sigma, double alpha, const double* k, const double* u, double* out,
                                                int64_t rows, int64_t cols, double inv_dx2, double inv_dy2, const int* skip) {
	if (skip && *skip) return;
	int64_t c = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= rows || c >= cols) return;
	int64_t i = r * cols + c;
	if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) {
		out[i] = u[i];
		return;
	}
	double kc = k[i], cv = u[i];
	double kw = hmean(kc, k[i - 1]), ke = hmean(kc, k[i + 1]);
	double kn = hmean(kc, k[i - cols]), ksv = hmean(kc, k[i + cols]);
	double fx = ke * (u[i + 1] - cv) - kw * (cv - u[i - 1]);
	double fy = ksv * (u[i + cols] - cv) - kn * (cv - u[i - cols]);
	double L = -(fx * inv_dx2 + fy * inv_dy2);
	out[i] = fma(alpha, L, sigma * cv);
}

// k_vark on the two-column traversal (even column counts): the same per-point
// arithmetic, identity boundary.
struct EpiOpVk {
	double sigma, alpha, ix2, iy2;
	double* out;
	template <class A>
	__device__ void pair(int64_t i, int64_t, int64_t, bool rin, bool ia, bool ib, const Nb& uA, const Nb& uB,
	                     const Nb& kA, const Nb& kB, A*) const {
		double2 o{uA.c, uB.c};
		double fs;
		if (rin && ia) o.x = fma(alpha, vk_lap(uA, kA, ix2, iy2, &fs), sigma * uA.c);
		if (rin && ib) o.y = fma(alpha, vk_lap(uB, kB, ix2, iy2, &fs), sigma * uB.c);
		*(double2*)(out + i) = o;
	}
};

// Multigrid transfer (ops.rs:369-423); prolong clamps the coarse index for
// odd fine sizes where the reference would index past the coarse array.
__global__ void k_restrict_ref(const double* f, double* c, int64_t fr, int64_t fc, int64_t cr, int64_t cc) {
	int64_t col = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= cr || col >= cc) return;
	int64_t a = 2 * r, b = 2 * col;
	int64_t a1 = a + 1 < fr - 1 ? a + 1 : fr - 1, b1 = b + 1 < fc - 1 ? b + 1 : fc - 1;
	c[r * cc + col] = 0.25 * (f[a * fc + b] + f[a1 * fc + b] + f[a * fc + b1] + f[a1 * fc + b1]);
}
__global__ void k_prolong_ref(const double* c, double* f, int64_t cr, int64_t cc, int64_t fr, int64_t fc) {
	int64_t col = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= fr || col >= fc) return;
	int64_t a = r / 2 < cr - 1 ? r / 2 : cr - 1, b = col / 2 < cc - 1 ? col / 2 : cc - 1;
	f[r * fc + col] += c[a * cc + b];
}

// Two-buffer damped Jacobi point update with the ops.rs:454-469 arithmetic.
__global__ __launch_bounds__(BLOCK) void k_jacobi(const double* xo, double* xn, const double* b, double omega,
                                                  double inv_diag, Ent ent, int64_t rows, int64_t cols) {
	int64_t c = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (r >= rows || c >= cols) return;
	int64_t i = r * cols + c;
	if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) {
		xn[i] = xo[i];
		return;
	}
	double ax = 0.0;
	for (int j = 0; j < ent.n; ++j) ax = fma(ent.w[j], xo[i + ent.dr[j] * cols + ent.dc[j]], ax);
	double res = b[i] - ax;
	xn[i] = fma(omega, inv_diag * res, xo[i]);
}

__global__ void k_zero_bnd(double* x, int64_t rows, int64_t cols) {
	int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < cols) { x[t] = 0.0; x[(rows - 1) * cols + t] = 0.0; }
	else if (t < cols + rows - 2) { int64_t r = 1 + (t - cols); x[r * cols] = 0.0; x[r * cols + cols - 1] = 0.0; }
}

dim3 pgrid(int64_t rows, int64_t cols) {
	return dim3((unsigned)((cols + WAVE - 1) / WAVE), (unsigned)((rows + NWAVES - 1) / NWAVES));
}

fx_status len_eq(const fx_buf* a, const fx_buf* b, const char* na, const char* nb) {
	FX_REQUIRE(a && b && a->ptr && b->ptr, FX_ERR_ARGUMENT, "null buffer (%s/%s)", na, nb);
	FX_REQUIRE(a->dtype == FX_F64 && b->dtype == FX_F64, FX_ERR_ARGUMENT, "fp64 buffers required (%s/%s)", na, nb);
	FX_REQUIRE(a->len == b->len, FX_ERR_DIMENSION, "dimension mismatch: expected %zu, got %zu", a->len, b->len);
	return FX_OK;
}
#define D(b) ((double*)(b)->ptr)
#define CD(b) ((const double*)(b)->ptr)

// ── static scratch of fx_dot_noalloc ──────────────────────────────────
// The block partials of the dot live in this file-scope device array instead
// of the backend's allocated partials buffer.  Each GPU holds its own copy for
// the life of the process.  In host code `scratch` names a host-side shadow,
// not device memory: kernels take the per-device address found by
// noalloc_setup() instead.  red_blocks() never launches more than 2048
// workgroups, so every input length fits in its 2048 · 16 bytes = 32 KiB.
constexpr unsigned NOALLOC_PARTIALS = 2048;
static __device__ __align__(256) Acc scratch[NOALLOC_PARTIALS];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

fx_status jacobi_scratch(fx_backend* be, size_t n, double** out) {
	static thread_local fx_backend* owner = nullptr;
	static thread_local double* buf = nullptr;
	static thread_local size_t cap = 0;
	if (owner != be || cap < n) {
		if (buf) { (void)hipStreamSynchronize(be->stream); fx::dev_free(buf); }
		buf = nullptr;
		cap = 0;
		void* p = nullptr;
		FX_CHECK(fx::dev_alloc(be, n * sizeof(double), &p));
		buf = (double*)p;
		cap = n;
		owner = be;
	}
	*out = buf;
	return FX_OK;
}

}  // namespace

// ── internal entry points used by the solvers ─────────────────────────
namespace fx {
fx_status k_copy(fx_backend* be, const double* s, double* d, size_t n) { return launch_elem(be, CopyF{s, d}, n, "copy"); }
fx_status k_fill(fx_backend* be, double* d, double v, size_t n) { return launch_elem(be, FillF{d, v}, n, "fill"); }
fx_status k_fill_f32(fx_backend* be, float* d, float v, size_t n) { return launch_elem(be, FillF32{d, v}, n, "fill32"); }
fx_status k_dot_host(fx_backend* be, const double* x, const double* y, size_t n, double* out) {
	return reduce_host<1>(be, DotT{x, y}, n, out, "dot");
}

// Looks up this device's address of `scratch` once (the backend's device is
// current: fx_backend_create has just selected it) and hands it to the backend.
fx_status noalloc_setup(fx_backend* be) {
	FX_REQUIRE(be->device >= 0 && be->device < NOALLOC_MAX_DEVICES, FX_ERR_NO_DEVICE,
	           "device %d is beyond the %d devices the no-allocation scratch table covers", be->device,
	           NOALLOC_MAX_DEVICES);
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_ptr[be->device]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		FX_REQUIRE(e == hipSuccess && p, FX_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", be->device,
		           hipGetErrorString(e));
		g_scratch_ptr[be->device] = p;
	}
	be->noalloc_scratch = g_scratch_ptr[be->device];
	return FX_OK;
}

fx_status k_apply_operator(fx_backend* be, const fx_operator* op, const double* x, double* y, const fx_grid* g,
                           const int* skip) {
	FX_REQUIRE(op, FX_ERR_ARGUMENT, "operator is null");
	int64_t rows = (int64_t)g->rows, cols = (int64_t)g->cols;
	size_t n = (size_t)rows * cols;
	if (op->kind == FX_OP_STENCIL) {
		FX_CHECK(check_bc(op->bc));
		return launch_stencil(be, x, EpiOp{op->coeff, y}, op->stencil, rows, cols, "apply_operator", skip);
	}
	if (op->kind == FX_OP_CONV_DIFF) {
		FX_REQUIRE(op->vx && op->vy, FX_ERR_ARGUMENT, "conv-diff operator needs vx and vy");
		FX_CHECK(check_buf(op->vx, n, FX_F64, "vx"));
		FX_CHECK(check_buf(op->vy, n, FX_F64, "vy"));
		if (op->kappa) FX_CHECK(check_buf(op->kappa, n, FX_F64, "kappa"));
		hipLaunchKernelGGL(k_conv_diff, pgrid(rows, cols), dim3(BLOCK), 0, be->stream, 2, op->coeff, op->sigma, x, y,
		                   op->kappa ? CD(op->kappa) : (const double*)nullptr, op->kappa_scalar, CD(op->vx), CD(op->vy),
		                   rows, cols, 1.0 / g->dx, 1.0 / g->dy, skip);
		return check_launch("conv_diff operator");
	}
	if (op->kind == FX_OP_VARK) {
		FX_REQUIRE(op->k, FX_ERR_ARGUMENT, "variable-coefficient operator needs k");
		FX_CHECK(check_buf(op->k, n, FX_F64, "k"));
		if (cols % 2 == 0) {
			const int64_t cb = (cols / 2 + BLOCK - 1) / BLOCK;
			int rpb = 64;
			while (rpb > 4 && cb * ((rows + rpb - 1) / rpb) < 1024) rpb /= 2;
			hipLaunchKernelGGL((k_pair5<0, true, PLoad, EpiOpVk>), fx::stream_grid(rows, cols / 2, rpb), dim3(BLOCK), 0,
			                   be->stream, PLoad{x}, CD(op->k),
			                   EpiOpVk{op->sigma, op->coeff, 1.0 / (g->dx * g->dx), 1.0 / (g->dy * g->dy), y}, rows, cols,
			                   rpb, RedScratch{}, skip);
			return check_launch("vark operator");
		}
		hipLaunchKernelGGL(k_vark, pgrid(rows, cols), dim3(BLOCK), 0, be->stream, op->sigma, op->coeff, CD(op->k), x, y,
		                   rows, cols, 1.0 / (g->dx * g->dx), 1.0 / (g->dy * g->dy), skip);
		return check_launch("vark operator");
	}
	set_error("unknown operator kind %d", op->kind);
	return FX_ERR_ARGUMENT;
}
}  // namespace fx

using namespace fx;

extern "C" {

fx_status fx_copy(fx_backend* be, const fx_buf* src, fx_buf* dst) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(len_eq(src, dst, "src", "dst"));
	return k_copy(be, CD(src), D(dst), src->len);
}
fx_status fx_fill(fx_backend* be, fx_buf* buf, double v) {
	FX_REQUIRE(be && buf && buf->ptr, FX_ERR_ARGUMENT, "null argument");
	if (buf->dtype == FX_F32) return k_fill_f32(be, (float*)buf->ptr, (float)v, buf->len);
	return k_fill(be, D(buf), v, buf->len);
}
fx_status fx_axpy(fx_backend* be, double a, const fx_buf* x, fx_buf* y) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(len_eq(x, y, "x", "y"));
	return launch_elem(be, AxpyF{a, CD(x), D(y)}, x->len, "axpy");
}
fx_status fx_scale(fx_backend* be, fx_buf* buf, double a) {
	FX_REQUIRE(be && buf && buf->ptr && buf->dtype == FX_F64, FX_ERR_ARGUMENT, "fp64 buffer required");
	return launch_elem(be, ScaleF{D(buf), a}, buf->len, "scale");
}
fx_status fx_waxpy(fx_backend* be, double a, const fx_buf* x, double b, const fx_buf* y, fx_buf* w) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(len_eq(x, y, "x", "y"));
	FX_CHECK(len_eq(x, w, "x", "w"));
	return launch_elem(be, WaxpyF{a, CD(x), b, CD(y), D(w)}, x->len, "waxpy");
}
fx_status fx_aypx(fx_backend* be, double a, const fx_buf* x, fx_buf* y) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(len_eq(x, y, "x", "y"));
	return launch_elem(be, AypxF{a, CD(x), D(y)}, x->len, "aypx");
}
static fx_status pointwise(fx_backend* be, int op, const fx_buf* x, const fx_buf* y, fx_buf* z) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(len_eq(x, y, "x", "y"));
	FX_CHECK(len_eq(x, z, "x", "z"));
	return launch_elem(be, PwF{op, CD(x), CD(y), D(z)}, x->len, "pointwise");
}
fx_status fx_pointwise_mult(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* z) { return pointwise(be, 0, x, y, z); }
fx_status fx_pointwise_div(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* z) { return pointwise(be, 1, x, y, z); }
fx_status fx_pointwise_max(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* z) { return pointwise(be, 2, x, y, z); }
fx_status fx_pointwise_min(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* z) { return pointwise(be, 3, x, y, z); }
fx_status fx_reciprocal(fx_backend* be, fx_buf* buf) {
	FX_REQUIRE(be && buf && buf->ptr && buf->dtype == FX_F64, FX_ERR_ARGUMENT, "fp64 buffer required");
	return launch_elem(be, RecipF{D(buf)}, buf->len, "reciprocal");
}
fx_status fx_abs_val(fx_backend* be, fx_buf* buf) {
	FX_REQUIRE(be && buf && buf->ptr && buf->dtype == FX_F64, FX_ERR_ARGUMENT, "fp64 buffer required");
	return launch_elem(be, AbsF{D(buf)}, buf->len, "abs");
}

fx_status fx_dot(fx_backend* be, const fx_buf* x, const fx_buf* y, double* out) {
	FX_REQUIRE(be && out, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(len_eq(x, y, "x", "y"));
	return reduce_host<1>(be, DotT{CD(x), CD(y)}, x->len, out, "dot");
}
// fx_dot's kernels with the block partials in the static scratch and the sum
// stored into out[0] on the device: no allocation, no host synchronisation.
fx_status fx_dot_noalloc(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* out) {
	FX_REQUIRE(be && out && out->ptr, FX_ERR_ARGUMENT, "null argument");
	FX_REQUIRE(out->dtype == FX_F64, FX_ERR_ARGUMENT, "out must be an fp64 buffer");
	FX_REQUIRE(out->len >= 1, FX_ERR_DIMENSION, "out holds %zu values; the dot needs 1", out->len);
	FX_CHECK(len_eq(x, y, "x", "y"));
	FX_REQUIRE(be->noalloc_scratch, FX_ERR_HIP, "the no-allocation scratch is not set up for device %d", be->device);
	const unsigned nb = red_blocks(x->len);
	FX_REQUIRE(nb <= NOALLOC_PARTIALS, FX_ERR_DIMENSION,
	           "a dot of %zu values needs %u block partials; the static scratch holds %u", x->len, nb,
	           NOALLOC_PARTIALS);
	const RedScratch rs{(Acc*)be->noalloc_scratch};
	hipLaunchKernelGGL((k_reduce<1, DotT>), dim3(nb), dim3(BLOCK), 0, be->stream, DotT{CD(x), CD(y)}, x->len, rs);
	hipLaunchKernelGGL((k_finish<1, FinStore<1>>), dim3(1), dim3(FIN_BLOCK), 0, be->stream, rs, nb,
	                   FinStore<1>{D(out)}, (const int*)nullptr);
	return check_launch("dot_noalloc");
}
fx_status fx_dot2(fx_backend* be, const fx_buf* a, const fx_buf* b, const fx_buf* c, const fx_buf* d, double* out) {
	FX_REQUIRE(be && out, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(len_eq(a, b, "a", "b"));
	FX_CHECK(len_eq(a, c, "a", "c"));
	FX_CHECK(len_eq(a, d, "a", "d"));
	return reduce_host<2>(be, Dot2T{CD(a), CD(b), CD(c), CD(d)}, a->len, out, "dot2");
}
fx_status fx_norm_l2(fx_backend* be, const fx_buf* x, double* out) {
	FX_REQUIRE(be && out && x && x->ptr && x->dtype == FX_F64, FX_ERR_ARGUMENT, "fp64 buffer required");
	double s = 0;
	FX_CHECK(reduce_host<1>(be, SqT{CD(x)}, x->len, &s, "norm_l2"));
	*out = std::sqrt(s);
	return FX_OK;
}
fx_status fx_reduce_sum(fx_backend* be, const fx_buf* x, double* out) {
	FX_REQUIRE(be && out && x && x->ptr && x->dtype == FX_F64, FX_ERR_ARGUMENT, "fp64 buffer required");
	return reduce_host<1>(be, SumT{CD(x)}, x->len, out, "reduce_sum");
}
static fx_status minmax(fx_backend* be, const fx_buf* x, double* out, bool mx) {
	FX_REQUIRE(be && out && x && x->ptr && x->dtype == FX_F64, FX_ERR_ARGUMENT, "fp64 buffer required");
	unsigned nb = red_blocks(x->len);
	FX_CHECK(ensure_partials(be, nb));
	if (mx) {
		hipLaunchKernelGGL(k_minmax<1>, dim3(nb), dim3(BLOCK), 0, be->stream, CD(x), x->len, red_scratch(be));
		hipLaunchKernelGGL(k_minmax_finish<1>, dim3(1), dim3(BLOCK), 0, be->stream, red_scratch(be), nb, be->dres);
	} else {
		hipLaunchKernelGGL(k_minmax<0>, dim3(nb), dim3(BLOCK), 0, be->stream, CD(x), x->len, red_scratch(be));
		hipLaunchKernelGGL(k_minmax_finish<0>, dim3(1), dim3(BLOCK), 0, be->stream, red_scratch(be), nb, be->dres);
	}
	FX_CHECK(check_launch("minmax"));
	FX_HIP(hipMemcpyAsync(be->hres, be->dres, sizeof(double), hipMemcpyDeviceToHost, be->stream));
	FX_HIP(hipStreamSynchronize(be->stream));
	*out = be->hres[0];
	return FX_OK;
}
fx_status fx_reduce_max(fx_backend* be, const fx_buf* x, double* out) { return minmax(be, x, out, true); }
fx_status fx_reduce_min(fx_backend* be, const fx_buf* x, double* out) { return minmax(be, x, out, false); }

// apply_stencil (ops.rs:77-127): boundary values from bc_value, interior from the stencil.
fx_status fx_apply_stencil(fx_backend* be, const fx_buf* in, fx_buf* out, const fx_grid* g, const fx_stencil* st,
                           const fx_boundaries* bc) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(check_bc(bc));
	FX_CHECK(check_grid(g));
	size_t n = g->rows * g->cols;
	FX_CHECK(check_buf(in, n, FX_F64, "input"));
	FX_CHECK(check_buf(out, n, FX_F64, "output"));
	FX_CHECK(stencil_args(st, nullptr, nullptr));
	int64_t rows = (int64_t)g->rows, cols = (int64_t)g->cols;
	int64_t per = 2 * cols + 2 * (rows - 2);
	hipLaunchKernelGGL(k_bc_perimeter, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, be->stream, CD(in), D(out),
	                   (const double*)nullptr, *bc, rows, cols, g->dx, g->dy);
	FX_CHECK(check_launch("bc_perimeter"));
	return launch_stencil(be, CD(in), EpiStore{D(out)}, st, rows, cols, "apply_stencil");
}

// apply_stencil_var (ops.rs:129-154): apply_stencil then out *= coeff everywhere.
fx_status fx_apply_stencil_var(fx_backend* be, const fx_buf* in, fx_buf* out, const fx_buf* coeff, const fx_grid* g,
                               const fx_stencil* st, const fx_boundaries* bc) {
	FX_CHECK(fx_apply_stencil(be, in, out, g, st, bc));
	FX_CHECK(check_buf(coeff, g->rows * g->cols, FX_F64, "coeff"));
	return launch_elem(be, MulF{D(out), CD(coeff)}, out->len, "apply_stencil_var");
}

// stencil_axpy (ops.rs:226-269): interior only, y untouched on the boundary.
fx_status fx_stencil_axpy(fx_backend* be, double alpha, const fx_buf* x, fx_buf* y, const fx_grid* g,
                          const fx_stencil* st, const fx_boundaries* bc) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(check_bc(bc));
	FX_CHECK(check_grid(g));
	size_t n = g->rows * g->cols;
	FX_CHECK(check_buf(x, n, FX_F64, "x"));
	FX_CHECK(check_buf(y, n, FX_F64, "y"));
	probe_begin(be, FX_PROBE_STENCIL_AXPY);
	fx_status s = launch_stencil(be, CD(x), EpiAxpy{alpha, D(y)}, st, (int64_t)g->rows, (int64_t)g->cols, "stencil_axpy");
	probe_end(be, FX_PROBE_STENCIL_AXPY);
	return s;
}

fx_status fx_restrict(fx_backend* be, const fx_buf* fine, fx_buf* coarse, const fx_grid* fg, const fx_grid* cg) {
	FX_REQUIRE(be && fg && cg, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(check_buf(fine, fg->rows * fg->cols, FX_F64, "fine"));
	FX_CHECK(check_buf(coarse, cg->rows * cg->cols, FX_F64, "coarse"));
	FX_REQUIRE(2 * (cg->rows - 1) < fg->rows && 2 * (cg->cols - 1) < fg->cols, FX_ERR_DIMENSION,
	           "coarse grid %llux%llu does not fit fine grid %llux%llu", (unsigned long long)cg->rows,
	           (unsigned long long)cg->cols, (unsigned long long)fg->rows, (unsigned long long)fg->cols);
	hipLaunchKernelGGL(k_restrict_ref, pgrid(cg->rows, cg->cols), dim3(BLOCK), 0, be->stream, CD(fine), D(coarse),
	                   (int64_t)fg->rows, (int64_t)fg->cols, (int64_t)cg->rows, (int64_t)cg->cols);
	return check_launch("restrict");
}

fx_status fx_prolong(fx_backend* be, const fx_buf* coarse, fx_buf* fine, const fx_grid* cg, const fx_grid* fg) {
	FX_REQUIRE(be && fg && cg, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(check_buf(fine, fg->rows * fg->cols, FX_F64, "fine"));
	FX_CHECK(check_buf(coarse, cg->rows * cg->cols, FX_F64, "coarse"));
	hipLaunchKernelGGL(k_prolong_ref, pgrid(fg->rows, fg->cols), dim3(BLOCK), 0, be->stream, CD(coarse), D(fine),
	                   (int64_t)cg->rows, (int64_t)cg->cols, (int64_t)fg->rows, (int64_t)fg->cols);
	return check_launch("prolong");
}

fx_status fx_weighted_jacobi(fx_backend* be, fx_buf* x, const fx_buf* b, double omega, const fx_grid* g,
                             const fx_stencil* st, const fx_boundaries* bc) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(check_grid(g));
	size_t n = g->rows * g->cols;
	FX_CHECK(check_buf(x, n, FX_F64, "x"));
	FX_CHECK(check_buf(b, n, FX_F64, "b"));
	Ent ent;
	FX_CHECK(stencil_args(st, nullptr, &ent));
	// ops.rs:443-447: first (0,0) entry, default −4.
	double diag = -4.0;
	for (uint32_t i = 0; i < st->n; ++i)
		if (st->entries[i].dr == 0 && st->entries[i].dc == 0) { diag = st->entries[i].weight; break; }
	double* tmp = nullptr;
	FX_CHECK(jacobi_scratch(be, n, &tmp));
	FX_CHECK(k_copy(be, CD(x), tmp, n));
	hipLaunchKernelGGL(k_jacobi, pgrid(g->rows, g->cols), dim3(BLOCK), 0, be->stream, (const double*)tmp, D(x), CD(b),
	                   omega, 1.0 / diag, ent, (int64_t)g->rows, (int64_t)g->cols);
	return check_launch("weighted_jacobi");
}

static fx_status conv_diff(fx_backend* be, int mode, double alpha, const fx_buf* u, fx_buf* out, const fx_buf* kappa,
                           const fx_buf* vx, const fx_buf* vy, const fx_grid* g) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(check_grid(g));
	size_t n = g->rows * g->cols;
	FX_CHECK(check_buf(u, n, FX_F64, "u"));
	FX_CHECK(check_buf(out, n, FX_F64, "output"));
	FX_CHECK(check_buf(kappa, n, FX_F64, "kappa"));
	FX_CHECK(check_buf(vx, n, FX_F64, "vx"));
	FX_CHECK(check_buf(vy, n, FX_F64, "vy"));
	hipLaunchKernelGGL(k_conv_diff, pgrid(g->rows, g->cols), dim3(BLOCK), 0, be->stream, mode, alpha, 1.0, CD(u), D(out),
	                   CD(kappa), 0.0, CD(vx), CD(vy), (int64_t)g->rows, (int64_t)g->cols, 1.0 / g->dx, 1.0 / g->dy,
	                   (const int*)nullptr);
	return check_launch("conv_diff");
}
fx_status fx_apply_conv_diff(fx_backend* be, const fx_buf* u, fx_buf* out, const fx_buf* kappa, const fx_buf* vx,
                             const fx_buf* vy, const fx_grid* g, const fx_boundaries* bc) {
	(void)bc;  // ops.rs:683-706: boundaries are forced to 0 regardless of `bc`.
	return conv_diff(be, 0, 0.0, u, out, kappa, vx, vy, g);
}
fx_status fx_conv_diff_axpy(fx_backend* be, double alpha, const fx_buf* u, fx_buf* out, const fx_buf* kappa,
                            const fx_buf* vx, const fx_buf* vy, const fx_grid* g, const fx_boundaries* bc) {
	(void)bc;
	return conv_diff(be, 1, alpha, u, out, kappa, vx, vy, g);
}

fx_status fx_zero_boundaries(fx_backend* be, fx_buf* buf, uint64_t rows, uint64_t cols) {
	FX_REQUIRE(be && rows >= 2 && cols >= 2, FX_ERR_ARGUMENT, "bad arguments");
	FX_CHECK(check_buf(buf, rows * cols, FX_F64, "buf"));
	int64_t per = (int64_t)cols + (int64_t)rows - 2;
	hipLaunchKernelGGL(k_zero_bnd, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, be->stream, D(buf), (int64_t)rows,
	                   (int64_t)cols);
	return check_launch("zero_boundaries");
}

fx_status fx_pipelined_cg_fused(fx_backend* be, double alpha, double beta, fx_buf* z, fx_buf* t, fx_buf* p, fx_buf* x,
                                fx_buf* r, fx_buf* w, const fx_buf* q) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(len_eq(z, t, "z", "t"));
	FX_CHECK(len_eq(z, p, "z", "p"));
	FX_CHECK(len_eq(z, x, "z", "x"));
	FX_CHECK(len_eq(z, r, "z", "r"));
	FX_CHECK(len_eq(z, w, "z", "w"));
	FX_CHECK(len_eq(z, q, "z", "q"));
	return launch_elem(be, PcgFusedF{alpha, beta, D(z), D(t), D(p), D(x), D(r), D(w), CD(q)}, z->len, "pipelined_cg_fused");
}

}  // extern "C"

namespace fx {
fx_status rk4_passes(fx_backend* be, double* u, double* const k[5], const double* src, double dt, double alpha,
                     const fx_stencil* st, const fx_boundaries* bc, const fx_grid* g) {
	const int64_t rows = (int64_t)g->rows, cols = (int64_t)g->cols;
	const int64_t per = 2 * cols + 2 * (rows - 2);
	const dim3 pg((unsigned)((per + 255) / 256));
	double *k1 = k[0], *k2 = k[1], *k3 = k[2], *sb = k[3], *sa = k[4];
	auto perim = [&](const double* in, double* out) {
		hipLaunchKernelGGL(k_bc_perimeter, pg, dim3(256), 0, be->stream, in, out, (const double*)nullptr, *bc, rows,
		                   cols, g->dx, g->dy);
		return check_launch("rk4 perimeter");
	};
	const double h2 = 0.5 * dt * alpha;
	FX_CHECK(launch_stencil(be, u, EpiRk1{h2, k1, sa}, st, rows, cols, "rk4 pass 1"));
	FX_CHECK(perim(u, k1));
	FX_CHECK(launch_stencil(be, sa, EpiRkMid{h2, k2, sb, u}, st, rows, cols, "rk4 pass 2"));
	FX_CHECK(perim(sa, k2));
	FX_CHECK(launch_stencil(be, sb, EpiRkMid{dt * alpha, k3, sa, u}, st, rows, cols, "rk4 pass 3"));
	FX_CHECK(perim(sb, k3));
	const double c = dt * alpha / 6.0;
	FX_CHECK(launch_stencil(be, sa, EpiRkLast{c, u, k1, k2, k3, src, dt}, st, rows, cols, "rk4 pass 4"));
	FX_CHECK(perim(sa, sb));
	hipLaunchKernelGGL(k_rk4_edge, pg, dim3(256), 0, be->stream, u, (const double*)k1, (const double*)k2,
	                   (const double*)k3, (const double*)sb, src, c, dt, rows, cols);
	return check_launch("rk4 edge");
}

fx_status explicit_stage(fx_backend* be, const double* x, double* y, double coeff, const double* src, double dt,
                         const double* avg, const fx_stencil* st, int64_t rows, int64_t cols) {
	return launch_stencil(be, x, EpiStage{coeff, y, src, dt, avg}, st, rows, cols, "explicit_stage");
}
}  // namespace fx

extern "C" {

fx_status fx_apply_operator(fx_backend* be, const fx_operator* op, const fx_buf* x, fx_buf* y, const fx_grid* g) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_CHECK(check_grid(g));
	size_t n = g->rows * g->cols;
	FX_CHECK(check_buf(x, n, FX_F64, "x"));
	FX_CHECK(check_buf(y, n, FX_F64, "y"));
	return k_apply_operator(be, op, CD(x), D(y), g);
}

}  // extern "C"
