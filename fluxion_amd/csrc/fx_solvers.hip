// fx_solvers.hip — Krylov solvers on the device.
//
// The reference drives every solver from the host: one backend call per
// vector operation and a blocking device-to-host copy per dot product
// (crates/runtime/src/cg.rs:118-157; backend-hip/src/backend.rs:356-413).
// Here the recurrence scalars live in device memory.  Each reduction is
// finished by the last workgroup of the kernel that produced its terms, which
// then evaluates the solver's scalar update (alpha, beta, convergence and
// breakdown tests, iteration count) on the device.  The host only enqueues
// kernels and polls a `done` flag a few iterations behind; kernels launched
// after `done` is set return immediately.
//
// The recurrences, their operation order and their stopping rules follow
//   CgSolver::solve / solve_preconditioned       cg.rs:96-256
//   PipelinedCgSolver::solve                     pipelined_cg.rs:96-200
//   BiCgStabSolver::solve / solve_preconditioned bicgstab.rs:90-346
//   GmresSolver::solve / solve_preconditioned    gmres.rs:116-490
#include "fx_mg.hpp"
#include "fx_mgfast.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

using namespace fxd;

namespace {

constexpr int HIST_MAX = 1 << 16;
constexpr int GM_MAX = 128;   // largest GMRES restart supported

// Solver scalars in device memory.
struct SState {
	double tol;
	double rs;          // CG: r·r; PCG/FCG: r·z; BiCGSTAB: rho
	double rr;          // r·r
	double alpha, beta, pap;
	double gprev, aprev;     // pipelined CG: gamma_prev, alpha_prev
	double omega, rho_new, snorm;
	double resid;
	int64_t iter, max_iter;
	int32_t done, reason;
	int32_t pend;        // BiCGSTAB: finish after the x correction
	int32_t kcur;        // GMRES: Arnoldi steps in the current cycle
	int32_t cycle_end;   // GMRES: cycle finished
	int32_t skip;        // done | pend | cycle_end: vector kernels return early
	int32_t gexit;       // GMRES: 1 = breakdown exit
	double* hist;
	int64_t hist_len;
	// GMRES (H column-major, (m+1) x m; gmres.rs:84-99)
	int32_t m;
	int32_t pad2;
	double gbeta, gcur, ghk;
	double *H, *cc, *ss, *grs, *y;
	// Learned preconditioner: lskip = done | lplain; lplain switches to the plain V-cycle once an
	// iteration whose direction came from a network application fails to cut the FP64 true residual
	// ‖b − A·x‖² (ltres, evaluated after every x/r update while the network is in use) below
	// stall2·(its value before that iteration).  lgate = 1 skips the network at this application
	// (schedule: from iteration `first`, every `every`-th, at most `max_apps` times; lapps counts them).
	int32_t lmode, lplain, lskip, lgate;
	double stall2;
	double ltres;
	int32_t lapps, pad3;
	// Host-visible copy of `done` (coherent pinned memory), set together with it so that the host's
	// enqueue loop can stop without a per-iteration device-to-host copy.
	int* hdone;
	// Mixed FCG with the x update deferred to the next step kernel: xpend = 1 while x += alpha·xp
	// (the last x/r update's direction) is still to be applied.
	int32_t xpend, pad4;
	const double* xp;
	// Host-visible iteration counter (next to hdone): FinPcgRR stores the iteration count once the
	// iteration's x/r update is done, so the enqueue loop can follow the device without events.
	int* hprog;
	double tres;   // ‖b − A·x‖ of the last true-residual check (FinTrueRes)
};

__device__ __forceinline__ void record(SState* s, int64_t k, double v) {
	if (s->hist && k >= 0 && k < s->hist_len) s->hist[k] = v;
}
__device__ __forceinline__ void set_done(SState* s) {
	s->done = 1;
	if (s->hdone) __hip_atomic_store(s->hdone, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void finish(SState* s, int64_t its, double res, int reason) {
	set_done(s);
	s->skip = 1;
	s->lskip = 1;
	s->reason = reason;
	s->iter = its;
	s->resid = res;
}

// Element-wise update + reduction; F::load(st) caches scalars, F::operator()(i, acc).
template <int NS, class F>
__global__ __launch_bounds__(BLOCK) void k_ered(F f, size_t n, RedScratch rs, const int* skip) {
	if (skip && *skip) return;
	f.load();
	Acc acc[NS];
#pragma unroll
	for (int k = 0; k < NS; ++k) acc[k] = Acc{0.0, 0.0};
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) f(i, acc);
	block_partial<NS>(acc, rs);
}
template <class F>
__global__ __launch_bounds__(BLOCK) void k_eupd(F f, size_t n, const int* skip) {
	if (skip && *skip) return;
	f.load();
	size_t stride = (size_t)gridDim.x * BLOCK;
	for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) f(i);
}
template <class Fin>
__global__ void k_scalar(Fin fin, const int* skip) {
	if (skip && *skip) return;
	fin();
}

// ── element-wise functors ──────────────────────────────────────────────
// r = b − ap (copy + axpy(−1): cg.rs:110-111), with r·r.
struct FResid {
	const double *b, *ap;
	double* r;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const {
		double v = b[i] - ap[i];
		r[i] = v;
		acc_add(a[0], v * v);
	}
};
// x += α·p; r −= α·ap (cg.rs:145-147), with r·r.
struct FXR {
	const SState* st;
	double* x;
	double* r;
	const double *p, *ap;
	double al;
	__device__ void load() { al = st->alpha; }
	__device__ void operator()(size_t i, Acc* a) const {
		x[i] = fma(al, p[i], x[i]);
		double v = fma(-al, ap[i], r[i]);
		r[i] = v;
		acc_add(a[0], v * v);
	}
};
struct FDot {
	const double *x, *y;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const { acc_add(a[0], x[i] * y[i]); }
};
struct FDot2 {
	const double *a0, *b0, *a1, *b1;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const {
		acc_add(a[0], a0[i] * b0[i]);
		acc_add(a[1], a1[i] * b1[i]);
	}
};
// p_new = src + β·p_old (scale + axpy(1, src): cg.rs:153-154), written to p_out.
struct FPUpd {
	const SState* st;
	const double *src, *pold;
	double* pout;
	double be;
	__device__ void load() { be = st->beta; }
	__device__ void operator()(size_t i) const { pout[i] = src[i] + be * pold[i]; }
};
struct FCopy {
	const double* s;
	double* d;
	__device__ void load() {}
	__device__ void operator()(size_t i) const { d[i] = s[i]; }
};
struct FZero {
	double* d;
	__device__ void load() {}
	__device__ void operator()(size_t i) const { d[i] = 0.0; }
};

// ── streaming stencil pieces for the fused CG step ─────────────────────
// Value source p_new = src + β·p_old, recomputed at every neighbour.
struct VPNew {
	const double *src, *pold;
	double be;
	__device__ double load(int64_t i) const { return src[i] + be * pold[i]; }
};
// ap = fma(coeff, lap, p) inside, ap = p on the boundary; writes p_new; p·ap.
struct EpiCG {
	double coeff;
	double *pout, *ap;
	__device__ void interior(int64_t i, double lap, double c, Acc* a) const {
		double v = fma(coeff, lap, c);
		pout[i] = c;
		ap[i] = v;
		acc_add(a[0], c * v);
	}
	__device__ void boundary(int64_t i, double c, Acc* a) const {
		pout[i] = c;
		ap[i] = c;
		acc_add(a[0], c * c);
	}
};

template <int NS, class V, class E>
__global__ __launch_bounds__(BLOCK) void k_stream_st(V v, E e, W9 w, int64_t rows, int64_t cols, int rpb,
                                                     RedScratch rs, const SState* st) {
	// Same traversal as fxd::k_stream5, reading β from the device state first.
	if (st->done) return;
	v.be = st->beta;
	int bx, by;
	xcd_tile(&bx, &by);
	const int64_t c = (int64_t)bx * BLOCK + threadIdx.x;
	const int64_t r0 = (int64_t)by * rpb;
	const int64_t r1 = r0 + rpb < rows ? r0 + rpb : rows;
	const bool col_ok = c < cols;
	const bool col_int = c > 0 && c < cols - 1;
	Acc acc[NS];
#pragma unroll
	for (int k = 0; k < NS; ++k) acc[k] = Acc{0.0, 0.0};
	double up = 0.0, cur = 0.0;
	if (col_ok) {
		if (r0 > 0) up = v.load((r0 - 1) * cols + c);
		cur = v.load(r0 * cols + c);
	}
	for (int64_t r = r0; r < r1; ++r) {
		const int64_t i = r * cols + c;
		double dn = 0.0;
		if (col_ok && r + 1 < rows) dn = v.load(i + cols);
		if (col_ok) {
			if (col_int && r > 0 && r < rows - 1) {
				const double wv = v.load(i - 1), ev = v.load(i + 1);
				e.interior(i, chain5(w, cur, up, dn, wv, ev), cur, acc);
			} else {
				e.boundary(i, cur, acc);
			}
		}
		up = cur;
		cur = dn;
	}
	block_partial<NS>(acc, rs);
}

// The same traversal for the variable-coefficient operator
// sigma·p + alpha·(−∇·(k∇p)) with harmonic face means (fx_ops.hip k_vark), k
// streamed alongside p; E::interior receives the operator value.
__device__ __forceinline__ double hmean_d(double a, double b) { return 2.0 * a * b / (a + b); }
// Column block bx, row block by of the traversal; the terms go to acc.
template <int NS, class V, class E>
__device__ __forceinline__ void stream_vk_body(int bx, int by, const V& v, const E& e, const double* __restrict__ k,
                                               double sigma, double alpha, double ix2, double iy2, int64_t rows,
                                               int64_t cols, int rpb, Acc* acc) {
	const int64_t c = (int64_t)bx * BLOCK + threadIdx.x;
	const int64_t r0 = (int64_t)by * rpb;
	const int64_t r1 = r0 + rpb < rows ? r0 + rpb : rows;
	const bool col_ok = c < cols;
	const bool col_int = c > 0 && c < cols - 1;
	double up = 0.0, cur = 0.0, kup = 1.0, kcur = 1.0;
	if (col_ok) {
		if (r0 > 0) {
			up = v.load((r0 - 1) * cols + c);
			kup = k[(r0 - 1) * cols + c];
		}
		cur = v.load(r0 * cols + c);
		kcur = k[r0 * cols + c];
	}
	for (int64_t r = r0; r < r1; ++r) {
		const int64_t i = r * cols + c;
		double dn = 0.0, kdn = 1.0;
		if (col_ok && r + 1 < rows) {
			dn = v.load(i + cols);
			kdn = k[i + cols];
		}
		if (col_ok) {
			if (col_int && r > 0 && r < rows - 1) {
				const double wv = v.load(i - 1), ev = v.load(i + 1);
				const double kw = hmean_d(kcur, k[i - 1]), ke = hmean_d(kcur, k[i + 1]);
				const double kn = hmean_d(kcur, kup), ks = hmean_d(kcur, kdn);
				const double fx = ke * (ev - cur) - kw * (cur - wv);
				const double fy = ks * (dn - cur) - kn * (cur - up);
				const double L = -(fx * ix2 + fy * iy2);
				e.interior(i, fma(alpha, L, sigma * cur), cur, acc);
			} else {
				e.boundary(i, cur, acc);
			}
		}
		up = cur;
		cur = dn;
		kup = kcur;
		kcur = kdn;
	}
}
template <int NS, class V, class E>
__global__ __launch_bounds__(BLOCK) void k_stream_vk(V v, E e, const double* __restrict__ k, double sigma, double alpha,
                                                     double ix2, double iy2, int64_t rows, int64_t cols, int rpb,
                                                     RedScratch rs, const SState* st) {
	if (st->done) return;
	v.be = st->beta;
	int bx, by;
	xcd_tile(&bx, &by);
	Acc acc[NS];
#pragma unroll
	for (int q = 0; q < NS; ++q) acc[q] = Acc{0.0, 0.0};
	stream_vk_body<NS>(bx, by, v, e, k, sigma, alpha, ix2, iy2, rows, cols, rpb, acc);
	block_partial<NS>(acc, rs);
}
// r = b − A·x for the variable-coefficient operator in one pass over x, k and
// b (the arithmetic of fx_ops.hip k_vark, identity boundary), with r·r; E
// receives each point's b − A·x.  Replaces the operator kernel plus the
// residual reduction, and the temporary A·x they pass between them.
struct VX {
	const double* x;
	double be;
	__device__ double load(int64_t i) const { return x[i]; }
};
template <class E>
__global__ __launch_bounds__(BLOCK) void k_resid_vk(VX v, E e, const double* __restrict__ k, double sigma,
                                                    double alpha, double ix2, double iy2, int64_t rows, int64_t cols,
                                                    int rpb, RedScratch rs) {
	int bx, by;
	xcd_tile(&bx, &by);
	Acc acc[1] = {Acc{0.0, 0.0}};
	stream_vk_body<1>(bx, by, v, e, k, sigma, alpha, ix2, iy2, rows, cols, rpb, acc);
	block_partial<1>(acc, rs);
}
// Writes r and its FP32 copy (FResid32's per-point work).
struct EpiRes32 {
	const double* b;
	double* r;
	float* r32;
	__device__ void put(int64_t i, double av, Acc* a) const {
		const double v = b[i] - av;
		r[i] = v;
		r32[i] = float(v);
		acc_add(a[0], v * v);
	}
	__device__ void interior(int64_t i, double av, double, Acc* a) const { put(i, av, a); }
	__device__ void boundary(int64_t i, double c, Acc* a) const { put(i, c, a); }
	__device__ void put2(int64_t i, double2 av, Acc* a) const {
		const double2 bv = *(const double2*)(b + i);
		const double2 v{bv.x - av.x, bv.y - av.y};
		*(double2*)(r + i) = v;
		*(float2*)(r32 + i) = float2{float(v.x), float(v.y)};
		acc_add(a[0], v.x * v.x);
		acc_add(a[0], v.y * v.y);
	}
};
// Only ‖b − A·x‖² (the true-residual check).
struct EpiResNorm {
	const double* b;
	__device__ void interior(int64_t i, double av, double, Acc* a) const {
		const double v = b[i] - av;
		acc_add(a[0], v * v);
	}
	__device__ void boundary(int64_t i, double c, Acc* a) const {
		const double v = b[i] - c;
		acc_add(a[0], v * v);
	}
	__device__ void put2(int64_t i, double2 av, Acc* a) const {
		const double2 bv = *(const double2*)(b + i);
		const double vx = bv.x - av.x, vy = bv.y - av.y;
		acc_add(a[0], vx * vx);
		acc_add(a[0], vy * vy);
	}
};

// ‖b − A·x‖ into st->tres, leaving the iteration state alone.
struct FinTrueRes {
	SState* st;
	__device__ void operator()(Acc* o) const { st->tres = sqrt(acc_value(o[0])); }
};

// k_pair5 value source p = src + β·p_old (VPNew's arithmetic), β read from the state.
struct VPNew2 {
	const double *src, *pold;
	const SState* st;
	double be;
	__device__ void prep() { be = st->beta; }
	__device__ double2 pair(int64_t i) const {
		const double2 s = *(const double2*)(src + i), p = *(const double2*)(pold + i);
		return double2{s.x + be * p.x, s.y + be * p.y};
	}
	__device__ double one(int64_t i) const { return src[i] + be * pold[i]; }
};
// EpiCGv on k_pair5's pairs: ap = A·p inside (k_vark's arithmetic), ap = p on the boundary; writes p; p·ap.
struct EpiCGvk {
	double sigma, alpha, ix2, iy2;
	double *pout, *ap;
	__device__ void pair(int64_t i, int64_t, int64_t, bool rin, bool ia, bool ib, const Nb& uA, const Nb& uB,
	                     const Nb& kA, const Nb& kB, Acc* a) const {
		double2 o{uA.c, uB.c};
		double fs;
		if (rin && ia) o.x = fma(alpha, vk_lap(uA, kA, ix2, iy2, &fs), sigma * uA.c);
		if (rin && ib) o.y = fma(alpha, vk_lap(uB, kB, ix2, iy2, &fs), sigma * uB.c);
		*(double2*)(pout + i) = double2{uA.c, uB.c};
		*(double2*)(ap + i) = o;
		acc_add(a[0], uA.c * o.x);
		acc_add(a[0], uB.c * o.y);
	}
};
// ap = A·p inside (operator value given), ap = p on the boundary; writes p_new; p·ap.
struct EpiCGv {
	double *pout, *ap;
	__device__ void interior(int64_t i, double av, double c, Acc* a) const {
		pout[i] = c;
		ap[i] = av;
		acc_add(a[0], c * av);
	}
	__device__ void boundary(int64_t i, double c, Acc* a) const {
		pout[i] = c;
		ap[i] = c;
		acc_add(a[0], c * c);
	}
};

// v of lane x − 1 (ctrl 0x138, wave_shr:1) or x + 1 (0x130, wave_shl:1) for a
// double, as two 32-bit DPP moves; 0 past the wave.
template <int CTRL>
__device__ __forceinline__ double dlane(double v) {
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
	return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// r = b − A·x for the variable-coefficient operator on grids with an even
// column count: two columns per thread and the DPP neighbour exchange of
// k_fcg_step2 below, the operator arithmetic of k_vark (identity boundary);
// E::put2 receives each pair's A·x.
template <class E>
__global__ __launch_bounds__(BLOCK) void k_resid2(const double* __restrict__ x, const double* __restrict__ k, E e,
                                                  double sigma, double alpha, double ix2, double iy2, int64_t rows,
                                                  int64_t cols, int rpb, RedScratch rs) {
	int bx, by;
	xcd_tile(&bx, &by);
	const int lane = (int)(threadIdx.x & (WAVE - 1));
	const int64_t c0 = 2 * ((int64_t)bx * BLOCK + threadIdx.x);
	const int64_t r0 = (int64_t)by * rpb;
	const int64_t r1 = r0 + rpb < rows ? r0 + rpb : rows;
	const bool ok = c0 < cols;
	const int64_t cl = ok ? c0 : cols - 2;
	const bool intA = c0 > 0, intB = c0 + 1 < cols - 1;
	const bool edgeW = lane == 0 && ok && c0 > 0, edgeE = lane == WAVE - 1 && ok && c0 + 2 < cols;
	struct Row {
		double2 x, k;
		double xw, xe, kw, ke;
	};
	auto ld = [&](int64_t r) {
		Row v;
		const int64_t i = r * cols + cl;
		v.x = *(const double2*)(x + i);
		v.k = *(const double2*)(k + i);
		v.xw = v.xe = 0.0;
		v.kw = v.ke = 1.0;
		if (edgeW) v.xw = x[i - 1], v.kw = k[i - 1];
		if (edgeE) v.xe = x[i + 2], v.ke = k[i + 2];
		return v;
	};
	const Row none{double2{0.0, 0.0}, double2{1.0, 1.0}, 0.0, 0.0, 1.0, 1.0};
	Acc acc[1] = {Acc{0.0, 0.0}};
	Row up = r0 > 0 ? ld(r0 - 1) : none;
	Row cur = ld(r0 < rows ? r0 : rows - 1);
	auto op = [&](double kc, double kn, double ks, double kwv, double kev, double c, double n, double s_, double w_,
	              double e_) {
		const double hw = hmean_d(kc, kwv), he = hmean_d(kc, kev);
		const double hn = hmean_d(kc, kn), hs = hmean_d(kc, ks);
		const double fx = he * (e_ - c) - hw * (c - w_);
		const double fy = hs * (s_ - c) - hn * (c - n);
		const double L = -(fx * ix2 + fy * iy2);
		return fma(alpha, L, sigma * c);
	};
	for (int64_t r = r0; r < r1; ++r) {
		const int64_t i = r * cols + c0;
		const Row dn = r + 1 < rows ? ld(r + 1) : none;
		double xw = dlane<0x138>(cur.x.y), xe = dlane<0x130>(cur.x.x);
		double kw = dlane<0x138>(cur.k.y), ke = dlane<0x130>(cur.k.x);
		if (lane == 0) xw = cur.xw, kw = cur.kw;
		if (lane == WAVE - 1) xe = cur.xe, ke = cur.ke;
		double2 av = cur.x;   // boundary rows and columns: identity
		if (r > 0 && r < rows - 1) {
			if (intA) av.x = op(cur.k.x, up.k.x, dn.k.x, kw, cur.k.y, cur.x.x, up.x.x, dn.x.x, xw, cur.x.y);
			if (intB) av.y = op(cur.k.y, up.k.y, dn.k.y, cur.k.x, ke, cur.x.y, up.x.y, dn.x.y, cur.x.x, xe);
		}
		if (ok) e.put2(i, av, acc);
		up = cur;
		cur = dn;
	}
	block_partial<1>(acc, rs);
}

// FCG step of the fused loop on grids with an even column count, two columns
// per thread (16-byte loads and stores): p = double(z) + β·p_old, ap = A·p
// (identity boundary rows), p·ap; after the first iteration (!FIRST) it also
// applies the previous iteration's deferred x += α·p_old (the x/r update then
// reads neither x nor p), the same fma on the same operands, only later.  A is
// the 5-point stencil operator p + coeff·S(p) (chain order of fxd::chain5) or,
// with VK, the variable-coefficient operator of k_stream_vk.  The west and
// east neighbours of a thread's pair come from the adjacent lanes (DPP); the
// first and last lane of a wave read the column beyond the wave with the
// row's other loads (the other lanes read their own column there, unused).
// Rows are loaded four at a time, unconditionally from clamped addresses, and
// combined (p, x + α·p_old) only when they are used, so that four rows' loads
// are in flight together.  Every lane
// walks the rows, so the cross-lane moves always see all 64 lanes.  Per-point
// arithmetic equals k_stream_st / k_stream_vk.
template <bool VK, bool FIRST>
__global__ __launch_bounds__(BLOCK) void k_fcg_step2(const float* __restrict__ z, const double* __restrict__ pold,
                                                     double* __restrict__ pnew, double* __restrict__ ap, W9 w,
                                                     double coeff, const double* __restrict__ k, double sigma,
                                                     double ix2, double iy2, int64_t rows, int64_t cols, int rpb,
                                                     RedScratch rs, const SState* st, double* __restrict__ x) {
	if (st->done) return;
	constexpr bool XU = !FIRST;   // a deferred x update is pending after every iteration but the first
	const double be = st->beta;
	const double al = st->alpha;
	int bx, by;
	xcd_tile(&bx, &by);
	const int lane = (int)(threadIdx.x & (WAVE - 1));
	const int64_t c0 = 2 * ((int64_t)bx * BLOCK + threadIdx.x);
	const int64_t r0 = (int64_t)by * rpb;
	const int64_t r1 = r0 + rpb < rows ? r0 + rpb : rows;
	const bool ok = c0 < cols;
	const int64_t cl = ok ? c0 : cols - 2;   // clamped pair column
	const bool intA = c0 > 0, intB = c0 + 1 < cols - 1;   // column interior tests (c0 + 1 > 0 always)
	// Column beside the pair: west of it for the first lane, east for the last (clamped; a column a
	// lane does not use, where it would leave the grid).
	const int64_t cb = lane == 0 ? (cl > 0 ? cl - 1 : cl) : (lane == WAVE - 1 ? (cl + 2 < cols ? cl + 2 : cl) : cl);
	struct Raw {
		float2 z;
		double2 po, k, x;
		float zb;
		double pob, kb;
	};
	auto ld = [&](int64_t r) {
		const int64_t rc = r < 0 ? 0 : (r < rows ? r : rows - 1);
		const int64_t i = rc * cols + cl, ib = rc * cols + cb;
		Raw v;
		v.z = *(const float2*)(z + i);
		v.zb = z[ib];
		if constexpr (FIRST) {
			v.po = double2{0.0, 0.0};   // p_old = 0
			v.pob = 0.0;
		} else {
			v.po = *(const double2*)(pold + i);
			v.pob = pold[ib];
		}
		if constexpr (VK) {
			v.k = *(const double2*)(k + i);
			v.kb = k[ib];
		} else {
			v.k = double2{1.0, 1.0};
			v.kb = 1.0;
		}
		v.x = XU ? *(const double2*)(x + i) : double2{0.0, 0.0};
		return v;
	};
	struct Row {
		double2 p, k, xn;   // p, k and (XU) x + α·p_old of the pair
		double pb, kb;      // p and k beside the pair
	};
	auto use = [&](const Raw& v) {
		Row o;
		o.p = double2{double(v.z.x) + be * v.po.x, double(v.z.y) + be * v.po.y};
		o.k = v.k;
		o.xn = double2{fma(al, v.po.x, v.x.x), fma(al, v.po.y, v.x.y)};
		o.pb = double(v.zb) + be * v.pob;
		o.kb = v.kb;
		return o;
	};
	Acc acc[1] = {Acc{0.0, 0.0}};
	Row up = use(ld(r0 - 1));   // rows outside the grid are only read by boundary rows, which do not use them
	Row cur = use(ld(r0));
	// Row r from up, cur and dn (row r + 1, formed from its loads).
	auto row = [&](int64_t r, const Raw& nx) {
		const int64_t i = r * cols + c0;
		const Row dn = use(nx);
		double pw = dlane<0x138>(cur.p.y), pe = dlane<0x130>(cur.p.x);
		double kw = dlane<0x138>(cur.k.y), ke = dlane<0x130>(cur.k.x);
		if (lane == 0) pw = cur.pb, kw = cur.kb;
		if (lane == WAVE - 1) pe = cur.pb, ke = cur.kb;
		const bool rint = r > 0 && r < rows - 1;
		double2 av = cur.p;   // boundary rows and columns: identity
		if (rint) {
			const double wv = intA ? pw : 0.0;
			const double ev = intB ? pe : 0.0;
			if constexpr (VK) {
				auto op = [&](double kc, double kn, double ks, double kwv, double kev, double c, double n, double s_,
				              double w_, double e_) {
					const double hw = hmean_d(kc, kwv), he = hmean_d(kc, kev);
					const double hn = hmean_d(kc, kn), hs = hmean_d(kc, ks);
					const double fx = he * (e_ - c) - hw * (c - w_);
					const double fy = hs * (s_ - c) - hn * (c - n);
					const double L = -(fx * ix2 + fy * iy2);
					return fma(coeff, L, sigma * c);
				};
				if (intA) av.x = op(cur.k.x, up.k.x, dn.k.x, kw, cur.k.y, cur.p.x, up.p.x, dn.p.x, wv, cur.p.y);
				if (intB) av.y = op(cur.k.y, up.k.y, dn.k.y, cur.k.x, ke, cur.p.y, up.p.y, dn.p.y, cur.p.x, ev);
			} else {
				if (intA) av.x = fma(coeff, chain5(w, cur.p.x, up.p.x, dn.p.x, wv, cur.p.y), cur.p.x);
				if (intB) av.y = fma(coeff, chain5(w, cur.p.y, up.p.y, dn.p.y, cur.p.x, ev), cur.p.y);
			}
		}
		if (ok) {
			store_nt(pnew + i, cur.p);
			store_nt(ap + i, av);
			if (XU) store_nt(x + i, cur.xn);
			acc_add(acc[0], cur.p.x * av.x);
			acc_add(acc[0], cur.p.y * av.y);
		}
		up = cur;
		cur = dn;
	};
	// SR rows per loop trip: the loads of the trip's SR next rows are issued together at its start and
	// used within it, so no loaded value is carried round the loop (a carried one is copied between
	// registers at the loop's end, which waits for every load in flight).
	constexpr int SR = 4;
	for (int64_t r = r0; r < r1; r += SR) {
		Raw q[SR];
#pragma unroll
		for (int j = 0; j < SR; ++j) q[j] = ld(r + 1 + j);
#pragma unroll
		for (int j = 0; j < SR; ++j)
			if (r + j < r1) row(r + j, q[j]);
	}
	block_partial<1>(acc, rs);
}

// r −= α·ap with r·r, and the FP32 copy of r, two points per thread per step
// (n even); the x update of the same iteration is deferred to the next FCG
// step kernel (or k_xflush).  Per-point arithmetic equals FXR32's r part.
__global__ __launch_bounds__(BLOCK) void k_fr32v(const SState* st, double* __restrict__ r, float* __restrict__ r32,
                                                 const double* __restrict__ ap, size_t n2, RedScratch rs,
                                                 const int* skip) {
	if (skip && *skip) return;
	const double al = st->alpha;
	Acc acc[1] = {Acc{0.0, 0.0}};
	const size_t stride = (size_t)gridDim.x * BLOCK;
#pragma unroll 4
	for (size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x; j < n2; j += stride) {
		const double2 rv = ((const double2*)r)[j], av = ((const double2*)ap)[j];
		const double2 rn{fma(-al, av.x, rv.x), fma(-al, av.y, rv.y)};
		((double2*)r)[j] = rn;
		((float2*)r32)[j] = float2{float(rn.x), float(rn.y)};
		acc_add(acc[0], rn.x * rn.x);
		acc_add(acc[0], rn.y * rn.y);
	}
	block_partial<1>(acc, rs);
}

// ── scalar epilogues ───────────────────────────────────────────────────
// cg.rs:113-125: rs_old = r·r; converged at iteration 0 if ‖r‖ < tol.
struct FinCGInit {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double rs = acc_value(o[0]);
		st->rs = rs;
		st->rr = rs;
		st->beta = 0.0;
		st->iter = 0;
		record(st, 0, sqrt(rs));
		if (st->max_iter == 0) finish(st, 0, sqrt(rs), FX_MAX_ITERATIONS);
		else if (sqrt(rs) < st->tol) finish(st, 0, sqrt(rs), FX_CONVERGED);
	}
};
// cg.rs:134-142 (CG) / 223-231 (PCG): α = rs/(p·Ap), breakdown below 1e−30.
struct FinAlpha {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double pap = acc_value(o[0]);
		st->pap = pap;
		if (fabs(pap) < 1e-30) finish(st, st->iter, sqrt(st->rr), FX_BREAKDOWN);
		else st->alpha = st->rs / pap;
	}
};
// cg.rs:149-156 + the check at the top of the next iteration (cg.rs:118-125).
struct FinCGBeta {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double rs_new = acc_value(o[0]);
		st->beta = rs_new / st->rs;
		st->rs = rs_new;
		st->rr = rs_new;
		int64_t it = st->iter + 1;
		st->iter = it;
		record(st, it, sqrt(rs_new));
		if (it >= st->max_iter) finish(st, it, sqrt(rs_new), FX_MAX_ITERATIONS);
		else if (sqrt(rs_new) < st->tol) finish(st, it, sqrt(rs_new), FX_CONVERGED);
		// Release: a host that sees the count also sees the done flag stored above.
		if (st->hprog) __hip_atomic_store(st->hprog, (int)it, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
	}
};
// PCG: r·r after the x/r update (cg.rs:205-211 of the next iteration).
struct FinPcgRR {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double rr = acc_value(o[0]);
		st->rr = rr;
		int64_t it = st->iter + 1;
		st->iter = it;
		record(st, it, sqrt(rr));
		if (it >= st->max_iter) finish(st, it, sqrt(rr), FX_MAX_ITERATIONS);
		else if (sqrt(rr) < st->tol) finish(st, it, sqrt(rr), FX_CONVERGED);
		// Release: a host that sees the count also sees the done flag stored above.
		if (st->hprog) __hip_atomic_store(st->hprog, (int)it, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
	}
};
// FinPcgRR after an r-only update whose x update (x += α·p) is deferred: the
// next step kernel applies it, or k_xflush once the loop has stopped.
struct FinPcgRRDefer {
	SState* st;
	const double* p;
	__device__ void operator()(Acc* o) const {
		FinPcgRR{st}(o);
		st->xpend = 1;
		st->xp = p;
	}
};
// FinAlpha after a step kernel that applied the deferred x update.
struct FinAlphaDefer {
	SState* st;
	__device__ void operator()(Acc* o) const {
		st->xpend = 0;
		FinAlpha{st}(o);
	}
};
// The deferred x update left when the loop stopped: x += α·xp (if any), two
// points per thread per step (n even), unrolled like k_fr32v.
__global__ __launch_bounds__(BLOCK) void k_xflush(const SState* st, double* __restrict__ x, size_t n2) {
	if (!st->xpend) return;
	const double al = st->alpha;
	const double2* p = (const double2*)st->xp;
	const size_t stride = (size_t)gridDim.x * BLOCK;
#pragma unroll 4
	for (size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x; j < n2; j += stride) {
		const double2 xv = ((const double2*)x)[j], pv = p[j];
		((double2*)x)[j] = double2{fma(al, pv.x, xv.x), fma(al, pv.y, xv.y)};
	}
}
// PCG init: r = b − Ax with r·r (no check yet: cg.rs:188-202 computes z first).  r is the true
// residual here: it also starts the learned preconditioner's true-residual record.
struct FinPcgInitRR {
	SState* st;
	__device__ void operator()(Acc* o) const {
		st->rr = acc_value(o[0]);
		st->ltres = st->rr;
		st->beta = 0.0;
		st->iter = 0;
	}
};
// Learned preconditioner's fallback test on the FP64 true residual ‖b − A·x‖² after an x/r update:
// the iteration just taken followed the preconditioner application that lgate describes, and only
// an iteration after a network application (lgate = 0) is judged (the V-cycle's own iterations need
// not reduce ‖b − A·x‖ monotonically, and often raise it in the first step).
struct FinLearnJudge {
	SState* st;
	__device__ void operator()(Acc* o) const {
		const double tr = acc_value(o[0]);
		if (st->lgate == 0 && !(tr < st->stall2 * st->ltres)) {
			st->lplain = 1;
			st->lskip = 1;
		}
		st->ltres = tr;
	}
};
// Whether the network runs at this preconditioner application (lgate = 0) or only the V-cycle.
struct FinLearnGate {
	SState* st;
	int every, max_apps, first;
	__device__ void operator()() const {
		const bool on = !st->done && !st->lplain && st->iter >= first && (st->iter - first) % every == 0 &&
		                (max_apps <= 0 || st->lapps < max_apps);
		st->lgate = on ? 0 : 1;
		if (on) st->lapps += 1;
	}
};
// PCG: rz_old = r·z at start, check ‖r‖ (cg.rs:202-211).
struct FinPcgInitRZ {
	SState* st;
	__device__ void operator()(Acc* o) const {
		st->rs = acc_value(o[0]);
		double rn = sqrt(st->rr);
		record(st, 0, rn);
		if (st->max_iter == 0) finish(st, 0, rn, FX_MAX_ITERATIONS);
		else if (rn < st->tol) finish(st, 0, rn, FX_CONVERGED);
	}
};
// PCG: β = rz_new/rz_old (cg.rs:240-245).
struct FinPcgBeta {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double rz = acc_value(o[0]);
		st->beta = rz / st->rs;
		st->rs = rz;
	}
};
// FCG: (r·z, z·Ap) -> α numerator and the flexible β = −(z·Ap)/(p·Ap).
struct FinFcgBeta {
	SState* st;
	__device__ void operator()(Acc* o) const {
		st->rs = acc_value(o[0]);
		st->beta = -acc_value(o[1]) / st->pap;
	}
};

// ── pipelined CG (pipelined_cg.rs:96-200) ──────────────────────────────
// Fused 6-way update in the CPU order (pipelined_cg.rs:182-199) followed by
// the next iteration's (γ, δ) = (r·r, r·w).
struct FPipeUpd {
	const SState* st;
	double *z, *t, *p, *x, *r, *w;
	const double* q;
	double a, b;
	__device__ void load() { a = st->alpha; b = st->beta; }
	__device__ void operator()(size_t i, Acc* acc) const {
		double zi = q[i] + b * z[i];
		double ti = w[i] + b * t[i];
		double pi = r[i] + b * p[i];
		x[i] = fma(a, pi, x[i]);
		double ri = fma(-a, ti, r[i]);
		double wi = fma(-a, zi, w[i]);
		r[i] = ri;
		w[i] = wi;
		z[i] = zi;
		t[i] = ti;
		p[i] = pi;
		acc_add(acc[0], ri * ri);
		acc_add(acc[1], ri * wi);
	}
};
struct FinPipe {
	SState* st;
	bool first;
	__device__ void operator()(Acc* o) const {
		double gamma = acc_value(o[0]), delta = acc_value(o[1]);
		int64_t k = first ? 0 : st->iter + 1;
		st->iter = k;
		st->rr = gamma;
		record(st, k, sqrt(gamma));
		if (k >= st->max_iter) { finish(st, k, sqrt(gamma), FX_MAX_ITERATIONS); return; }
		if (sqrt(gamma) < st->tol) { finish(st, k, sqrt(gamma), FX_CONVERGED); return; }
		double gp = first ? INFINITY : st->gprev, ap = first ? INFINITY : st->aprev;
		double beta = gamma / gp;
		double alpha = gamma / (delta - beta * gamma / ap);
		st->beta = beta;
		st->alpha = alpha;
		st->gprev = gamma;
		st->aprev = alpha;
	}
};

// ── preconditioned pipelined CG (Ghysels & Vanroose 2014, Alg. 3) ──────
struct FPPUpd {
	const SState* st;
	double *z, *q, *s, *p, *x, *r, *u, *w;
	const double *nn, *m;
	double a, b;
	__device__ void load() { a = st->alpha; b = st->beta; }
	__device__ void operator()(size_t i, Acc* acc) const {
		double zi = nn[i] + b * z[i];
		double qi = m[i] + b * q[i];
		double si = w[i] + b * s[i];
		double pi = u[i] + b * p[i];
		x[i] = fma(a, pi, x[i]);
		double ri = fma(-a, si, r[i]);
		double ui = fma(-a, qi, u[i]);
		double wi = fma(-a, zi, w[i]);
		z[i] = zi; q[i] = qi; s[i] = si; p[i] = pi;
		r[i] = ri; u[i] = ui; w[i] = wi;
		acc_add(acc[0], ri * ui);
		acc_add(acc[1], wi * ui);
		acc_add(acc[2], ri * ri);
	}
};
struct FDot3 {
	const double *r, *u, *w;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const {
		acc_add(a[0], r[i] * u[i]);
		acc_add(a[1], w[i] * u[i]);
		acc_add(a[2], r[i] * r[i]);
	}
};
// (γ, δ, ‖r‖²) -> check, then α, β (the new β/α are applied by the next update).
struct FinPP {
	SState* st;
	bool first;
	__device__ void operator()(Acc* o) const {
		double gamma = acc_value(o[0]), delta = acc_value(o[1]), rr = acc_value(o[2]);
		int64_t k = first ? 0 : st->iter + 1;
		st->iter = k;
		st->rr = rr;
		record(st, k, sqrt(rr));
		if (k >= st->max_iter) { finish(st, k, sqrt(rr), FX_MAX_ITERATIONS); return; }
		if (sqrt(rr) < st->tol) { finish(st, k, sqrt(rr), FX_CONVERGED); return; }
		double beta, alpha;
		if (first) { beta = 0.0; alpha = gamma / delta; }
		else { beta = gamma / st->gprev; alpha = gamma / (delta - beta * gamma / st->aprev); }
		if (!isfinite(alpha)) { finish(st, k, sqrt(rr), FX_BREAKDOWN); return; }
		st->beta = beta;
		st->alpha = alpha;
		st->gprev = gamma;
		st->aprev = alpha;
	}
};

// ── pipelined flexible CG around an FP32 multigrid (extension) ─────────
// One fused reduction per iteration, as in pipelined_cg.rs:130-159, but with the preconditioner applied
// to the current residual every iteration (u = M(r), w = A·u) instead of carried by recurrences, so an
// inexact (FP32) M does not corrupt the recurrence: from (r·u, w·u, u·s_old, r·r) the flexible
// β = −(u·s_old)/(p_old·s_old), p·s = w·u + β·(u·s_old) (A symmetric) and α = (r·u)/(p·s);
// p = u + β·p, s = w + β·s (= A·p), x += α·p, r −= α·s.
// w = A·u (identity boundary rows) with the four dots.
struct EpiPF {
	const double *r, *s;
	double* w;
	__device__ void interior(int64_t i, double av, double c, Acc* a) const {
		w[i] = av;
		const double ri = r[i];
		acc_add(a[0], ri * c);
		acc_add(a[1], av * c);
		acc_add(a[2], c * s[i]);
		acc_add(a[3], ri * ri);
	}
	__device__ void boundary(int64_t i, double c, Acc* a) const { interior(i, c, c, a); }
};
// EpiPF for the 5-point stencil traversal, which hands over S(u): A·u = fma(coeff, S(u), u).
struct EpiPFs {
	double coeff;
	EpiPF e;
	__device__ void interior(int64_t i, double lap, double c, Acc* a) const { e.interior(i, fma(coeff, lap, c), c, a); }
	__device__ void boundary(int64_t i, double c, Acc* a) const { e.boundary(i, c, a); }
};
// The same reductions over a stored w (operators without a fused traversal).
struct FDotPF {
	const double *r, *u, *w, *s;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const {
		const double ri = r[i], ui = u[i];
		acc_add(a[0], ri * ui);
		acc_add(a[1], w[i] * ui);
		acc_add(a[2], ui * s[i]);
		acc_add(a[3], ri * ri);
	}
};
// Top of iteration k: ‖r‖ check (pipelined_cg.rs:134-141 order: budget, then tolerance), then α, β.
struct FinPF {
	SState* st;
	bool first;
	__device__ void operator()(Acc* o) const {
		const double gamma = acc_value(o[0]), delta = acc_value(o[1]), eps = acc_value(o[2]), rr = acc_value(o[3]);
		const int64_t k = first ? 0 : st->iter + 1;
		st->iter = k;
		st->rr = rr;
		record(st, k, sqrt(rr));
		if (k >= st->max_iter) { finish(st, k, sqrt(rr), FX_MAX_ITERATIONS); return; }
		if (sqrt(rr) < st->tol) { finish(st, k, sqrt(rr), FX_CONVERGED); return; }
		const double beta = first ? 0.0 : -eps / st->pap;
		const double eta = first ? delta : delta + beta * eps;
		if (!(fabs(eta) >= 1e-30)) { finish(st, k, sqrt(rr), FX_BREAKDOWN); return; }
		st->beta = beta;
		st->alpha = gamma / eta;
		st->pap = eta;
	}
};
// p = u + β·p; s = w + β·s; x += α·p; r −= α·s.
struct FPFUpd {
	const SState* st;
	double *p, *s, *x, *r;
	const double *u, *w;
	double a, b;
	__device__ void load() { a = st->alpha; b = st->beta; }
	__device__ void operator()(size_t i) const {
		const double pi = u[i] + b * p[i];
		const double si = w[i] + b * s[i];
		x[i] = fma(a, pi, x[i]);
		r[i] = fma(-a, si, r[i]);
		p[i] = pi;
		s[i] = si;
	}
};

// ── learned preconditioner pieces ─────────────────────────────────────
// r1 = r − A·z (in place over the A·z buffer), with Σ r1².
struct FLr1 {
	const double* r;
	double* az;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const {
		double v = r[i] - az[i];
		az[i] = v;
		acc_add(a[0], v * v);
	}
};
// s = rms(r1) (1 when r1 = 0).
struct FinScale {
	double* s;
	double n;
	__device__ void operator()(Acc* o) const {
		double v = sqrt(acc_value(o[0]) / n);
		*s = v > 0.0 ? v : 1.0;
	}
};
// Network input [p][2] = (r1/s, log k).
struct FPackIn {
	const double* r1;
	const double* logk;
	const double* s;
	float* x;
	double inv;
	__device__ void load() { inv = 1.0 / *s; }
	__device__ void operator()(size_t i) const {
		x[2 * i] = float(r1[i] * inv);
		x[2 * i + 1] = logk ? float(logk[i]) : 0.0f;
	}
};
// c = s·h²·y on interior points, 0 on the boundary (y = the network output, one value per point).
struct FNetC {
	double* c;
	const float* y;
	const double* s;
	double h2;
	int64_t rows, cols;
	double sc;
	__device__ void load() { sc = *s * h2; }
	__device__ void operator()(size_t i) const {
		int64_t r = (int64_t)i / cols, col = (int64_t)i % cols;
		c[i] = (r == 0 || col == 0 || r == rows - 1 || col == cols - 1) ? 0.0 : sc * double(y[i]);
	}
};
// α* = (c·r1)/(c·Ac), the step along c that minimises the A-norm of the error left after the V-cycle
// (r1 = A·e1), clipped at 0: adding α*·c never increases that error.
struct FinStep {
	double* a;
	__device__ void operator()(Acc* o) const {
		const double cr = acc_value(o[0]), cac = acc_value(o[1]);
		*a = (cac > 0.0 && cr > 0.0) ? cr / cac : 0.0;
	}
};
// z += α*·c.
struct FAddStep {
	double* z;
	const double* c;
	const double* a;
	double al;
	__device__ void load() { al = *a; }
	__device__ void operator()(size_t i) const { z[i] = fma(al, c[i], z[i]); }
};
// z += s·h²·y on interior points (y = the network output, one value per point).
struct FAddNet {
	double* z;
	const float* y;
	const double* s;
	double h2;
	int64_t rows, cols;
	double sc;
	__device__ void load() { sc = *s * h2; }
	__device__ void operator()(size_t i) const {
		int64_t r = (int64_t)i / cols, c = (int64_t)i % cols;
		if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) return;
		z[i] += sc * double(y[i]);
	}
};

// ── fused mixed-precision FCG pieces ───────────────────────────────────
// p_new = z + β·p_old with z held in FP32 (exact widening, same value as the FP64 copy).
struct VPNewF {
	const float* z;
	const double* pold;
	double be;
	__device__ double load(int64_t i) const { return double(z[i]) + be * pold[i]; }
};
// r = b − ap with r·r, and the FP32 copy of r for the V-cycle.
struct FResid32 {
	const double *b, *ap;
	double* r;
	float* r32;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const {
		double v = b[i] - ap[i];
		r[i] = v;
		r32[i] = float(v);
		acc_add(a[0], v * v);
	}
};
// x += α·p; r −= α·ap with r·r, and the FP32 copy of r.
struct FXR32 {
	const SState* st;
	double* x;
	double* r;
	float* r32;
	const double *p, *ap;
	double al;
	__device__ void load() { al = st->alpha; }
	__device__ void operator()(size_t i, Acc* a) const {
		x[i] = fma(al, p[i], x[i]);
		double v = fma(-al, ap[i], r[i]);
		r[i] = v;
		r32[i] = float(v);
		acc_add(a[0], v * v);
	}
};
// Level-0 epilogues of the fused V-cycle: r·z (start) and (r·z, z·Ap) (iterations).
// The FCG x/r update's residual part, carried by the level-0 wave restriction
// of the fused V-cycle (fmg::upd_ok): r_new = r − α·A·p with k_fr32v's
// arithmetic at every point, read from r and stored to a second vector (other
// waves still read r around the points a wave owns), and r_new·r_new reduced
// into FinPcgRRDefer, whose deferred x update is along p.
struct RUpd {
	static constexpr bool ON = true;
	const double *r, *ap;
	double* rn;
	SState* st;
	const double* p;
	double al;
	struct Pf {
		double r, ap;
	};
	__device__ void begin() { al = st->alpha; }
	__device__ Pf load_at(size_t o, unsigned c) const { return Pf{(r + o)[c], (ap + o)[c]}; }
	__device__ double rnew(const Pf& v) const { return fma(-al, v.ap, v.r); }
	__device__ void store(size_t i, double v) const { store_nt(rn + i, v); }
	using Fin = FinPcgRRDefer;
	Fin fin() const { return Fin{st, p}; }
};

// RHS_FROM_R: the level-0 rhs of the V-cycle these epilogues end is float(r)
// (written by the residual and x/r-update kernels), so the wave post kernel
// takes it from the residual it loads for the reduction.
struct EpiRZ {
	static constexpr int NS = 1;
	static constexpr bool RHS_FROM_R = true;
	const double* r;
	SState* st;
	__device__ void term(int64_t j, float z, Acc* a) const { acc_add(a[0], r[j] * double(z)); }
	struct Pf {
		double r;
	};
	__device__ Pf load(int64_t j) const { return Pf{r[j]}; }
	__device__ void use(const Pf& p, float z, Acc* a) const { acc_add(a[0], p.r * double(z)); }
	// Row-offset form (uniform row offset o, lane column c) and plain per-lane sums (wave kernels).
	__device__ Pf load_at(size_t o, unsigned c) const { return Pf{(r + o)[c]}; }
	__device__ void use_lane(const Pf& p, float z, double* s) const { s[0] += p.r * double(z); }
	__device__ float rhs(const Pf& p) const { return float(p.r); }
	using Fin = FinPcgInitRZ;
	Fin fin() const { return Fin{st}; }
};
struct EpiFcg {
	static constexpr int NS = 2;
	static constexpr bool RHS_FROM_R = true;
	const double *r, *ap;
	SState* st;
	__device__ void term(int64_t j, float z, Acc* a) const {
		double zd = double(z);
		acc_add(a[0], r[j] * zd);
		acc_add(a[1], zd * ap[j]);
	}
	struct Pf {
		double r, ap;
	};
	__device__ Pf load(int64_t j) const { return Pf{r[j], ap[j]}; }
	__device__ void use(const Pf& p, float z, Acc* a) const {
		const double zd = double(z);
		acc_add(a[0], p.r * zd);
		acc_add(a[1], zd * p.ap);
	}
	__device__ Pf load_at(size_t o, unsigned c) const { return Pf{(r + o)[c], (ap + o)[c]}; }
	__device__ void use_lane(const Pf& p, float z, double* s) const {
		const double zd = double(z);
		s[0] += p.r * zd;
		s[1] += zd * p.ap;
	}
	__device__ float rhs(const Pf& p) const { return float(p.r); }
	using Fin = FinFcgBeta;
	Fin fin() const { return Fin{st}; }
};

// ── BiCGSTAB (bicgstab.rs:90-346) ──────────────────────────────────────
// p = r + β·(p − ω·v)  (axpy(−ω, v), scale(β), axpy(1, r): bicgstab.rs:134-137)
struct FBiP {
	const SState* st;
	double* p;
	const double *r, *v;
	double be, om;
	__device__ void load() { be = st->beta; om = st->omega; }
	__device__ void operator()(size_t i) const {
		double t = fma(-om, v[i], p[i]);
		p[i] = r[i] + be * t;
	}
};
// s = r − α·v (copy + axpy: bicgstab.rs:156-158), with s·s.
struct FBiS {
	const SState* st;
	double* s;
	const double *r, *v;
	double al;
	__device__ void load() { al = st->alpha; }
	__device__ void operator()(size_t i, Acc* a) const {
		double t = fma(-al, v[i], r[i]);
		s[i] = t;
		acc_add(a[0], t * t);
	}
};
// x += α·p̂ + ω·ŝ; r = s − ω·t (bicgstab.rs:190-196), then (r·r, r̂·r) for the next iteration.
// While `pend` is set only x += α·p̂ is applied (bicgstab.rs:160-172, 180-188).
struct FBiX {
	const SState* st;
	double *x, *r;
	const double *pa, *sa, *s, *t, *rh;
	double al, om;
	int pend;
	__device__ void load() { al = st->alpha; om = st->omega; pend = st->pend; }
	__device__ void operator()(size_t i, Acc* a) const {
		double xv = fma(al, pa[i], x[i]);
		if (pend) { x[i] = xv; return; }
		x[i] = fma(om, sa[i], xv);
		double rv = fma(-om, t[i], s[i]);
		r[i] = rv;
		acc_add(a[0], rv * rv);
		acc_add(a[1], rh[i] * rv);
	}
};
// Top of iteration i (bicgstab.rs:115-132): ‖r‖ check, ρ_new = r̂·r, β.
struct FinBiTop {
	SState* st;
	bool first;
	__device__ void operator()(Acc* o) const {
		if (!first && st->pend) {   // finish the early exit with the values recorded by FinBiS / FinBiOm
			set_done(st);
			st->skip = 1;
			return;
		}
		double rr = acc_value(o[0]), rho_new = acc_value(o[1]);
		if (!first) st->rs = st->rho_new;   // rho = rho_new (bicgstab.rs:198)
		int64_t i = first ? 0 : st->iter + 1;
		st->iter = i;
		st->rr = rr;
		double rn = sqrt(rr);
		record(st, i, rn);
		if (i >= st->max_iter) { finish(st, i, rn, FX_MAX_ITERATIONS); return; }
		if (rn < st->tol) { finish(st, i, rn, FX_CONVERGED); return; }
		if (fabs(rho_new) < 1e-30) { finish(st, i, rn, FX_BREAKDOWN); return; }
		st->rho_new = rho_new;
		st->beta = (rho_new / st->rs) * (st->alpha / st->omega);
	}
};
// α = ρ_new / (r̂·v) (bicgstab.rs:145-154).
struct FinBiAlpha {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double rv = acc_value(o[0]);
		if (fabs(rv) < 1e-30) { finish(st, st->iter, sqrt(st->rr), FX_BREAKDOWN); return; }
		st->alpha = st->rho_new / rv;
	}
};
// ‖s‖ < tol -> x += α·p̂ and stop (bicgstab.rs:160-172).
struct FinBiS {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double sn = sqrt(acc_value(o[0]));
		st->snorm = sn;
		if (sn < st->tol) {
			st->pend = 1;
			st->skip = 1;
			st->reason = FX_CONVERGED;
			st->resid = sn;
		}
	}
};
// ω = (t·s)/(t·t); breakdown -> x += α·p̂ and stop (bicgstab.rs:178-189).
struct FinBiOm {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double ts = acc_value(o[0]), tt = acc_value(o[1]);
		if (fabs(tt) < 1e-30) {
			st->pend = 1;
			st->skip = 1;
			st->reason = FX_BREAKDOWN;
			st->resid = st->snorm;
			return;
		}
		st->omega = ts / tt;
	}
};
struct FinNone {
	__device__ void operator()(Acc*) const {}
};
struct FinBiInit {
	SState* st;
	__device__ void operator()(Acc*) const {
		st->rs = 1.0;      // rho
		st->alpha = 1.0;
		st->omega = 1.0;
		st->iter = 0;
	}
};

// ── GMRES(m) (gmres.rs:116-490) ────────────────────────────────────────
// v0 = b − A·x (scale(−1) + axpy(1, b): gmres.rs:127-134) with ‖v0‖².
struct FGmR0 {
	const double* b;
	double* v;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const {
		double t = b[i] + (-v[i]);
		v[i] = t;
		acc_add(a[0], t * t);
	}
};
struct FSumSq {
	const double* v;
	__device__ void load() {}
	__device__ void operator()(size_t i, Acc* a) const { acc_add(a[0], v[i] * v[i]); }
};
// v *= 1/β (gmres.rs:146) or v = w·(1/h) (gmres.rs:183-186); the factor is 1/state value.
struct FGmScale {
	const double* den;
	const double* src;
	double* dst;
	double inv;
	__device__ void load() { inv = 1.0 / *den; }
	__device__ void operator()(size_t i) const { dst[i] = src[i] * inv; }
};
// MGS (gmres.rs:171-176): w −= h_{j−1}·v_{j−1} (if j > 0) then w·v_j.
struct FGmMgs {
	double* w;
	const double *vprev, *vj;
	const double* hprev;
	double h;
	int has_prev;
	__device__ void load() { h = has_prev ? *hprev : 0.0; }
	__device__ void operator()(size_t i, Acc* a) const {
		double t = w[i];
		if (has_prev) {
			t = fma(-h, vprev[i], t);
			w[i] = t;
		}
		acc_add(a[0], t * (vj ? vj[i] : t));
	}
};
__device__ __forceinline__ double& Hm(SState* s, int i, int j) { return s->H[j * (s->m + 1) + i]; }
// Start of a cycle (gmres.rs:136-150): β = ‖v0‖.
struct FinGmStart {
	SState* st;
	__device__ void operator()(Acc* o) const {
		double beta = sqrt(acc_value(o[0]));
		record(st, st->iter, beta);
		if (beta < st->tol) { finish(st, st->iter, beta, FX_CONVERGED); return; }
		// gmres.rs:155-164: the iteration budget is checked before the first Arnoldi step.
		if (st->iter >= st->max_iter) { finish(st, st->iter, beta, FX_MAX_ITERATIONS); return; }
		st->gbeta = beta;
		st->gcur = beta;
		for (int i = 0; i <= st->m; ++i) st->grs[i] = 0.0;
		st->grs[0] = beta;
		for (int i = 0; i < (st->m + 1) * st->m; ++i) st->H[i] = 0.0;
		st->kcur = 0;
		st->cycle_end = 0;
		st->gexit = 0;
		st->skip = 0;
	}
};
struct FinGmH {
	SState* st;
	int j, k;
	__device__ void operator()(Acc* o) const { Hm(st, j, k) = acc_value(o[0]); }
};
// End of an Arnoldi step (gmres.rs:178-247): h_{k+1,k}, Givens, |g_{k+1}|, stopping tests.
struct FinGmStep {
	SState* st;
	int k;
	__device__ void operator()(Acc* o) const {
		double hk = sqrt(acc_value(o[0]));
		Hm(st, k + 1, k) = hk;
		st->ghk = hk;
		for (int j = 0; j < k; ++j) {
			double hj = Hm(st, j, k), hj1 = Hm(st, j + 1, k), cj = st->cc[j], sj = st->ss[j];
			Hm(st, j, k) = fma(cj, hj, sj * hj1);
			Hm(st, j + 1, k) = fma(cj, hj1, -sj * hj);
		}
		double a = Hm(st, k, k), b = Hm(st, k + 1, k);
		double r = hypot(a, b);
		if (fabs(r) < 1e-30) {
			st->gexit = 1;
			st->cycle_end = 1;
			st->skip = 1;
			return;
		}
		st->cc[k] = a / r;
		st->ss[k] = b / r;
		Hm(st, k, k) = r;
		Hm(st, k + 1, k) = 0.0;
		double gk = st->grs[k];
		st->grs[k] = st->cc[k] * gk;
		st->grs[k + 1] = -st->ss[k] * gk;
		st->gcur = fabs(st->grs[k + 1]);
		st->iter += 1;
		st->kcur = k + 1;
		record(st, st->iter, st->gcur);
		if (st->gcur < st->tol || hk <= 1e-30 || k + 1 >= st->m || st->iter >= st->max_iter) {
			st->cycle_end = 1;
			st->skip = 1;
		}
	}
};
// v_{k+1} = w/h only when h > 1e−30 (gmres.rs:181-187).
struct FGmNorm {
	const SState* st;
	const double* w;
	double* dst;
	double inv;
	int ok;
	__device__ void load() { ok = st->ghk > 1e-30; inv = ok ? 1.0 / st->ghk : 0.0; }
	__device__ void operator()(size_t i) const {
		if (ok) dst[i] = w[i] * inv;
	}
};
// Back-substitution (gmres.rs:461-481).
struct FinGmBack {
	SState* st;
	__device__ void operator()() const {
		int k = st->kcur, mp1 = st->m + 1;
		for (int i = 0; i < k; ++i) st->y[i] = 0.0;
		for (int i = k - 1; i >= 0; --i) {
			double s = st->grs[i];
			for (int j = i + 1; j < k; ++j) s -= st->H[j * mp1 + i] * st->y[j];
			double diag = st->H[i * mp1 + i];
			if (fabs(diag) < 1e-30) break;
			st->y[i] = s / diag;
		}
	}
};
// x += Σ y_i·v_i in order (gmres.rs:483-487).
struct VecList {
	const double* v[GM_MAX + 1];
};
struct FGmX {
	const SState* st;
	double* x;
	VecList vl;
	int k;
	const double* y;
	__device__ void load() { k = st->kcur; y = st->y; }
	__device__ void operator()(size_t i) const {
		double t = x[i];
		for (int j = 0; j < k; ++j) t = fma(y[j], vl.v[j][i], t);
		x[i] = t;
	}
};
// After build_solution (gmres.rs:250-283): stop or restart.
struct FinGmEnd {
	SState* st;
	__device__ void operator()() const {
		if (st->gexit) { finish(st, st->iter, st->gcur, FX_BREAKDOWN); return; }
		if (st->gcur < st->tol) { finish(st, st->iter, st->gcur, FX_CONVERGED); return; }
		if (st->iter >= st->max_iter) { finish(st, st->iter, st->gcur, FX_MAX_ITERATIONS); return; }
		st->cycle_end = 0;
		st->skip = 0;
	}
};
struct FinMaxCheck {
	SState* st;
	__device__ void operator()() const {
		if (st->iter >= st->max_iter) finish(st, st->iter, st->gcur, FX_MAX_ITERATIONS);
	}
};

}  // namespace

// ════════════════════════════════════════════════════════════════════════
struct fx_solver {
	fx_backend* be = nullptr;
	int kind = 0;
	fx_grid g{};
	size_t n = 0;
	uint32_t restart = 0;
	double tol = 0;
	uint64_t max_iter = 0;
	std::vector<double*> v;      // work vectors
	std::vector<double*> basis;  // GMRES Krylov basis (m+1)
	std::vector<double*> zbasis; // FGMRES preconditioned directions z_k = M⁻¹v_k (m)
	SState* dst = nullptr;       // device state
	SState* hst = nullptr;       // pinned host copy
	double* dhist = nullptr;
	double* dgm = nullptr;       // GMRES H, cc, ss, grs, y
	int* hdone = nullptr;        // coherent pinned host ints: [0] copy of the device `done` flag, [1] iteration count
	hipEvent_t ev[8] = {};
	// Residual history of the last pass: hist_n entries in dhist, copied to hist when first asked for.
	mutable std::vector<double> hist;
	mutable bool hist_stale = false;
	int64_t hist_n = 0;
	double* vec(int i) {
		while ((int)v.size() <= i) {
			void* p = nullptr;
			if (fx::dev_alloc(be, n * sizeof(double), &p) != FX_OK) return nullptr;
			v.push_back((double*)p);
		}
		return v[i];
	}
};

// Learned preconditioner z = V(r) + s·h²·N([r1/s, log k]), r1 = r − A·V(r), s = rms(r1).
struct fx_learned {
	fx_backend* be = nullptr;
	fx_unet* net = nullptr;
	fx_multigrid* mg = nullptr;
	const fx_buf* logk = nullptr;
	fx_grid g{};
	double stall = 0.9;
	double* s_dev = nullptr;   // [0] rms(r1), [1] the step α* of the network correction
	double* r1 = nullptr;
	double* c = nullptr;       // the network correction
	double* ac = nullptr;      // A·c
	float* xin = nullptr;
	double* tr = nullptr;      // b − A·x of the fallback test
	// Network schedule: from iteration `first` (default 0), every `every`-th, at most `max_apps` times.
	int every = 1, max_apps = 0, first = 0;
	int fell_back = 0;
};

namespace {

using fx::check_launch;

const char* solver_name(int kind) {
	switch (kind) {
	case FX_SOLVER_CG: return "CG";
	case FX_SOLVER_PIPELINED_CG: return "PipelinedCG";
	case FX_SOLVER_BICGSTAB: return "BiCGSTAB";
	case FX_SOLVER_GMRES: return "GMRES";
	case FX_SOLVER_FCG: return "FCG";
	case FX_SOLVER_FGMRES: return "FGMRES";
	default: return "?";
	}
}

struct Ctx {
	fx_solver* s;
	fx_backend* be;
	hipStream_t st;
	const fx_operator* op;
	const fx_precond* pc;
	size_t n;
	int64_t rows, cols;
	SState* d;
	const int* done;   // &d->done
	const int* skip;   // &d->skip
	bool fused5;       // operator is the reference 5-point stencil kind
	W9 w9;

	unsigned eb() const { return fx::elem_blocks(n); }
	unsigned rb() const { return fx::red_blocks(n); }

	template <int NS, class F, class Fin>
	fx_status ered(F f, Fin fin, const char* what, const int* sk) {
		unsigned nb = rb();
		FX_CHECK(fx::ensure_partials(be, (size_t)nb * NS));
		hipLaunchKernelGGL((k_ered<NS, F>), dim3(nb), dim3(BLOCK), 0, st, f, n, fx::red_scratch(be), sk);
		hipLaunchKernelGGL((k_finish<NS, Fin>), dim3(1), dim3(FIN_BLOCK), 0, st, fx::red_scratch(be), nb, fin, sk);
		return check_launch(what);
	}
	template <int NS, class F, class Fin>
	fx_status ered(F f, Fin fin, const char* what) { return ered<NS>(f, fin, what, skip); }
	template <class F>
	fx_status eupd(F f, const char* what, const int* sk) {
		hipLaunchKernelGGL(k_eupd<F>, dim3(eb()), dim3(BLOCK), 0, st, f, n, sk);
		return check_launch(what);
	}
	template <class F>
	fx_status eupd(F f, const char* what) { return eupd(f, what, skip); }
	template <class Fin>
	fx_status scalar(Fin fin, const char* what, const int* sk) {
		hipLaunchKernelGGL(k_scalar<Fin>, dim3(1), dim3(1), 0, st, fin, sk);
		return check_launch(what);
	}
	fx_status apply_op(const double* x, double* y, const int* sk) {
		fx_grid g = s->g;
		return fx::k_apply_operator(be, op, x, y, &g, sk);
	}
	fx_status apply_op(const double* x, double* y) { return apply_op(x, y, skip); }
	// r = b − A·x in one pass (variable-coefficient operator), then fin on r·r.
	template <class E, class Fin>
	fx_status resid_vk(const double* x, E e, Fin fin, const char* what) {
		const double ix2 = 1.0 / (s->g.dx * s->g.dx), iy2 = 1.0 / (s->g.dy * s->g.dy);
		const double* kp = (const double*)op->k->ptr;
		dim3 grid;
		if (cols % 2 == 0) {
			// Two columns per thread; row blocks as tall as 1024 workgroups allow.
			const int64_t cb = (cols / 2 + BLOCK - 1) / BLOCK;
			int rpb = 64;
			while (rpb > 4 && cb * ((rows + rpb - 1) / rpb) < 1024) rpb /= 2;
			grid = fx::stream_grid(rows, cols / 2, rpb);
			FX_CHECK(fx::ensure_partials(be, (size_t)grid.x * grid.y));
			hipLaunchKernelGGL((k_resid2<E>), grid, dim3(BLOCK), 0, st, x, kp, e, op->sigma, op->coeff, ix2, iy2, rows,
			                   cols, rpb, fx::red_scratch(be));
		} else {
			const int rpb = stream_rpb(rows, cols);
			grid = fx::stream_grid(rows, cols, rpb);
			FX_CHECK(fx::ensure_partials(be, (size_t)grid.x * grid.y));
			hipLaunchKernelGGL((k_resid_vk<E>), grid, dim3(BLOCK), 0, st, VX{x, 0.0}, e, kp, op->sigma, op->coeff,
			                   ix2, iy2, rows, cols, rpb, fx::red_scratch(be));
		}
		hipLaunchKernelGGL((k_finish<1, Fin>), dim3(1), dim3(FIN_BLOCK), 0, st, fx::red_scratch(be),
		                   grid.x * grid.y, fin, nullptr);
		return check_launch(what);
	}
	// z = M·r, z zeroed first (cg.rs:193-194); identity copies (preconditioner.rs:29-37).
	fx_status precond(double* z, const double* r, const int* sk) {
		if (!pc || pc->kind == FX_PC_IDENTITY) return eupd(FCopy{r, z}, "pc identity", sk);
		if (pc->kind == FX_PC_MULTIGRID) {
			FX_CHECK(eupd(FZero{z}, "pc zero", sk));
			return fx::mg_apply((fx_multigrid*)pc->handle, z, r, sk);
		}
		if (pc->kind == FX_PC_LEARNED) return learned((fx_learned*)pc->handle, z, r, sk, &d->lskip);
		fx::set_error("preconditioner kind %d not supported here", pc->kind);
		return FX_ERR_UNSUPPORTED;
	}
	fx_status precond(double* z, const double* r) { return precond(z, r, skip); }

	// z = V(r), then the network correction unless *lsk is set.
	fx_status learned(fx_learned* L, double* z, const double* r, const int* sk, const int* lskip) {
		(void)lskip;
		FX_CHECK(eupd(FZero{z}, "pc zero", sk));
		FX_CHECK(fx::mg_apply(L->mg, z, r, sk));
		FX_CHECK(scalar(FinLearnGate{d, L->every, L->max_apps, L->first}, "learned gate", nullptr));
		const int* lsk = &d->lgate;
		FX_CHECK(apply_op(z, L->r1, lsk));
		FX_CHECK(ered<1>(FLr1{r, L->r1}, FinScale{L->s_dev, (double)n}, "learned r1", lsk));
		hipLaunchKernelGGL(k_eupd<FPackIn>, dim3(eb()), dim3(BLOCK), 0, st,
		                   FPackIn{L->r1, L->logk ? (const double*)L->logk->ptr : nullptr, L->s_dev, L->xin, 0.0}, n, lsk);
		FX_CHECK(check_launch("learned pack"));
		const float* y = nullptr;
		FX_CHECK(fx::unet_infer(L->net, L->xin, 1, (int)rows, (int)cols, lsk, &y));
		hipLaunchKernelGGL(k_eupd<FNetC>, dim3(eb()), dim3(BLOCK), 0, st,
		                   FNetC{L->c, y, L->s_dev, s->g.dx * s->g.dy, rows, cols, 0.0}, n, lsk);
		FX_CHECK(check_launch("learned correction"));
		FX_CHECK(apply_op(L->c, L->ac, lsk));
		FX_CHECK(ered<2>(FDot2{L->c, L->r1, L->c, L->ac}, FinStep{L->s_dev + 1}, "learned step", lsk));
		return eupd(FAddStep{z, L->c, L->s_dev + 1, 0.0}, "learned add", lsk);
	}

	// p_out = src + β·p_old; ap = A·p_out; p·ap -> α (FinAlpha).
	fx_status cg_step1(const double* src, const double* pold, double* pout, double* ap) {
		if (fused5) {
			int rpb = stream_rpb(rows, cols);
			dim3 grid = fx::stream_grid(rows, cols, rpb);
			FX_CHECK(fx::ensure_partials(be, (size_t)grid.x * grid.y));
			hipLaunchKernelGGL((k_stream_st<1, VPNew, EpiCG>), grid, dim3(BLOCK), 0, st, VPNew{src, pold, 0.0},
			                   EpiCG{op->coeff, pout, ap}, w9, rows, cols, rpb, fx::red_scratch(be), (const SState*)d);
			hipLaunchKernelGGL((k_finish<1, FinAlpha>), dim3(1), dim3(FIN_BLOCK), 0, st, fx::red_scratch(be),
			                   grid.x * grid.y, FinAlpha{d}, done);
			return check_launch("cg fused step");
		}
		if (op->kind == FX_OP_VARK && cols % 2 == 0) {
			// p = src + β·p_old, ap = A·p and p·ap in one two-column pass over src, p_old and k.
			const int64_t cb = (cols / 2 + BLOCK - 1) / BLOCK;
			int rpb = 64;
			while (rpb > 4 && cb * ((rows + rpb - 1) / rpb) < 1024) rpb /= 2;
			const dim3 grid = fx::stream_grid(rows, cols / 2, rpb);
			FX_CHECK(fx::ensure_partials(be, (size_t)grid.x * grid.y));
			hipLaunchKernelGGL((k_pair5<1, true, VPNew2, EpiCGvk>), grid, dim3(BLOCK), 0, st, VPNew2{src, pold, d, 0.0},
			                   (const double*)op->k->ptr,
			                   EpiCGvk{op->sigma, op->coeff, 1.0 / (s->g.dx * s->g.dx), 1.0 / (s->g.dy * s->g.dy), pout, ap},
			                   rows, cols, rpb, fx::red_scratch(be), done);
			hipLaunchKernelGGL((k_finish<1, FinAlpha>), dim3(1), dim3(FIN_BLOCK), 0, st, fx::red_scratch(be),
			                   grid.x * grid.y, FinAlpha{d}, done);
			return check_launch("cg fused step (variable coefficient)");
		}
		FX_CHECK(eupd(FPUpd{d, src, pold, pout, 0.0}, "p update"));
		FX_CHECK(apply_op(pout, ap));
		return ered<1>(FDot{pout, ap}, FinAlpha{d}, "p.ap");
	}
};

fx_status make_ctx(fx_solver* s, const fx_operator* op, const fx_precond* pc, Ctx* c) {
	c->s = s;
	c->be = s->be;
	c->st = s->be->stream;
	c->op = op;
	c->pc = pc;
	c->n = s->n;
	c->rows = (int64_t)s->g.rows;
	c->cols = (int64_t)s->g.cols;
	c->d = s->dst;
	c->done = &s->dst->done;
	c->skip = &s->dst->skip;
	c->fused5 = false;
	if (op->kind == FX_OP_STENCIL) {
		FX_CHECK(fx::check_bc(op->bc));
		FX_CHECK(fx::stencil_args(op->stencil, &c->w9, nullptr));
		c->fused5 = op->stencil->n == 5;
	}
	FX_REQUIRE(!pc || pc->kind == FX_PC_IDENTITY || pc->kind == FX_PC_MULTIGRID || pc->kind == FX_PC_LEARNED,
	           FX_ERR_UNSUPPORTED, "preconditioner kind %d is not supported", pc ? pc->kind : -1);
	if (pc && pc->kind == FX_PC_LEARNED) {
		fx_learned* L = (fx_learned*)pc->handle;
		FX_REQUIRE(L && L->g.rows == s->g.rows && L->g.cols == s->g.cols, FX_ERR_DIMENSION,
		           "learned preconditioner grid does not match the solver grid");
	}
	if (pc && pc->kind == FX_PC_MULTIGRID) {
		fx_multigrid* mg = (fx_multigrid*)pc->handle;
		FX_REQUIRE(mg && mg->lv[0].g.rows == s->g.rows && mg->lv[0].g.cols == s->g.cols, FX_ERR_DIMENSION,
		           "multigrid grid does not match the solver grid");
	}
	return FX_OK;
}

// Resets the device state for a new solve.
fx_status reset_state(fx_solver* s, int64_t max_iter, const fx_precond* pc = nullptr) {
	SState h;
	std::memset(&h, 0, sizeof h);
	if (pc && pc->kind == FX_PC_LEARNED) {
		h.lmode = 1;
		double st = ((fx_learned*)pc->handle)->stall;
		h.stall2 = st * st;
	}
	h.tol = s->tol;
	h.max_iter = max_iter;
	__atomic_store_n(s->hdone, 0, __ATOMIC_RELEASE);
	__atomic_store_n(s->hdone + 1, 0, __ATOMIC_RELEASE);
	void* hd = nullptr;
	FX_HIP(hipHostGetDevicePointer(&hd, s->hdone, 0));
	h.hdone = (int*)hd;
	h.hprog = (int*)hd + 1;
	h.hist = s->dhist;
	h.hist_len = HIST_MAX;
	h.m = (int32_t)s->restart;
	if (s->dgm) {
		size_t m = s->restart;
		h.H = s->dgm;
		h.cc = h.H + (m + 1) * m;
		h.ss = h.cc + m;
		h.grs = h.ss + m;
		h.y = h.grs + m + 1;
	}
	*s->hst = h;
	// hst is pinned and the host next writes it only through fetch_stats (a copy ordered after this
	// one on the stream), so the copy needs no synchronisation here.
	FX_HIP(hipMemcpyAsync(s->dst, s->hst, sizeof(SState), hipMemcpyHostToDevice, s->be->stream));
	return FX_OK;
}

// A step failed part-way: let the kernels already queued finish before returning, so that none of
// them stores to the solver's host-visible flags after the next solve has reset them.
fx_status drain_on_error(hipStream_t st, fx_status e) {
	(void)hipStreamSynchronize(st);
	return e;
}

// Enqueues `step(i)` until the device reports done, polling `look` steps behind.
template <class Step>
fx_status drive(fx_solver* s, Step step, int64_t max_steps, int look = 2) {
	constexpr int RING = 8;
	hipStream_t st = s->be->stream;
	for (int64_t i = 0;; ++i) {
		if (i > max_steps + look + 2) break;
		if (const fx_status e = step(i); e != FX_OK) return drain_on_error(st, e);
		FX_HIP(hipEventRecord(s->ev[i % RING], st));
		if (i >= look) {
			// Once step i − look has completed, its kernels' stores to *hdone are visible here.
			FX_HIP(hipEventSynchronize(s->ev[(i - look) % RING]));
			if (__atomic_load_n(s->hdone, __ATOMIC_ACQUIRE)) break;
		}
	}
	FX_HIP(hipStreamSynchronize(st));
	return FX_OK;
}

// drive() for loops whose step i stores the iteration count i + 1 in host
// memory (FinPcgRR, FinPcgRRDefer, FinCGBeta): after enqueuing step i it waits,
// without stream events (each of which stalls the queue for a few
// microseconds), until step i − lag has stored its count.  With lag 0 (the
// count stored before the step's preconditioner) the rest of step i is still
// queued while step i + 1 is enqueued, so the device does not run dry and once
// the solve is done no further step has been enqueued; a step that stores its
// count at its end takes lag 2, and the steps enqueued past the last return
// early on the done flag.  An emptied stream (an error, or a step that ends early) also ends
// the wait; the stream is queried only once the count has not moved for
// QUIET_US, since each query leaves a marker in the queue that holds the next
// kernel back by a few microseconds.
template <class Step>
fx_status drive_counted(fx_solver* s, Step step, int64_t max_steps, int lag = 0) {
	using clk = std::chrono::steady_clock;
	constexpr int64_t QUIET_US = 2000;
	hipStream_t st = s->be->stream;
	const volatile int* done = s->hdone;
	const volatile int* prog = s->hdone + 1;
	for (int64_t i = 0; i <= max_steps + 2 + lag; ++i) {
		if (const fx_status e = step(i); e != FX_OK) return drain_on_error(st, e);
		const clk::time_point t0 = clk::now();
		for (unsigned spin = 0;; ++spin) {
			if (__atomic_load_n(prog, __ATOMIC_ACQUIRE) >= i + 1 - lag || __atomic_load_n(done, __ATOMIC_ACQUIRE))
				break;
			if ((spin & 63) == 63 &&
			    std::chrono::duration_cast<std::chrono::microseconds>(clk::now() - t0).count() >= QUIET_US) {
				const hipError_t q = hipStreamQuery(st);
				if (q == hipSuccess) break;
				FX_REQUIRE(q == hipErrorNotReady, FX_ERR_HIP, "solver stream failed");
			}
		}
		if (__atomic_load_n(done, __ATOMIC_ACQUIRE)) break;
	}
	// No synchronisation: the caller's next work (x flush, true residual, statistics copy) is ordered
	// after the loop's kernels on the stream, and the statistics copy synchronises.
	return FX_OK;
}

fx_status fetch_stats(fx_solver* s, fx_solve_stats* out) {
	FX_HIP(hipMemcpyAsync(s->hst, s->dst, sizeof(SState), hipMemcpyDeviceToHost, s->be->stream));
	FX_HIP(hipStreamSynchronize(s->be->stream));
	s->hist_n = std::max<int64_t>(std::min<int64_t>(s->hst->iter + 1, HIST_MAX), 0);
	s->hist_stale = true;   // copied by fx_solver_history, when asked for
	if (out) {
		out->iterations = (uint64_t)s->hst->iter;
		out->residual = s->hst->resid;
		out->reason = s->hst->reason;
		out->true_residual = NAN;
	}
	FX_REQUIRE(s->hst->done, FX_ERR_HIP, "solver did not reach a stopping state");
	return FX_OK;
}

// ‖b − A·x‖₂ in FP64 (overwrites work vector `tmp`).
fx_status true_residual(Ctx& c, const double* x, const double* b, double* tmp, double* out) {
	if (c.op->kind == FX_OP_VARK) {
		FX_CHECK(c.resid_vk(x, EpiResNorm{b}, FinPcgInitRR{c.d}, "true residual"));
	} else {
		FX_CHECK(c.apply_op(x, tmp, nullptr));
		FX_CHECK(c.ered<1>(FResid{b, tmp, tmp}, FinPcgInitRR{c.d}, "true residual", nullptr));
	}
	FX_HIP(hipMemcpyAsync(c.s->hst, c.d, sizeof(SState), hipMemcpyDeviceToHost, c.st));
	FX_HIP(hipStreamSynchronize(c.st));
	*out = std::sqrt(c.s->hst->rr);
	return FX_OK;
}

// ── CG / PCG (cg.rs:96-256) ───────────────────────────────────────────
fx_status solve_cg(Ctx& c, double* x, const double* b, fx_solve_stats* stats) {
	fx_solver* s = c.s;
	double *r = s->vec(0), *p0 = s->vec(1), *p1 = s->vec(2), *ap = s->vec(3);
	FX_REQUIRE(r && p0 && p1 && ap, FX_ERR_ALLOC, "work vector allocation failed");
	FX_CHECK(reset_state(s, (int64_t)s->max_iter, c.pc));
	FX_CHECK(c.apply_op(x, ap, nullptr));
	FX_CHECK(fx::k_fill(c.be, p0, 0.0, c.n));
	double* pp[2] = {p0, p1};
	if (!c.pc) {
		FX_CHECK(c.ered<1>(FResid{b, ap, r}, FinCGInit{c.d}, "cg init", nullptr));
		FX_CHECK(drive_counted(s, [&](int64_t i) -> fx_status {
			double* pold = pp[i & 1];
			double* pnew = pp[(i + 1) & 1];
			FX_CHECK(c.cg_step1(r, pold, pnew, ap));
			return c.ered<1>(FXR{c.d, x, r, pnew, ap, 0.0}, FinCGBeta{c.d}, "cg x/r update");
		}, (int64_t)s->max_iter, 2));
	} else {
		double* z = s->vec(4);
		FX_REQUIRE(z, FX_ERR_ALLOC, "work vector allocation failed");
		FX_CHECK(c.ered<1>(FResid{b, ap, r}, FinPcgInitRR{c.d}, "pcg init", nullptr));
		FX_CHECK(c.precond(z, r));
		FX_CHECK(c.ered<1>(FDot{r, z}, FinPcgInitRZ{c.d}, "pcg r.z"));
		FX_CHECK(drive_counted(s, [&](int64_t i) -> fx_status {
			double* pold = pp[i & 1];
			double* pnew = pp[(i + 1) & 1];
			FX_CHECK(c.cg_step1(z, pold, pnew, ap));
			FX_CHECK(c.ered<1>(FXR{c.d, x, r, pnew, ap, 0.0}, FinPcgRR{c.d}, "pcg x/r update"));
			FX_CHECK(c.precond(z, r));
			return c.ered<1>(FDot{r, z}, FinPcgBeta{c.d}, "pcg r.z");
		}, (int64_t)s->max_iter, 1));
	}
	FX_CHECK(fetch_stats(s, stats));
	if (stats) FX_CHECK(true_residual(c, x, b, ap, &stats->true_residual));
	return FX_OK;
}

// ── FCG: flexible PCG with FP64 true-residual refinement (extension) ──
fx_status solve_fcg(Ctx& c, double* x, const double* b, fx_solve_stats* stats) {
	// With a fused FP32 multigrid and the 5-point stencil operator an iteration is
	// the fused p-update/A·p/p·Ap step, the x/r update (which also writes the FP32
	// residual) and the fused V-cycle, whose level-0 kernel produces z and the
	// (r·z, z·Ap) reduction.
	fx_solver* s = c.s;
	double *r = s->vec(0), *p0 = s->vec(1), *p1 = s->vec(2), *ap = s->vec(3), *z = s->vec(4);
	FX_REQUIRE(r && p0 && p1 && ap && z, FX_ERR_ALLOC, "work vector allocation failed");
	fx_multigrid* fm = nullptr;
	const bool vk = c.op->kind == FX_OP_VARK;
	if (c.pc && c.pc->kind == FX_PC_MULTIGRID && (c.fused5 || vk)) {
		fx_multigrid* m = (fx_multigrid*)c.pc->handle;
		if (m->opt.fused) fm = m;
	}
	uint64_t used = 0;
	fx_solve_stats st{};
	double* pp[2] = {p0, p1};
	double prev_tr = INFINITY;
	bool at_floor = false;
	for (int pass = 0; pass < 8; ++pass) {
		FX_CHECK(reset_state(s, (int64_t)(s->max_iter - used), c.pc));
		// The fused even-column step takes p_old = 0 in its first iteration itself.
		if (!fm || c.cols % 2 != 0) FX_CHECK(fx::k_fill(c.be, p0, 0.0, c.n));
		if (fm) {
			float* r32 = (float*)fm->lv[0].rhs;
			const float* z32 = (const float*)fm->lv[0].tmp;
			if (vk) {
				FX_CHECK(c.resid_vk(x, EpiRes32{b, r, r32}, FinPcgInitRR{c.d}, "fcg init"));
			} else {
				FX_CHECK(c.apply_op(x, ap, nullptr));
				FX_CHECK(c.ered<1>(FResid32{b, ap, r, r32}, FinPcgInitRR{c.d}, "fcg init", nullptr));
			}
			FX_CHECK(fmg::fast_cycle(fm, c.skip, EpiRZ{r, c.d}));
			const bool even = c.cols % 2 == 0;
			// With the residual update carried by the V-cycle's level-0 restriction, r alternates
			// between two vectors (the restriction reads one while it writes the other).
			const bool rupd = even && fmg::upd_ok(fm);
			double* r2 = rupd ? s->vec(5) : nullptr;
			FX_REQUIRE(!rupd || r2, FX_ERR_ALLOC, "work vector allocation failed");
			FX_CHECK(drive_counted(s, [&](int64_t i) -> fx_status {
				double* pold = pp[i & 1];
				double* pnew = pp[(i + 1) & 1];
				if (even) {
					// Two columns per thread: half as many threads per row.  Taller row blocks (fewer
					// halo rows) while the grid keeps at least 1024 workgroups.
					const int64_t cb = (c.cols / 2 + BLOCK - 1) / BLOCK;
					int rpb = 64;
					while (rpb > 4 && cb * ((c.rows + rpb - 1) / rpb) < 1024) rpb /= 2;
					dim3 grid = fx::stream_grid(c.rows, c.cols / 2, rpb);
					FX_CHECK(fx::ensure_partials(c.be, (size_t)grid.x * grid.y));
					const double* kp = vk ? (const double*)c.op->k->ptr : nullptr;
					const double ix2 = 1.0 / (s->g.dx * s->g.dx), iy2 = 1.0 / (s->g.dy * s->g.dy);
					fx::probe_begin(c.be, FX_PROBE_FCG_STEP);
					auto step = [&](auto kern) {
						hipLaunchKernelGGL(kern, grid, dim3(BLOCK), 0, c.st, z32, pold, pnew, ap, c.w9, c.op->coeff, kp,
						                   vk ? c.op->sigma : 0.0, ix2, iy2, c.rows, c.cols, rpb, fx::red_scratch(c.be),
						                   (const SState*)c.d, x);
					};
					if (vk) i == 0 ? step(k_fcg_step2<true, true>) : step(k_fcg_step2<true, false>);
					else i == 0 ? step(k_fcg_step2<false, true>) : step(k_fcg_step2<false, false>);
					fx::probe_end(c.be, FX_PROBE_FCG_STEP);
					hipLaunchKernelGGL((k_finish<1, FinAlphaDefer>), dim3(1), dim3(FIN_BLOCK), 0, c.st,
					                   fx::red_scratch(c.be), grid.x * grid.y, FinAlphaDefer{c.d}, c.done);
					FX_CHECK(check_launch("fcg fused step"));
					if (rupd) {
						double* rold = (i & 1) ? r2 : r;
						double* rnew = (i & 1) ? r : r2;
						return fmg::fast_cycle(fm, c.skip, EpiFcg{rnew, ap, c.d}, RUpd{rold, ap, rnew, c.d, pnew, 0.0});
					}
					const size_t n2 = c.n / 2;
					// About four pairs per thread, unrolled, so each thread keeps several loads in flight.
					const unsigned nb = (unsigned)std::max<size_t>(1, std::min<size_t>(16384, (n2 + 4 * BLOCK - 1) / (4 * BLOCK)));
					FX_CHECK(fx::ensure_partials(c.be, nb));
					fx::probe_begin(c.be, FX_PROBE_FCG_UPDATE);
					hipLaunchKernelGGL(k_fr32v, dim3(nb), dim3(BLOCK), 0, c.st, (const SState*)c.d, r, r32,
					                   (const double*)ap, n2, fx::red_scratch(c.be), c.skip);
					fx::probe_end(c.be, FX_PROBE_FCG_UPDATE);
					hipLaunchKernelGGL((k_finish<1, FinPcgRRDefer>), dim3(1), dim3(FIN_BLOCK), 0, c.st,
					                   fx::red_scratch(c.be), nb, FinPcgRRDefer{c.d, pnew}, c.skip);
					FX_CHECK(check_launch("fcg x/r update"));
					return fmg::fast_cycle(fm, c.skip, EpiFcg{r, ap, c.d});
				}
				int rpb = stream_rpb(c.rows, c.cols);
				dim3 grid = fx::stream_grid(c.rows, c.cols, rpb);
				FX_CHECK(fx::ensure_partials(c.be, (size_t)grid.x * grid.y));
				if (vk)
					hipLaunchKernelGGL((k_stream_vk<1, VPNewF, EpiCGv>), grid, dim3(BLOCK), 0, c.st,
					                   VPNewF{z32, pold, 0.0}, EpiCGv{pnew, ap}, (const double*)c.op->k->ptr, c.op->sigma,
					                   c.op->coeff, 1.0 / (s->g.dx * s->g.dx), 1.0 / (s->g.dy * s->g.dy), c.rows, c.cols,
					                   rpb, fx::red_scratch(c.be), (const SState*)c.d);
				else
					hipLaunchKernelGGL((k_stream_st<1, VPNewF, EpiCG>), grid, dim3(BLOCK), 0, c.st,
					                   VPNewF{z32, pold, 0.0}, EpiCG{c.op->coeff, pnew, ap}, c.w9, c.rows, c.cols, rpb,
					                   fx::red_scratch(c.be), (const SState*)c.d);
				hipLaunchKernelGGL((k_finish<1, FinAlpha>), dim3(1), dim3(FIN_BLOCK), 0, c.st, fx::red_scratch(c.be),
				                   grid.x * grid.y, FinAlpha{c.d}, c.done);
				FX_CHECK(check_launch("fcg fused step"));
				FX_CHECK(c.ered<1>(FXR32{c.d, x, r, r32, pnew, ap, 0.0}, FinPcgRR{c.d}, "fcg x/r update"));
				return fmg::fast_cycle(fm, c.skip, EpiFcg{r, ap, c.d});
			}, (int64_t)(s->max_iter - used)));
			if (c.n % 2 == 0) {   // only the even-column loop defers its x update
				const size_t n2 = c.n / 2;
				const unsigned nb = (unsigned)std::max<size_t>(1, std::min<size_t>(16384, (n2 + 4 * BLOCK - 1) / (4 * BLOCK)));
				hipLaunchKernelGGL(k_xflush, dim3(nb), dim3(BLOCK), 0, c.st, (const SState*)c.d, x, n2);
			}
			FX_CHECK(check_launch("fcg deferred x update"));
		} else {
			FX_CHECK(c.apply_op(x, ap, nullptr));
			FX_CHECK(c.ered<1>(FResid{b, ap, r}, FinPcgInitRR{c.d}, "fcg init", nullptr));
			FX_CHECK(c.precond(z, r));
			FX_CHECK(c.ered<1>(FDot{r, z}, FinPcgInitRZ{c.d}, "fcg r.z"));
			FX_CHECK(drive(s, [&](int64_t i) -> fx_status {
				double* pold = pp[i & 1];
				double* pnew = pp[(i + 1) & 1];
				FX_CHECK(c.cg_step1(z, pold, pnew, ap));
				FX_CHECK(c.ered<1>(FXR{c.d, x, r, pnew, ap, 0.0}, FinPcgRR{c.d}, "fcg x/r update"));
				if (c.pc && c.pc->kind == FX_PC_LEARNED) {
					// While the network is in use, judge it by the FP64 true residual of this iterate.
					double* tr = ((fx_learned*)c.pc->handle)->tr;
					const int* lsk = &c.d->lskip;
					FX_CHECK(c.apply_op(x, tr, lsk));
					FX_CHECK(c.ered<1>(FResid{b, tr, tr}, FinLearnJudge{c.d}, "learned true residual", lsk));
				}
				FX_CHECK(c.precond(z, r));
				return c.ered<2>(FDot2{r, z, z, ap}, FinFcgBeta{c.d}, "fcg r.z z.ap");
			}, (int64_t)(s->max_iter - used)));
		}
		double tr = 0;
		if (fm && vk) {
			// The true residual is reduced into the device state before the stats are copied: one
			// host synchronisation for both.
			FX_CHECK(c.resid_vk(x, EpiResNorm{b}, FinTrueRes{c.d}, "true residual"));
			FX_CHECK(fetch_stats(s, &st));
			tr = s->hst->tres;
		} else {
			FX_CHECK(fetch_stats(s, &st));
			FX_CHECK(true_residual(c, x, b, ap, &tr));
		}
		if (c.pc && c.pc->kind == FX_PC_LEARNED) ((fx_learned*)c.pc->handle)->fell_back |= s->hst->lplain;
		used += st.iterations;
		st.true_residual = tr;
		if (st.reason != FX_CONVERGED || tr < s->tol || used >= s->max_iter) break;
		// A restart that did not halve the true residual has reached the FP64
		// attainable accuracy (about eps·‖A‖·‖x‖); further passes only repeat it.
		if (!(tr < 0.5 * prev_tr)) { at_floor = true; break; }
		prev_tr = tr;
	}
	st.iterations = used;
	// Converged means the recurrence met the tolerance (CgSolver's criterion, cg.rs:118-124) and the
	// true residual is below it or at the FP64 floor; the true residual is reported either way.
	if (st.reason == FX_CONVERGED && !(st.true_residual < s->tol) && !at_floor) st.reason = FX_MAX_ITERATIONS;
	if (stats) *stats = st;
	return FX_OK;
}

// ── PipelinedCgSolver::solve (pipelined_cg.rs:96-167) ─────────────────
fx_status solve_pipelined(Ctx& c, double* x, const double* b, fx_solve_stats* stats) {
	fx_solver* s = c.s;
	double *r = s->vec(0), *p = s->vec(1), *w = s->vec(2), *t = s->vec(3), *z = s->vec(4), *q = s->vec(5);
	FX_REQUIRE(r && p && w && t && z && q, FX_ERR_ALLOC, "work vector allocation failed");
	FX_CHECK(reset_state(s, (int64_t)s->max_iter));
	FX_CHECK(c.apply_op(x, q, nullptr));
	FX_CHECK(c.ered<1>(FResid{b, q, r}, FinPcgInitRR{c.d}, "pipecg init", nullptr));
	FX_CHECK(c.apply_op(r, w, nullptr));
	FX_CHECK(fx::k_fill(c.be, p, 0.0, c.n));
	FX_CHECK(fx::k_fill(c.be, t, 0.0, c.n));
	FX_CHECK(fx::k_fill(c.be, z, 0.0, c.n));
	FX_CHECK(c.ered<2>(FDot2{r, r, r, w}, FinPipe{c.d, true}, "pipecg dot2"));
	FX_CHECK(drive(s, [&](int64_t) -> fx_status {
		// q = A·w (pipelined_cg.rs:143-147), fused update + next (γ, δ).
		FX_CHECK(c.apply_op(w, q));
		return c.ered<2>(FPipeUpd{c.d, z, t, p, x, r, w, q, 0.0, 0.0}, FinPipe{c.d, false}, "pipecg update");
	}, (int64_t)s->max_iter));
	FX_CHECK(fetch_stats(s, stats));
	if (stats) FX_CHECK(true_residual(c, x, b, q, &stats->true_residual));
	return FX_OK;
}

// ── preconditioned pipelined CG (extension, Ghysels & Vanroose Alg. 3) ─
fx_status solve_ppcg(Ctx& c, double* x, const double* b, fx_solve_stats* stats) {
	fx_solver* s = c.s;
	double *r = s->vec(0), *u = s->vec(1), *w = s->vec(2), *m = s->vec(3), *nn = s->vec(4), *z = s->vec(5),
	       *q = s->vec(6), *sv = s->vec(7), *p = s->vec(8);
	FX_REQUIRE(r && u && w && m && nn && z && q && sv && p, FX_ERR_ALLOC, "work vector allocation failed");
	FX_CHECK(reset_state(s, (int64_t)s->max_iter));
	FX_CHECK(c.apply_op(x, sv, nullptr));
	FX_CHECK(c.ered<1>(FResid{b, sv, r}, FinPcgInitRR{c.d}, "ppcg init", nullptr));
	FX_CHECK(c.precond(u, r));
	FX_CHECK(c.apply_op(u, w, nullptr));
	for (double* v : {z, q, sv, p}) FX_CHECK(fx::k_fill(c.be, v, 0.0, c.n));
	FX_CHECK(c.ered<3>(FDot3{r, u, w}, FinPP{c.d, true}, "ppcg dot3"));
	FX_CHECK(drive(s, [&](int64_t) -> fx_status {
		FX_CHECK(c.precond(m, w));
		FX_CHECK(c.apply_op(m, nn));
		return c.ered<3>(FPPUpd{c.d, z, q, sv, p, x, r, u, w, nn, m, 0.0, 0.0}, FinPP{c.d, false}, "ppcg update");
	}, (int64_t)s->max_iter));
	FX_CHECK(fetch_stats(s, stats));
	if (stats) FX_CHECK(true_residual(c, x, b, m, &stats->true_residual));
	return FX_OK;
}

// ── pipelined flexible CG with FP64 true-residual refinement (extension) ─
// For an FP32 multigrid, whose rounding breaks the Ghysels–Vanroose recurrences (u = M·r carried by
// u −= α·M·s) after a few digits.  Passes restart from x while the FP64 true residual is above the
// tolerance and keeps halving, as solve_fcg does.
fx_status solve_pfcg(Ctx& c, double* x, const double* b, fx_solve_stats* stats) {
	fx_solver* s = c.s;
	double *r = s->vec(0), *u = s->vec(1), *w = s->vec(2), *p = s->vec(3), *sv = s->vec(4);
	FX_REQUIRE(r && u && w && p && sv, FX_ERR_ALLOC, "work vector allocation failed");
	const bool vk = c.op->kind == FX_OP_VARK;
	uint64_t used = 0;
	fx_solve_stats st{};
	double prev_tr = INFINITY;
	bool at_floor = false;
	for (int pass = 0; pass < 8; ++pass) {
		FX_CHECK(reset_state(s, (int64_t)(s->max_iter - used)));
		FX_CHECK(c.apply_op(x, w, nullptr));
		FX_CHECK(c.ered<1>(FResid{b, w, r}, FinPcgInitRR{c.d}, "pfcg init", nullptr));
		FX_CHECK(fx::k_fill(c.be, p, 0.0, c.n));
		FX_CHECK(fx::k_fill(c.be, sv, 0.0, c.n));
		FX_CHECK(drive(s, [&](int64_t i) -> fx_status {
			FX_CHECK(c.precond(u, r));
			if (c.fused5 || vk) {
				const int rpb = stream_rpb(c.rows, c.cols);
				dim3 grid = fx::stream_grid(c.rows, c.cols, rpb);
				FX_CHECK(fx::ensure_partials(c.be, (size_t)grid.x * grid.y * 4));
				if (vk)
					hipLaunchKernelGGL((k_stream_vk<4, VX, EpiPF>), grid, dim3(BLOCK), 0, c.st, VX{u, 0.0},
					                   EpiPF{r, sv, w}, (const double*)c.op->k->ptr, c.op->sigma, c.op->coeff,
					                   1.0 / (s->g.dx * s->g.dx), 1.0 / (s->g.dy * s->g.dy), c.rows, c.cols, rpb,
					                   fx::red_scratch(c.be), (const SState*)c.d);
				else
					hipLaunchKernelGGL((k_stream_st<4, VX, EpiPFs>), grid, dim3(BLOCK), 0, c.st, VX{u, 0.0},
					                   EpiPFs{c.op->coeff, EpiPF{r, sv, w}}, c.w9, c.rows, c.cols, rpb,
					                   fx::red_scratch(c.be), (const SState*)c.d);
				hipLaunchKernelGGL((k_finish<4, FinPF>), dim3(1), dim3(FIN_BLOCK), 0, c.st, fx::red_scratch(c.be),
				                   grid.x * grid.y, FinPF{c.d, i == 0}, c.done);
				FX_CHECK(check_launch("pfcg operator"));
			} else {
				FX_CHECK(c.apply_op(u, w));
				FX_CHECK(c.ered<4>(FDotPF{r, u, w, sv}, FinPF{c.d, i == 0}, "pfcg dots"));
			}
			return c.eupd(FPFUpd{c.d, p, sv, x, r, u, w, 0.0, 0.0}, "pfcg update");
		}, (int64_t)(s->max_iter - used)));
		double tr = 0;
		FX_CHECK(fetch_stats(s, &st));
		FX_CHECK(true_residual(c, x, b, w, &tr));
		used += st.iterations;
		st.true_residual = tr;
		if (st.reason != FX_CONVERGED || tr < s->tol || used >= s->max_iter) break;
		if (!(tr < 0.5 * prev_tr)) { at_floor = true; break; }
		prev_tr = tr;
	}
	st.iterations = used;
	if (st.reason == FX_CONVERGED && !(st.true_residual < s->tol) && !at_floor) st.reason = FX_MAX_ITERATIONS;
	if (stats) *stats = st;
	return FX_OK;
}

// ── BiCgStabSolver (bicgstab.rs:90-346) ───────────────────────────────
fx_status solve_bicgstab(Ctx& c, double* x, const double* b, fx_solve_stats* stats) {
	fx_solver* s = c.s;
	double *r = s->vec(0), *rh = s->vec(1), *p = s->vec(2), *v = s->vec(3), *sv = s->vec(4), *t = s->vec(5),
	       *ph = s->vec(6), *sh = s->vec(7);
	FX_REQUIRE(r && rh && p && v && sv && t && ph && sh, FX_ERR_ALLOC, "work vector allocation failed");
	const bool pre = c.pc != nullptr;
	const double* pa = pre ? ph : p;
	const double* sa = pre ? sh : sv;
	FX_CHECK(reset_state(s, (int64_t)s->max_iter));
	// r = b − A·x; r̂ = r; v = p = 0 (bicgstab.rs:98-113).
	FX_CHECK(c.apply_op(x, v, nullptr));
	FX_CHECK(c.ered<1>(FResid{b, v, r}, FinBiInit{c.d}, "bicgstab init", nullptr));
	FX_CHECK(c.eupd(FCopy{r, rh}, "copy", nullptr));
	FX_CHECK(fx::k_fill(c.be, v, 0.0, c.n));
	FX_CHECK(fx::k_fill(c.be, p, 0.0, c.n));
	FX_CHECK(c.ered<2>(FDot2{r, r, rh, r}, FinBiTop{c.d, true}, "bicgstab top", nullptr));
	FX_CHECK(drive(s, [&](int64_t) -> fx_status {
		FX_CHECK(c.eupd(FBiP{c.d, p, r, v, 0.0, 0.0}, "bicgstab p"));
		if (pre) FX_CHECK(c.precond(ph, p));
		FX_CHECK(c.apply_op(pa, v));
		FX_CHECK(c.ered<1>(FDot{rh, v}, FinBiAlpha{c.d}, "bicgstab rh.v"));
		FX_CHECK(c.ered<1>(FBiS{c.d, sv, r, v, 0.0}, FinBiS{c.d}, "bicgstab s"));
		if (pre) FX_CHECK(c.precond(sh, sv));
		FX_CHECK(c.apply_op(sa, t));
		FX_CHECK(c.ered<2>(FDot2{t, sv, t, t}, FinBiOm{c.d}, "bicgstab omega"));
		return c.ered<2>(FBiX{c.d, x, r, pa, sa, sv, t, rh, 0.0, 0.0, 0}, FinBiTop{c.d, false}, "bicgstab x/r", c.done);
	}, (int64_t)s->max_iter));
	FX_CHECK(fetch_stats(s, stats));
	if (stats) FX_CHECK(true_residual(c, x, b, t, &stats->true_residual));
	return FX_OK;
}

// ── GmresSolver (gmres.rs:116-490) ────────────────────────────────────
fx_status solve_gmres(Ctx& c, double* x, const double* b, fx_solve_stats* stats) {
	fx_solver* s = c.s;
	const int m = (int)s->restart;
	// Flexible GMRES preconditions on the right (z_k = M⁻¹v_k); GMRES on the left (gmres.rs:294-454).
	const bool flex = s->kind == FX_SOLVER_FGMRES && c.pc != nullptr;
	const bool pre = c.pc != nullptr && !flex;
	double* w = s->vec(0);
	FX_REQUIRE(w && (int)s->basis.size() == m + 1, FX_ERR_ALLOC, "work vector allocation failed");
	FX_REQUIRE(!flex || (int)s->zbasis.size() == m, FX_ERR_ALLOC, "work vector allocation failed");
	auto& vv = s->basis;
	auto& zz = s->zbasis;
	FX_CHECK(reset_state(s, (int64_t)s->max_iter));
	VecList vl{};
	for (int i = 0; i < m + (flex ? 0 : 1); ++i) vl.v[i] = flex ? zz[i] : vv[i];
	const SState* dd = c.d;
	FX_CHECK(drive(s, [&](int64_t) -> fx_status {
		// Cycle start (gmres.rs:124-157 / 305-337).
		if (!pre) {
			FX_CHECK(c.apply_op(x, vv[0], c.done));
			FX_CHECK(c.ered<1>(FGmR0{b, vv[0]}, FinGmStart{c.d}, "gmres r0", c.done));
		} else {
			FX_CHECK(c.apply_op(x, w, c.done));
			FX_CHECK(c.ered<1>(FGmR0{b, w}, FinNone{}, "gmres r0", c.done));
			FX_CHECK(c.precond(vv[0], w, c.done));
			FX_CHECK(c.ered<1>(FSumSq{vv[0]}, FinGmStart{c.d}, "gmres beta", c.done));
		}
		FX_CHECK(c.eupd(FGmScale{&dd->gbeta, vv[0], vv[0], 0.0}, "gmres v0 scale"));
		for (int k = 0; k < m; ++k) {
			if (flex) {
				FX_CHECK(c.precond(zz[k], vv[k]));
				FX_CHECK(c.apply_op(zz[k], w));
			} else {
				FX_CHECK(c.apply_op(vv[k], w));
			}
			if (pre) {
				FX_CHECK(c.precond(vv[k + 1], w));
				FX_CHECK(c.eupd(FCopy{vv[k + 1], w}, "gmres copy"));
			}
			// H(i, k) lives at dgm[k·(m+1) + i] (column-major, gmres.rs:84-99).
			const double* Hk = s->dgm + (size_t)k * (m + 1);
			for (int j = 0; j <= k; ++j)
				FX_CHECK(c.ered<1>(FGmMgs{w, j ? vv[j - 1] : nullptr, vv[j], j ? Hk + (j - 1) : nullptr, 0.0, j > 0},
				                   FinGmH{c.d, j, k}, "gmres mgs"));
			FX_CHECK(c.ered<1>(FGmMgs{w, vv[k], nullptr, Hk + k, 0.0, 1}, FinGmStep{c.d, k}, "gmres h"));
			FX_CHECK(c.eupd(FGmNorm{c.d, w, vv[k + 1], 0.0, 0}, "gmres normalize"));
		}
		FX_CHECK(c.scalar(FinGmBack{c.d}, "gmres back", c.done));
		FX_CHECK(c.eupd(FGmX{c.d, x, vl, 0, nullptr}, "gmres x", c.done));
		return c.scalar(FinGmEnd{c.d}, "gmres end", c.done);
	}, (int64_t)s->max_iter + 1, 1));
	FX_CHECK(fetch_stats(s, stats));
	if (stats) FX_CHECK(true_residual(c, x, b, w, &stats->true_residual));
	return FX_OK;
}

// ══════════════════════════════════════════════════════════════════════
// Batched multi-problem solver (extension; no reference counterpart).
//
// Many independent −∇·(k∇u) = f problems on grids of uneven sizes, each solved
// by FP64 flexible CG around its own fused FP32 coefficient-aware V-cycle, all
// advancing together.  Every kernel of an iteration is one launch over a
// prebuilt work list of (problem, level, tile) items, so an iteration is a
// fixed launch sequence over all problems; it is captured once in a hipGraph
// and replayed until every problem has converged.  Each problem has its own
// device scalar state (SState) and done flag; blocks of finished problems
// return at once, and each problem's reductions are finished by its own
// workgroup of a per-problem finish kernel.  Launch shapes that leave every
// per-point result unchanged (rows per block of the step kernel, points per
// block of the update kernel) are chosen by timing candidates at creation.
// ══════════════════════════════════════════════════════════════════════

struct BProb {
	int64_t off;          // first point of the problem in the concatenated FP64 vectors
	int32_t rows, cols;
	double ix2, iy2;      // 1/dx², 1/dy² of the operator
	float* r32;           // level-0 rhs of the problem's hierarchy (FP32 residual)
	const float* z32;     // level-0 result of the problem's V-cycle
	int32_t stp0, nstp;   // its block range in the step list
	int32_t upd0, nupd;   // … in the update list
	int32_t pst0, npst;   // … in the level-0 post list
	int32_t lev0, nlev;   // its levels in the level table
	int32_t sweeps, pad;  // coarsest-level sweeps
};
struct BLev {
	float* rhs;
	float* tmp;
	fmg::LWf L;
	fx_xtab X;
	int32_t rows, cols, cr, cc;   // this level and the next coarser one
	// Wave-streaming launch shapes (wave work lists): post strips × segments of psh rows, restriction
	// strips × segments of rscr coarse rows.
	int32_t pns, pnseg, psh, rns, rnseg, rscr;
};
struct BItem {
	int32_t p, l, bx, by;         // problem, level-table index, tile
};

constexpr int BUPD_THREAD = 8;    // points per thread of one update chunk before tuning

// p = double(z) + β·p_old, ap = A·p, p·ap (the traversal of k_stream_vk) for one (problem, tile).
__global__ __launch_bounds__(BLOCK) void k_b_step(const BItem* __restrict__ items, const BProb* __restrict__ pr,
                                                  const SState* __restrict__ st, const double* __restrict__ k,
                                                  const double* __restrict__ pold, double* __restrict__ pnew,
                                                  double* __restrict__ ap, double sigma, double alpha, int rpb,
                                                  RedScratch rs) {
	const BItem it = items[blockIdx.x];
	if (st[it.p].skip) return;
	const BProb P = pr[it.p];
	VPNewF v{P.z32, pold + P.off, st[it.p].beta};
	EpiCGv e{pnew + P.off, ap + P.off};
	Acc acc[1] = {Acc{0.0, 0.0}};
	stream_vk_body<1>(it.bx, it.by, v, e, k + P.off, sigma, alpha, P.ix2, P.iy2, P.rows, P.cols, rpb, acc);
	block_partial<1>(acc, rs);
}

// x += α·p; r −= α·ap; r32 = float(r); r·r (FXR32) over one chunk of a problem.
__global__ __launch_bounds__(BLOCK) void k_b_update(const BItem* __restrict__ items, const BProb* __restrict__ pr,
                                                    const SState* __restrict__ st, double* __restrict__ x,
                                                    double* __restrict__ r, const double* __restrict__ p,
                                                    const double* __restrict__ ap, int chunk, RedScratch rs) {
	const BItem it = items[blockIdx.x];
	if (st[it.p].skip) return;
	const BProb P = pr[it.p];
	const double al = st[it.p].alpha;
	const int64_t n = (int64_t)P.rows * P.cols;
	const int64_t i0 = (int64_t)it.bx * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
	Acc acc[1] = {Acc{0.0, 0.0}};
	for (int64_t i = i0 + threadIdx.x; i < i1; i += BLOCK) {
		const int64_t g = P.off + i;
		x[g] = fma(al, p[g], x[g]);
		const double v = fma(-al, ap[g], r[g]);
		r[g] = v;
		P.r32[i] = float(v);
		acc_add(acc[0], v * v);
	}
	block_partial<1>(acc, rs);
}

// r = b − A·x, r32 = float(r), r·r (mode 0); or only ‖b − A·x‖² (mode 1), over one chunk.
__global__ __launch_bounds__(BLOCK) void k_b_resid(const BItem* __restrict__ items, const BProb* __restrict__ pr,
                                                   const double* __restrict__ k, const double* __restrict__ x,
                                                   const double* __restrict__ b, double* __restrict__ r, double sigma,
                                                   double alpha, int chunk, int mode, RedScratch rs) {
	const BItem it = items[blockIdx.x];
	const BProb P = pr[it.p];
	const int64_t n = (int64_t)P.rows * P.cols;
	const int64_t i0 = (int64_t)it.bx * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
	const double* kk = k + P.off;
	const double* u = x + P.off;
	Acc acc[1] = {Acc{0.0, 0.0}};
	for (int64_t i = i0 + threadIdx.x; i < i1; i += BLOCK) {
		const int rr = (int)(i / P.cols), c = (int)(i % P.cols);
		double av;
		if (rr == 0 || c == 0 || rr == P.rows - 1 || c == P.cols - 1) {
			av = u[i];
		} else {
			const double kc = kk[i], cv = u[i];
			const double kw = hmean_d(kc, kk[i - 1]), ke = hmean_d(kc, kk[i + 1]);
			const double kn = hmean_d(kc, kk[i - P.cols]), ks = hmean_d(kc, kk[i + P.cols]);
			const double fx = ke * (u[i + 1] - cv) - kw * (cv - u[i - 1]);
			const double fy = ks * (u[i + P.cols] - cv) - kn * (cv - u[i - P.cols]);
			av = fma(alpha, -(fx * P.ix2 + fy * P.iy2), sigma * cv);
		}
		const double v = b[P.off + i] - av;
		if (mode == 0) {
			r[P.off + i] = v;
			P.r32[i] = float(v);
		}
		acc_add(acc[0], v * v);
	}
	block_partial<1>(acc, rs);
}

__global__ __launch_bounds__(BLOCK) void k_b_zero(const BItem* __restrict__ items, const BProb* __restrict__ pr,
                                                  double* __restrict__ v, int chunk) {
	const BItem it = items[blockIdx.x];
	const BProb P = pr[it.p];
	const int64_t n = (int64_t)P.rows * P.cols;
	const int64_t i0 = (int64_t)it.bx * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
	for (int64_t i = i0 + threadIdx.x; i < i1; i += BLOCK) v[P.off + i] = 0.0;
}

// V-cycle pieces over (problem, level, tile) items.
template <bool VK, int NPRE>
__global__ __launch_bounds__(BLOCK) void k_b_restrict(const BItem* __restrict__ items, const BLev* __restrict__ lv,
                                                      const SState* __restrict__ st, const float* __restrict__ dummy) {
	const BItem it = items[blockIdx.x];
	if (st[it.p].skip) return;
	const BLev& L = lv[it.l];
	fmg::frestrict_tile<VK, NPRE>(it.bx, it.by, L.rhs, lv[it.l + 1].rhs, L.L, L.X, L.rows, L.cols, L.cr, L.cc);
}
__global__ __launch_bounds__(1024) void k_b_coarse(const BProb* __restrict__ pr, const BLev* __restrict__ lv,
                                                   const SState* __restrict__ st) {
	const int p = blockIdx.x;
	if (st[p].skip) return;
	const BLev& L = lv[pr[p].lev0 + pr[p].nlev - 1];
	fmg::fcoarse_body(L.rhs, L.tmp, L.L, pr[p].sweeps, L.rows, L.cols);
}
template <bool VK, int NPRE, int NPOST>
__global__ __launch_bounds__(BLOCK) void k_b_post(const BItem* __restrict__ items, const BLev* __restrict__ lv,
                                                  const SState* __restrict__ st) {
	const BItem it = items[blockIdx.x];
	if (st[it.p].skip) return;
	const BLev& L = lv[it.l];
	Acc acc[1];
	fmg::fpost_tile<VK, NPRE, NPOST, fmg::NoEpi>(it.bx, it.by, L.rhs, lv[it.l + 1].tmp, L.tmp, L.L, L.X, L.rows, L.cols,
	                                             L.cr, L.cc, fmg::NoEpi{}, acc);
}
// Level-0 post with the FCG reductions of each problem: (r·z, z·Ap) at iterations, r·z at the start.
template <bool VK, int NPRE, int NPOST, int NS>
__global__ __launch_bounds__(BLOCK) void k_b_post0(const BItem* __restrict__ items, const BProb* __restrict__ pr,
                                                   const BLev* __restrict__ lv, SState* __restrict__ st,
                                                   const double* __restrict__ r, const double* __restrict__ ap,
                                                   RedScratch rs) {
	const BItem it = items[blockIdx.x];
	if (st[it.p].skip) return;
	const BLev& L = lv[it.l];
	const BProb& P = pr[it.p];
	Acc acc[NS];
#pragma unroll
	for (int q = 0; q < NS; ++q) acc[q] = Acc{0.0, 0.0};
	if constexpr (NS == 2)
		fmg::fpost_tile<VK, NPRE, NPOST, EpiFcg>(it.bx, it.by, L.rhs, lv[it.l + 1].tmp, L.tmp, L.L, L.X, L.rows, L.cols,
		                                         L.cr, L.cc, EpiFcg{r + P.off, ap + P.off, st + it.p}, acc);
	else
		fmg::fpost_tile<VK, NPRE, NPOST, EpiRZ>(it.bx, it.by, L.rhs, lv[it.l + 1].tmp, L.tmp, L.L, L.X, L.rows, L.cols,
		                                        L.cr, L.cc, EpiRZ{r + P.off, st + it.p}, acc);
	block_partial<NS>(acc, rs);
}

// The same V-cycle pieces on the wave-streaming level kernels: an item is a block of NWAVES consecutive
// waves (strip-major within the segment order of fmg::k_wpost / k_wrestrict) of one problem's level;
// waves past the level's last return at once (those of a level-0 post block still join its partial sums).
template <bool VK, int NPRE>
__global__ __launch_bounds__(BLOCK) void k_b_wrestrict(const BItem* __restrict__ items, const BLev* __restrict__ lv,
                                                       const SState* __restrict__ st) {
	const BItem it = items[blockIdx.x];
	if (st[it.p].skip) return;
	const BLev& L = lv[it.l];
	const int gw = it.bx + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
	if (gw >= L.rns * L.rnseg) return;
	double lsum = 0.0;
	fmg::wrestrict_wave<VK, NPRE>(L.rhs, lv[it.l + 1].rhs, L.L, L.X, L.rows, L.cols, L.cr, L.cc, L.rns, L.rnseg,
	                              gw % L.rns, gw / L.rns, L.rscr, fmg::NoUpd{}, lsum);
}
template <bool VK, int NPRE, int NPOST>
__global__ __launch_bounds__(BLOCK) void k_b_wpost(const BItem* __restrict__ items, const BLev* __restrict__ lv,
                                                   const SState* __restrict__ st) {
	const BItem it = items[blockIdx.x];
	if (st[it.p].skip) return;
	const BLev& L = lv[it.l];
	const int gw = it.bx + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
	if (gw >= L.pns * L.pnseg) return;
	double lsum[1] = {0.0};
	fmg::NoEpi e;
	fmg::wpost_wave<VK, NPRE, NPOST>(L.rhs, lv[it.l + 1].tmp, L.tmp, L.L, L.X, L.rows, L.cols, L.cr, L.cc, gw % L.pns,
	                                 gw / L.pns, L.psh, e, lsum);
}
template <bool VK, int NPRE, int NPOST, int NS>
__global__ __launch_bounds__(BLOCK) void k_b_wpost0(const BItem* __restrict__ items, const BProb* __restrict__ pr,
                                                    const BLev* __restrict__ lv, SState* __restrict__ st,
                                                    const double* __restrict__ r, const double* __restrict__ ap,
                                                    RedScratch rs) {
	const BItem it = items[blockIdx.x];
	if (st[it.p].skip) return;   // the whole block belongs to one problem
	const BLev& L = lv[it.l];
	const BProb& P = pr[it.p];
	const int gw = it.bx + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
	double lsum[NS];
#pragma unroll
	for (int q = 0; q < NS; ++q) lsum[q] = 0.0;
	if (gw < L.pns * L.pnseg) {
		if constexpr (NS == 2) {
			EpiFcg e{r + P.off, ap + P.off, st + it.p};
			fmg::wpost_wave<VK, NPRE, NPOST>(L.rhs, lv[it.l + 1].tmp, L.tmp, L.L, L.X, L.rows, L.cols, L.cr, L.cc,
			                                 gw % L.pns, gw / L.pns, L.psh, e, lsum);
		} else {
			EpiRZ e{r + P.off, st + it.p};
			fmg::wpost_wave<VK, NPRE, NPOST>(L.rhs, lv[it.l + 1].tmp, L.tmp, L.L, L.X, L.rows, L.cols, L.cr, L.cc,
			                                 gw % L.pns, gw / L.pns, L.psh, e, lsum);
		}
	}
	Acc acc[NS];
#pragma unroll
	for (int q = 0; q < NS; ++q) acc[q] = Acc{lsum[q], 0.0};
	block_partial<NS>(acc, rs);
}

// Per-problem finish: workgroup p merges its blocks' partials (fixed order) and runs Fin on its state.
template <int NS, class Fin>
__global__ __launch_bounds__(BLOCK) void k_b_finish(const BProb* __restrict__ pr, SState* __restrict__ st,
                                                    RedScratch rs, int which, int force) {
	const int p = blockIdx.x;
	if (!force && st[p].skip) return;
	const BProb& P = pr[p];
	const int b0 = which == 0 ? P.stp0 : which == 1 ? P.upd0 : P.pst0;
	const int nb = which == 0 ? P.nstp : which == 1 ? P.nupd : P.npst;
	Acc a[NS];
#pragma unroll
	for (int k = 0; k < NS; ++k) a[k] = Acc{0.0, 0.0};
	for (int b = threadIdx.x; b < nb; b += BLOCK) {
#pragma unroll
		for (int k = 0; k < NS; ++k) acc_merge(a[k], rs.partials[(size_t)(b0 + b) * NS + k]);
	}
	block_reduce<NS>(a);
	if (threadIdx.x == 0) Fin{st + p}(a);
}
// ‖b − A·x‖ of each problem into its `resid` slot (after the solve).
struct FinTrue {
	SState* st;
	__device__ void operator()(Acc* o) const { st->omega = sqrt(acc_value(o[0])); }
};
// Count of problems still running.
__global__ void k_b_active(const SState* __restrict__ st, int nprob, int* out) {
	__shared__ int cnt;
	if (threadIdx.x == 0) cnt = 0;
	__syncthreads();
	int c = 0;
	for (int p = threadIdx.x; p < nprob; p += blockDim.x) c += st[p].done ? 0 : 1;
	atomicAdd(&cnt, c);
	__syncthreads();
	if (threadIdx.x == 0) *out = cnt;
}

}  // namespace

struct fx_batch {
	fx_backend* be = nullptr;
	int nprob = 0;
	int64_t ntot = 0;
	fx_mg_options opt{};
	std::vector<fx_grid> grids;
	std::vector<fx_multigrid*> mgs;
	std::vector<BProb> hp;
	const fx_buf* k = nullptr;
	// device tables
	BProb* dp = nullptr;
	BLev* dl = nullptr;
	SState* dst = nullptr;
	BItem *step = nullptr, *upd = nullptr, *post0 = nullptr;
	std::vector<BItem*> restr, post;         // per level depth d (restrict d → d+1; post at d ≥ 1)
	std::vector<int> nrestr, npost;
	std::vector<size_t> restr_lds;
	// Wave-streaming level kernels (when every level's restriction fits a wave): work lists of
	// NWAVES-wave blocks per level depth, the level-0 post blocks, and the segment heights in use.
	bool wave = false;
	std::vector<BLev> hl;                   // host copy of the level table
	std::vector<BItem*> wrestr, wpost;
	std::vector<int> nwrestr, nwpost;
	BItem* wpost0 = nullptr;
	int nwpost0 = 0, wsh = 16, wscr = 8;
	int nstep = 0, nupd = 0, npost0 = 0, maxdepth = 0;
	int rpb = 16, chunk = BLOCK * BUPD_THREAD;
	// work vectors
	double *x = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr, *ap = nullptr;
	fxd::Acc* partials = nullptr;
	int* dactive = nullptr;
	int* hactive = nullptr;
	hipGraphExec_t exec = nullptr;
	hipGraph_t graph = nullptr;
	int graph_iters = 2;
};

namespace {

void batch_free(fx_batch* B) {
	if (!B) return;
	if (B->be) (void)hipStreamSynchronize(B->be->stream);
	if (B->exec) (void)hipGraphExecDestroy(B->exec);
	if (B->graph) (void)hipGraphDestroy(B->graph);
	for (auto* m : B->mgs) fx_multigrid_destroy(m);
	fx::dev_free(B->dp);
	fx::dev_free(B->dl);
	fx::dev_free(B->dst);
	fx::dev_free(B->step);
	fx::dev_free(B->upd);
	fx::dev_free(B->post0);
	for (auto* q : B->restr) fx::dev_free(q);
	for (auto* q : B->post) fx::dev_free(q);
	for (auto* q : B->wrestr) fx::dev_free(q);
	for (auto* q : B->wpost) fx::dev_free(q);
	fx::dev_free(B->wpost0);
	fx::dev_free(B->x);
	fx::dev_free(B->r);
	fx::dev_free(B->p0);
	fx::dev_free(B->p1);
	fx::dev_free(B->ap);
	fx::dev_free(B->partials);
	fx::dev_free(B->dactive);
	if (B->hactive) (void)hipHostFree(B->hactive);
	delete B;
}

template <class T>
fx_status upload_vec(fx_backend* be, const std::vector<T>& v, T** out) {
	void* p = nullptr;
	FX_CHECK(fx::dev_alloc(be, std::max<size_t>(1, v.size()) * sizeof(T), &p));
	// The copy goes on the backend's stream, behind dev_alloc's zero fill: a copy on the null stream is
	// not ordered after work on this non-blocking stream, and the fill could land on the table afterwards.
	// It is finished before returning, as v may be a temporary.
	if (!v.empty()) {
		FX_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, be->stream));
		FX_HIP(hipStreamSynchronize(be->stream));
	}
	*out = (T*)p;
	return FX_OK;
}

// Step-list items for a rows-per-block choice; block ranges go into hp.
std::vector<BItem> step_items(fx_batch* B, int rpb) {
	std::vector<BItem> v;
	for (int p = 0; p < B->nprob; ++p) {
		const fx_grid& g = B->grids[p];
		B->hp[p].stp0 = (int32_t)v.size();
		const int nbx = (int)((g.cols + BLOCK - 1) / BLOCK), nby = (int)((g.rows + rpb - 1) / rpb);
		for (int by = 0; by < nby; ++by)
			for (int bx = 0; bx < nbx; ++bx) v.push_back(BItem{p, 0, bx, by});
		B->hp[p].nstp = (int32_t)v.size() - B->hp[p].stp0;
	}
	return v;
}
std::vector<BItem> upd_items(fx_batch* B, int chunk) {
	std::vector<BItem> v;
	for (int p = 0; p < B->nprob; ++p) {
		const fx_grid& g = B->grids[p];
		const int64_t n = (int64_t)g.rows * g.cols;
		B->hp[p].upd0 = (int32_t)v.size();
		for (int64_t c = 0; c * chunk < n; ++c) v.push_back(BItem{p, 0, (int32_t)c, 0});
		B->hp[p].nupd = (int32_t)v.size() - B->hp[p].upd0;
	}
	return v;
}

// Wave work lists for post segments of `sh` fine rows and restriction segments of `scr` coarse rows:
// sets the levels' launch shapes (B->hl, uploaded to B->dl), the per-depth lists and the level-0 post
// blocks' ranges per problem (hp[p].pst0 / npst).  Frees the previous lists.
fx_status wave_lists(fx_batch* B, int sh, int scr) {
	fx_backend* be = B->be;
	for (auto* q : B->wrestr) fx::dev_free(q);
	for (auto* q : B->wpost) fx::dev_free(q);
	fx::dev_free(B->wpost0);
	B->wrestr.assign(B->maxdepth, nullptr);
	B->wpost.assign(B->maxdepth, nullptr);
	B->wpost0 = nullptr;
	B->nwrestr.assign(B->maxdepth, 0);
	B->nwpost.assign(B->maxdepth, 0);
	for (int p = 0; p < B->nprob; ++p) {
		const BProb& P = B->hp[p];
		for (int l = 0; l + 1 < P.nlev; ++l) {
			BLev& L = B->hl[P.lev0 + l];
			L.psh = sh;
			L.pns = (int32_t)fmg::wave_strips_post(L.cols);
			L.pnseg = (int32_t)((L.rows + sh - 1) / sh);
			L.rscr = scr;
			L.rns = (int32_t)fmg::wave_strips_restrict(L.cc);
			L.rnseg = (int32_t)fmg::wave_segs_restrict(L.cr, scr);
		}
	}
	std::vector<BItem> p0;
	for (int d = 0; d < B->maxdepth; ++d) {
		std::vector<BItem> ri, pi;
		for (int p = 0; p < B->nprob; ++p) {
			const BProb& P = B->hp[p];
			if (d + 1 >= P.nlev) continue;
			const BLev& L = B->hl[P.lev0 + d];
			const int nr = L.rns * L.rnseg, np = L.pns * L.pnseg;
			for (int w0 = 0; w0 < nr; w0 += NWAVES) ri.push_back(BItem{p, P.lev0 + d, w0, 0});
			std::vector<BItem>& dst = d == 0 ? p0 : pi;
			if (d == 0) B->hp[p].pst0 = (int32_t)p0.size();
			for (int w0 = 0; w0 < np; w0 += NWAVES) dst.push_back(BItem{p, P.lev0 + d, w0, 0});
			if (d == 0) B->hp[p].npst = (int32_t)p0.size() - B->hp[p].pst0;
		}
		B->nwrestr[d] = (int)ri.size();
		B->nwpost[d] = (int)pi.size();
		FX_CHECK(upload_vec(be, ri, &B->wrestr[d]));
		FX_CHECK(upload_vec(be, pi, &B->wpost[d]));
	}
	B->nwpost0 = (int)p0.size();
	FX_CHECK(upload_vec(be, p0, &B->wpost0));
	FX_HIP(hipMemcpyAsync(B->dl, B->hl.data(), B->hl.size() * sizeof(BLev), hipMemcpyHostToDevice, be->stream));
	FX_HIP(hipStreamSynchronize(be->stream));
	B->wsh = sh;
	B->wscr = scr;
	return FX_OK;
}

template <bool VK, int NPRE, int NPOST>
fx_status batch_vcycle_wave(fx_batch* B, bool start) {
	hipStream_t st = B->be->stream;
	for (int d = 0; d + 1 <= B->maxdepth - 1; ++d) {
		if (B->nwrestr[d] == 0) continue;
		hipLaunchKernelGGL((k_b_wrestrict<VK, NPRE>), dim3(B->nwrestr[d]), dim3(BLOCK), 0, st, B->wrestr[d], B->dl,
		                   B->dst);
	}
	hipLaunchKernelGGL(k_b_coarse, dim3(B->nprob), dim3(1024), 0, st, B->dp, B->dl, B->dst);
	for (int d = B->maxdepth - 2; d >= 1; --d) {
		if (B->nwpost[d] == 0) continue;
		hipLaunchKernelGGL((k_b_wpost<VK, NPRE, NPOST>), dim3(B->nwpost[d]), dim3(BLOCK), 0, st, B->wpost[d], B->dl,
		                   B->dst);
	}
	RedScratch rs{B->partials};
	if (start) {
		hipLaunchKernelGGL((k_b_wpost0<VK, NPRE, NPOST, 1>), dim3(B->nwpost0), dim3(BLOCK), 0, st, B->wpost0, B->dp,
		                   B->dl, B->dst, (const double*)B->r, (const double*)B->ap, rs);
		hipLaunchKernelGGL((k_b_finish<1, FinPcgInitRZ>), dim3(B->nprob), dim3(BLOCK), 0, st, B->dp, B->dst, rs, 2, 0);
	} else {
		hipLaunchKernelGGL((k_b_wpost0<VK, NPRE, NPOST, 2>), dim3(B->nwpost0), dim3(BLOCK), 0, st, B->wpost0, B->dp,
		                   B->dl, B->dst, (const double*)B->r, (const double*)B->ap, rs);
		hipLaunchKernelGGL((k_b_finish<2, FinFcgBeta>), dim3(B->nprob), dim3(BLOCK), 0, st, B->dp, B->dst, rs, 2, 0);
	}
	return fx::check_launch("batched V-cycle (wave kernels)");
}

template <bool VK, int NPRE, int NPOST>
fx_status batch_vcycle(fx_batch* B, bool start) {
	if (B->wave) return batch_vcycle_wave<VK, NPRE, NPOST>(B, start);
	hipStream_t st = B->be->stream;
	for (int d = 0; d + 1 <= B->maxdepth - 1; ++d) {
		if (B->nrestr[d] == 0) continue;
		hipLaunchKernelGGL((k_b_restrict<VK, NPRE>), dim3(B->nrestr[d]), dim3(BLOCK), (unsigned)B->restr_lds[d], st,
		                   B->restr[d], B->dl, B->dst, (const float*)nullptr);
	}
	hipLaunchKernelGGL(k_b_coarse, dim3(B->nprob), dim3(1024), 0, st, B->dp, B->dl, B->dst);
	for (int d = B->maxdepth - 2; d >= 1; --d) {
		if (B->npost[d] == 0) continue;
		hipLaunchKernelGGL((k_b_post<VK, NPRE, NPOST>), dim3(B->npost[d]), dim3(BLOCK), 0, st, B->post[d], B->dl, B->dst);
	}
	RedScratch rs{B->partials};
	if (start) {
		hipLaunchKernelGGL((k_b_post0<VK, NPRE, NPOST, 1>), dim3(B->npost0), dim3(BLOCK), 0, st, B->post0, B->dp, B->dl,
		                   B->dst, (const double*)B->r, (const double*)B->ap, rs);
		hipLaunchKernelGGL((k_b_finish<1, FinPcgInitRZ>), dim3(B->nprob), dim3(BLOCK), 0, st, B->dp, B->dst, rs, 2, 0);
	} else {
		hipLaunchKernelGGL((k_b_post0<VK, NPRE, NPOST, 2>), dim3(B->npost0), dim3(BLOCK), 0, st, B->post0, B->dp, B->dl,
		                   B->dst, (const double*)B->r, (const double*)B->ap, rs);
		hipLaunchKernelGGL((k_b_finish<2, FinFcgBeta>), dim3(B->nprob), dim3(BLOCK), 0, st, B->dp, B->dst, rs, 2, 0);
	}
	return fx::check_launch("batched V-cycle");
}

fx_status batch_vcycle_any(fx_batch* B, bool start) {
	const int a = B->opt.n_pre, b = B->opt.n_post;
	if (a == 2 && b == 2) return batch_vcycle<true, 2, 2>(B, start);
	if (a == 1 && b == 1) return batch_vcycle<true, 1, 1>(B, start);
	if (a == 1 && b == 2) return batch_vcycle<true, 1, 2>(B, start);
	return batch_vcycle<true, 2, 1>(B, start);
}

// One FCG iteration of every running problem: p_new = z + β·p_old, A·p, α; x/r update and the
// convergence test; V-cycle with (r·z, z·Ap), β.  Even and odd iterations swap p0 and p1.
fx_status batch_iteration(fx_batch* B, int parity) {
	hipStream_t st = B->be->stream;
	RedScratch rs{B->partials};
	double* pold = parity ? B->p1 : B->p0;
	double* pnew = parity ? B->p0 : B->p1;
	hipLaunchKernelGGL(k_b_step, dim3(B->nstep), dim3(BLOCK), 0, st, B->step, B->dp, (const SState*)B->dst,
	                   (const double*)B->k->ptr, (const double*)pold, pnew, B->ap, 0.0, 1.0, B->rpb, rs);
	hipLaunchKernelGGL((k_b_finish<1, FinAlpha>), dim3(B->nprob), dim3(BLOCK), 0, st, B->dp, B->dst, rs, 0, 0);
	hipLaunchKernelGGL(k_b_update, dim3(B->nupd), dim3(BLOCK), 0, st, B->upd, B->dp, (const SState*)B->dst, B->x, B->r,
	                   (const double*)pnew, (const double*)B->ap, B->chunk, rs);
	hipLaunchKernelGGL((k_b_finish<1, FinPcgRR>), dim3(B->nprob), dim3(BLOCK), 0, st, B->dp, B->dst, rs, 1, 0);
	FX_CHECK(fx::check_launch("batched FCG step"));
	return batch_vcycle_any(B, false);
}

}  // namespace

using namespace fx;

extern "C" {

fx_status fx_solver_create(fx_backend* be, int32_t kind, const fx_grid* grid, uint32_t restart, double tol,
                           uint64_t max_iter, fx_solver** out) {
	FX_REQUIRE(be && out, FX_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	FX_CHECK(check_grid(grid));
	FX_REQUIRE(kind >= FX_SOLVER_CG && kind <= FX_SOLVER_FGMRES, FX_ERR_ARGUMENT, "unknown solver kind %d", kind);
	const bool gm = kind == FX_SOLVER_GMRES || kind == FX_SOLVER_FGMRES;
	FX_REQUIRE(!gm || (restart >= 1 && restart <= GM_MAX), FX_ERR_ARGUMENT, "GMRES restart must be in 1..%d",
	           GM_MAX);
	fx_solver* s = new fx_solver();
	s->be = be;
	s->kind = kind;
	s->g = *grid;
	s->n = grid->rows * grid->cols;
	s->restart = restart;
	s->tol = tol;
	s->max_iter = max_iter;
	auto fail = [&](fx_status st) {
		fx_solver_destroy(s);
		return st;
	};
	void* p = nullptr;
	fx_status st;
	if ((st = dev_alloc(be, sizeof(SState), &p)) != FX_OK) return fail(st);
	s->dst = (SState*)p;
	if ((st = dev_alloc(be, sizeof(double) * HIST_MAX, &p)) != FX_OK) return fail(st);
	s->dhist = (double*)p;
	if (hipHostMalloc((void**)&s->hst, sizeof(SState), hipHostMallocDefault) != hipSuccess) return fail(FX_ERR_ALLOC);
	if (hipHostMalloc((void**)&s->hdone, 2 * sizeof(int), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
		return fail(FX_ERR_ALLOC);
	for (auto& e : s->ev)
		if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(FX_ERR_HIP);
	// Work vectors up front, as the reference constructors do (cg.rs:79-84).
	int nvec = kind == FX_SOLVER_CG || kind == FX_SOLVER_FCG ? 5
	         : kind == FX_SOLVER_PIPELINED_CG                ? 6
	         : kind == FX_SOLVER_BICGSTAB                    ? 8
	                                                         : 1;
	for (int i = 0; i < nvec; ++i)
		if (!s->vec(i)) return fail(FX_ERR_ALLOC);
	if (gm) {
		for (uint32_t i = 0; i <= restart; ++i) {
			if (dev_alloc(be, s->n * sizeof(double), &p) != FX_OK) return fail(FX_ERR_ALLOC);
			s->basis.push_back((double*)p);
		}
		for (uint32_t i = 0; kind == FX_SOLVER_FGMRES && i < restart; ++i) {
			if (dev_alloc(be, s->n * sizeof(double), &p) != FX_OK) return fail(FX_ERR_ALLOC);
			s->zbasis.push_back((double*)p);
		}
		size_t m = restart;
		if (dev_alloc(be, sizeof(double) * ((m + 1) * m + 3 * m + 1 + m), &p) != FX_OK) return fail(FX_ERR_ALLOC);
		s->dgm = (double*)p;
	}
	if (hipStreamSynchronize(be->stream) != hipSuccess) return fail(FX_ERR_HIP);
	*out = s;
	return FX_OK;
}

void fx_solver_destroy(fx_solver* s) {
	if (!s) return;
	if (s->be) (void)hipStreamSynchronize(s->be->stream);
	for (double* p : s->v) dev_free(p);
	for (double* p : s->basis) dev_free(p);
	for (double* p : s->zbasis) dev_free(p);
	dev_free(s->dst);
	dev_free(s->dhist);
	dev_free(s->dgm);
	if (s->hst) (void)hipHostFree(s->hst);
	if (s->hdone) (void)hipHostFree(s->hdone);
	for (auto& e : s->ev)
		if (e) (void)hipEventDestroy(e);
	delete s;
}

const char* fx_solver_name(const fx_solver* s) { return s ? solver_name(s->kind) : "?"; }

size_t fx_solver_history(const fx_solver* s, double* out, size_t len) {
	if (!s) return 0;
	if (s->hist_stale) {
		s->hist.assign((size_t)s->hist_n, 0.0);
		if (s->hist_n > 0 && (hipMemcpyAsync(s->hist.data(), s->dhist, sizeof(double) * s->hist_n, hipMemcpyDeviceToHost,
		                                     s->be->stream) != hipSuccess ||
		                      hipStreamSynchronize(s->be->stream) != hipSuccess)) {
			s->hist.clear();
			return 0;
		}
		s->hist_stale = false;
	}
	size_t n = std::min(len, s->hist.size());
	if (out && n) std::memcpy(out, s->hist.data(), n * sizeof(double));
	return s->hist.size();
}

fx_status fx_solver_solve_op(fx_solver* s, fx_buf* x, const fx_buf* b, const fx_operator* op, const fx_precond* pc,
                             fx_solve_stats* stats) {
	FX_REQUIRE(s && op, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(check_buf(x, s->n, FX_F64, "x"));
	FX_CHECK(check_buf(b, s->n, FX_F64, "b"));
	Ctx c;
	FX_CHECK(make_ctx(s, op, pc, &c));
	double* xp = (double*)x->ptr;
	const double* bp = (const double*)b->ptr;
	switch (s->kind) {
	case FX_SOLVER_CG: return solve_cg(c, xp, bp, stats);
	case FX_SOLVER_FCG: return solve_fcg(c, xp, bp, stats);
	case FX_SOLVER_PIPELINED_CG:
		if (!pc) return solve_pipelined(c, xp, bp, stats);
		// An FP32 multigrid gets the flexible pipelined form; exact preconditioners Ghysels–Vanroose.
		if (pc->kind == FX_PC_MULTIGRID && ((fx_multigrid*)pc->handle)->f32) return solve_pfcg(c, xp, bp, stats);
		return solve_ppcg(c, xp, bp, stats);
	case FX_SOLVER_BICGSTAB: return solve_bicgstab(c, xp, bp, stats);
	case FX_SOLVER_GMRES:
	case FX_SOLVER_FGMRES: return solve_gmres(c, xp, bp, stats);
	default: break;
	}
	set_error("solver kind %d not implemented", s->kind);
	return FX_ERR_UNSUPPORTED;
}

fx_status fx_learned_create(fx_backend* be, fx_unet* net, fx_multigrid* mg, const fx_buf* logk, const fx_grid* grid,
                            double stall, fx_learned** out) {
	FX_REQUIRE(be && net && mg && out, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(check_grid(grid));
	FX_REQUIRE(grid->rows % 8 == 0 && grid->cols % 8 == 0, FX_ERR_ARGUMENT,
	           "the network needs grid sides divisible by 8, got %llux%llu", (unsigned long long)grid->rows,
	           (unsigned long long)grid->cols);
	FX_REQUIRE(mg->lv[0].g.rows == grid->rows && mg->lv[0].g.cols == grid->cols, FX_ERR_DIMENSION,
	           "multigrid grid does not match");
	size_t n = grid->rows * grid->cols;
	if (logk) FX_CHECK(check_buf(logk, n, FX_F64, "logk"));
	fx_learned* L = new fx_learned();
	L->be = be;
	L->net = net;
	L->mg = mg;
	L->logk = logk;
	L->g = *grid;
	L->stall = stall;
	void* p;
	fx_status st;
	if ((st = dev_alloc(be, 2 * sizeof(double), &p)) != FX_OK) { fx_learned_destroy(L); return st; }
	L->s_dev = (double*)p;
	if ((st = dev_alloc(be, n * sizeof(double), &p)) != FX_OK) { fx_learned_destroy(L); return st; }
	L->r1 = (double*)p;
	if ((st = dev_alloc(be, n * sizeof(double), &p)) != FX_OK) { fx_learned_destroy(L); return st; }
	L->c = (double*)p;
	if ((st = dev_alloc(be, n * sizeof(double), &p)) != FX_OK) { fx_learned_destroy(L); return st; }
	L->ac = (double*)p;
	if ((st = dev_alloc(be, n * 2 * sizeof(float), &p)) != FX_OK) { fx_learned_destroy(L); return st; }
	L->xin = (float*)p;
	if ((st = dev_alloc(be, n * sizeof(double), &p)) != FX_OK) { fx_learned_destroy(L); return st; }
	L->tr = (double*)p;
	*out = L;
	return FX_OK;
}

void fx_learned_destroy(fx_learned* L) {
	if (!L) return;
	if (L->be) (void)hipStreamSynchronize(L->be->stream);
	dev_free(L->s_dev);
	dev_free(L->r1);
	dev_free(L->c);
	dev_free(L->ac);
	dev_free(L->xin);
	dev_free(L->tr);
	delete L;
}

fx_status fx_learned_set_schedule(fx_learned* L, int32_t every, int32_t max_apps, int32_t first) {
	FX_REQUIRE(L, FX_ERR_ARGUMENT, "learned preconditioner is null");
	FX_REQUIRE(every >= 1 && max_apps >= 0 && first >= 0, FX_ERR_ARGUMENT,
	           "every must be >= 1, max_apps >= 0 and first >= 0");
	L->every = every;
	L->max_apps = max_apps;
	L->first = first;
	return FX_OK;
}

int32_t fx_learned_fell_back(fx_learned* L) {
	if (!L) return 0;
	int v = L->fell_back;
	L->fell_back = 0;
	return v;
}

static fx_operator stencil_op(double coeff, const fx_stencil* st, const fx_boundaries* bc) {
	fx_operator op{};
	op.kind = FX_OP_STENCIL;
	op.coeff = coeff;
	op.sigma = 1.0;
	op.stencil = st;
	op.bc = bc;
	return op;
}

fx_status fx_solver_solve(fx_solver* s, fx_buf* x, const fx_buf* b, double coeff, const fx_stencil* st,
                          const fx_boundaries* bc, fx_solve_stats* stats) {
	fx_operator op = stencil_op(coeff, st, bc);
	return fx_solver_solve_op(s, x, b, &op, nullptr, stats);
}

fx_status fx_solver_solve_preconditioned(fx_solver* s, fx_buf* x, const fx_buf* b, double coeff, const fx_stencil* st,
                                         const fx_boundaries* bc, const fx_precond* pc, fx_solve_stats* stats) {
	FX_REQUIRE(pc, FX_ERR_ARGUMENT, "preconditioner is null");
	fx_operator op = stencil_op(coeff, st, bc);
	return fx_solver_solve_op(s, x, b, &op, pc, stats);
}

fx_status fx_batch_create(fx_backend* be, uint32_t nprob, const fx_grid* grids, const fx_buf* k,
                          const fx_mg_options* opt, fx_batch** out) {
	FX_REQUIRE(be && grids && k && opt && out && nprob > 0, FX_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	FX_REQUIRE(opt->precision == FX_F32 && opt->fused && opt->smoother == FX_SMOOTH_JACOBI &&
	               opt->transfer == FX_TRANSFER_FW_BILINEAR,
	           FX_ERR_UNSUPPORTED, "the batched solver runs the fused FP32 V-cycle (FX_F32, fused, Jacobi, FW)");
	int64_t ntot = 0;
	for (uint32_t p = 0; p < nprob; ++p) {
		FX_CHECK(check_grid(&grids[p]));
		FX_REQUIRE(grids[p].rows < (1u << 30) && grids[p].cols < (1u << 30), FX_ERR_GRID, "grid too large");
		ntot += (int64_t)grids[p].rows * grids[p].cols;
	}
	FX_CHECK(check_buf(k, (size_t)ntot, FX_F64, "k"));
	fx_batch* B = new fx_batch();
	B->be = be;
	B->nprob = (int)nprob;
	B->ntot = ntot;
	B->opt = *opt;
	B->k = k;
	B->grids.assign(grids, grids + nprob);
	B->hp.resize(nprob);
	fx_status s = FX_OK;
	auto fail = [&](fx_status e) { batch_free(B); return e; };
	// Per-problem hierarchies (coefficient-aware FP32 levels built from a view of k).
	fx_boundaries zd;
	zd.top = zd.bottom = zd.left = zd.right = fx_bc{FX_BC_DIRICHLET, 0.0, 0.0, 0.0};
	std::vector<BLev> lev;
	int64_t off = 0;
	for (int p = 0; p < B->nprob; ++p) {
		const fx_grid& g = grids[p];
		fx_buf view;
		view.be = be;
		view.ptr = (double*)k->ptr + off;
		view.len = (size_t)g.rows * g.cols;
		view.dtype = FX_F64;
		fx_mg_options o = *opt;
		o.k = &view;
		fx_multigrid* mg = nullptr;
		if ((s = fx_multigrid_create_ex(be, &g, &zd, &o, &mg)) != FX_OK) return fail(s);
		mg->opt.k = nullptr;
		B->mgs.push_back(mg);
		BProb& P = B->hp[p];
		P.off = off;
		P.rows = (int32_t)g.rows;
		P.cols = (int32_t)g.cols;
		P.ix2 = 1.0 / (g.dx * g.dx);
		P.iy2 = 1.0 / (g.dy * g.dy);
		P.r32 = (float*)mg->lv[0].rhs;
		P.z32 = (const float*)mg->lv[0].tmp;
		P.lev0 = (int32_t)lev.size();
		P.nlev = (int32_t)mg->lv.size();
		P.sweeps = opt->n_pre + opt->n_post + opt->extra_coarse;
		for (size_t l = 0; l < mg->lv.size(); ++l) {
			const auto& L = mg->lv[l];
			BLev d{};
			d.rhs = (float*)L.rhs;
			d.tmp = (float*)L.tmp;
			d.L = fmg::lwf(mg, L);
			d.X = L.xt;
			d.rows = (int32_t)L.g.rows;
			d.cols = (int32_t)L.g.cols;
			d.cr = l + 1 < mg->lv.size() ? (int32_t)mg->lv[l + 1].g.rows : 0;
			d.cc = l + 1 < mg->lv.size() ? (int32_t)mg->lv[l + 1].g.cols : 0;
			lev.push_back(d);
		}
		B->maxdepth = std::max(B->maxdepth, P.nlev);
		off += (int64_t)g.rows * g.cols;
	}
	// Work lists: V-cycle tiles per level depth and the level-0 post tiles.
	B->restr.assign(B->maxdepth, nullptr);
	B->post.assign(B->maxdepth, nullptr);
	B->nrestr.assign(B->maxdepth, 0);
	B->npost.assign(B->maxdepth, 0);
	B->restr_lds.assign(B->maxdepth, 0);
	std::vector<BItem> p0items;
	for (int d = 0; d < B->maxdepth; ++d) {
		std::vector<BItem> ri, pi;
		for (int p = 0; p < B->nprob; ++p) {
			const BProb& P = B->hp[p];
			if (d + 1 >= P.nlev) continue;   // level d is the coarsest (or absent) for this problem
			const BLev& L = lev[P.lev0 + d];
			const dim3 rg = fmg::rgrid(L.cr, L.cc), tg = fmg::tgrid(L.rows, L.cols);
			for (unsigned by = 0; by < rg.y; ++by)
				for (unsigned bx = 0; bx < rg.x; ++bx) ri.push_back(BItem{p, P.lev0 + d, (int32_t)bx, (int32_t)by});
			B->restr_lds[d] = std::max(B->restr_lds[d], fmg::frestrict_lds(L.X, true));
			std::vector<BItem>& dst = d == 0 ? p0items : pi;
			if (d == 0) B->hp[p].pst0 = (int32_t)p0items.size();
			for (unsigned by = 0; by < tg.y; ++by)
				for (unsigned bx = 0; bx < tg.x; ++bx) dst.push_back(BItem{p, P.lev0 + d, (int32_t)bx, (int32_t)by});
			if (d == 0) B->hp[p].npst = (int32_t)p0items.size() - B->hp[p].pst0;
		}
		if (B->restr_lds[d] > 64 * 1024) {
			set_error("restriction tile too large for level depth %d", d);
			return fail(FX_ERR_UNSUPPORTED);
		}
		B->nrestr[d] = (int)ri.size();
		B->npost[d] = (int)pi.size();
		if ((s = upload_vec(be, ri, &B->restr[d])) != FX_OK) return fail(s);
		if ((s = upload_vec(be, pi, &B->post[d])) != FX_OK) return fail(s);
	}
	for (int p = 0; p < B->nprob; ++p) {
		if (B->hp[p].nlev < 2) {
			set_error("problem %d is too small for a two-level hierarchy", p);
			return fail(FX_ERR_GRID);
		}
	}
	B->npost0 = (int)p0items.size();
	if ((s = upload_vec(be, p0items, &B->post0)) != FX_OK) return fail(s);
	if ((s = upload_vec(be, lev, &B->dl)) != FX_OK) return fail(s);
	// The wave-streaming level kernels serve the batch when every level's restriction fits a wave.
	B->hl = lev;
	B->wave = true;
	for (int p = 0; p < B->nprob && B->wave; ++p)
		for (int l = 0; l + 1 < B->hp[p].nlev; ++l)
			if (!fmg::wrestrict_ok(lev[B->hp[p].lev0 + l].X)) B->wave = false;
	// Work vectors and state.
	void* q = nullptr;
	double** vecs[5] = {&B->x, &B->r, &B->p0, &B->p1, &B->ap};
	for (double** v : vecs) {
		if ((s = fx::dev_alloc(be, (size_t)ntot * sizeof(double), &q)) != FX_OK) return fail(s);
		*v = (double*)q;
	}
	if ((s = fx::dev_alloc(be, (size_t)B->nprob * sizeof(SState), &q)) != FX_OK) return fail(s);
	B->dst = (SState*)q;
	if ((s = fx::dev_alloc(be, sizeof(int), &q)) != FX_OK) return fail(s);
	B->dactive = (int*)q;
	if (hipHostMalloc((void**)&B->hactive, sizeof(int), hipHostMallocDefault) != hipSuccess) {
		set_error("hipHostMalloc failed");
		return fail(FX_ERR_ALLOC);
	}
	// Launch-shape tuning: rows per block of the step kernel and points per block of the update kernel,
	// each timed on this batch with zero states (β = α = 0, nothing skipped; only scratch vectors move).
	FX_HIP(hipMemsetAsync(B->dst, 0, (size_t)B->nprob * sizeof(SState), be->stream));
	size_t maxparts = (size_t)B->npost0 * 2;
	static constexpr int WSH[4] = {8, 16, 32, 64}, WSCR[4] = {4, 8, 16, 32};   // wave segment heights tried
	if (B->wave) {
		size_t nb = 0;   // level-0 post blocks at the shortest segments
		for (int p = 0; p < B->nprob; ++p) {
			const fx_grid& g = B->grids[p];
			nb += (size_t)((fmg::wave_strips_post(g.cols) * ((g.rows + WSH[0] - 1) / WSH[0]) + NWAVES - 1) / NWAVES);
		}
		maxparts = std::max(maxparts, nb * 2);
	}
	for (int rpb : {4, 8, 16, 32}) maxparts = std::max(maxparts, step_items(B, rpb).size());
	for (int ch : {BLOCK * 4, BLOCK * 8, BLOCK * 16}) maxparts = std::max(maxparts, upd_items(B, ch).size());
	if ((s = fx::dev_alloc(be, maxparts * sizeof(fxd::Acc), &q)) != FX_OK) return fail(s);
	B->partials = (fxd::Acc*)q;
	if ((s = fx::dev_alloc(be, (size_t)B->nprob * sizeof(BProb), &q)) != FX_OK) return fail(s);
	B->dp = (BProb*)q;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	FX_HIP(hipEventCreate(&e0));
	FX_HIP(hipEventCreate(&e1));
	auto timed = [&](auto launch, float* ms) -> fx_status {
		FX_HIP(hipMemcpyAsync(B->dp, B->hp.data(), B->hp.size() * sizeof(BProb), hipMemcpyHostToDevice, be->stream));
		launch();
		FX_HIP(hipEventRecord(e0, be->stream));
		for (int i = 0; i < 3; ++i) launch();
		FX_HIP(hipEventRecord(e1, be->stream));
		FX_HIP(hipEventSynchronize(e1));
		FX_CHECK(fx::check_launch("batch tuning"));
		FX_HIP(hipEventElapsedTime(ms, e0, e1));
		return FX_OK;
	};
	float best = 1e30f;
	for (int rpb : {4, 8, 16, 32}) {
		std::vector<BItem> it = step_items(B, rpb);
		BItem* d = nullptr;
		if ((s = upload_vec(be, it, &d)) != FX_OK) return fail(s);
		float ms = 0.0f;
		s = timed([&] {
			hipLaunchKernelGGL(k_b_step, dim3((unsigned)it.size()), dim3(BLOCK), 0, be->stream, (const BItem*)d,
			                   (const BProb*)B->dp, (const SState*)B->dst, (const double*)k->ptr, (const double*)B->p0,
			                   B->p1, B->ap, 0.0, 1.0, rpb, RedScratch{B->partials});
		}, &ms);
		fx::dev_free(d);
		if (s != FX_OK) return fail(s);
		if (ms < best) {
			best = ms;
			B->rpb = rpb;
		}
	}
	best = 1e30f;
	for (int ch : {BLOCK * 4, BLOCK * 8, BLOCK * 16}) {
		std::vector<BItem> it = upd_items(B, ch);
		BItem* d = nullptr;
		if ((s = upload_vec(be, it, &d)) != FX_OK) return fail(s);
		float ms = 0.0f;
		s = timed([&] {
			hipLaunchKernelGGL(k_b_update, dim3((unsigned)it.size()), dim3(BLOCK), 0, be->stream, (const BItem*)d,
			                   (const BProb*)B->dp, (const SState*)B->dst, B->x, B->r, (const double*)B->p1,
			                   (const double*)B->ap, ch, RedScratch{B->partials});
		}, &ms);
		fx::dev_free(d);
		if (s != FX_OK) return fail(s);
		if (ms < best) {
			best = ms;
			B->chunk = ch;
		}
	}
	if (B->wave) {
		// Segment heights of the wave level kernels, one pair for the whole batch: the V-cycle timed with
		// zero states for each candidate (taller segments re-read fewer lag rows, shorter ones give more
		// waves).
		best = 1e30f;
		int pick = 1;
		for (int c = 0; c < 4; ++c) {
			if ((s = wave_lists(B, WSH[c], WSCR[c])) != FX_OK) return fail(s);
			float ms = 0.0f;
			s = timed([&] { (void)batch_vcycle_any(B, false); }, &ms);
			if (s != FX_OK) return fail(s);
			if (ms < best) best = ms, pick = c;
		}
		if ((s = wave_lists(B, WSH[pick], WSCR[pick])) != FX_OK) return fail(s);
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	std::vector<BItem> sit = step_items(B, B->rpb), uit = upd_items(B, B->chunk);
	B->nstep = (int)sit.size();
	B->nupd = (int)uit.size();
	if ((s = upload_vec(be, sit, &B->step)) != FX_OK) return fail(s);
	if ((s = upload_vec(be, uit, &B->upd)) != FX_OK) return fail(s);
	FX_HIP(hipMemcpyAsync(B->dp, B->hp.data(), B->hp.size() * sizeof(BProb), hipMemcpyHostToDevice, be->stream));
	// The iteration graph: graph_iters FCG iterations of every problem, then the count of running problems.
	FX_HIP(hipStreamSynchronize(be->stream));
	FX_HIP(hipStreamBeginCapture(be->stream, hipStreamCaptureModeRelaxed));
	fx_status cs = FX_OK;
	for (int it = 0; it < B->graph_iters && cs == FX_OK; ++it) cs = batch_iteration(B, it & 1);
	hipLaunchKernelGGL(k_b_active, dim3(1), dim3(BLOCK), 0, be->stream, (const SState*)B->dst, B->nprob, B->dactive);
	hipError_t ce = hipStreamEndCapture(be->stream, &B->graph);
	if (cs != FX_OK) return fail(cs);
	if (ce != hipSuccess) {
		set_error("graph capture failed: %s", hipGetErrorString(ce));
		return fail(FX_ERR_HIP);
	}
	if (hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0) != hipSuccess) {
		set_error("graph instantiation failed");
		return fail(FX_ERR_HIP);
	}
	*out = B;
	return FX_OK;
}

fx_status fx_batch_solve(fx_batch* B, fx_buf* x, const fx_buf* b, const double* tol, uint64_t max_iter,
                         fx_solve_stats* stats) {
	FX_REQUIRE(B && tol && stats, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(check_buf(x, (size_t)B->ntot, FX_F64, "x"));
	FX_CHECK(check_buf(b, (size_t)B->ntot, FX_F64, "b"));
	fx_backend* be = B->be;
	hipStream_t st = be->stream;
	const size_t bytes = (size_t)B->ntot * sizeof(double);
	FX_HIP(hipMemcpyAsync(B->x, x->ptr, bytes, hipMemcpyDeviceToDevice, st));
	std::vector<SState> h((size_t)B->nprob);
	std::memset(h.data(), 0, h.size() * sizeof(SState));
	for (int p = 0; p < B->nprob; ++p) {
		FX_REQUIRE(tol[p] >= 0.0, FX_ERR_ARGUMENT, "negative tolerance for problem %d", p);
		h[p].tol = tol[p];
		h[p].max_iter = (int64_t)max_iter;
	}
	FX_HIP(hipMemcpyAsync(B->dst, h.data(), h.size() * sizeof(SState), hipMemcpyHostToDevice, st));
	RedScratch rs{B->partials};
	// r = b − A·x, r·r; p = 0; z = V(r) with r·z and the iteration-0 test.
	hipLaunchKernelGGL(k_b_resid, dim3(B->nupd), dim3(BLOCK), 0, st, (const BItem*)B->upd, (const BProb*)B->dp,
	                   (const double*)B->k->ptr, (const double*)B->x, (const double*)b->ptr, B->r, 0.0, 1.0, B->chunk, 0,
	                   rs);
	hipLaunchKernelGGL((k_b_finish<1, FinPcgInitRR>), dim3(B->nprob), dim3(BLOCK), 0, st, B->dp, B->dst, rs, 1, 1);
	hipLaunchKernelGGL(k_b_zero, dim3(B->nupd), dim3(BLOCK), 0, st, (const BItem*)B->upd, (const BProb*)B->dp, B->p0,
	                   B->chunk);
	FX_CHECK(fx::check_launch("batch init"));
	FX_CHECK(batch_vcycle_any(B, true));
	hipLaunchKernelGGL(k_b_active, dim3(1), dim3(BLOCK), 0, st, (const SState*)B->dst, B->nprob, B->dactive);
	FX_HIP(hipMemcpyAsync(B->hactive, B->dactive, sizeof(int), hipMemcpyDeviceToHost, st));
	FX_HIP(hipStreamSynchronize(st));
	// Replay the captured iterations until no problem is running (every problem stops by max_iter).
	for (uint64_t done = 0; *B->hactive > 0 && done <= max_iter; done += (uint64_t)B->graph_iters) {
		FX_HIP(hipGraphLaunch(B->exec, st));
		FX_HIP(hipMemcpyAsync(B->hactive, B->dactive, sizeof(int), hipMemcpyDeviceToHost, st));
		FX_HIP(hipStreamSynchronize(st));
	}
	// FP64 true residual of every problem.
	hipLaunchKernelGGL(k_b_resid, dim3(B->nupd), dim3(BLOCK), 0, st, (const BItem*)B->upd, (const BProb*)B->dp,
	                   (const double*)B->k->ptr, (const double*)B->x, (const double*)b->ptr, B->r, 0.0, 1.0, B->chunk, 1,
	                   rs);
	hipLaunchKernelGGL((k_b_finish<1, FinTrue>), dim3(B->nprob), dim3(BLOCK), 0, st, B->dp, B->dst, rs, 1, 1);
	FX_CHECK(fx::check_launch("batch true residual"));
	FX_HIP(hipMemcpyAsync(x->ptr, B->x, bytes, hipMemcpyDeviceToDevice, st));
	FX_HIP(hipMemcpyAsync(h.data(), B->dst, h.size() * sizeof(SState), hipMemcpyDeviceToHost, st));
	FX_HIP(hipStreamSynchronize(st));
	for (int p = 0; p < B->nprob; ++p) {
		stats[p].iterations = (uint64_t)h[p].iter;
		stats[p].residual = h[p].resid;
		stats[p].reason = h[p].done ? h[p].reason : FX_MAX_ITERATIONS;
		stats[p].pad = 0;
		stats[p].true_residual = h[p].omega;
	}
	return FX_OK;
}

void fx_batch_destroy(fx_batch* B) { batch_free(B); }

fx_status fx_batch_info(const fx_batch* B, int32_t* rows_per_block, int32_t* points_per_block, uint64_t* points,
                        int32_t* graph_iterations) {
	FX_REQUIRE(B, FX_ERR_ARGUMENT, "null batch");
	if (rows_per_block) *rows_per_block = B->rpb;
	if (points_per_block) *points_per_block = B->chunk;
	if (points) *points = (uint64_t)B->ntot;
	if (graph_iterations) *graph_iterations = B->graph_iters;
	return FX_OK;
}

fx_status fx_batch_level_kernels(const fx_batch* B, int32_t* wave, int32_t* post_rows, int32_t* restrict_rows) {
	FX_REQUIRE(B, FX_ERR_ARGUMENT, "null batch");
	if (wave) *wave = B->wave ? 1 : 0;
	if (post_rows) *post_rows = B->wave ? B->wsh : 0;
	if (restrict_rows) *restrict_rows = B->wave ? B->wscr : 0;
	return FX_OK;
}

}  // extern "C"
