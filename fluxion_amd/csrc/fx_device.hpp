// fx_device.hpp — device-side building blocks shared by the gfx950 kernels.
//
// All HIP sources are compiled with -ffp-contract=off: a fused multiply-add
// appears only where fma()/fmaf() is written, which is exactly where the
// reference CPU backend writes `mul_add` (crates/backend-cpu/src/ops.rs).
// Stencil and BLAS-1 results are therefore bit-identical to the reference's
// arithmetic; only reductions differ (a different, more accurate summation).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fxd {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;            // 4 waves
constexpr int NWAVES = BLOCK / WAVE;

// Stencil weights by offset (ops.rs:908-920 extraction, last match wins).
struct W9 {
	double c, n, s, w, e, nw, ne, sw, se;
};
// Generic radius-1 stencil in entry order (ops.rs:872-883 loop order).
struct Ent {
	int n;
	int dr[32];
	int dc[32];
	double w[32];
};

// ops.rs:924-935 / 1093-1106: w_c·c + (w_n·n + (w_s·s + (w_w·w + w_e·e))).
__device__ __forceinline__ double chain5(const W9& k, double c, double n, double s, double w, double e) {
	return fma(k.c, c, fma(k.n, n, fma(k.s, s, fma(k.w, w, k.e * e))));
}
__device__ __forceinline__ float chain5f(const float* k, float c, float n, float s, float w, float e) {
	return fmaf(k[0], c, fmaf(k[1], n, fmaf(k[2], s, fmaf(k[3], w, k[4] * e))));
}
// ops.rs:984-1007 / 1058-1083.
__device__ __forceinline__ double chain9(const W9& k, const double* x, int64_t i, int64_t cols) {
	return fma(k.c, x[i], fma(k.n, x[i - cols], fma(k.s, x[i + cols], fma(k.w, x[i - 1],
	       fma(k.e, x[i + 1], fma(k.nw, x[i - cols - 1], fma(k.ne, x[i - cols + 1],
	       fma(k.sw, x[i + cols - 1], k.se * x[i + cols + 1]))))))));
}
__device__ __forceinline__ double chain_generic(const Ent& es, const double* x, int64_t i, int64_t cols) {
	double acc = 0.0;
	for (int j = 0; j < es.n; ++j) acc = fma(es.w[j], x[i + es.dr[j] * cols + es.dc[j]], acc);
	return acc;
}

// ── compensated (TwoSum) accumulation ────────────────────────────────
// value = hi + lo.  Deterministic for a fixed launch shape.
struct Acc {
	double hi, lo;
};
__device__ __forceinline__ void acc_add(Acc& a, double x) {
	double s = a.hi + x;
	double bp = s - a.hi;
	a.lo += (a.hi - (s - bp)) + (x - bp);
	a.hi = s;
}
__device__ __forceinline__ void acc_merge(Acc& a, const Acc& b) {
	double s = a.hi + b.hi;
	double bp = s - a.hi;
	a.lo = (a.lo + b.lo) + ((a.hi - (s - bp)) + (b.hi - bp));
	a.hi = s;
}
__device__ __forceinline__ double acc_value(const Acc& a) { return a.hi + a.lo; }

__device__ __forceinline__ int tid_linear() { return threadIdx.x + threadIdx.y * blockDim.x; }
__device__ __forceinline__ int bid_linear() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ __forceinline__ int nblocks_linear() { return gridDim.x * gridDim.y; }

template <int NS>
__device__ __forceinline__ void wave_reduce(Acc (&v)[NS]) {
#pragma unroll
	for (int off = WAVE / 2; off > 0; off >>= 1) {
#pragma unroll
		for (int k = 0; k < NS; ++k) {
			Acc o;
			o.hi = __shfl_xor(v[k].hi, off);
			o.lo = __shfl_xor(v[k].lo, off);
			acc_merge(v[k], o);
		}
	}
}

// Block reduction over BLOCK threads; result valid in thread 0.
template <int NS>
__device__ __forceinline__ void block_reduce(Acc (&v)[NS]) {
	__shared__ Acc sh[NS][NWAVES];
	wave_reduce<NS>(v);
	const int t = tid_linear();
	if ((t & (WAVE - 1)) == 0) {
#pragma unroll
		for (int k = 0; k < NS; ++k) sh[k][t / WAVE] = v[k];
	}
	__syncthreads();
	if (t == 0) {
#pragma unroll
		for (int k = 0; k < NS; ++k) {
			Acc a = sh[k][0];
			for (int w = 1; w < NWAVES; ++w) acc_merge(a, sh[k][w]);
			v[k] = a;
		}
	}
	__syncthreads();
}

// Workgroups are dealt round-robin to the eight XCDs in dispatch order (x
// fastest).  xcd_remap maps the hardware block index to a logical tile index so
// that each XCD receives one contiguous run of tiles (neighbouring tiles then
// share that XCD's L2 instead of re-reading halos from HBM); a bijection on
// [0, nb).  xcd_tile applies it to a 2-D grid, row-major.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
	constexpr int NX = 8;
	const int q = nb / NX, r = nb % NX;
	const int i = b % NX, k = b / NX;
	return i * q + (i < r ? i : r) + k;
}
__device__ __forceinline__ void xcd_tile(int* bx, int* by) {
	const int lid = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
	*bx = lid % gridDim.x;
	*by = lid / gridDim.x;
}

// Reduction scratch handed to every reducing kernel.
struct RedScratch {
	Acc* partials;        // [nblocks * NS]
};

// Non-temporal stores, for results the next kernels do not read back from
// the caches (a later kernel streams them from HBM): they leave the L2 and
// the MALL to the data that is reused, which measurably speeds up the kernel
// that follows a large writer.
typedef double fx_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(double* p, double2 v) {
	__builtin_nontemporal_store(fx_d2v{v.x, v.y}, (fx_d2v*)p);
}
__device__ __forceinline__ void store_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }

// Writes this block's partial sums to partials[block * NS + k].  The global
// sums are formed by k_finish in the next launch on the same stream: the
// kernel boundary orders the partial stores before their loads, so no
// device-scope fence or atomic is needed (on gfx950 such a fence writes back
// the XCD's L2, which serialises a grid of thousands of workgroups).
template <int NS>
__device__ __forceinline__ void block_partial(Acc (&v)[NS], const RedScratch& rs) {
	block_reduce<NS>(v);
	if (tid_linear() == 0) {
#pragma unroll
		for (int k = 0; k < NS; ++k) rs.partials[bid_linear() * NS + k] = v[k];
	}
}

// One workgroup of FIN_BLOCK threads merges `nb` block partials in a fixed
// order (thread t takes blocks t, t+FIN_BLOCK, …, four loads in flight, then
// the wave and workgroup trees) and runs fin(out) in thread 0.  Deterministic
// for a fixed nb.  Skipped, like the kernel that wrote the partials, when
// *skip is set.
constexpr int FIN_BLOCK = 1024;
template <int NS, class Fin>
__global__ __launch_bounds__(FIN_BLOCK) void k_finish(RedScratch rs, unsigned nb, Fin fin, const int* skip) {
	if (skip && *skip) return;
	const int t = threadIdx.x;
	Acc a[NS];
#pragma unroll
	for (int k = 0; k < NS; ++k) a[k] = Acc{0.0, 0.0};
	unsigned b = t;
	for (; b + 3 * FIN_BLOCK < nb; b += 4 * FIN_BLOCK) {
		Acc p[4][NS];
#pragma unroll
		for (int u = 0; u < 4; ++u)
#pragma unroll
			for (int k = 0; k < NS; ++k) p[u][k] = rs.partials[(size_t)(b + u * FIN_BLOCK) * NS + k];
#pragma unroll
		for (int u = 0; u < 4; ++u)
#pragma unroll
			for (int k = 0; k < NS; ++k) acc_merge(a[k], p[u][k]);
	}
	for (; b < nb; b += FIN_BLOCK) {
#pragma unroll
		for (int k = 0; k < NS; ++k) acc_merge(a[k], rs.partials[(size_t)b * NS + k]);
	}
	wave_reduce<NS>(a);
	__shared__ Acc sh[NS][FIN_BLOCK / WAVE];
	if ((t & (WAVE - 1)) == 0) {
#pragma unroll
		for (int k = 0; k < NS; ++k) sh[k][t / WAVE] = a[k];
	}
	__syncthreads();
	if (t == 0) {
#pragma unroll
		for (int k = 0; k < NS; ++k) {
			Acc v = sh[k][0];
			for (int w = 1; w < FIN_BLOCK / WAVE; ++w) acc_merge(v, sh[k][w]);
			a[k] = v;
		}
		fin(a);
	}
}

// ── streaming 5-point stencil ────────────────────────────────────────
// One thread per column, marching down `rpb` rows and keeping the north /
// centre / south values in registers; west/east come from the cache.  The
// value source `V` and the epilogue `E` make the same traversal serve
// stencil_axpy, A·v, residuals, smoothers and the fused CG updates.
//   V::load(i)            -> value of the (possibly derived) input at i
//   E::interior(i, lap, c, acc)   E::boundary(i, c, acc)
//   F::operator()(Acc*)   runs once, in the last block, on the global sums
struct NoFinal {
	template <class A>
	__device__ void operator()(A*) const {}
};

template <int NS, class V, class E, class F>
__global__ __launch_bounds__(BLOCK) void k_stream5(V v, E e, F fin, W9 w, int64_t rows, int64_t cols, int rpb,
                                                   RedScratch rs, const int* done) {
	if (done && *done) return;
	int bx, by;
	xcd_tile(&bx, &by);
	const int64_t c = (int64_t)bx * BLOCK + threadIdx.x;
	const int64_t r0 = (int64_t)by * rpb;
	const int64_t r1 = r0 + rpb < rows ? r0 + rpb : rows;
	const bool col_ok = c < cols;
	const bool col_int = c > 0 && c < cols - 1;
	Acc acc[NS > 0 ? NS : 1];
#pragma unroll
	for (int k = 0; k < (NS > 0 ? NS : 1); ++k) acc[k] = Acc{0.0, 0.0};
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
	if constexpr (NS > 0) block_partial<NS>(acc, rs);
}

// v of lane x − 1 (CTRL 0x138, wave_shr:1) or x + 1 (0x130, wave_shl:1) for a
// double, as two 32-bit DPP moves; 0 past the wave.
template <int CTRL>
__device__ __forceinline__ double dlane64(double v) {
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
	return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// k_stream5 for an even column count, two columns per thread: one 16-byte load
// of the pair per row instead of a load per point and its west and east
// neighbours; the neighbours come from the adjacent lanes by DPP (the wave's
// first and last lanes read the column beyond the wave).  The epilogue gets
// both points at once (E::pair) and moves its own vectors 16 bytes at a time.  Rows past the grid
// are read from the last row (clamped) and never used.  Every lane walks the
// rows, so the cross-lane moves see all 64 lanes.  chain5 is evaluated on the
// same operands per point, so the results equal k_stream5's.
template <class E>
__global__ __launch_bounds__(BLOCK) void k_stream5d(const double* __restrict__ x, E e, W9 w, int64_t rows,
                                                    int64_t cols, int rpb, const int* done) {
	if (done && *done) return;
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
		double2 v;
		double w, e;
	};
	auto ld = [&](int64_t r) {
		Row o;
		const int64_t i = r * cols + cl;
		o.v = *(const double2*)(x + i);
		o.w = edgeW ? x[i - 1] : 0.0;
		o.e = edgeE ? x[i + 2] : 0.0;
		return o;
	};
	Row up = ld(r0 > 0 ? r0 - 1 : r0);
	Row cur = ld(r0);
	for (int64_t r = r0; r < r1; ++r) {
		const Row dn = ld(r + 1 < rows ? r + 1 : r);
		double wv = dlane64<0x138>(cur.v.y), ev = dlane64<0x130>(cur.v.x);
		if (lane == 0) wv = cur.w;
		if (lane == WAVE - 1) ev = cur.e;
		if (ok) {
			const bool rin = r > 0 && r < rows - 1;
			e.pair(r * cols + c0, rin && intA, chain5(w, cur.v.x, up.v.x, dn.v.x, wv, cur.v.y), rin && intB,
			       chain5(w, cur.v.y, up.v.y, dn.v.y, cur.v.x, ev), cur.v);
		}
		up = cur;
		cur = dn;
	}
}

// A point's five-point neighbourhood: centre, north (row above), south, west, east.
struct Nb {
	double c, n, s, w, e;
};
__device__ __forceinline__ double hmean2(double a, double b) { return 2.0 * a * b / (a + b); }
// −∇·(k∇u) at a point with harmonic face means, as ops-level kernels form it:
// L = −(fx·ix2 + fy·iy2) with fx = k_e·(u_e − u) − k_w·(u − u_w), fy likewise;
// *fsum gets (k_w + k_e)·ix2 + (k_n + k_s)·iy2 (the diagonal before scale and sigma).
__device__ __forceinline__ double vk_lap(const Nb& u, const Nb& k, double ix2, double iy2, double* fsum) {
	const double kw = hmean2(k.c, k.w), ke = hmean2(k.c, k.e);
	const double kn = hmean2(k.c, k.n), ks = hmean2(k.c, k.s);
	const double fx = ke * (u.e - u.c) - kw * (u.c - u.w);
	const double fy = ks * (u.s - u.c) - kn * (u.c - u.n);
	*fsum = (kw + ke) * ix2 + (kn + ks) * iy2;
	return -(fx * ix2 + fy * iy2);
}

// Two-column streaming traversal for FP64 grids with an even column count,
// the shape of k_stream5d carrying an optional coefficient field k (VK) beside
// the values: per row one 16-byte load of the value pair (V::pair) and, with
// VK, of the k pair; west and east neighbours by DPP from the adjacent lanes
// (the first and last lane of a wave read the column beyond the wave).  The
// epilogue receives both points' neighbourhoods (Nb of values and of k) with
// their row and column flags and does all stores, 16 bytes at a time:
//   V::prep()                    once, before any load (reads device scalars)
//   V::pair(i) -> double2        values at i, i + 1 (i even)
//   V::one(i)  -> double         value at i
//   E::pair(i, r, c0, rin, ia, ib, uA, uB, kA, kB, acc)
// Rows past the grid are read from the last row (clamped) and never used.
// Every lane walks the rows, so the cross-lane moves always see all 64 lanes.
// With NS > 0 the epilogue's sums go through block_partial.
template <int NS, bool VK, class V, class E>
__global__ __launch_bounds__(BLOCK) void k_pair5(V v, const double* __restrict__ k, E e, int64_t rows, int64_t cols,
                                                 int rpb, RedScratch rs, const int* skip) {
	if (skip && *skip) return;
	v.prep();
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
		const int64_t i = (r < rows ? r : rows - 1) * cols + cl;
		Row o;
		o.x = v.pair(i);
		o.xw = edgeW ? v.one(i - 1) : 0.0;
		o.xe = edgeE ? v.one(i + 2) : 0.0;
		if constexpr (VK) {
			o.k = *(const double2*)(k + i);
			o.kw = edgeW ? k[i - 1] : 1.0;
			o.ke = edgeE ? k[i + 2] : 1.0;
		} else {
			o.k = double2{1.0, 1.0};
			o.kw = o.ke = 1.0;
		}
		return o;
	};
	Acc acc[NS > 0 ? NS : 1];
#pragma unroll
	for (int q = 0; q < (NS > 0 ? NS : 1); ++q) acc[q] = Acc{0.0, 0.0};
	Row up = ld(r0 > 0 ? r0 - 1 : r0);
	Row cur = ld(r0);
	for (int64_t r = r0; r < r1; ++r) {
		const Row dn = ld(r + 1);
		double xw = dlane64<0x138>(cur.x.y), xe = dlane64<0x130>(cur.x.x);
		double kw = 1.0, ke = 1.0;
		if constexpr (VK) {
			kw = dlane64<0x138>(cur.k.y);
			ke = dlane64<0x130>(cur.k.x);
		}
		if (lane == 0) xw = cur.xw, kw = cur.kw;
		if (lane == WAVE - 1) xe = cur.xe, ke = cur.ke;
		if (ok) {
			const Nb uA{cur.x.x, up.x.x, dn.x.x, xw, cur.x.y}, uB{cur.x.y, up.x.y, dn.x.y, cur.x.x, xe};
			const Nb kA{cur.k.x, up.k.x, dn.k.x, kw, cur.k.y}, kB{cur.k.y, up.k.y, dn.k.y, cur.k.x, ke};
			e.pair(r * cols + c0, r, c0, r > 0 && r < rows - 1, intA, intB, uA, uB, kA, kB, acc);
		}
		up = cur;
		cur = dn;
	}
	if constexpr (NS > 0) block_partial<NS>(acc, rs);
}
// Value sources for k_pair5.
struct PLoad {
	const double* x;
	__device__ void prep() {}
	__device__ double2 pair(int64_t i) const { return *(const double2*)(x + i); }
	__device__ double one(int64_t i) const { return x[i]; }
};

// Rows per block for the streaming kernels: enough blocks to fill 256 CUs
// several times over, long enough columns to amortise the two halo rows.
inline int stream_rpb(int64_t rows, int64_t cols) {
	int64_t cb = (cols + BLOCK - 1) / BLOCK;
	int rpb = 32;
	while (rpb > 4 && cb * ((rows + rpb - 1) / rpb) < 2048) rpb /= 2;
	return rpb;
}

}  // namespace fxd
