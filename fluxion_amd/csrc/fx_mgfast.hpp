// fx_mgfast.hpp — fused FP32 V-cycle kernels (no reference counterpart).
//
// The V-cycle of multigrid.rs:188-230 with two-buffer Jacobi, linear same-domain
// transfers and a zero initial guess, regrouped into three kernels per level:
//   k_fpre           n_pre (1 or 2) sweeps from u = 0: reads b, writes u
//   k_fres_restrict  coarse rhs = R(b − A·u): reads u, b, writes the coarse rhs
//   k_fpost          u += P·e_c then n_post (1 or 2) sweeps in an LDS tile:
//                    reads u, b, e_c, writes the level's result
// and one single-workgroup kernel for the coarsest level's sweeps.  Each point
// is computed with exactly the operations of the unfused kernels (fx_mg.hip),
// so the fused cycle is bit-identical to the oracle's FP32 V-cycle.
#pragma once
#include "fx_mg.hpp"

namespace {
namespace fmg {

using namespace fxd;

struct LWf {
	float w[5];
	float inv_diag;
	float omega;
	// Variable-coefficient level operator when k != nullptr (fx_mg.hip vk_apply).
	const float* k;
	float ix2, iy2, sigma, scale;
};

__device__ __forceinline__ float hmf(float a, float b) { return 2.0f * a * b / (a + b); }
// sigma·u + scale·(−∇·(k∇u)) from the five u and k values of a point; *diag gets the diagonal.
__device__ __forceinline__ float vkf(const LWf& L, float kc, float kn, float ks, float kw, float ke, float c, float n,
                                     float s, float w, float e, float* diag) {
	float hw = hmf(kc, kw), he = hmf(kc, ke), hn = hmf(kc, kn), hs = hmf(kc, ks);
	float fx = he * (e - c) - hw * (c - w);
	float fy = hs * (s - c) - hn * (c - n);
	float Lv = -(fx * L.ix2 + fy * L.iy2);
	*diag = fmaf(L.scale, (hw + he) * L.ix2 + (hn + hs) * L.iy2, L.sigma);
	return fmaf(L.scale, Lv, L.sigma * c);
}
// Variable-coefficient Jacobi point update (fx_mg.hip k_jac with k).
__device__ __forceinline__ float jack(const LWf& L, float kc, float kn, float ks, float kw, float ke, float c, float n,
                                      float s, float w, float e, float b) {
	float d;
	float au = vkf(L, kc, kn, ks, kw, ke, c, n, s, w, e, &d);
	return fmaf(L.omega, (b - au) / d, c);
}

struct Lin1 {
	int64_t i0;
	double w;
};
__device__ __forceinline__ Lin1 lin1(int64_t i, int64_t nf, int64_t nc) {
	int64_t num = i * (nc - 1);
	int64_t den = nf - 1;
	return {num / den, double(num % den) * (1.0 / double(den))};
}
__device__ __forceinline__ double lin_weight(int64_t i, int64_t I, int64_t nf, int64_t nc) {
	Lin1 l = lin1(i, nf, nc);
	if (l.i0 == I) return 1.0 - l.w;
	if (l.i0 + 1 == I) return l.w;
	return 0.0;
}
__device__ __forceinline__ void lin_range(int64_t I, int64_t nf, int64_t nc, int64_t* lo, int64_t* hi) {
	int64_t den = nc - 1;
	int64_t a = (I - 1) * (nf - 1);
	*lo = a <= 0 ? 1 : (a + den - 1) / den;
	*hi = ((I + 1) * (nf - 1)) / den;
	if (*lo < 1) *lo = 1;
	if (*hi > nf - 2) *hi = nf - 2;
}

// Jacobi point update (fx_mg.hip k_jac): ax over (0,0),(−1,0),(1,0),(0,−1),(0,1).
__device__ __forceinline__ float jac(const LWf& L, float c, float n, float s, float w, float e, float b) {
	float ax = 0.0f;
	ax = fmaf(L.w[0], c, ax);
	ax = fmaf(L.w[1], n, ax);
	ax = fmaf(L.w[2], s, ax);
	ax = fmaf(L.w[3], w, ax);
	ax = fmaf(L.w[4], e, ax);
	float res = b - ax;
	return fmaf(L.omega, L.inv_diag * res, c);
}
// First sweep from u = 0 (interior); boundary values stay 0.
__device__ __forceinline__ float jac0(const LWf& L, float b) {
	float ax = 0.0f;
	ax = fmaf(L.w[0], 0.0f, ax);
	ax = fmaf(L.w[1], 0.0f, ax);
	ax = fmaf(L.w[2], 0.0f, ax);
	ax = fmaf(L.w[3], 0.0f, ax);
	ax = fmaf(L.w[4], 0.0f, ax);
	float res = b - ax;
	return fmaf(L.omega, L.inv_diag * res, 0.0f);
}

#define FMG_PT                                                              \
	int64_t c = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));   \
	int64_t r = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);       \
	if (r >= rows || c >= cols) return;                                     \
	int64_t i = r * cols + c;

// n_pre sweeps (1 or 2) from a zero initial guess.
__global__ __launch_bounds__(BLOCK) void k_fpre(const float* __restrict__ b, float* __restrict__ u, LWf L, int npre,
                                                int64_t rows, int64_t cols, const int* skip) {
	if (skip && *skip) return;
	FMG_PT;
	if (r == 0 || c == 0 || r == rows - 1 || c == cols - 1) { u[i] = 0.0f; return; }
	if (L.k) {
		const float* k = L.k;
		auto first = [&](int64_t rr, int64_t cc, int64_t j) -> float {
			if (rr == 0 || cc == 0 || rr == rows - 1 || cc == cols - 1) return 0.0f;
			return jack(L, k[j], k[j - cols], k[j + cols], k[j - 1], k[j + 1], 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, b[j]);
		};
		float uc = first(r, c, i);
		if (npre == 1) { u[i] = uc; return; }
		u[i] = jack(L, k[i], k[i - cols], k[i + cols], k[i - 1], k[i + 1], uc, first(r - 1, c, i - cols),
		            first(r + 1, c, i + cols), first(r, c - 1, i - 1), first(r, c + 1, i + 1), b[i]);
		return;
	}
	if (npre == 1) { u[i] = jac0(L, b[i]); return; }
	auto u1 = [&](int64_t rr, int64_t cc, int64_t j) -> float {
		return (rr == 0 || cc == 0 || rr == rows - 1 || cc == cols - 1) ? 0.0f : jac0(L, b[j]);
	};
	float bc = b[i];
	float uc = jac0(L, bc);
	u[i] = jac(L, uc, u1(r - 1, c, i - cols), u1(r + 1, c, i + cols), u1(r, c - 1, i - 1), u1(r, c + 1, i + 1), bc);
}

// Coarse rhs = restriction of the fine residual rhs − S(u) (fx_mg.hip k_lstencil mode 1 + k_restrict lin).
__global__ __launch_bounds__(BLOCK) void k_fres_restrict(const float* __restrict__ u, const float* __restrict__ b,
                                                         float* __restrict__ bc_out, LWf L, int64_t fr, int64_t fc,
                                                         int64_t cr, int64_t cc, const int* skip) {
	if (skip && *skip) return;
	int64_t col = (int64_t)blockIdx.x * WAVE + (threadIdx.x & (WAVE - 1));
	int64_t R = (int64_t)blockIdx.y * NWAVES + (threadIdx.x / WAVE);
	if (R >= cr || col >= cc) return;
	if (R == 0 || col == 0 || R == cr - 1 || col == cc - 1) { bc_out[R * cc + col] = 0.0f; return; }
	const float sy = float(double(cr - 1) / double(fr - 1)), sx = float(double(cc - 1) / double(fc - 1));
	int64_t rlo, rhi, clo, chi;
	lin_range(R, fr, cr, &rlo, &rhi);
	lin_range(col, fc, cc, &clo, &chi);
	float acc = 0.0f;
	for (int64_t ii = rlo; ii <= rhi; ++ii) {
		float wy = float(lin_weight(ii, R, fr, cr));
		float row = 0.0f;
		for (int64_t jj = clo; jj <= chi; ++jj) {
			int64_t q = ii * fc + jj;
			float s;
			if (L.k) {
				float d;
				const float* k = L.k;
				s = vkf(L, k[q], k[q - fc], k[q + fc], k[q - 1], k[q + 1], u[q], u[q - fc], u[q + fc], u[q - 1],
				        u[q + 1], &d);
			} else {
				s = fmaf(L.w[0], u[q], fmaf(L.w[1], u[q - fc], fmaf(L.w[2], u[q + fc], fmaf(L.w[3], u[q - 1],
				    L.w[4] * u[q + 1]))));
			}
			float res = fmaf(1.0f, b[q], -s);
			row += float(lin_weight(jj, col, fc, cc)) * res;
		}
		acc += wy * row;
	}
	bc_out[R * cc + col] = acc * (sy * sx);
}

// Coarsest level: `sweeps` Jacobi sweeps from zero in one workgroup (n ≤ 4096 points).
__global__ __launch_bounds__(1024) void k_fcoarse(const float* __restrict__ b, float* __restrict__ out, LWf L, int sweeps,
                                                  int rows, int cols, const int* skip) {
	if (skip && *skip) return;
	__shared__ float sa[4096], sb[4096];
	const int n = rows * cols;
	for (int q = threadIdx.x; q < n; q += blockDim.x) sa[q] = 0.0f;
	__syncthreads();
	float* cur = sa;
	float* nxt = sb;
	for (int it = 0; it < sweeps; ++it) {
		for (int q = threadIdx.x; q < n; q += blockDim.x) {
			int rr = q / cols, cc = q % cols;
			if (rr == 0 || cc == 0 || rr == rows - 1 || cc == cols - 1) { nxt[q] = cur[q]; continue; }
			if (L.k) {
				const float* k = L.k;
				nxt[q] = jack(L, k[q], k[q - cols], k[q + cols], k[q - 1], k[q + 1], cur[q], cur[q - cols],
				              cur[q + cols], cur[q - 1], cur[q + 1], b[q]);
			} else {
				nxt[q] = jac(L, cur[q], cur[q - cols], cur[q + cols], cur[q - 1], cur[q + 1], b[q]);
			}
		}
		__syncthreads();
		float* t = cur; cur = nxt; nxt = t;
	}
	for (int q = threadIdx.x; q < n; q += blockDim.x) out[q] = cur[q];
}

// Post-smoothing tile: 64 columns × 16 rows per 256-thread workgroup.
constexpr int PTX = 64, PTY = 16;

struct NoEpi {
	static constexpr int NS = 0;
	__device__ void term(int64_t, float, Acc*) const {}
	template <class A> __device__ void fin(A*) const {}
};

// u_fin = Jacobi^npost(u + P·e) on the tile.  Epi adds per-point reduction
// terms for the level-0 result and finishes them in the last workgroup.
template <class Epi>
__global__ __launch_bounds__(BLOCK) void k_fpost(const float* __restrict__ u, const float* __restrict__ b,
                                                 const float* __restrict__ ec, float* __restrict__ out, LWf L, int npost,
                                                 int64_t rows, int64_t cols, int64_t cr, int64_t cc, Epi epi,
                                                 RedScratch rs, const int* skip) {
	if (skip && *skip) return;
	constexpr int H = 2;
	constexpr int RX = PTX + 2 * H, RY = PTY + 2 * H;
	__shared__ float s3[RY][RX];
	__shared__ float s4[RY][RX];
	__shared__ float sbv[RY][RX];
	__shared__ float skv[RY][RX];
	__shared__ int64_t ry0[RY], ry1[RY], rx0[RX], rx1[RX];
	__shared__ float rwy[RY], rwx[RX];
	const int64_t r0 = (int64_t)blockIdx.y * PTY - H, c0 = (int64_t)blockIdx.x * PTX - H;
	const int t = threadIdx.x;
	// Interpolation tables for the region's rows and columns.
	for (int q = t; q < RY; q += BLOCK) {
		int64_t rr = r0 + q;
		if (rr > 0 && rr < rows - 1) {
			Lin1 l = lin1(rr, rows, cr);
			ry0[q] = l.i0;
			ry1[q] = l.i0 + 1 < cr ? l.i0 + 1 : l.i0;
			rwy[q] = float(l.w);
		}
	}
	for (int q = t; q < RX; q += BLOCK) {
		int64_t cl = c0 + q;
		if (cl > 0 && cl < cols - 1) {
			Lin1 l = lin1(cl, cols, cc);
			rx0[q] = l.i0;
			rx1[q] = l.i0 + 1 < cc ? l.i0 + 1 : l.i0;
			rwx[q] = float(l.w);
		}
	}
	__syncthreads();
	// u3 = u + P·e (fx_mg.hip k_prolong lin), b into LDS.
	for (int q = t; q < RX * RY; q += BLOCK) {
		int qy = q / RX, qx = q % RX;
		int64_t rr = r0 + qy, cl = c0 + qx;
		float v = 0.0f, bv = 0.0f;
		if (rr >= 0 && rr < rows && cl >= 0 && cl < cols) {
			int64_t j = rr * cols + cl;
			v = u[j];
			bv = b[j];
			if (rr > 0 && rr < rows - 1 && cl > 0 && cl < cols - 1) {
				float wy = rwy[qy], vy = 1.0f - wy, wx = rwx[qx], vx = 1.0f - wx;
				int64_t a0 = ry0[qy], a1 = ry1[qy], b0 = rx0[qx], b1 = rx1[qx];
				float top = vx * ec[a0 * cc + b0] + wx * ec[a0 * cc + b1];
				float bot = vx * ec[a1 * cc + b0] + wx * ec[a1 * cc + b1];
				v += vy * top + wy * bot;
			}
		}
		s3[qy][qx] = v;
		sbv[qy][qx] = bv;
		if (L.k) skv[qy][qx] = (rr >= 0 && rr < rows && cl >= 0 && cl < cols) ? L.k[rr * cols + cl] : 1.0f;
	}
	__syncthreads();
	float (*last)[RX] = s3;
	if (npost == 2) {
		// First sweep on the region minus its outer ring.
		for (int q = t; q < (RX - 2) * (RY - 2); q += BLOCK) {
			int qy = 1 + q / (RX - 2), qx = 1 + q % (RX - 2);
			int64_t rr = r0 + qy, cl = c0 + qx;
			float v = s3[qy][qx];
			if (rr > 0 && rr < rows - 1 && cl > 0 && cl < cols - 1) {
				if (L.k)
					v = jack(L, skv[qy][qx], skv[qy - 1][qx], skv[qy + 1][qx], skv[qy][qx - 1], skv[qy][qx + 1], v,
					         s3[qy - 1][qx], s3[qy + 1][qx], s3[qy][qx - 1], s3[qy][qx + 1], sbv[qy][qx]);
				else
					v = jac(L, v, s3[qy - 1][qx], s3[qy + 1][qx], s3[qy][qx - 1], s3[qy][qx + 1], sbv[qy][qx]);
			}
			s4[qy][qx] = v;
		}
		__syncthreads();
		last = s4;
	}
	// Final sweep on the tile.
	Acc acc[Epi::NS > 0 ? Epi::NS : 1];
	for (int k = 0; k < (Epi::NS > 0 ? Epi::NS : 1); ++k) acc[k] = Acc{0.0, 0.0};
	for (int q = t; q < PTX * PTY; q += BLOCK) {
		int qy = H + q / PTX, qx = H + q % PTX;
		int64_t rr = r0 + qy, cl = c0 + qx;
		if (rr >= rows || cl >= cols) continue;
		float v = last[qy][qx];
		if (rr > 0 && rr < rows - 1 && cl > 0 && cl < cols - 1) {
			if (L.k)
				v = jack(L, skv[qy][qx], skv[qy - 1][qx], skv[qy + 1][qx], skv[qy][qx - 1], skv[qy][qx + 1], v,
				         last[qy - 1][qx], last[qy + 1][qx], last[qy][qx - 1], last[qy][qx + 1], sbv[qy][qx]);
			else
				v = jac(L, v, last[qy - 1][qx], last[qy + 1][qx], last[qy][qx - 1], last[qy][qx + 1], sbv[qy][qx]);
		}
		int64_t j = rr * cols + cl;
		out[j] = v;
		epi.term(j, v, acc);
	}
	if constexpr (Epi::NS > 0) {
		Acc res[Epi::NS];
		if (block_partial_and_last<Epi::NS>(acc, rs, res) && tid_linear() == 0) epi.fin(res);
	}
}

inline dim3 pgrid(int64_t rows, int64_t cols) {
	return dim3((unsigned)((cols + WAVE - 1) / WAVE), (unsigned)((rows + NWAVES - 1) / NWAVES));
}
inline dim3 tgrid(int64_t rows, int64_t cols) {
	return dim3((unsigned)((cols + PTX - 1) / PTX), (unsigned)((rows + PTY - 1) / PTY));
}

inline LWf lwf(const fx_multigrid* mg, const fx_multigrid::Level& L) {
	LWf o;
	for (int e = 0; e < 5; ++e) o.w[e] = float(L.w[e]);
	o.inv_diag = float(L.inv_diag);
	o.omega = float(mg->opt.omega);
	o.k = (const float*)L.k;
	o.ix2 = float(1.0 / (L.g.dx * L.g.dx));
	o.iy2 = float(1.0 / (L.g.dy * L.g.dy));
	o.sigma = float(mg->opt.sigma);
	o.scale = float(mg->opt.scale);
	return o;
}

// Whether the fused cycle can run this hierarchy.
inline bool fused_ok(const fx_multigrid* mg) {
	const auto& o = mg->opt;
	if (!mg->f32 || o.smoother != FX_SMOOTH_JACOBI || o.transfer != FX_TRANSFER_FW_BILINEAR) return false;
	if (o.n_pre < 1 || o.n_pre > 2 || o.n_post < 1 || o.n_post > 2) return false;
	if (mg->lv.size() < 2) return false;
	const fx_grid& g = mg->lv.back().g;
	return g.rows * g.cols <= 4096;
}

// One fused V-cycle with a zero initial guess.  The level-0 rhs must already be
// in lv[0].rhs (FP32); the result is written to lv[0].tmp, with epi0's
// per-point terms reduced over level 0 by the last workgroup of its kernel.
template <class Epi>
fx_status fast_cycle(fx_multigrid* mg, const int* skip, Epi epi0) {
	fx_backend* be = mg->be;
	hipStream_t st = be->stream;
	auto& lv = mg->lv;
	const int nl = (int)lv.size();
	const auto& o = mg->opt;
	for (int l = 0; l < nl - 1; ++l) {
		const fx_grid& g = lv[l].g;
		const fx_grid& cg = lv[l + 1].g;
		LWf w = lwf(mg, lv[l]);
		hipLaunchKernelGGL(k_fpre, pgrid(g.rows, g.cols), dim3(BLOCK), 0, st, (const float*)lv[l].rhs, (float*)lv[l].u, w,
		                   o.n_pre, (int64_t)g.rows, (int64_t)g.cols, skip);
		hipLaunchKernelGGL(k_fres_restrict, pgrid(cg.rows, cg.cols), dim3(BLOCK), 0, st, (const float*)lv[l].u,
		                   (const float*)lv[l].rhs, (float*)lv[l + 1].rhs, w, (int64_t)g.rows, (int64_t)g.cols,
		                   (int64_t)cg.rows, (int64_t)cg.cols, skip);
	}
	{
		const auto& L = lv[nl - 1];
		hipLaunchKernelGGL(k_fcoarse, dim3(1), dim3(1024), 0, st, (const float*)L.rhs, (float*)L.tmp, lwf(mg, L),
		                   o.n_pre + o.n_post + o.extra_coarse, (int)L.g.rows, (int)L.g.cols, skip);
	}
	for (int l = nl - 2; l >= 1; --l) {
		const fx_grid& g = lv[l].g;
		const fx_grid& cg = lv[l + 1].g;
		hipLaunchKernelGGL(k_fpost<NoEpi>, tgrid(g.rows, g.cols), dim3(BLOCK), 0, st, (const float*)lv[l].u,
		                   (const float*)lv[l].rhs, (const float*)lv[l + 1].tmp, (float*)lv[l].tmp, lwf(mg, lv[l]), o.n_post,
		                   (int64_t)g.rows, (int64_t)g.cols, (int64_t)cg.rows, (int64_t)cg.cols, NoEpi{},
		                   RedScratch{nullptr, nullptr}, skip);
	}
	const fx_grid& g = lv[0].g;
	const fx_grid& cg = lv[1].g;
	dim3 tg = tgrid(g.rows, g.cols);
	RedScratch rs{nullptr, nullptr};
	if (Epi::NS > 0) {
		FX_CHECK(fx::ensure_partials(be, (size_t)tg.x * tg.y * Epi::NS));
		rs = fx::red_scratch(be);
	}
	hipLaunchKernelGGL(k_fpost<Epi>, tg, dim3(BLOCK), 0, st, (const float*)lv[0].u, (const float*)lv[0].rhs,
	                   (const float*)lv[1].tmp, (float*)lv[0].tmp, lwf(mg, lv[0]), o.n_post, (int64_t)g.rows,
	                   (int64_t)g.cols, (int64_t)cg.rows, (int64_t)cg.cols, epi0, rs, skip);
	return fx::check_launch("fused V-cycle");
}

}  // namespace fmg
}  // namespace
