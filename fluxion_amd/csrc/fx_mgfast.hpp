// fx_mgfast.hpp — fused FP32 V-cycle kernels (no reference counterpart).
//
// The V-cycle of multigrid.rs:188-230 with two-buffer Jacobi, linear same-domain
// transfers and a zero initial guess, regrouped into two LDS-tiled kernels per
// level:
//   k_frestrict  coarse rhs = R(b − A·J^npre(0; b)): reads b (and k), writes the
//                coarse rhs.  Pre-smoothing, residual and restriction happen in
//                LDS; each fine residual is computed once per tile.
//   k_fpost      out = J^npost(J^npre(0; b) + P·e_c): reads b (and k) and the
//                coarse correction, writes the level's result.  The pre-smoothed
//                iterate is recomputed from b in LDS rather than stored, so a
//                level's iterate never travels through HBM.
// plus one single-workgroup kernel for the coarsest level's sweeps.  Each point
// is computed with exactly the operations of the unfused kernels (fx_mg.hip),
// so the fused cycle is bit-identical to the oracle's FP32 V-cycle.
#pragma once
#include <type_traits>

#include "fx_mg.hpp"

namespace {
namespace fmg {

using namespace fxd;

struct LWf {
	float w[5];
	float inv_diag;
	float omega;
	// Variable-coefficient level operator when k != nullptr (fx_mg.hip vk_apply);
	// kx / ky: its face coefficients and wd: ω/diagonal on interior points (fused hierarchies).
	const float* k;
	const float* kx;
	const float* ky;
	const float* wd;
	const float* kp;   // {kx, ky, wd} packed per point (the wave kernels' row loads)
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
// vkf from the four face coefficients hw = hm(kc, kw), he, hn, hs (same arithmetic).
__device__ __forceinline__ float vkff(const LWf& L, float hw, float he, float hn, float hs, float c, float n, float s,
                                      float w, float e, float* diag) {
	float fx = he * (e - c) - hw * (c - w);
	float fy = hs * (s - c) - hn * (c - n);
	float Lv = -(fx * L.ix2 + fy * L.iy2);
	*diag = fmaf(L.scale, (hw + he) * L.ix2 + (hn + hs) * L.iy2, L.sigma);
	return fmaf(L.scale, Lv, L.sigma * c);
}
// Variable-coefficient Jacobi point update u + (ω/d)·(b − A·u) (fx_mg.hip k_jac with k).
__device__ __forceinline__ float jackf(const LWf& L, float hw, float he, float hn, float hs, float c, float n, float s,
                                       float w, float e, float b) {
	float d;
	float au = vkff(L, hw, he, hn, hs, c, n, s, w, e, &d);
	return fmaf(L.omega / d, b - au, c);
}
// jackf with ω/d read from the level's wd plane (k_wdiag: the same ω/d).
__device__ __forceinline__ float jackw(const LWf& L, float hw, float he, float hn, float hs, float wd, float c, float n,
                                       float s, float w, float e, float b) {
	float d;
	float au = vkff(L, hw, he, hn, hs, c, n, s, w, e, &d);
	return fmaf(wd, b - au, c);
}
// Variable-coefficient Jacobi point update (fx_mg.hip k_jac with k).
__device__ __forceinline__ float jack(const LWf& L, float kc, float kn, float ks, float kw, float ke, float c, float n,
                                      float s, float w, float e, float b) {
	float d;
	float au = vkf(L, kc, kn, ks, kw, ke, c, n, s, w, e, &d);
	return fmaf(L.omega / d, b - au, c);
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
// First sweep from u = 0 (interior); boundary values stay 0.  With finite
// weights the operator of u = 0 sums to +0, so b − A·0 is b (a zero of either
// sign gives +0 after the final fma either way) and the sweep is exactly
// ω·(d⁻¹·b) + 0.
__device__ __forceinline__ float jac0(const LWf& L, float b) {
	return fmaf(L.omega, L.inv_diag * b, 0.0f);
}
// S(u) of the level operator (fx_mg.hip k_lstencil interior).
__device__ __forceinline__ float sop(const LWf& L, float c, float n, float s, float w, float e) {
	return fmaf(L.w[0], c, fmaf(L.w[1], n, fmaf(L.w[2], s, fmaf(L.w[3], w, L.w[4] * e))));
}
// Restriction weight of fine index i toward coarse index I (the transpose of
// the linear interpolation), as float(1 − w) or float(w).
__device__ __forceinline__ float xw(const int32_t* i0, const double* w, int64_t i, int64_t I) {
	const int64_t a = i0[i];
	if (a == I) return float(1.0 - w[i]);
	if (a + 1 == I) return float(w[i]);
	return 0.0f;
}

// ── k_frestrict ──────────────────────────────────────────────────────
// Coarse tile RCX × RCY (one coarse point per thread, RCX·RCY ≤ BLOCK); the
// fine window it gathers from (about 2·RCX + 7 ≤ 64 columns, one wave per row)
// is staged in LDS with a halo of 2 (pre-smoothing + stencil), together with
// the window's transfer weights.
constexpr int RCX = 28, RCY = 8;
static_assert(RCX * RCY <= BLOCK, "one coarse point per thread");

// Dynamic LDS bytes of k_frestrict for a level.
inline size_t frestrict_lds(const fx_xtab& X, bool vk) {
	size_t f = 3 * (size_t)X.ry * X.rx;                  // b, two work planes
	if (vk) f += (size_t)X.ry * (X.rx + 1) + (size_t)(X.ry + 1) * X.rx;  // west/east and north/south faces
	f += 2 * (size_t)(X.ry + X.rx);                      // window weights (1 − w, w)
	size_t i = (size_t)(X.ry + X.rx) + 2 * (RCY + RCX);  // window i0, tile lo/hi
	return f * sizeof(float) + i * sizeof(int32_t);
}

// One coarse tile (bx, by) of the residual + restriction (dynamic LDS frestrict_lds()).
template <bool VK, int NPRE>
__device__ __forceinline__ void frestrict_tile(int bx, int by, const float* __restrict__ b, float* __restrict__ bco,
                                               const LWf& L, const fx_xtab& X, int64_t fr, int64_t fc, int64_t cr,
                                               int64_t cc) {
	extern __shared__ float smem[];
	const int RY = X.ry, RX = X.rx, FX = RX + 1;
	float* sb = smem;
	float* sa = sb + RY * RX;
	float* su = sa + RY * RX;
	float* sfx = su + RY * RX;                   // VK: [RY][RX + 1] face left of region column x
	float* sfy = sfx + (VK ? RY * FX : 0);       // VK: [RY + 1][RX] face above region row y
	float* wlr = sfy + (VK ? (RY + 1) * RX : 0); // [RY] float(1 − w) of the window's fine rows
	float* whr = wlr + RY;                       // [RY] float(w)
	float* wlc = whr + RY;                       // [RX]
	float* whc = wlc + RX;
	int32_t* i0r = (int32_t*)(whc + RX);         // [RY]
	int32_t* i0c = i0r + RY;                     // [RX]
	int32_t* tlr = i0c + RX;                     // [RCY] lo/hi of the tile's coarse rows
	int32_t* thr = tlr + RCY;
	int32_t* tlc = thr + RCY;                    // [RCX]
	int32_t* thc = tlc + RCX;
	const int t = threadIdx.x;
	const int64_t I0 = (int64_t)by * RCY, J0 = (int64_t)bx * RCX;
	const int ly = t / RCX, lx = t % RCX;
	const int64_t I = I0 + ly, J = J0 + lx;
	const bool mine = t < RCX * RCY && I < cr && J < cc;
	const int64_t Ia = I0 > 1 ? I0 : 1, Ib = I0 + RCY - 1 < cr - 2 ? I0 + RCY - 1 : cr - 2;
	const int64_t Ja = J0 > 1 ? J0 : 1, Jb = J0 + RCX - 1 < cc - 2 ? J0 + RCX - 1 : cc - 2;
	int64_t F0 = 0, F1 = -1, G0 = 0, G1 = -1;
	if (Ia <= Ib && Ja <= Jb) {
		F0 = X.lor[Ia];
		F1 = X.hir[Ib];
		G0 = X.loc[Ja];
		G1 = X.hic[Jb];
	}
	if (F1 < F0 || G1 < G0) {  // only boundary coarse points in this tile
		if (mine) bco[I * cc + J] = 0.0f;
		return;
	}
	const int64_t R0 = F0 - 2, C0 = G0 - 2;  // region origin (fine indices)
	const int ny = (int)(F1 - F0) + 5, nx = (int)(G1 - G0) + 5;
	const int tx = t & (WAVE - 1), ty = t / WAVE;
	for (int qy = ty; qy < ny; qy += NWAVES) {
		const int64_t gr = R0 + qy;
		for (int qx = tx; qx < nx; qx += WAVE) {
			const int64_t gc = C0 + qx;
			const bool in = gr >= 0 && gr < fr && gc >= 0 && gc < fc;
			sb[qy * RX + qx] = in ? b[gr * fc + gc] : 0.0f;
		}
	}
	if constexpr (VK) {
		// sfx[y][x] = face between region columns x − 1 and x; sfy[y][x] = between rows y − 1 and y.
		for (int qy = ty; qy < ny; qy += NWAVES) {
			const int64_t gr = R0 + qy;
			for (int qx = tx; qx < nx + 1; qx += WAVE) {
				const int64_t gc = C0 + qx - 1;
				const bool in = gr >= 0 && gr < fr && gc >= 0 && gc < fc - 1;
				sfx[qy * FX + qx] = in ? L.kx[gr * fc + gc] : 1.0f;
			}
		}
		for (int qy = ty; qy < ny + 1; qy += NWAVES) {
			const int64_t gr = R0 + qy - 1;
			for (int qx = tx; qx < nx; qx += WAVE) {
				const int64_t gc = C0 + qx;
				const bool in = gr >= 0 && gr < fr - 1 && gc >= 0 && gc < fc;
				sfy[qy * RX + qx] = in ? L.ky[gr * fc + gc] : 1.0f;
			}
		}
	}
	// Transfer weights of the window's fine rows/columns and the tile's ranges.
	for (int q = t; q < ny - 4; q += BLOCK) {
		const int64_t f = F0 + q;
		const double w = X.wr[f];
		i0r[q] = X.i0r[f];
		wlr[q] = float(1.0 - w);
		whr[q] = float(w);
	}
	for (int q = t; q < nx - 4; q += BLOCK) {
		const int64_t f = G0 + q;
		const double w = X.wc[f];
		i0c[q] = X.i0c[f];
		wlc[q] = float(1.0 - w);
		whc[q] = float(w);
	}
	if (t < RCY) {
		const int64_t c = I0 + t;
		tlr[t] = c < cr ? X.lor[c] : 0;
		thr[t] = c < cr ? X.hir[c] : -1;
	} else if (t >= WAVE && t < WAVE + RCX) {
		const int64_t c = J0 + (t - WAVE);
		tlc[t - WAVE] = c < cc ? X.loc[c] : 0;
		thc[t - WAVE] = c < cc ? X.hic[c] : -1;
	}
	__syncthreads();
	// 1 ≤ row ≤ fr − 2 and 1 ≤ col ≤ fc − 2, as 32-bit unsigned range tests.
	const int iR0 = (int)R0 - 1, iC0 = (int)C0 - 1;
	const unsigned ur = (unsigned)(fr - 2), uc = (unsigned)(fc - 2);
	auto interior = [&](int qy, int qx) { return (unsigned)(iR0 + qy) < ur && (unsigned)(iC0 + qx) < uc; };
	// Face coefficients of (qy, qx) in the (west, east, north, south) order of vkff.
#define FMG_F4(y, x) sfx[(y) * FX + (x)], sfx[(y) * FX + (x) + 1], sfy[(y) * RX + (x)], sfy[((y) + 1) * RX + (x)]
	// Pre-smoothing sweep 1 from zero over the whole region.
	float* s1 = NPRE == 2 ? sa : su;
	for (int qy = ty; qy < ny; qy += NWAVES) {
		for (int qx = tx; qx < nx; qx += WAVE) {
			const int q = qy * RX + qx;
			float v = 0.0f;
			if (interior(qy, qx)) {
				if constexpr (VK) v = jackf(L, FMG_F4(qy, qx), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, sb[q]);
				else v = jac0(L, sb[q]);
			}
			s1[q] = v;
		}
	}
	__syncthreads();
	if constexpr (NPRE == 2) {
		for (int qy = 1 + ty; qy < ny - 1; qy += NWAVES) {
			for (int qx = 1 + tx; qx < nx - 1; qx += WAVE) {
				const int q = qy * RX + qx;
				float v = 0.0f;
				if (interior(qy, qx)) {
					if constexpr (VK) v = jackf(L, FMG_F4(qy, qx), sa[q], sa[q - RX], sa[q + RX], sa[q - 1], sa[q + 1], sb[q]);
					else v = jac(L, sa[q], sa[q - RX], sa[q + RX], sa[q - 1], sa[q + 1], sb[q]);
				}
				su[q] = v;
			}
		}
		__syncthreads();
	}
	// Residual b − S(u) on the fine window (all fine-interior points).
	for (int qy = 2 + ty; qy < ny - 2; qy += NWAVES) {
		for (int qx = 2 + tx; qx < nx - 2; qx += WAVE) {
			const int q = qy * RX + qx;
			float s;
			if constexpr (VK) {
				float d;
				s = vkff(L, FMG_F4(qy, qx), su[q], su[q - RX], su[q + RX], su[q - 1], su[q + 1], &d);
			} else {
				s = sop(L, su[q], su[q - RX], su[q + RX], su[q - 1], su[q + 1]);
			}
			sa[q] = fmaf(1.0f, sb[q], -s);
		}
	}
#undef FMG_F4
	__syncthreads();
	if (!mine) return;
	if (I == 0 || J == 0 || I == cr - 1 || J == cc - 1) {
		bco[I * cc + J] = 0.0f;
		return;
	}
	// Window-relative ranges: fine row ii ↔ window index ii − F0, LDS row ii − R0.
	const int rlo = tlr[ly] - (int)F0, rhi = thr[ly] - (int)F0;
	const int clo = tlc[lx] - (int)G0, chi = thc[lx] - (int)G0;
	const int iI = (int)I, iJ = (int)J;
	float acc = 0.0f;
	for (int ii = rlo; ii <= rhi; ++ii) {
		const int ar = i0r[ii];
		const float wy = ar == iI ? wlr[ii] : (ar + 1 == iI ? whr[ii] : 0.0f);
		const float* rr = sa + (ii + 2) * RX + 2;
		float row = 0.0f;
		for (int jj = clo; jj <= chi; ++jj) {
			const int ac = i0c[jj];
			const float wx = ac == iJ ? wlc[jj] : (ac + 1 == iJ ? whc[jj] : 0.0f);
			row += wx * rr[jj];
		}
		acc += wy * row;
	}
	bco[I * cc + J] = acc * (X.sy * X.sx);
}

template <bool VK, int NPRE>
__global__ __launch_bounds__(BLOCK) void k_frestrict(const float* __restrict__ b, float* __restrict__ bco, LWf L,
                                                     fx_xtab X, int64_t fr, int64_t fc, int64_t cr, int64_t cc,
                                                     const int* skip) {
	if (skip && *skip) return;
	frestrict_tile<VK, NPRE>(blockIdx.x, blockIdx.y, b, bco, L, X, fr, fc, cr, cc);
}

// Coarsest level: `sweeps` Jacobi sweeps from zero in one workgroup (n ≤ 4096 points).
__device__ __forceinline__ void fcoarse_body(const float* __restrict__ b, float* __restrict__ out, const LWf& L,
                                             int sweeps, int rows, int cols) {
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
__global__ __launch_bounds__(1024) void k_fcoarse(const float* __restrict__ b, float* __restrict__ out, LWf L, int sweeps,
                                                  int rows, int cols, const int* skip) {
	if (skip && *skip) return;
	fcoarse_body(b, out, L, sweeps, rows, cols);
}

// ── k_fpost ──────────────────────────────────────────────────────────
// Tile PTX × PTY of fine points per 256-thread workgroup; b (and k) are staged
// with a halo of PH = 3 ≥ npre + npost − 1 (k one ring more).
constexpr int PTX = 64, PTY = 16, PH = 3;
constexpr int PRY = PTY + 2 * PH, PRX = PTX + 2 * PH;
static_assert(PTX == WAVE, "one tile column per lane in the final sweep");

// The fused residual update's memory traffic on scratch vectors, for timing it
// (the autotuner): r and A·p read, r_new stored, α = 0.
struct TuneUpd {
	static constexpr bool ON = true;
	const double *r, *ap;
	double* rn;
	struct Pf {
		double r, ap;
	};
	__device__ void begin() {}
	__device__ Pf load_at(size_t o, unsigned c) const { return Pf{(r + o)[c], (ap + o)[c]}; }
	__device__ double rnew(const Pf& v) const { return fma(-0.0, v.ap, v.r); }
	__device__ void store(size_t i, double v) const { rn[i] = v; }
};
// A solver epilogue's memory traffic and reductions on scratch vectors (r and
// A·p read, the rhs taken from r), for timing level 0's post kernel.
struct TuneEpi {
	static constexpr int NS = 2;
	static constexpr bool RHS_FROM_R = true;
	const double *r, *ap;
	struct Pf {
		double r, ap;
	};
	__device__ Pf load_at(size_t o, unsigned c) const { return Pf{(r + o)[c], (ap + o)[c]}; }
	__device__ void use_lane(const Pf& p, float z, double* s) const {
		s[0] += p.r * double(z);
		s[1] += double(z) * p.ap;
	}
	__device__ float rhs(const Pf& p) const { return float(p.r); }
};
// No residual update in the wave restriction (its rhs is read as it is).
struct NoUpd {
	static constexpr bool ON = false;
	struct Pf {};
	__device__ void begin() {}
	__device__ Pf load_at(size_t, unsigned) const { return Pf{}; }
};
struct NoEpi {
	static constexpr int NS = 0;
	static constexpr bool RHS_FROM_R = false;  // the wave kernel reads the level-0 rhs itself
	struct Pf {};
	__device__ void term(int64_t, float, Acc*) const {}
	__device__ Pf load(int64_t) const { return Pf{}; }
	__device__ void use(const Pf&, float, Acc*) const {}
	__device__ Pf load_at(size_t, unsigned) const { return Pf{}; }
	__device__ void use_lane(const Pf&, float, double*) const {}
};

// Calls f(y, x) for every point of the NY × NX block at (OY, OX) of an LDS region.
template <int OY, int OX, int NY, int NX, class F>
__device__ __forceinline__ void region(F&& f) {
	for (int q = threadIdx.x; q < NY * NX; q += BLOCK) f(OY + q / NX, OX + q % NX);
}

// out = J^NPOST(J^NPRE(0; b) + P·e) on tile (bx, by).  Epi adds per-point
// reduction terms of the level-0 result to acc.
template <bool VK, int NPRE, int NPOST, class Epi>
__device__ __forceinline__ void fpost_tile(int bx, int by, const float* __restrict__ b, const float* __restrict__ ec,
                                           float* __restrict__ out, const LWf& L, const fx_xtab& X, int64_t rows,
                                           int64_t cols, int64_t cr, int64_t cc, const Epi& epi, Acc* acc) {
	static_assert(NPRE + NPOST - 1 <= PH, "halo too small");
	// Offsets into the region: the pre-smoothed iterate is needed on the tile ±
	// NPOST (OU), its first sweep one ring wider (O1), b on the widest (OB).
	constexpr int OU = PH - NPOST, O1 = OU - 1, OB = NPRE == 2 ? O1 : OU;
	__shared__ float sb[PRY][PRX];
	__shared__ float sa[PRY][PRX];
	__shared__ float su[PRY][PRX];
	__shared__ float sfx[VK ? PRY : 1][VK ? PRX + 1 : 1];   // face left of region column x
	__shared__ float sfy[VK ? PRY + 1 : 1][VK ? PRX : 1];   // face above region row y
	__shared__ int32_t ry0[PRY], ry1[PRY], rx0[PRX], rx1[PRX];
	__shared__ float rwy[PRY], rwx[PRX];
	const int64_t r0 = (int64_t)by * PTY - PH, c0 = (int64_t)bx * PTX - PH;
	const int t = threadIdx.x;
	// Interpolation tables for the region's rows and columns.
	for (int q = t; q < PRY; q += BLOCK) {
		const int64_t rr = r0 + q;
		if (rr > 0 && rr < rows - 1) {
			const int32_t a = X.i0r[rr];
			ry0[q] = a;
			ry1[q] = a + 1 < cr ? a + 1 : a;
			rwy[q] = float(X.wr[rr]);
		}
	}
	for (int q = t; q < PRX; q += BLOCK) {
		const int64_t cl = c0 + q;
		if (cl > 0 && cl < cols - 1) {
			const int32_t a = X.i0c[cl];
			rx0[q] = a;
			rx1[q] = a + 1 < cc ? a + 1 : a;
			rwx[q] = float(X.wc[cl]);
		}
	}
	region<OB, OB, PRY - 2 * OB, PRX - 2 * OB>([&](int y, int x) {
		const int64_t rr = r0 + y, cl = c0 + x;
		const bool in = rr >= 0 && rr < rows && cl >= 0 && cl < cols;
		sb[y][x] = in ? b[rr * cols + cl] : 0.0f;
	});
	if constexpr (VK) {
		region<OB, OB, PRY - 2 * OB, PRX + 1 - 2 * OB>([&](int y, int x) {
			const int64_t rr = r0 + y, cl = c0 + x - 1;
			const bool in = rr >= 0 && rr < rows && cl >= 0 && cl < cols - 1;
			sfx[y][x] = in ? L.kx[rr * cols + cl] : 1.0f;
		});
		region<OB, OB, PRY + 1 - 2 * OB, PRX - 2 * OB>([&](int y, int x) {
			const int64_t rr = r0 + y - 1, cl = c0 + x;
			const bool in = rr >= 0 && rr < rows - 1 && cl >= 0 && cl < cols;
			sfy[y][x] = in ? L.ky[rr * cols + cl] : 1.0f;
		});
	}
	__syncthreads();
	// 1 ≤ row ≤ rows − 2 and 1 ≤ col ≤ cols − 2, as 32-bit unsigned range tests.
	const int ir0 = (int)r0 - 1, ic0 = (int)c0 - 1;
	const unsigned ur = (unsigned)(rows - 2), uc = (unsigned)(cols - 2);
	auto interior = [&](int qy, int qx) { return (unsigned)(ir0 + qy) < ur && (unsigned)(ic0 + qx) < uc; };
	// Face coefficients of (y, x) in the (west, east, north, south) order of vkff.
#define FMG_F4(y, x) sfx[y][x], sfx[y][(x) + 1], sfy[y][x], sfy[(y) + 1][x]
	if constexpr (NPRE == 2) {
		region<O1, O1, PRY - 2 * O1, PRX - 2 * O1>([&](int y, int x) {
			float v = 0.0f;
			if (interior(y, x)) {
				if constexpr (VK) v = jackf(L, FMG_F4(y, x), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, sb[y][x]);
				else v = jac0(L, sb[y][x]);
			}
			sa[y][x] = v;
		});
		__syncthreads();
	}
	// u = J^NPRE(0; b), then u += P·e (fx_mg.hip k_prolong lin) on interior points.
	region<OU, OU, PRY - 2 * OU, PRX - 2 * OU>([&](int y, int x) {
		float v = 0.0f;
		if (interior(y, x)) {
			if constexpr (NPRE == 2) {
				if constexpr (VK)
					v = jackf(L, FMG_F4(y, x), sa[y][x], sa[y - 1][x], sa[y + 1][x], sa[y][x - 1], sa[y][x + 1], sb[y][x]);
				else
					v = jac(L, sa[y][x], sa[y - 1][x], sa[y + 1][x], sa[y][x - 1], sa[y][x + 1], sb[y][x]);
			} else {
				if constexpr (VK) v = jackf(L, FMG_F4(y, x), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, sb[y][x]);
				else v = jac0(L, sb[y][x]);
			}
			const float wy = rwy[y], vy = 1.0f - wy, wx = rwx[x], vx = 1.0f - wx;
			const int64_t a0 = ry0[y], a1 = ry1[y], b0 = rx0[x], b1 = rx1[x];
			const float top = vx * ec[a0 * cc + b0] + wx * ec[a0 * cc + b1];
			const float bot = vx * ec[a1 * cc + b0] + wx * ec[a1 * cc + b1];
			v += vy * top + wy * bot;
		}
		su[y][x] = v;
	});
	__syncthreads();
	float (*last)[PRX] = su;
	if constexpr (NPOST == 2) {
		region<PH - 1, PH - 1, PTY + 2, PTX + 2>([&](int y, int x) {
			float v = su[y][x];
			if (interior(y, x)) {
				if constexpr (VK)
					v = jackf(L, FMG_F4(y, x), v, su[y - 1][x], su[y + 1][x], su[y][x - 1], su[y][x + 1], sb[y][x]);
				else
					v = jac(L, v, su[y - 1][x], su[y + 1][x], su[y][x - 1], su[y][x + 1], sb[y][x]);
			}
			sa[y][x] = v;
		});
		__syncthreads();
		last = sa;
	}
	// Final sweep on the tile.
	region<PH, PH, PTY, PTX>([&](int y, int x) {
		const int64_t rr = r0 + y, cl = c0 + x;
		if (rr >= rows || cl >= cols) return;
		float v = last[y][x];
		if (interior(y, x)) {
			if constexpr (VK)
				v = jackf(L, FMG_F4(y, x), v, last[y - 1][x], last[y + 1][x], last[y][x - 1], last[y][x + 1], sb[y][x]);
			else
				v = jac(L, v, last[y - 1][x], last[y + 1][x], last[y][x - 1], last[y][x + 1], sb[y][x]);
		}
		const int64_t j = rr * cols + cl;
		out[j] = v;
		epi.term(j, v, acc);
	});
#undef FMG_F4
}

// out = J^NPOST(J^NPRE(0; b) + P·e) on the tile.  Epi adds per-point reduction
// terms of the level-0 result (block partials; k_finish forms the sums).
template <bool VK, int NPRE, int NPOST, class Epi>
__global__ __launch_bounds__(BLOCK) void k_fpost(const float* __restrict__ b, const float* __restrict__ ec,
                                                 float* __restrict__ out, LWf L, fx_xtab X, int64_t rows, int64_t cols,
                                                 int64_t cr, int64_t cc, Epi epi, RedScratch rs, const int* skip) {
	if (skip && *skip) return;
	Acc acc[Epi::NS > 0 ? Epi::NS : 1];
	for (int k = 0; k < (Epi::NS > 0 ? Epi::NS : 1); ++k) acc[k] = Acc{0.0, 0.0};
	fpost_tile<VK, NPRE, NPOST, Epi>(blockIdx.x, blockIdx.y, b, ec, out, L, X, rows, cols, cr, cc, epi, acc);
	if constexpr (Epi::NS > 0) block_partial<Epi::NS>(acc, rs);
}

// ── k_spost: streaming post kernel ───────────────────────────────────
// The work of k_fpost reorganised for HBM streaming.  A workgroup owns a strip
// of SSW = 58 output columns (a region of SRW = 64 columns, one wave wide, with
// a halo of 3 each side) and a segment of SSH output rows, which it walks down
// in chunks of SCH rows.  Each stage (b, first pre-sweep, pre-smoothed iterate
// + P·e, first post-sweep, final sweep) runs a fixed number of rows ahead of
// the next and keeps its recent rows in an LDS ring, so no row is computed
// twice inside a segment; one extra chunk before the segment fills the rings.
// The global loads of the next chunk are issued before the current chunk's
// stages run.  Per-point arithmetic is that of k_fpost.
constexpr int SRW = 64, SHX = 3, SSW = SRW - 2 * SHX, SCH = 8, SRB = 16;
static_assert(SCH == 2 * NWAVES, "two rows per wave per chunk");

template <bool VK, int NPRE, int NPOST, class Epi>
__global__ __launch_bounds__(BLOCK) void k_spost(const float* __restrict__ b, const float* __restrict__ ec,
                                                 float* __restrict__ out, LWf L, fx_xtab X, int rows, int cols, int cr,
                                                 int cc, int nstrips, int sh, Epi epi, RedScratch rs, const int* skip) {
	static_assert(NPRE + NPOST - 1 <= SHX, "halo too small");
	if (skip && *skip) return;
	// Rows each stage runs ahead of the output rows of the chunk.
	constexpr int LP1 = 1, LU = NPOST, LU1 = NPOST + 1, LB = NPRE == 2 ? LU1 : LU;
	__shared__ float rb[SRB][SRW];
	__shared__ float r1[NPRE == 2 ? SRB : 1][SRW];
	__shared__ float ru[SRB][SRW];
	__shared__ float rp[NPOST == 2 ? SRB : 1][SRW];
	__shared__ float rfx[VK ? SRB : 1][VK ? SRW + 1 : 1];  // face west of region column x (x = 0..SRW)
	__shared__ float rfy[VK ? SRB : 1][VK ? SRW : 1];      // face south of the row
	const int lid = xcd_remap(blockIdx.x, gridDim.x);
	const int strip = lid % nstrips, seg = lid / nstrips;
	const int c0 = strip * SSW - SHX;
	const int S0 = seg * sh, S1 = S0 + sh < rows ? S0 + sh : rows;
	const int t = threadIdx.x, x = t & (WAVE - 1), w = t / WAVE;
	const int gc = c0 + x;
	const bool colin = gc >= 0 && gc < cols;
	const unsigned ucols = (unsigned)(cols - 2), urows = (unsigned)(rows - 2);
	const bool colint = (unsigned)(gc - 1) < ucols;
	auto slot = [](int r) { return (r + 8 * SRB) & (SRB - 1); };
	auto interior = [&](int r) { return colint && (unsigned)(r - 1) < urows; };
	// Column interpolation data of this lane (fixed for the workgroup).
	int b0 = 0, b1 = 0;
	float wx = 0.0f, vx = 0.0f;
	if (colint) {
		const int a = X.i0c[gc];
		b0 = a;
		b1 = a + 1 < cc ? a + 1 : a;
		wx = float(X.wc[gc]);
		vx = 1.0f - wx;
	}
	// Loads of row r into registers (b, the west face of this lane's column and, for lane 0,
	// the face east of the region, the south face) and their store into the rings.
	struct Row {
		float b, fx, fx2, fy;
	};
	auto load_row = [&](int r) {
		Row v{0.0f, 1.0f, 1.0f, 1.0f};
		if (r >= 0 && r < rows) {
			const size_t o = (size_t)r * cols;
			if (colin) v.b = b[o + gc];
			if constexpr (VK) {
				const int cw = gc - 1;
				if (cw >= 0 && cw < cols - 1) v.fx = L.kx[o + cw];
				if (x == 0) {
					const int ce = c0 + SRW - 1;
					if (ce >= 0 && ce < cols - 1) v.fx2 = L.kx[o + ce];
				}
				if (colin && r < rows - 1) v.fy = L.ky[o + gc];
			}
		}
		return v;
	};
	auto store_row = [&](int r, const Row& v) {
		const int s = slot(r);
		rb[s][x] = v.b;
		if constexpr (VK) {
			rfx[s][x] = v.fx;
			if (x == 0) rfx[s][SRW] = v.fx2;
			rfy[s][x] = v.fy;
		}
	};
	// Face coefficients of row r at lane column xx (west, east, north, south).
#define FMG_SF4(r, xx) rfx[slot(r)][xx], rfx[slot(r)][(xx) + 1], rfy[slot((r) - 1)][xx], rfy[slot(r)][xx]
	Acc acc[Epi::NS > 0 ? Epi::NS : 1];
	for (int k = 0; k < (Epi::NS > 0 ? Epi::NS : 1); ++k) acc[k] = Acc{0.0, 0.0};
	const int R0 = S0 - SCH;
	// The first chunk also needs the row above its loads (north faces, and the stages' upper halo).
	if (w == 0) store_row(R0 + LB - 1, load_row(R0 + LB - 1));
	Row nx0 = load_row(R0 + LB + w), nx1 = load_row(R0 + LB + w + NWAVES);
	for (int R = R0; R < S1; R += SCH) {
		__syncthreads();  // the previous chunk's stages are done with the ring rows reused below
		store_row(R + LB + w, nx0);
		store_row(R + LB + w + NWAVES, nx1);
		// This chunk's coarse-correction values (rows of the u stage) and epilogue operands,
		// issued before the next chunk's loads so that waiting on them never waits on those.
		float e00[2], e01[2], e10[2], e11[2];
		float wyv[2];
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const int r = R + LU + w + h * NWAVES;
			e00[h] = e01[h] = e10[h] = e11[h] = 0.0f;
			wyv[h] = 0.0f;
			if (interior(r)) {
				const int a0 = X.i0r[r], a1 = a0 + 1 < cr ? a0 + 1 : a0;
				wyv[h] = float(X.wr[r]);
				e00[h] = ec[(size_t)a0 * cc + b0];
				e01[h] = ec[(size_t)a0 * cc + b1];
				e10[h] = ec[(size_t)a1 * cc + b0];
				e11[h] = ec[(size_t)a1 * cc + b1];
			}
		}
		typename Epi::Pf pf[2];
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const int r = R + w + h * NWAVES;
			if (r >= S0 && r < S1 && x >= SHX && x < SRW - SHX && colin) pf[h] = epi.load((int64_t)r * cols + gc);
		}
		nx0 = load_row(R + SCH + LB + w);
		nx1 = load_row(R + SCH + LB + w + NWAVES);
		__syncthreads();
		if constexpr (NPRE == 2) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				const int r = R + LU1 + w + h * NWAVES;
				float v = 0.0f;
				if (interior(r)) {
					if constexpr (VK) v = jackf(L, FMG_SF4(r, x), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, rb[slot(r)][x]);
					else v = jac0(L, rb[slot(r)][x]);
				}
				r1[slot(r)][x] = v;
			}
			__syncthreads();
		}
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const int r = R + LU + w + h * NWAVES;
			float v = 0.0f;
			if (interior(r) && x > 0 && x < SRW - 1) {
				const int s = slot(r);
				if constexpr (NPRE == 2) {
					const int sn = slot(r - 1), ss = slot(r + 1);
					if constexpr (VK)
						v = jackf(L, FMG_SF4(r, x), r1[s][x], r1[sn][x], r1[ss][x], r1[s][x - 1], r1[s][x + 1], rb[s][x]);
					else
						v = jac(L, r1[s][x], r1[sn][x], r1[ss][x], r1[s][x - 1], r1[s][x + 1], rb[s][x]);
				} else {
					if constexpr (VK) v = jackf(L, FMG_SF4(r, x), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, rb[s][x]);
					else v = jac0(L, rb[s][x]);
				}
				const float wy = wyv[h], vy = 1.0f - wy;
				const float top = vx * e00[h] + wx * e01[h];
				const float bot = vx * e10[h] + wx * e11[h];
				v += vy * top + wy * bot;
			}
			ru[slot(r)][x] = v;
		}
		__syncthreads();
		float (*last)[SRW] = ru;
		if constexpr (NPOST == 2) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				const int r = R + LP1 + w + h * NWAVES;
				const int s = slot(r);
				float v = ru[s][x];
				if (interior(r) && x > 1 && x < SRW - 2) {
					const int sn = slot(r - 1), ss = slot(r + 1);
					if constexpr (VK)
						v = jackf(L, FMG_SF4(r, x), v, ru[sn][x], ru[ss][x], ru[s][x - 1], ru[s][x + 1], rb[s][x]);
					else
						v = jac(L, v, ru[sn][x], ru[ss][x], ru[s][x - 1], ru[s][x + 1], rb[s][x]);
				}
				rp[s][x] = v;
			}
			__syncthreads();
			last = rp;
		}
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const int r = R + w + h * NWAVES;
			if (r < S0 || r >= S1 || x < SHX || x >= SRW - SHX || !colin) continue;
			const int s = slot(r);
			float v = last[s][x];
			if (interior(r)) {
				const int sn = slot(r - 1), ss = slot(r + 1);
				if constexpr (VK)
					v = jackf(L, FMG_SF4(r, x), v, last[sn][x], last[ss][x], last[s][x - 1], last[s][x + 1], rb[s][x]);
				else
					v = jac(L, v, last[sn][x], last[ss][x], last[s][x - 1], last[s][x + 1], rb[s][x]);
			}
			const size_t j = (size_t)r * cols + gc;
			out[j] = v;
			epi.use(pf[h], v, acc);
		}
	}
#undef FMG_SF4
	if constexpr (Epi::NS > 0) block_partial<Epi::NS>(acc, rs);
}

inline unsigned spost_blocks(int64_t rows, int64_t cols, int sh) {
	return (unsigned)(((cols + SSW - 1) / SSW) * ((rows + sh - 1) / sh));
}
// Segment height (rows, a multiple of SCH) for a level: the tallest of 128, 64, 32 that still gives
// about 2048 workgroups; 0 when even 32 gives fewer than 1024 (the tiled kernel is used then).
inline int spost_height(int64_t rows, int64_t cols) {
	int sh = 128;
	while (sh > 32 && spost_blocks(rows, cols, sh) < 2048) sh /= 2;
	return spost_blocks(rows, cols, sh) >= 1024 ? sh : 0;
}

// ── k_srestrict: streaming residual + restriction ─────────────────────
// The work of k_frestrict reorganised like k_spost.  A workgroup owns RCX
// coarse columns (a fine window of at most SRW columns, one wave wide) and a
// segment of SCR coarse rows, and walks the fine rows down in chunks of SCH:
// b (and faces) → first pre-sweep → pre-smoothed iterate → residual, each stage
// a few rows ahead of the next in LDS rings; after each chunk every coarse row
// whose fine rows are all available is gathered.  Per-point arithmetic and the
// gather order are those of k_frestrict.

template <bool VK, int NPRE>
__global__ __launch_bounds__(BLOCK) void k_srestrict(const float* __restrict__ b, float* __restrict__ bco, LWf L,
                                                     fx_xtab X, int fr, int fc, int cr, int cc, int nstrips,
                                                     int scr, const int* skip) {
	if (skip && *skip) return;
	constexpr int LR = 0, LU = 1, LU1 = 2, LB = NPRE == 2 ? LU1 : LU;
	__shared__ float rb[SRB][SRW];
	__shared__ float r1[NPRE == 2 ? SRB : 1][SRW];
	__shared__ float ru[SRB][SRW];
	__shared__ float rr[SRB][SRW];
	__shared__ float rfx[VK ? SRB : 1][VK ? SRW + 1 : 1];
	__shared__ float rfy[VK ? SRB : 1][VK ? SRW : 1];
	__shared__ int32_t ci0[SRW];
	__shared__ float cwl[SRW], cwh[SRW];
	__shared__ int32_t clo[RCX], chi[RCX];
	__shared__ float rsr[SRB][RCX];   // per fine row: Σ_jj w(jj, J)·res(row, jj) for the strip's coarse columns
	__shared__ int32_t ri0[SRB];      // per fine row: coarse row below it and its two weights
	__shared__ float rwl[SRB], rwh[SRB];
	const int lid = xcd_remap(blockIdx.x, gridDim.x);
	const int strip = lid % nstrips, seg = lid / nstrips;
	const int J0 = strip * RCX, I0 = seg * scr;
	const int Jend = J0 + RCX < cc ? J0 + RCX : cc, Iend = I0 + scr < cr ? I0 + scr : cr;
	const int t = threadIdx.x, x = t & (WAVE - 1), w = t / WAVE;
	// Boundary coarse rows and columns are zero.
	for (int q = t; q < (Jend - J0); q += BLOCK) {
		if (I0 == 0) bco[(size_t)J0 + q] = 0.0f;
		if (Iend == cr) bco[(size_t)(cr - 1) * cc + J0 + q] = 0.0f;
	}
	const int Ia = I0 > 1 ? I0 : 1, Ib = Iend - 1 < cr - 2 ? Iend - 1 : cr - 2;
	const int Ja = J0 > 1 ? J0 : 1, Jb = Jend - 1 < cc - 2 ? Jend - 1 : cc - 2;
	if (Ia > Ib) return;
	const int G0 = Ja <= Jb ? X.loc[Ja] : 0, G1 = Ja <= Jb ? X.hic[Jb] : -1;
	const int F0 = X.lor[Ia], F1 = X.hir[Ib];
	if (G1 < G0 || F1 < F0) {  // nothing to gather: the strip's interior rows are zero
		const int nc = Jend - J0;
		for (int q = t; q < (Ib - Ia + 1) * nc; q += BLOCK) bco[(size_t)(Ia + q / nc) * cc + J0 + q % nc] = 0.0f;
		return;
	}
	const int C0 = G0 - 2;  // fine column of lane 0
	const int gc = C0 + x;
	const bool colin = gc >= 0 && gc < fc;
	const unsigned ucols = (unsigned)(fc - 2), urows = (unsigned)(fr - 2);
	const bool colint = (unsigned)(gc - 1) < ucols;
	auto slot = [](int r) { return (r + 8 * SRB) & (SRB - 1); };
	auto interior = [&](int r) { return colint && (unsigned)(r - 1) < urows; };
	// Column weights of the lane's fine column and the strip's coarse column ranges.
	if (colint) {
		const double wv = X.wc[gc];
		ci0[x] = X.i0c[gc];
		cwl[x] = float(1.0 - wv);
		cwh[x] = float(wv);
	} else if (w == 0) {
		ci0[x] = -2;
		cwl[x] = cwh[x] = 0.0f;
	}
	if (t >= WAVE && t < WAVE + RCX) {
		const int J = J0 + t - WAVE;
		clo[t - WAVE] = J < cc ? X.loc[J] : 0;
		chi[t - WAVE] = J < cc ? X.hic[J] : -1;
	}
	struct Row {
		float b, fx, fx2, fy;
	};
	auto load_row = [&](int r) {
		Row v{0.0f, 1.0f, 1.0f, 1.0f};
		if (r >= 0 && r < fr) {
			const size_t o = (size_t)r * fc;
			if (colin) v.b = b[o + gc];
			if constexpr (VK) {
				const int cw = gc - 1;
				if (cw >= 0 && cw < fc - 1) v.fx = L.kx[o + cw];
				if (x == 0) {
					const int ce = C0 + SRW - 1;
					if (ce >= 0 && ce < fc - 1) v.fx2 = L.kx[o + ce];
				}
				if (colin && r < fr - 1) v.fy = L.ky[o + gc];
			}
		}
		return v;
	};
	auto store_row = [&](int r, const Row& v) {
		const int s = slot(r);
		rb[s][x] = v.b;
		if constexpr (VK) {
			rfx[s][x] = v.fx;
			if (x == 0) rfx[s][SRW] = v.fx2;
			rfy[s][x] = v.fy;
		}
	};
#define FMG_SF4(r, xx) rfx[slot(r)][xx], rfx[slot(r)][(xx) + 1], rfy[slot((r) - 1)][xx], rfy[slot(r)][xx]
	const int R0 = F0 - SCH;
	if (w == 0) store_row(R0 + LB - 1, load_row(R0 + LB - 1));
	Row nx0 = load_row(R0 + LB + w), nx1 = load_row(R0 + LB + w + NWAVES);
	int Inext = Ia;  // next coarse row to gather
	const float scale = X.sy * X.sx;
	for (int R = R0; Inext <= Ib; R += SCH) {
		__syncthreads();
		store_row(R + LB + w, nx0);
		store_row(R + LB + w + NWAVES, nx1);
		nx0 = load_row(R + SCH + LB + w);
		nx1 = load_row(R + SCH + LB + w + NWAVES);
		__syncthreads();
		if constexpr (NPRE == 2) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				const int r = R + LU1 + w + h * NWAVES;
				float v = 0.0f;
				if (interior(r)) {
					if constexpr (VK) v = jackf(L, FMG_SF4(r, x), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, rb[slot(r)][x]);
					else v = jac0(L, rb[slot(r)][x]);
				}
				r1[slot(r)][x] = v;
			}
			__syncthreads();
		}
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const int r = R + LU + w + h * NWAVES;
			const int s = slot(r);
			float v = 0.0f;
			if constexpr (NPRE == 2) {
				if (interior(r) && x > 0 && x < SRW - 1) {
					const int sn = slot(r - 1), ss = slot(r + 1);
					if constexpr (VK)
						v = jackf(L, FMG_SF4(r, x), r1[s][x], r1[sn][x], r1[ss][x], r1[s][x - 1], r1[s][x + 1], rb[s][x]);
					else
						v = jac(L, r1[s][x], r1[sn][x], r1[ss][x], r1[s][x - 1], r1[s][x + 1], rb[s][x]);
				}
			} else {
				if (interior(r)) {
					if constexpr (VK) v = jackf(L, FMG_SF4(r, x), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, rb[s][x]);
					else v = jac0(L, rb[s][x]);
				}
			}
			ru[s][x] = v;
		}
		__syncthreads();
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const int r = R + LR + w + h * NWAVES;
			const int s = slot(r);
			float res = 0.0f;
			if (interior(r) && x > 0 && x < SRW - 1) {
				const int sn = slot(r - 1), ss = slot(r + 1);
				float sv;
				if constexpr (VK) {
					float d;
					sv = vkff(L, FMG_SF4(r, x), ru[s][x], ru[sn][x], ru[ss][x], ru[s][x - 1], ru[s][x + 1], &d);
				} else {
					sv = sop(L, ru[s][x], ru[sn][x], ru[ss][x], ru[s][x - 1], ru[s][x + 1]);
				}
				res = fmaf(1.0f, rb[s][x], -sv);
			}
			rr[s][x] = res;
		}
		__syncthreads();
		// Row pass: each fine row of the chunk against each coarse column of the strip (the inner sums of
		// the restriction, in the fine-column order of k_frestrict), plus the row's interpolation data.
		const int ncol = Jend - J0;
		if (t < SCH * RCX) {
			const int hr = t / RCX, jx = t % RCX;
			const int r = R + LR + hr, J = J0 + jx;
			const int s = slot(r);
			if (jx < ncol && J > 0 && J < cc - 1) {
				const int lo = clo[jx] - C0, hi = chi[jx] - C0;
				const float* row = rr[s];
				float rs = 0.0f;
				for (int jj = lo; jj <= hi; ++jj) {
					const int ac = ci0[jj];
					const float wxv = ac == J ? cwl[jj] : (ac + 1 == J ? cwh[jj] : 0.0f);
					rs += wxv * row[jj];
				}
				rsr[s][jx] = rs;
			}
			if (jx == 0) {
				const bool in = (unsigned)(r - 1) < urows;
				const double wr = in ? X.wr[r] : 0.0;
				ri0[s] = in ? X.i0r[r] : -2;
				rwl[s] = float(1.0 - wr);
				rwh[s] = float(wr);
			}
		}
		__syncthreads();
		// Column pass: coarse rows whose fine rows lo..hi have all been computed.
		const int last = R + SCH - 1;
		int Iready = Inext;
		while (Iready <= Ib && X.hir[Iready] <= last) ++Iready;
		const int nrow = Iready - Inext;
		for (int q = t; q < nrow * ncol; q += BLOCK) {
			const int I = Inext + q / ncol, jx = q % ncol, J = J0 + jx;
			float outv = 0.0f;
			if (J > 0 && J < cc - 1) {
				const int rlo = X.lor[I], rhi = X.hir[I];
				float acc = 0.0f;
				for (int ii = rlo; ii <= rhi; ++ii) {
					const int s = slot(ii);
					const int ar = ri0[s];
					const float wy = ar == I ? rwl[s] : (ar + 1 == I ? rwh[s] : 0.0f);
					acc += wy * rsr[s][jx];
				}
				outv = acc * scale;
			}
			bco[(size_t)I * cc + J] = outv;
		}
		Inext = Iready;
	}
#undef FMG_SF4
}

inline unsigned srestrict_blocks(int64_t cr, int64_t cc, int scr) {
	return (unsigned)(((cc + RCX - 1) / RCX) * ((cr + scr - 1) / scr));
}
// Coarse rows per segment (64, 32 or 16) for about 2048 workgroups; 0 when the level's fine window
// is wider than a wave or even 16 rows give fewer than 1024 workgroups (tiled kernel then).
inline int srestrict_height(const fx_xtab& X, int64_t cr, int64_t cc) {
	if (X.rx > SRW) return 0;
	int scr = 64;
	while (scr > 16 && srestrict_blocks(cr, cc, scr) < 2048) scr /= 2;
	return srestrict_blocks(cr, cc, scr) >= 1024 ? scr : 0;
}

inline dim3 rgrid(int64_t cr, int64_t cc) {
	return dim3((unsigned)((cc + RCX - 1) / RCX), (unsigned)((cr + RCY - 1) / RCY));
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
	o.kx = (const float*)L.kx;
	o.ky = (const float*)L.ky;
	o.wd = (const float*)L.wd;
	o.kp = (const float*)L.kp;
	o.ix2 = float(1.0 / (L.g.dx * L.g.dx));
	o.iy2 = float(1.0 / (L.g.dy * L.g.dy));
	o.sigma = float(mg->opt.sigma);
	o.scale = float(mg->opt.scale);
	return o;
}

// ── wave-streaming level kernels (k_wrestrict, k_wpost) ───────────────
// The work of k_frestrict / k_fpost for a whole level, with no LDS and no
// workgroup barrier.  Each wave is independent: it owns a strip of SRW = 64
// fine columns (one per lane) and a segment of rows, which it walks down one
// row per step.  Every stage (first pre-sweep, pre-smoothed iterate + P·e,
// post-sweeps, residual) runs a fixed number of rows behind the row being
// loaded and keeps its last three rows in registers; west and east neighbours
// come from the adjacent lanes through DPP wavefront shifts.  Rows are loaded
// several steps ahead from clamped (always in-bounds) addresses, so the loads are
// unconditional and the steps straight-line code.  Per-point arithmetic, and
// the restriction's summation order, are those of the unfused kernels.
constexpr int WNMAX = 5;  // most fine columns (or rows) a coarse point gathers from (2·(h/H)⁻¹ + 1 for H ≈ 2h)

// v of lane x − 1 (lane 0: 0) and of lane x + 1 (lane 63: 0), wave-wide.
__device__ __forceinline__ float lane_west(float v) {
	return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_east(float v) {
	return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// One row of a level's inputs at lane column gc: b and, for variable
// coefficients, the faces west (kx[gc − 1]), east (kx[gc]), north (ky one row
// up) and south (ky[gc]) of the point, and ω/d.  wload reads b and the packed
// {east face, south face, ω/d} unconditionally from clamped addresses (a
// uniform row offset plus the lane's 32-bit column): two loads per row.  wtake completes
// the row when it is consumed: the west face is the east face of lane x − 1
// and the north face the south face of the row above (the previous step's
// row, still in the ring).  Lane 0's west face (0) and the north face of a
// wave's first row (1) are therefore not the grid's, but every value that
// depends on them already depends on a west neighbour or a row above that
// the wave does not hold: halo lanes and lag rows, which the wave discards.
// wfix, applied when the row is consumed (not when it is loaded, so that the
// compiler cannot tie the load to a branch and wait on it at once), puts 0
// and 1 outside the grid, as the LDS kernels stage them.
struct WRow {
	float b, kw, ke, kn, ks, wd;
};
struct __attribute__((aligned(4))) KP3 {
	float kx, ky, wd;
};
template <bool VK, bool LOADB = true>
__device__ __forceinline__ WRow wload(const float* __restrict__ b, const LWf& L, int r, int rows, int cols,
                                      unsigned c) {
	const int rc = r < 0 ? 0 : (r < rows ? r : rows - 1);
	const size_t o = (size_t)rc * (size_t)cols;
	WRow v;
	v.b = LOADB ? (b + o)[c] : 0.0f;
	if constexpr (VK) {
		const KP3 q = ((const KP3*)(L.kp + 3 * o))[c];
		v.ke = q.kx;
		v.ks = q.ky;
		v.wd = q.wd;
		v.kw = v.kn = 1.0f;
	} else {
		v.kw = v.ke = v.kn = v.ks = 1.0f;
		v.wd = 0.0f;
	}
	return v;
}
template <bool VK>
__device__ __forceinline__ WRow wtake(WRow v, const WRow& above) {
	if constexpr (VK) {
		v.kw = lane_west(v.ke);
		v.kn = above.ks;
	}
	return v;
}
template <bool VK>
__device__ __forceinline__ WRow wfix(const WRow& v, int r, int rows, bool colin, bool kwok, bool keok) {
	const bool rin = r >= 0 && r < rows;
	WRow o;
	o.b = rin && colin ? v.b : 0.0f;
	if constexpr (VK) {
		o.kw = rin && kwok ? v.kw : 1.0f;
		o.ke = rin && keok ? v.ke : 1.0f;
		o.kn = colin && r >= 1 && r <= rows - 1 ? v.kn : 1.0f;
		o.ks = rin && colin && r < rows - 1 ? v.ks : 1.0f;
		o.wd = v.wd;  // used on interior points only
	} else {
		o.kw = o.ke = o.kn = o.ks = 1.0f;
		o.wd = 0.0f;
	}
	return o;
}

// Row rings.  A wave's steps are unrolled by the input ring's depth D (8), and
// row t of unrolled step k sits in slot (k − i) mod W of a W-slot ring when the
// current row is t + i, a compile-time index, so advancing a row moves no data.
// The inputs' ring has W = D slots: the stages read the last few rows in place
// and each slot is refilled, D − (rows kept) steps ahead, once its row is no
// longer read.  The stages' own rings have W = 4 slots (D is a multiple of 4).
constexpr int WD_POST = 8;      // input ring depth of k_wpost
constexpr int WD_RESTRICT = 8;  // and of k_wrestrict
constexpr int WR = 4;
template <int W, int K, int I>
struct Slot {
	static constexpr int v = ((K - I) % W + W) % W;
};
// Jacobi update / operator value of the point whose inputs are v, from its
// centre c, north n, south s, west w and east e values.
template <bool VK>
__device__ __forceinline__ float wjac(const LWf& L, const WRow& v, float c, float n, float s, float w, float e) {
	if constexpr (VK) return jackw(L, v.kw, v.ke, v.kn, v.ks, v.wd, c, n, s, w, e, v.b);
	else return jac(L, c, n, s, w, e, v.b);
}
// First sweep from u = 0: with finite faces the operator of u = 0 is a zero
// and (b − A·0) is b up to the sign of a zero, which the final fma with +0
// removes, so jackw(…, 0, 0, 0, 0, 0, b) is exactly (ω/d)·b + 0.
template <bool VK>
__device__ __forceinline__ float wjac0(const LWf& L, const WRow& v) {
	if constexpr (VK) return fmaf(v.wd, v.b, 0.0f);
	else return jac0(L, v.b);
}
template <bool VK>
__device__ __forceinline__ float wop(const LWf& L, const WRow& v, float c, float n, float s, float w, float e) {
	if constexpr (VK) {
		float d;
		return vkff(L, v.kw, v.ke, v.kn, v.ks, c, n, s, w, e, &d);
	} else {
		return sop(L, c, n, s, w, e);
	}
}

// lin1(r, nf, nc) of consecutive rows r (clamped to [0, nf − 1]): the coarse
// row i0 below fine row r and the remainder rem of r·(nc − 1) = i0·(nf − 1) + rem,
// advanced by exact integer steps, so w = double(rem)·(1/(nf − 1)) equals the
// transfer table's weight without a memory access.
struct LinRow {
	int r, i0, rem, nf, step, den;
	__device__ void init(int r0, int nf_, int nc) {
		nf = nf_, step = nc - 1, den = nf_ - 1, r = r0;
		const int rc = r0 < 0 ? 0 : (r0 < nf_ ? r0 : nf_ - 1);
		const int64_t num = (int64_t)rc * step;
		i0 = (int)(num / den);
		rem = (int)(num % den);
	}
	// Coarsening levels have nc ≤ nf, so step < den and one subtraction suffices.
	__device__ void next() {
		++r;
		if (r >= 1 && r <= nf - 1) {
			rem += step;
			if (rem >= den) rem -= den, ++i0;
		}
	}
	__device__ double w(double inv) const { return double(rem) * inv; }
};

// Waves of a wave-streaming launch over a level: strips × segments.
inline int64_t wave_strips_post(int64_t cols) { return (cols + SSW - 1) / SSW; }
inline int64_t wave_strips_restrict(int64_t cc) { return (cc - 2 + RCX - 1) / RCX; }  // interior coarse columns
inline int64_t wave_segs_restrict(int64_t cr, int scr) { return (cr - 2 + scr - 1) / scr; }   // interior coarse rows
inline unsigned wave_blocks(int64_t nwaves) { return (unsigned)((nwaves + NWAVES - 1) / NWAVES); }

// Runs step(integral_constant<k>, t, cl) for t = t0 … t1 − 1, k = (t − t0) mod
// W: whole blocks of W steps without conditions (so the compiler's memory-wait
// counts carry across the loop), then the remainder.  cl (true_type or
// false_type) is passed through: a wave whose rows t0 … t1 − 1 and columns are
// all interior runs its steps without the grid-edge fixes.  The choice is made
// once per wave, so that each variant is a loop nest of its own (one loop
// holding both would need the registers of both).
template <class F, class CL, int... K>
__device__ __forceinline__ void run_block(F& step, int b, CL cl, std::integer_sequence<int, K...>) {
	(step(std::integral_constant<int, K>{}, b + K, cl), ...);
}
template <class F, class CL, int... K>
__device__ __forceinline__ void run_rest(F& step, int b, int t1, CL cl, std::integer_sequence<int, K...>) {
	((b + K < t1 ? step(std::integral_constant<int, K>{}, b + K, cl) : void()), ...);
}
template <int W, class F, class CL>
__device__ __forceinline__ void run_steps(F&& step, int t0, int t1, CL cl) {
	int b = t0;
	for (; b + W <= t1; b += W) run_block(step, b, cl, std::make_integer_sequence<int, W>{});
	run_rest(step, b, t1, cl, std::make_integer_sequence<int, W>{});
}

// out = J^NPOST(J^NPRE(0; b) + P·e) over the level (k_fpost's result), with Epi's
// per-point reduction terms (summed per lane in FP64, then merged with
// compensation across lanes and workgroups).  Wave gw takes strip gw % nstrips
// and row segment gw / nstrips of sh rows.  Stage values are computed on every
// lane and row and then selected, so a step is branch-free up to the store.
// One wave of k_wpost: strip `strip` of SSW owned columns and row segment `seg` of sh rows of the level;
// per-lane reduction terms of Epi accumulate into lsum.
template <bool VK, int NPRE, int NPOST, class Epi>
__device__ __forceinline__ void wpost_wave(const float* __restrict__ b, const float* __restrict__ ec,
                                           float* __restrict__ out, const LWf& L, const fx_xtab& X, int rows, int cols,
                                           int cr, int cc, int strip, int seg, int sh, Epi& epi, double* lsum) {
	constexpr int LU = NPRE - 1;     // rows the iterate u = J^npre(0; b) + P·e lags the loaded row
	constexpr int LS = LU + 1;       // first post-sweep (NPOST == 2)
	constexpr int LZ = LU + NPOST;   // the level's result
	constexpr int WP = WD_POST;
	static_assert(LZ <= SHX && LZ < WR && LZ < WP && WP % WR == 0, "halo or rings too small");
	// With Epi::RHS_FROM_R the rhs is float(r) of the epilogue's FP64 residual
	// (which it loads anyway, for the same row and column), not a load of b.
	constexpr bool LOADB = !Epi::RHS_FROM_R;
	const int x = (int)(threadIdx.x & (WAVE - 1));
	{
		const int gc = strip * SSW - SHX + x;
		const unsigned uc = (unsigned)(gc < 0 ? 0 : (gc < cols ? gc : cols - 1));
		const bool colin = gc >= 0 && gc < cols;
		const bool colint = gc >= 1 && gc <= cols - 2;
		const bool kwok = gc >= 1 && gc <= cols - 1, keok = gc >= 0 && gc <= cols - 2;
		const bool outlane = colin && x >= SHX && x < SRW - SHX;
		const bool wclean = strip * SSW - SHX >= 1 && strip * SSW - SHX + SRW - 1 <= cols - 2;  // all lanes interior
		const int S0 = seg * sh, S1 = S0 + sh < rows ? S0 + sh : rows;
		// Column interpolation data of this lane.
		unsigned b0 = 0, b1 = 0;
		float wx = 0.0f, vx = 0.0f;
		if (colint) {
			const int a = X.i0c[gc];
			b0 = (unsigned)a;
			b1 = (unsigned)(a + 1 < cc ? a + 1 : a);
			wx = float(X.wc[gc]);
			vx = 1.0f - wx;
		}
		auto inter = [&](int r) { return colint && r >= 1 && r <= rows - 2; };
		// Coarse-correction operands of a row (read WP steps before the row's iterate is formed): the
		// coarse row a0 below the fine row, and the values of the row a0 + 1 above it.  The values of
		// row a0 are those of row a0 + 1 of an earlier fine row (or of the wave's first), since a0
		// advances by at most one per fine row, so a row costs two loads, not four.
		struct EV {
			float e10, e11, wy;
			int a0;
		};
		const double rinv = 1.0 / double(rows - 1);
		LinRow le;  // row of the next coarse-correction load
		auto eload = [&]() {
			const int a0 = le.i0;
			const int a1 = a0 + 1 < cr ? a0 + 1 : a0;
			const float* e1 = ec + (size_t)a1 * (size_t)cc;
			const EV v{e1[b0], e1[b1], float(le.w(rinv)), a0};
			le.next();
			return v;
		};
		auto pload = [&](int r) {
			const int rc = r < 0 ? 0 : (r < rows ? r : rows - 1);
			return epi.load_at((size_t)rc * (size_t)cols, uc);
		};
		const int T0 = S0 - LZ, T1 = S1 + LZ;  // steps: rows loaded
		// The wave's whole walk, instantiated once with and once without the
		// grid-edge fixes (cl), so that each variant gets its own registers.
		auto walk = [&](auto cl) {
			le.init(T0 - LU, rows, cr);
			// Row a0 of the current fine row (ca), its values (e00, e01) and row a0 + 1's of the last one.
			int ca = le.i0;
			float e00 = (ec + (size_t)ca * (size_t)cc)[b0], e01 = (ec + (size_t)ca * (size_t)cc)[b1];
			float p10 = 0.0f, p11 = 0.0f;
			// Input ring: row T0 + j in slot j mod WP; the first WP − LZ rows now, the rest as slots free up.
			WRow rb[WP];
			EV evb[WR];   // coarse-correction operands, WR rows ahead (cache hits mostly: a short lead suffices)
			typename Epi::Pf pfb[WP];
#pragma unroll
			for (int k = 0; k < WP; ++k) {
				rb[k] = k < WP - LZ ? wload<VK, LOADB>(b, L, T0 + k, rows, cols, uc) : WRow{0.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0.0f};
				if (k < WR) evb[k] = eload();
				pfb[k] = pload(T0 + k - LZ);
			}
			float U1[WR] = {0.0f, 0.0f, 0.0f, 0.0f}, U[WR] = {0.0f, 0.0f, 0.0f, 0.0f}, S[WR] = {0.0f, 0.0f, 0.0f, 0.0f};
			auto step = [&](auto kc, int t, auto cl) {
				constexpr int k = decltype(kc)::value, q = k % WR;
				constexpr bool CL = decltype(cl)::value;
				auto in = [&](int r) { return CL || inter(r); };
				{
					WRow v = wtake<VK>(rb[k], rb[Slot<WP, k, 1>::v]);
					// Row t's residual was loaded with the row: it is in the epilogue slot LZ steps on.
					if constexpr (!LOADB) v.b = epi.rhs(pfb[Slot<WP, k, -LZ>::v]);
					rb[k] = CL ? v : wfix<VK>(v, t, rows, colin, kwok, keok);
				}
				// Inputs of row t − i: rb[Slot<WP, k, i>].
				float u;  // iterate at row t − LU
				if constexpr (NPRE == 2) {
					const float u1 = wjac0<VK>(L, rb[k]);
					U1[q] = in(t) ? u1 : 0.0f;
					const float c = U1[Slot<WR, q, 1>::v];
					const float wv = lane_west(c), evv = lane_east(c);
					u = wjac<VK>(L, rb[Slot<WP, k, 1>::v], c, U1[Slot<WR, q, 2>::v], U1[q], wv, evv);
				} else {
					u = wjac0<VK>(L, rb[k]);
				}
				{
					const EV& ev = evb[q];
					const bool adv = ev.a0 != ca;   // the coarse row below moved up by one
					e00 = adv ? p10 : e00;
					e01 = adv ? p11 : e01;
					ca = ev.a0;
					p10 = ev.e10, p11 = ev.e11;
					const float wy = ev.wy, vy = 1.0f - wy;
					const float top = vx * e00 + wx * e01;
					const float bot = vx * ev.e10 + wx * ev.e11;
					u += vy * top + wy * bot;
				}
				U[q] = in(t - LU) ? u : 0.0f;
				const float* C = U;
				if constexpr (NPOST == 2) {
					const float c = U[Slot<WR, q, 1>::v];
					const float wv = lane_west(c), evv = lane_east(c);
					const float sv = wjac<VK>(L, rb[Slot<WP, k, LS>::v], c, U[Slot<WR, q, 2>::v], U[q], wv, evv);
					S[q] = in(t - LS) ? sv : c;
					C = S;
				}
				const float c = C[Slot<WR, q, 1>::v];
				const float wv = lane_west(c), evv = lane_east(c);
				const float zv = wjac<VK>(L, rb[Slot<WP, k, LZ>::v], c, C[Slot<WR, q, 2>::v], C[q], wv, evv);
				const int rz = t - LZ;
				const float z = in(rz) ? zv : c;
				if (outlane && rz >= S0 && rz < S1) {
					// The solver's level 0 (an epilogue): z is read back only by the next FCG step.
					if constexpr (Epi::NS > 0) store_nt((out + (size_t)rz * (size_t)cols) + (unsigned)gc, z);
					else (out + (size_t)rz * (size_t)cols)[(unsigned)gc] = z;
					epi.use_lane(pfb[k], z, lsum);
				}
				// Refill: the slot of row t − LZ (no longer read) with row t + WP − LZ, and this step's operands.
				// Rows past the walk's last are never used: the last one is read again instead (a cache
				// hit), so that a segment does not fetch the next segment's rows from memory.
				const int tn = t + WP - LZ < T1 ? t + WP - LZ : T1 - 1;
				rb[Slot<WP, k, LZ>::v] = wload<VK, LOADB>(b, L, tn, rows, cols, uc);
				evb[q] = eload();
				pfb[k] = pload(tn);
		};
		run_steps<WP>(step, T0, T1, cl);
		};
		if (wclean && T0 >= 1 && T1 <= rows - 1) walk(std::true_type{});
		else walk(std::false_type{});
	}
}
template <bool VK, int NPRE, int NPOST, class Epi>
__global__ __launch_bounds__(BLOCK) void k_wpost(const float* __restrict__ b, const float* __restrict__ ec,
                                                 float* __restrict__ out, LWf L, fx_xtab X, int rows, int cols, int cr,
                                                 int cc, int nstrips, int nseg, int sh, Epi epi, RedScratch rs,
                                                 const int* skip) {
	if (skip && *skip) return;
	constexpr int NA = Epi::NS > 0 ? Epi::NS : 1;
	double lsum[NA];
#pragma unroll
	for (int k = 0; k < NA; ++k) lsum[k] = 0.0;
	const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
	const int gw = xcd_remap(blockIdx.x, gridDim.x) * NWAVES + w;
	if (gw < nstrips * nseg)
		wpost_wave<VK, NPRE, NPOST>(b, ec, out, L, X, rows, cols, cr, cc, gw % nstrips, gw / nstrips, sh, epi, lsum);
	if constexpr (Epi::NS > 0) {
		Acc acc[NA];
#pragma unroll
		for (int k = 0; k < NA; ++k) acc[k] = Acc{lsum[k], 0.0};
		block_partial<Epi::NS>(acc, rs);
	}
}

// coarse rhs = R(b − A·J^npre(0; b)) over the level (k_frestrict's result).
// Wave gw takes RCX interior coarse columns Ja … Jb (strip gw % nstrips; lane
// l < RCX owns column Ja + l) and scr interior coarse rows Ia … Ib (segment
// gw / nstrips); the first and last strips and segments also write the zero
// boundary.  Its 64 lanes hold the fine window from two columns left of the
// strip's first gathered fine column.  Each fine row's residual is summed over
// the coarse column's fine columns in ascending order (cross-lane permutes),
// then added, in ascending fine-row order, to the (at most two) coarse rows the
// fine row contributes to: the order of k_restrict's loops.
//
// With Upd::ON the rhs is not read: the wave forms it, row by row, as
// float(r_new) with r_new = r − α·A·p of the FCG x/r update (Upd loads r and
// A·p), stores r_new at the points it owns, and sums r_new² over them into
// lsum.  Ownership partitions the grid: strip s owns fine columns from
// ⌈C0(s)⌉₈ (its lane-0 column rounded up to a multiple of 8, so that owners'
// stores cover whole 64-byte pieces) to the next strip's, which lie in its
// window while consecutive strips' first columns are at most 57 apart;
// segment g owns fine rows from its first gathered row to the next segment's.
template <bool VK, int NPRE, class Upd>
__device__ __forceinline__ void wrestrict_wave(const float* __restrict__ b, float* __restrict__ bco, const LWf& L,
                                               const fx_xtab& X, int fr, int fc, int cr, int cc, int nstrips, int nseg,
                                               int strip, int seg, int scr, const Upd& upd, double& lsum) {
	constexpr int LU = NPRE - 1, LR = NPRE;  // rows the pre-smoothed iterate and the residual lag the loaded row
	constexpr int WQ = WD_RESTRICT;
	static_assert(LR < WR && LR < WQ && WQ % WR == 0, "rings too small");
	const int x = (int)(threadIdx.x & (WAVE - 1));
	const bool sfirst = strip == 0, slast = strip == nstrips - 1, gfirst = seg == 0, glast = seg == nseg - 1;
	const int Ja = 1 + strip * RCX, Jb = Ja + RCX - 1 < cc - 2 ? Ja + RCX - 1 : cc - 2;
	const int Ia = 1 + seg * scr, Ib = Ia + scr - 1 < cr - 2 ? Ia + scr - 1 : cr - 2;
	const int J = Ja + x;
	const bool jint = x < RCX && J <= Jb;  // the lane's interior coarse column
	// Zero boundary rows (first / last segment) and columns (first / last strip).
	if (gfirst || glast) {
		const size_t row = gfirst ? 0 : (size_t)(cr - 1) * cc;
		if (jint) bco[row + J] = 0.0f;
		if (x == 0 && sfirst) bco[row] = 0.0f;
		if (x == 0 && slast) bco[row + cc - 1] = 0.0f;
		if (gfirst && glast && jint) bco[(size_t)(cr - 1) * cc + J] = 0.0f;
		if (gfirst && glast && x == 0 && sfirst) bco[(size_t)(cr - 1) * cc] = 0.0f;
		if (gfirst && glast && x == 0 && slast) bco[(size_t)(cr - 1) * cc + cc - 1] = 0.0f;
	}
	if ((sfirst || slast) && x < 2) {
		const bool left = x == 0;
		if ((left && sfirst) || (!left && slast))
			for (int I = Ia; I <= Ib; ++I) bco[(size_t)I * cc + (left ? 0 : cc - 1)] = 0.0f;
	}
	const int G0 = X.loc[Ja];
	const int F0 = X.lor[Ia], F1 = X.hir[Ib];
	if (G0 > X.hic[Jb] || F0 > F1) {  // nothing to gather (tiny grids): the strip's interior rows are zero
		if (jint)
			for (int I = Ia; I <= Ib; ++I) bco[(size_t)I * cc + J] = 0.0f;
		return;
	}
	const int C0 = G0 - 2;  // fine column of lane 0
	const int gc = C0 + x;
	const unsigned uc = (unsigned)(gc < 0 ? 0 : (gc < fc ? gc : fc - 1));
	// Points whose r_new this wave stores (Upd::ON).
	const int ocl = sfirst ? 0 : (C0 + 7) & ~7;
	const int och = slast ? fc : (X.loc[Ja + RCX] - 2 + 7) & ~7;
	const bool ownc = gc >= ocl && gc < och;
	const int orl = gfirst ? 0 : F0, orh = glast ? fr : X.lor[Ia + scr];
	const bool colin = gc >= 0 && gc < fc;
	const bool wclean = C0 >= 1 && C0 + SRW - 1 <= fc - 2;  // all lanes interior
	const bool colint = gc >= 1 && gc <= fc - 2;
	const bool kwok = gc >= 1 && gc <= fc - 1, keok = gc >= 0 && gc <= fc - 2;
	auto inter = [&](int r) { return colint && r >= 1 && r <= fr - 2; };
	// The owned coarse column's fine columns lo..hi, their weights and source lanes.
	int cnt = 0;
	float wk[WNMAX];
	int sk[WNMAX];
	{
		const int lo = jint ? X.loc[J] : 0, hi = jint ? X.hic[J] : -1;
		cnt = hi - lo + 1;
#pragma unroll
		for (int k = 0; k < WNMAX; ++k) {
			const bool v = k < cnt;
			const int jj = v ? lo + k : 0;
			const int ac = v ? X.i0c[jj] : -4;
			const double wv = v ? X.wc[jj] : 0.0;
			wk[k] = ac == J ? float(1.0 - wv) : (ac + 1 == J ? float(wv) : 0.0f);
			sk[k] = v ? jj - C0 : 0;
		}
		// Unused terms read the first gathered column with weight 0: a finite
		// value times 0 is a zero, and adding a zero to the row sum (which
		// starts at +0 and so is never −0) leaves it unchanged.
#pragma unroll
		for (int k = 1; k < WNMAX; ++k)
			if (k >= cnt) sk[k] = sk[0];
	}
	const float scale = X.sy * X.sx;
	int Icur = X.i0r[F0];  // coarse row accumulated in alo (ahi: Icur + 1)
	float alo = 0.0f, ahi = 0.0f;
	auto emit = [&](int I, float v) {
		if (jint && I >= Ia && I <= Ib) bco[(size_t)I * cc + J] = v * scale;
	};
	const int T0 = F0 - LR, T1 = F1 + LR + 1;  // steps: rows loaded
	const double rinv = 1.0 / double(fr - 1);
	// The wave's whole walk, instantiated once with and once without the
	// grid-edge fixes (cl), so that each variant gets its own registers.
	auto walk = [&](auto cl) {
		LinRow lr;  // fine row whose residual the step produces
		lr.init(T0 - LR, fr, cr);
		constexpr bool LOADB = !Upd::ON;
		auto uload = [&](int r) {
			const int rc = r < 0 ? 0 : (r < fr ? r : fr - 1);
			return upd.load_at((size_t)rc * (size_t)fc, uc);
		};
		WRow rb[WQ];
		typename Upd::Pf ub[WR];  // Upd::ON: row t's r and A·p in slot k mod WR, loaded WR steps ahead
#pragma unroll
		for (int k = 0; k < WQ; ++k) {
			rb[k] = k < WQ - LR ? wload<VK, LOADB>(b, L, T0 + k, fr, fc, uc) : WRow{0.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0.0f};
			if constexpr (Upd::ON)
				if (k < WR) ub[k] = uload(T0 + k);
		}
		float U1[WR] = {0.0f, 0.0f, 0.0f, 0.0f}, Uv[WR] = {0.0f, 0.0f, 0.0f, 0.0f};
		auto step = [&](auto kc, int t, auto cl) {
			constexpr int k = decltype(kc)::value, q = k % WR;
			constexpr bool CL = decltype(cl)::value;
			auto in = [&](int r) { return CL || inter(r); };
			{
				WRow v = wtake<VK>(rb[k], rb[Slot<WQ, k, 1>::v]);
				if constexpr (Upd::ON) {
					const double rn = upd.rnew(ub[q]);
					v.b = float(rn);
					if (ownc && t >= orl && t < orh) {
						upd.store((size_t)t * (size_t)fc + (unsigned)gc, rn);
						lsum += rn * rn;
					}
					ub[q] = uload(t + WR < T1 ? t + WR : T1 - 1);   // rows past the walk: never used
				}
				rb[k] = CL ? v : wfix<VK>(v, t, fr, colin, kwok, keok);
			}
			float u;  // pre-smoothed iterate at row t − LU
			if constexpr (NPRE == 2) {
				const float u1 = wjac0<VK>(L, rb[k]);
				U1[q] = in(t) ? u1 : 0.0f;
				const float c = U1[Slot<WR, q, 1>::v];
				const float wv = lane_west(c), evv = lane_east(c);
				u = wjac<VK>(L, rb[Slot<WQ, k, 1>::v], c, U1[Slot<WR, q, 2>::v], U1[q], wv, evv);
			} else {
				u = wjac0<VK>(L, rb[k]);
			}
			Uv[q] = in(t - LU) ? u : 0.0f;
			// Residual at fine row i = t − LR.
			const int i = t - LR;
			const float c = Uv[Slot<WR, q, 1>::v];
			const float wv = lane_west(c), evv = lane_east(c);
			const WRow& ri = rb[Slot<WQ, k, LR>::v];
			const float sv = wop<VK>(L, ri, c, Uv[Slot<WR, q, 2>::v], Uv[q], wv, evv);
			const float res = in(i) ? fmaf(1.0f, ri.b, -sv) : 0.0f;
			// Refill the slot of row t − LR (no longer read) with row t + WQ − LR (rows past the walk's
			// last are never used: the last is read again, a cache hit, instead of the next segment's).
			rb[Slot<WQ, k, LR>::v] = wload<VK, LOADB>(b, L, t + WQ - LR < T1 ? t + WQ - LR : T1 - 1, fr, fc, uc);
			// Row sum over the owned coarse column's fine columns.
			float row = 0.0f;
#pragma unroll
			for (int k2 = 0; k2 < WNMAX; ++k2) row += wk[k2] * __shfl(res, sk[k2]);
			const int a = lr.i0;
			const double wrv = lr.w(rinv);
			lr.next();
			if (i >= F0 && i <= F1) {
				// Coarse rows advance by at most one per fine row (nc ≤ nf).
				if (Icur < a) {
					emit(Icur, alo);
					alo = ahi;
					ahi = 0.0f;
					++Icur;
				}
				alo += float(1.0 - wrv) * row;
				ahi += float(wrv) * row;
			}
		};
		run_steps<WQ>(step, T0, T1, cl);
		emit(Icur, alo);
		emit(Icur + 1, ahi);
	};
	if (wclean && T0 >= 1 && T1 <= fr - 1) walk(std::true_type{});
	else walk(std::false_type{});
}
template <bool VK, int NPRE, class Upd>
__global__ __launch_bounds__(BLOCK) void k_wrestrict(const float* __restrict__ b, float* __restrict__ bco, LWf L,
                                                     fx_xtab X, int fr, int fc, int cr, int cc, int nstrips, int nseg,
                                                     int scr, Upd upd, RedScratch rs, const int* skip) {
	if (skip && *skip) return;
	double lsum = 0.0;
	upd.begin();
	const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
	const int gw = xcd_remap(blockIdx.x, gridDim.x) * NWAVES + w;
	if (gw < nstrips * nseg)
		wrestrict_wave<VK, NPRE>(b, bco, L, X, fr, fc, cr, cc, nstrips, nseg, gw % nstrips, gw / nstrips, scr, upd, lsum);
	if constexpr (Upd::ON) {
		Acc acc[1] = {Acc{lsum, 0.0}};
		block_partial<1>(acc, rs);
	}
}

// ── k_ftail: the V-cycle below a small level, in one workgroup ─────────
// The levels from the first one whose whole sub-hierarchy fits in LDS down to
// the coarsest are cycled by one workgroup of TAIL_BLOCK threads, phase by
// phase with workgroup barriers in between: pre-sweeps, residual, restriction
// (fine-row sums over each coarse column's fine columns, then the weighted
// sums of those over each coarse row's fine rows, the order of k_restrict),
// the coarsest level's sweeps, prolongation and post-sweeps.  Every plane
// (face coefficients, ω/d, b, iterates) and transfer table lives in LDS, so a
// phase costs a few LDS round trips instead of a kernel launch.  Per-point
// arithmetic is that of the other fused kernels, so the result is
// bit-identical.
constexpr int TAIL_MAXL = 6;
// Plane points per thread staged by k_ftail (≥ ⌈plane length / TAIL_BLOCK⌉; the
// LDS limit alone keeps the plane length below 7 (VK) or 14 blocks).
constexpr int tail_su(bool vk) { return vk ? 7 : 14; }
constexpr int TAIL_BLOCK = 1024;
constexpr size_t TAIL_LDS_MAX = 160 * 1024;
struct TailLevel {
	LWf L;
	fx_xtab X;          // transfer to the next tail level (levels above the coarsest)
	int rows, cols;
	int off;            // offset of the level in each point plane
	int tr, tc;         // offsets of its fine-row / fine-column tables (levels above the coarsest)
	int tcr, tcc;       // offsets of the next level's coarse-row / coarse-column ranges
};
struct TailArgs {
	TailLevel lv[TAIL_MAXL];
	int nl;             // tail levels; lv[nl − 1] is the hierarchy's coarsest
	int sweeps;         // coarsest-level sweeps from zero
	int npts;           // points over all tail levels (plane length)
	int nrs;            // row-sum scratch length: max over levels of rows · (next level's cols)
	int ntab;           // table entries (fine rows + fine cols of levels above the coarsest)
	int nrng;           // range entries (coarse rows + coarse cols of levels below the first)
	const float* rhs0;  // rhs of the first tail level (global)
	float* out0;        // result of the first tail level (global)
};
// LDS bytes of k_ftail for the given plane, scratch and table sizes.
inline size_t tail_lds(bool vk, int npts, int nrs, int ntab, int nrng) {
	return (size_t)(vk ? 6 : 3) * npts * sizeof(float) + (size_t)nrs * sizeof(float) +
	       (size_t)ntab * (sizeof(int32_t) + 2 * sizeof(float)) + (size_t)nrng * 2 * sizeof(int32_t);
}

template <bool VK, int NPRE, int NPOST>
__global__ __launch_bounds__(TAIL_BLOCK) void k_ftail(TailArgs A, const int* skip) {
	if (skip && *skip) return;
	extern __shared__ float tsm[];
	const int P = A.npts;
	float* pb = tsm;            // b
	float* pu = pb + P;         // iterate (pre-smoothed, then corrected)
	float* pv = pu + P;         // work plane
	float* kxp = pv + P;        // VK: faces and ω/d
	float* kyp = kxp + (VK ? P : 0);
	float* wdp = kyp + (VK ? P : 0);
	float* rsum = VK ? wdp + P : kxp;  // restriction row sums
	int32_t* ti0 = (int32_t*)(rsum + A.nrs);   // per fine row / column: coarse index below it
	float* twl = (float*)(ti0 + A.ntab);       // its weights float(1 − w), float(w)
	float* twh = twl + A.ntab;
	int32_t* rlo = (int32_t*)(twh + A.ntab);   // per coarse row / column: fine range lo..hi
	int32_t* rhi = rlo + A.nrng;
	const int t = threadIdx.x;
	// Stage the planes and tables, level by level (a uniform loop, so that the
	// level's fields are scalar loads); within a level each thread issues all of
	// its global loads before its first LDS store, so that their latencies overlap.
	for (int l = 0; l < A.nl; ++l) {
		const TailLevel& T = A.lv[l];
		const int n = T.rows * T.cols;
		float sv[tail_su(VK)][VK ? 4 : 1];
#pragma unroll
		for (int j = 0; j < tail_su(VK); ++j) {
			const int q = t + TAIL_BLOCK * j;
			if (q < n) {
				sv[j][0] = l == 0 ? A.rhs0[q] : 0.0f;   // b of the other levels comes from the restriction
				if constexpr (VK) {
					sv[j][1] = T.L.kx[q];
					sv[j][2] = T.L.ky[q];
					sv[j][3] = T.L.wd[q];
				}
			}
		}
		const bool tab = l + 1 < A.nl;
		const int cr = tab ? A.lv[l + 1].rows : 0, cc = tab ? A.lv[l + 1].cols : 0;
		int ir = 0, ic = 0, lr = 0, hr = 0, lc = 0, hc = 0;
		double wrv = 0.0, wcv = 0.0;
		if (tab && t < T.rows) ir = T.X.i0r[t], wrv = T.X.wr[t];
		if (tab && t < T.cols) ic = T.X.i0c[t], wcv = T.X.wc[t];
		if (tab && t < cr) lr = T.X.lor[t], hr = T.X.hir[t];
		if (tab && t < cc) lc = T.X.loc[t], hc = T.X.hic[t];
#pragma unroll
		for (int j = 0; j < tail_su(VK); ++j) {
			const int q = t + TAIL_BLOCK * j;
			if (q < n) {
				pb[T.off + q] = sv[j][0];
				if constexpr (VK) {
					kxp[T.off + q] = sv[j][1];
					kyp[T.off + q] = sv[j][2];
					wdp[T.off + q] = sv[j][3];
				}
			}
		}
		if (tab && t < T.rows) ti0[T.tr + t] = ir, twl[T.tr + t] = float(1.0 - wrv), twh[T.tr + t] = float(wrv);
		if (tab && t < T.cols) ti0[T.tc + t] = ic, twl[T.tc + t] = float(1.0 - wcv), twh[T.tc + t] = float(wcv);
		if (tab && t < cr) rlo[T.tcr + t] = lr, rhi[T.tcr + t] = hr;
		if (tab && t < cc) rlo[T.tcc + t] = lc, rhi[T.tcc + t] = hc;
	}
	__syncthreads();
	// One Jacobi point update of level T at q from plane `in` (interior points only).
	auto jacq = [&](const TailLevel& T, const float* in, int q, float c) {
		const int cols = T.cols;
		const float* u = in + T.off;
		if constexpr (VK) {
			const int o = T.off + q;
			return jackw(T.L, kxp[o - 1], kxp[o], kyp[o - cols], kyp[o], wdp[o], c, u[q - cols], u[q + cols], u[q - 1],
			             u[q + 1], pb[o]);
		} else {
			return jac(T.L, c, u[q - cols], u[q + cols], u[q - 1], u[q + 1], pb[T.off + q]);
		}
	};
	// Point loops: thread t takes points q = t + TAIL_BLOCK·j of the row-major plane, so that every
	// thread has work on levels narrower than a wave too.  The row is ⌊(q + ½)/cols⌋ in FP32: exact
	// while rows·cols < 2²¹ (the product's rounding error stays below the ½/cols margin).
	auto pts = [&](int rows, int cols, auto f) {
		const int n = rows * cols;
		const float inv = 1.0f / float(cols);
		for (int q = t; q < n; q += TAIL_BLOCK) {
			const int r = int((float(q) + 0.5f) * inv);
			f(r, q - r * cols, q);
		}
	};
	auto inner = [](const TailLevel& T, int r, int c) { return r > 0 && c > 0 && r < T.rows - 1 && c < T.cols - 1; };
	// ── down: pre-smoothing, residual, restriction ──
	for (int l = 0; l + 1 < A.nl; ++l) {
		const TailLevel& T = A.lv[l];
		const TailLevel& C = A.lv[l + 1];
		// First sweep from zero into pv (NPRE == 2) or pu (NPRE == 1).
		float* first = NPRE == 2 ? pv : pu;
		pts(T.rows, T.cols, [&](int r, int c, int q) {
			float v = 0.0f;
			if (inner(T, r, c)) {
				if constexpr (VK) {
					const int o = T.off + q;
					v = jackw(T.L, kxp[o - 1], kxp[o], kyp[o - T.cols], kyp[o], wdp[o], 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, pb[o]);
				} else {
					v = jac0(T.L, pb[T.off + q]);
				}
			}
			first[T.off + q] = v;
		});
		__syncthreads();
		if constexpr (NPRE == 2) {
			pts(T.rows, T.cols, [&](int r, int c, int q) {
				pu[T.off + q] = inner(T, r, c) ? jacq(T, pv, q, pv[T.off + q]) : 0.0f;
			});
			__syncthreads();
		}
		// Residual into pv.
		pts(T.rows, T.cols, [&](int r, int c, int q) {
			float res = 0.0f;
			if (inner(T, r, c)) {
				const float* u = pu + T.off;
				const int cols = T.cols;
				float sv;
				if constexpr (VK) {
					const int o = T.off + q;
					float d;
					sv = vkff(T.L, kxp[o - 1], kxp[o], kyp[o - cols], kyp[o], u[q], u[q - cols], u[q + cols], u[q - 1],
					          u[q + 1], &d);
				} else {
					sv = sop(T.L, u[q], u[q - cols], u[q + cols], u[q - 1], u[q + 1]);
				}
				res = fmaf(1.0f, pb[T.off + q], -sv);
			}
			pv[T.off + q] = res;
		});
		__syncthreads();
		// Row pass: rsum[i·cc + J] = Σ_jj w(jj, J)·res(i, jj) over J's fine columns.
		const int cr = C.rows, cc = C.cols;
		pts(T.rows, cc, [&](int i, int J, int q) {
			if (i < 1 || i > T.rows - 2 || J < 1 || J > cc - 2) return;
			const int lo = rlo[T.tcc + J], hi = rhi[T.tcc + J];
			const float* row = pv + T.off + i * T.cols;
			float rs = 0.0f;
			for (int jj = lo; jj <= hi; ++jj) {
				const int ac = ti0[T.tc + jj];
				const float wxv = ac == J ? twl[T.tc + jj] : (ac + 1 == J ? twh[T.tc + jj] : 0.0f);
				rs += wxv * row[jj];
			}
			rsum[q] = rs;
		});
		__syncthreads();
		// Column pass: the coarse rhs.
		const float scale = T.X.sy * T.X.sx;
		pts(cr, cc, [&](int I, int J, int q) {
			float outv = 0.0f;
			if (I > 0 && I < cr - 1 && J > 0 && J < cc - 1) {
				const int lo = rlo[T.tcr + I], hi = rhi[T.tcr + I];
				float acc = 0.0f;
				for (int ii = lo; ii <= hi; ++ii) {
					const int ar = ti0[T.tr + ii];
					const float wy = ar == I ? twl[T.tr + ii] : (ar + 1 == I ? twh[T.tr + ii] : 0.0f);
					acc += wy * rsum[ii * cc + J];
				}
				outv = acc * scale;
			}
			pb[C.off + q] = outv;
		});
		__syncthreads();
	}
	// ── coarsest: sweeps from zero (k_fcoarse), result in pu ──
	{
		const TailLevel& T = A.lv[A.nl - 1];
		pts(T.rows, T.cols, [&](int, int, int q) { pu[T.off + q] = 0.0f; });
		__syncthreads();
		float* cur = pu;
		float* nxt = pv;
		for (int it = 0; it < A.sweeps; ++it) {
			pts(T.rows, T.cols, [&](int r, int c, int q) {
				const float v = cur[T.off + q];
				nxt[T.off + q] = inner(T, r, c) ? jacq(T, cur, q, v) : v;
			});
			__syncthreads();
			float* x = cur;
			cur = nxt;
			nxt = x;
		}
		if (cur != pu) {
			pts(T.rows, T.cols, [&](int, int, int q) { pu[T.off + q] = cur[T.off + q]; });
			__syncthreads();
		}
	}
	// ── up: prolongation and post-smoothing, each level's result in pu ──
	for (int l = A.nl - 2; l >= 0; --l) {
		const TailLevel& T = A.lv[l];
		const TailLevel& C = A.lv[l + 1];
		const int cr = C.rows, cc = C.cols;
		const float* e = pu + C.off;
		pts(T.rows, T.cols, [&](int r, int c, int q) {
			if (!inner(T, r, c)) return;
			const int a0 = ti0[T.tr + r], a1 = a0 + 1 < cr ? a0 + 1 : a0;
			const int b0 = ti0[T.tc + c], b1 = b0 + 1 < cc ? b0 + 1 : b0;
			const float wy = twh[T.tr + r], vy = 1.0f - wy;
			const float wx = twh[T.tc + c], vx = 1.0f - wx;
			const float top = vx * e[a0 * cc + b0] + wx * e[a0 * cc + b1];
			const float bot = vx * e[a1 * cc + b0] + wx * e[a1 * cc + b1];
			pu[T.off + q] += vy * top + wy * bot;
		});
		__syncthreads();
		const float* last = pu;
		if constexpr (NPOST == 2) {
			pts(T.rows, T.cols, [&](int r, int c, int q) {
				const float v = pu[T.off + q];
				pv[T.off + q] = inner(T, r, c) ? jacq(T, pu, q, v) : v;
			});
			__syncthreads();
			last = pv;
		}
		float* dst = l == 0 ? A.out0 : pu + T.off;
		if (NPOST == 1 && l > 0) dst = pv + T.off;
		pts(T.rows, T.cols, [&](int r, int c, int q) {
			const float v = last[T.off + q];
			dst[q] = inner(T, r, c) ? jacq(T, last, q, v) : v;
		});
		__syncthreads();
		if (NPOST == 1 && l > 0) {
			pts(T.rows, T.cols, [&](int, int, int q) { pu[T.off + q] = pv[T.off + q]; });
			__syncthreads();
		}
	}
}

// First level of the tail kernel (≥ 1: level 0 carries the solver's reductions),
// the smallest index whose sub-hierarchy fits TAIL_MAXL levels and the LDS;
// nl − 1 when only the coarsest level does (then k_fcoarse runs alone).
// Fills *A (except rhs0 / out0) for that level.
inline int tail_plan(const fx_multigrid* mg, bool vk, TailArgs* A, size_t* lds) {
	const auto& lv = mg->lv;
	const int nl = (int)lv.size();
	for (int lt = 1; lt < nl - 1; ++lt) {
		if (nl - lt > TAIL_MAXL) continue;
		int npts = 0, nrs = 0, ntab = 0, nrng = 0;
		for (int l = lt; l < nl; ++l) {
			npts += (int)(lv[l].g.rows * lv[l].g.cols);
			if (l + 1 < nl) {
				nrs = std::max<int>(nrs, (int)(lv[l].g.rows * lv[l + 1].g.cols));
				ntab += (int)(lv[l].g.rows + lv[l].g.cols);
				nrng += (int)(lv[l + 1].g.rows + lv[l + 1].g.cols);
			}
		}
		const size_t b = tail_lds(vk, npts, nrs, ntab, nrng);
		bool fits = b <= TAIL_LDS_MAX;
		for (int l = lt; l < nl; ++l)   // staging: a level's points in tail_su blocks, its table rows and columns in one
			fits = fits && lv[l].g.rows * lv[l].g.cols <= (size_t)tail_su(vk) * TAIL_BLOCK && lv[l].g.rows <= TAIL_BLOCK &&
			       lv[l].g.cols <= TAIL_BLOCK;
		if (!fits) continue;
		A->nl = nl - lt;
		A->sweeps = mg->opt.n_pre + mg->opt.n_post + mg->opt.extra_coarse;
		A->npts = npts, A->nrs = nrs, A->ntab = ntab, A->nrng = nrng;
		int off = 0, tab = 0, rng = 0;
		for (int l = lt; l < nl; ++l) {
			TailLevel& T = A->lv[l - lt];
			T.L = lwf(mg, lv[l]);
			T.rows = (int)lv[l].g.rows, T.cols = (int)lv[l].g.cols;
			T.off = off;
			off += T.rows * T.cols;
			if (l + 1 < nl) {
				T.X = lv[l].xt;
				T.tr = tab, tab += T.rows;
				T.tc = tab, tab += T.cols;
				T.tcr = rng, rng += (int)lv[l + 1].g.rows;
				T.tcc = rng, rng += (int)lv[l + 1].g.cols;
			} else {
				T.X = fx_xtab{};
				T.tr = T.tc = T.tcr = T.tcc = 0;
			}
		}
		*lds = b;
		return lt;
	}
	return nl - 1;
}

// Launches of the wave kernels.
template <bool VK, int NPRE, int NPOST, class Epi>
void launch_wpost(hipStream_t st, unsigned blocks, const float* b, const float* ec, float* out, const LWf& L,
                  const fx_xtab& X, int rows, int cols, int cr, int cc, int ns, int nseg, int sh, Epi epi, RedScratch rs,
                  const int* skip) {
	hipLaunchKernelGGL((k_wpost<VK, NPRE, NPOST, Epi>), dim3(blocks), dim3(BLOCK), 0, st, b, ec, out, L, X, rows, cols,
	                   cr, cc, ns, nseg, sh, epi, rs, skip);
}
template <bool VK, int NPRE, class Upd = NoUpd>
void launch_wrestrict(hipStream_t st, unsigned blocks, const float* b, float* bco, const LWf& L, const fx_xtab& X,
                      int fr, int fc, int cr, int cc, int ns, int nseg, int scr, const int* skip, Upd upd = Upd{},
                      RedScratch rs = RedScratch{nullptr}) {
	hipLaunchKernelGGL((k_wrestrict<VK, NPRE, Upd>), dim3(blocks), dim3(BLOCK), 0, st, b, bco, L, X, fr, fc, cr, cc, ns,
	                   nseg, scr, upd, rs, skip);
}

// Segment heights of the wave-streaming kernels for a level: the tallest of
// 256 … 8 fine rows (post) or 128 … 4 coarse rows (restrict) that still gives
// WAVE_TARGET waves (eight per CU), unless a tuned height is set.
constexpr int64_t WAVE_TARGET = 2048;
// A tuned height of WROWS_TILED (levels ≥ 1 only) runs the level on the LDS-tile kernels instead.
constexpr int WROWS_TILED = -1;
inline int wpost_rows(int64_t rows, int64_t cols) {
	const int64_t ns = wave_strips_post(cols);
	int sh = 256;
	while (sh > 8 && ns * ((rows + sh - 1) / sh) < WAVE_TARGET) sh /= 2;
	return sh;
}
inline int wrestrict_rows(int64_t cr, int64_t cc) {
	const int64_t ns = wave_strips_restrict(cc);
	int s = 128;
	while (s > 4 && ns * wave_segs_restrict(cr, s) < WAVE_TARGET) s /= 2;
	return s;
}
// The wave restriction needs each strip's fine window (two columns of halo
// each side) within one wave and at most WNMAX fine columns per coarse column.
inline bool wrestrict_ok(const fx_xtab& X) { return X.wx <= SRW && X.gmax <= WNMAX; }

// Whether the fused cycle can run this hierarchy.
inline bool fused_ok(const fx_multigrid* mg) {
	const auto& o = mg->opt;
	if (!mg->f32 || o.smoother != FX_SMOOTH_JACOBI || o.transfer != FX_TRANSFER_FW_BILINEAR) return false;
	if (o.n_pre < 1 || o.n_pre > 2 || o.n_post < 1 || o.n_post > 2) return false;
	if (mg->lv.size() < 2) return false;
	for (size_t l = 0; l + 1 < mg->lv.size(); ++l)
		if (!mg->lv[l].xbuf || frestrict_lds(mg->lv[l].xt, mg->lv[l].k != nullptr) > 64 * 1024) return false;
	const fx_grid& g = mg->lv.back().g;
	return g.rows * g.cols <= 4096;
}

// Whether level 0's wave restriction can also carry the FCG residual update
// (Upd::ON): wave kernels, a level large enough that every strip and segment
// gathers, and consecutive strips at most 57 fine columns apart (column
// ownership in 64-byte pieces).
inline bool upd_ok(const fx_multigrid* mg) {
	if (mg->kernels != FX_MGK_WAVE || mg->lv.size() < 2) return false;
	const auto& L = mg->lv[0];
	return wrestrict_ok(L.xt) && L.xt.adv <= 57 && L.g.rows >= 17 && L.g.cols >= 17;
}

template <bool VK, int NPRE, int NPOST, class Epi, class Upd>
fx_status fast_cycle_t(fx_multigrid* mg, const int* skip, Epi epi0, Upd upd0) {
	fx_backend* be = mg->be;
	hipStream_t st = be->stream;
	auto& lv = mg->lv;
	const int nl = (int)lv.size();
	const auto& o = mg->opt;
	auto tuned = [](const std::vector<int>& v, int l) { return l < (int)v.size() ? v[l] : 0; };
	// Levels lt … nl − 1 run in the single-workgroup tail kernel (wave variant only).
	TailArgs ta{};
	size_t tlds = 0;
	const int lt = mg->kernels == FX_MGK_WAVE ? tail_plan(mg, VK, &ta, &tlds) : nl - 1;
	for (int l = 0; l < lt; ++l) {
		const fx_grid& g = lv[l].g;
		const fx_grid& cg = lv[l + 1].g;
		if (l == 0) fx::probe_begin(be, FX_PROBE_VC_RESTRICT0);
		bool fused0 = false;
		unsigned fin_blocks = 0;   // the fused update's partial sums, merged after the probe ends
		if constexpr (Upd::ON) {
			if (l == 0) {
				// Level 0 forms its rhs from the FCG residual update (the caller checked upd_ok).
				int scr = mg->wrows_restrict_upd > 0 ? mg->wrows_restrict_upd : tuned(mg->wrows_restrict, l);
				if (scr <= 0) scr = wrestrict_rows(cg.rows, cg.cols);
				const int64_t ns = wave_strips_restrict(cg.cols), nseg = wave_segs_restrict(cg.rows, scr);
				const unsigned blocks = wave_blocks(ns * nseg);
				FX_CHECK(fx::ensure_partials(be, blocks));
				launch_wrestrict<VK, NPRE>(st, blocks, (const float*)lv[l].rhs, (float*)lv[l + 1].rhs, lwf(mg, lv[l]),
				                           lv[l].xt, (int)g.rows, (int)g.cols, (int)cg.rows, (int)cg.cols, (int)ns,
				                           (int)nseg, scr, skip, upd0, fx::red_scratch(be));
				fin_blocks = blocks;
				fused0 = true;
			}
		}
		if (fused0) {
		} else if (mg->kernels == FX_MGK_WAVE && wrestrict_ok(lv[l].xt) && tuned(mg->wrows_restrict, l) >= 0) {
			int scr = tuned(mg->wrows_restrict, l);
			if (scr <= 0) scr = wrestrict_rows(cg.rows, cg.cols);
			const int64_t ns = wave_strips_restrict(cg.cols), nseg = wave_segs_restrict(cg.rows, scr);
			launch_wrestrict<VK, NPRE>(st, wave_blocks(ns * nseg), (const float*)lv[l].rhs, (float*)lv[l + 1].rhs,
			                           lwf(mg, lv[l]), lv[l].xt, (int)g.rows, (int)g.cols, (int)cg.rows, (int)cg.cols,
			                           (int)ns, (int)nseg, scr, skip);
		} else {
			const int scr = mg->kernels == FX_MGK_RING ? srestrict_height(lv[l].xt, cg.rows, cg.cols) : 0;
			if (scr > 0)
				hipLaunchKernelGGL((k_srestrict<VK, NPRE>), dim3(srestrict_blocks(cg.rows, cg.cols, scr)), dim3(BLOCK), 0,
				                   st, (const float*)lv[l].rhs, (float*)lv[l + 1].rhs, lwf(mg, lv[l]), lv[l].xt,
				                   (int)g.rows, (int)g.cols, (int)cg.rows, (int)cg.cols,
				                   (int)((cg.cols + RCX - 1) / RCX), scr, skip);
			else
				hipLaunchKernelGGL((k_frestrict<VK, NPRE>), rgrid(cg.rows, cg.cols), dim3(BLOCK),
				                   (unsigned)frestrict_lds(lv[l].xt, VK), st, (const float*)lv[l].rhs,
				                   (float*)lv[l + 1].rhs, lwf(mg, lv[l]), lv[l].xt, (int64_t)g.rows, (int64_t)g.cols,
				                   (int64_t)cg.rows, (int64_t)cg.cols, skip);
		}
		if (l == 0) fx::probe_end(be, FX_PROBE_VC_RESTRICT0);
		if constexpr (Upd::ON) {
			if (fin_blocks)
				hipLaunchKernelGGL((k_finish<1, typename Upd::Fin>), dim3(1), dim3(FIN_BLOCK), 0, st,
				                   fx::red_scratch(be), fin_blocks, upd0.fin(), skip);
		}
	}
	if (lt < nl - 1) {
		ta.rhs0 = (const float*)lv[lt].rhs;
		ta.out0 = (float*)lv[lt].tmp;
		static bool lds_set[2][3][3] = {};
		bool& set = lds_set[VK][NPRE][NPOST];
		if (!set) {
			FX_HIP(hipFuncSetAttribute((const void*)k_ftail<VK, NPRE, NPOST>,
			                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)TAIL_LDS_MAX));
			set = true;
		}
		hipLaunchKernelGGL((k_ftail<VK, NPRE, NPOST>), dim3(1), dim3(TAIL_BLOCK), (unsigned)tlds, st, ta, skip);
	} else {
		const auto& L = lv[nl - 1];
		hipLaunchKernelGGL(k_fcoarse, dim3(1), dim3(1024), 0, st, (const float*)L.rhs, (float*)L.tmp, lwf(mg, L),
		                   o.n_pre + o.n_post + o.extra_coarse, (int)L.g.rows, (int)L.g.cols, skip);
	}
	// Launch shape of a level's post kernel: {variant, blocks, wave strips, segments, rows per segment}.
	struct PostShape {
		int kind;
		unsigned blocks;
		int ns, nseg, sh;
	};
	auto post_shape = [&](int l) {
		const fx_grid& g = lv[l].g;
		if (mg->kernels == FX_MGK_WAVE && tuned(mg->wrows_post, l) >= 0) {
			int sh = l == 0 && Epi::NS > 0 && mg->wrows_post_epi > 0 ? mg->wrows_post_epi : tuned(mg->wrows_post, l);
			if (sh <= 0) sh = wpost_rows(g.rows, g.cols);
			const int64_t ns = wave_strips_post(g.cols), nseg = (g.rows + sh - 1) / sh;
			return PostShape{FX_MGK_WAVE, wave_blocks(ns * nseg), (int)ns, (int)nseg, sh};
		}
		const int sh = mg->kernels == FX_MGK_RING ? spost_height(g.rows, g.cols) : 0;
		if (sh > 0) return PostShape{FX_MGK_RING, spost_blocks(g.rows, g.cols, sh), (int)((g.cols + SSW - 1) / SSW), 0, sh};
		const dim3 tg = tgrid(g.rows, g.cols);
		return PostShape{FX_MGK_TILED, tg.x * tg.y, 0, 0, 0};
	};
	auto post = [&](int l, auto epi, RedScratch rs, const PostShape& ps) {
		const fx_grid& g = lv[l].g;
		const fx_grid& cg = lv[l + 1].g;
		if (ps.kind == FX_MGK_WAVE) {
			launch_wpost<VK, NPRE, NPOST>(st, ps.blocks, (const float*)lv[l].rhs, (const float*)lv[l + 1].tmp,
			                              (float*)lv[l].tmp, lwf(mg, lv[l]), lv[l].xt, (int)g.rows, (int)g.cols,
			                              (int)cg.rows, (int)cg.cols, ps.ns, ps.nseg, ps.sh, epi, rs, skip);
		} else if (ps.kind == FX_MGK_RING) {
			hipLaunchKernelGGL((k_spost<VK, NPRE, NPOST, decltype(epi)>), dim3(ps.blocks), dim3(BLOCK), 0, st,
			                   (const float*)lv[l].rhs, (const float*)lv[l + 1].tmp, (float*)lv[l].tmp, lwf(mg, lv[l]),
			                   lv[l].xt, (int)g.rows, (int)g.cols, (int)cg.rows, (int)cg.cols, ps.ns, ps.sh, epi, rs, skip);
		} else {
			const dim3 tg = tgrid(g.rows, g.cols);
			hipLaunchKernelGGL((k_fpost<VK, NPRE, NPOST, decltype(epi)>), tg, dim3(BLOCK), 0, st, (const float*)lv[l].rhs,
			                   (const float*)lv[l + 1].tmp, (float*)lv[l].tmp, lwf(mg, lv[l]), lv[l].xt, (int64_t)g.rows,
			                   (int64_t)g.cols, (int64_t)cg.rows, (int64_t)cg.cols, epi, rs, skip);
		}
	};
	for (int l = lt - 1; l >= 1; --l) post(l, NoEpi{}, RedScratch{nullptr}, post_shape(l));
	const PostShape ps0 = post_shape(0);
	RedScratch rs{nullptr};
	if (Epi::NS > 0) {
		FX_CHECK(fx::ensure_partials(be, (size_t)ps0.blocks * Epi::NS));
		rs = fx::red_scratch(be);
	}
	fx::probe_begin(be, FX_PROBE_VC_POST0);
	post(0, epi0, rs, ps0);
	fx::probe_end(be, FX_PROBE_VC_POST0);
	if constexpr (Epi::NS > 0)
		hipLaunchKernelGGL((k_finish<Epi::NS, typename Epi::Fin>), dim3(1), dim3(FIN_BLOCK), 0, st, rs, ps0.blocks,
		                   epi0.fin(), skip);
	return fx::check_launch("fused V-cycle");
}
// Autotuner of the wave kernels' segment heights: for every level, each
// candidate height is timed with HIP events over `reps` launches of the level's
// kernel on the level's own buffers (the kernels' speed does not depend on the
// values), and the fastest is kept in wrows_post / wrows_restrict; on levels
// ≥ 1 the LDS-tile kernels compete too (WROWS_TILED when they win).  *ms_out
// (if given) receives the sum over levels of the chosen per-launch times.
template <bool VK, int NPRE, int NPOST>
fx_status autotune_t(fx_multigrid* mg, int reps, double* ms_out) {
	fx_backend* be = mg->be;
	hipStream_t st = be->stream;
	auto& lv = mg->lv;
	const int nl = (int)lv.size();
	mg->wrows_post.assign(nl, 0);
	mg->wrows_restrict.assign(nl, 0);
	mg->wrows_restrict_upd = 0;
	mg->wrows_post_epi = 0;
	hipEvent_t e0, e1;
	FX_HIP(hipEventCreate(&e0));
	FX_HIP(hipEventCreate(&e1));
	double total = 0.0;
	fx_status status = FX_OK;
	auto time_it = [&](auto launch, float* ms) -> fx_status {
		launch();
		FX_HIP(hipEventRecord(e0, st));
		for (int i = 0; i < reps; ++i) launch();
		FX_HIP(hipEventRecord(e1, st));
		FX_HIP(hipEventSynchronize(e1));
		FX_HIP(hipEventElapsedTime(ms, e0, e1));
		*ms /= (float)reps;
		return fx::check_launch("multigrid autotune");
	};
	for (int l = 0; l + 1 < nl && status == FX_OK; ++l) {
		const fx_grid& g = lv[l].g;
		const fx_grid& cg = lv[l + 1].g;
		const LWf lw = lwf(mg, lv[l]);
		float best = INFINITY;
		int pick = 0;
		for (int sh = 2; sh <= 512 && status == FX_OK; sh *= 2) {
			if (sh > 2 && sh / 2 >= g.rows) break;
			const int64_t ns = wave_strips_post(g.cols), nseg = (g.rows + sh - 1) / sh;
			float ms = 0.0f;
			status = time_it([&] {
				launch_wpost<VK, NPRE, NPOST>(st, wave_blocks(ns * nseg), (const float*)lv[l].rhs,
				                              (const float*)lv[l + 1].tmp, (float*)lv[l].tmp, lw, lv[l].xt, (int)g.rows,
				                              (int)g.cols, (int)cg.rows, (int)cg.cols, (int)ns, (int)nseg, sh, NoEpi{},
				                              RedScratch{nullptr}, nullptr);
			}, &ms);
			if (ms < best) best = ms, pick = sh;
		}
		if (l > 0 && status == FX_OK) {
			// Below level 0 a level's data sits in the caches and a wave's walk down its segment is
			// latency-bound; the LDS-tile kernel issues all its loads at once and may win there.
			float ms = 0.0f;
			status = time_it([&] {
				hipLaunchKernelGGL((k_fpost<VK, NPRE, NPOST, NoEpi>), tgrid(g.rows, g.cols), dim3(BLOCK), 0, st,
				                   (const float*)lv[l].rhs, (const float*)lv[l + 1].tmp, (float*)lv[l].tmp, lw,
				                   lv[l].xt, (int64_t)g.rows, (int64_t)g.cols, (int64_t)cg.rows, (int64_t)cg.cols,
				                   NoEpi{}, RedScratch{nullptr}, nullptr);
			}, &ms);
			if (ms < best) best = ms, pick = WROWS_TILED;
		}
		mg->wrows_post[l] = pick;
		total += best;
		if (!wrestrict_ok(lv[l].xt)) continue;
		best = INFINITY;
		pick = 0;
		for (int scr = 1; scr <= 256 && status == FX_OK; scr *= 2) {
			if (scr > 1 && scr / 2 >= cg.rows - 2) break;
			const int64_t ns = wave_strips_restrict(cg.cols), nseg = wave_segs_restrict(cg.rows, scr);
			float ms = 0.0f;
			status = time_it([&] {
				launch_wrestrict<VK, NPRE>(st, wave_blocks(ns * nseg), (const float*)lv[l].rhs, (float*)lv[l + 1].rhs, lw,
				                           lv[l].xt, (int)g.rows, (int)g.cols, (int)cg.rows, (int)cg.cols, (int)ns,
				                           (int)nseg, scr, nullptr);
			}, &ms);
			if (ms < best) best = ms, pick = scr;
		}
		if (l > 0 && status == FX_OK) {
			float ms = 0.0f;
			status = time_it([&] {
				hipLaunchKernelGGL((k_frestrict<VK, NPRE>), rgrid(cg.rows, cg.cols), dim3(BLOCK),
				                   (unsigned)frestrict_lds(lv[l].xt, VK), st, (const float*)lv[l].rhs,
				                   (float*)lv[l + 1].rhs, lw, lv[l].xt, (int64_t)g.rows, (int64_t)g.cols,
				                   (int64_t)cg.rows, (int64_t)cg.cols, nullptr);
			}, &ms);
			if (ms < best) best = ms, pick = WROWS_TILED;
		}
		mg->wrows_restrict[l] = pick;
		total += best;
		if (l == 0 && upd_ok(mg)) {
			// Level 0's restriction as the FCG loop runs it, with the residual update (three scratch
			// vectors of the level's size; the bytes per row change, and with them the best height).
			const size_t n0 = (size_t)g.rows * g.cols;
			void* sc = nullptr;
			status = fx::dev_alloc(be, 3 * n0 * sizeof(double), &sc);
			if (status != FX_OK) break;
			const TuneUpd tu{(const double*)sc, (const double*)sc + n0, (double*)sc + 2 * n0};
			float bu = INFINITY;
			for (int scr = 1; scr <= 256 && status == FX_OK; scr *= 2) {
				if (scr > 1 && scr / 2 >= cg.rows - 2) break;
				const int64_t ns = wave_strips_restrict(cg.cols), nseg = wave_segs_restrict(cg.rows, scr);
				const unsigned blocks = wave_blocks(ns * nseg);
				status = fx::ensure_partials(be, blocks);
				if (status != FX_OK) break;
				float ms = 0.0f;
				status = time_it([&] {
					launch_wrestrict<VK, NPRE>(st, blocks, (const float*)lv[l].rhs, (float*)lv[l + 1].rhs, lw, lv[l].xt,
					                           (int)g.rows, (int)g.cols, (int)cg.rows, (int)cg.cols, (int)ns, (int)nseg,
					                           scr, nullptr, tu, fx::red_scratch(be));
				}, &ms);
				if (ms < bu) bu = ms, mg->wrows_restrict_upd = scr;
			}
			// Level 0's post kernel as the FCG loop runs it, with the solver's epilogue traffic.
			const TuneEpi te{(const double*)sc, (const double*)sc + n0};
			bu = INFINITY;
			for (int sh = 2; sh <= 512 && status == FX_OK; sh *= 2) {
				if (sh > 2 && sh / 2 >= g.rows) break;
				const int64_t ns = wave_strips_post(g.cols), nseg = (g.rows + sh - 1) / sh;
				const unsigned blocks = wave_blocks(ns * nseg);
				status = fx::ensure_partials(be, (size_t)blocks * TuneEpi::NS);
				if (status != FX_OK) break;
				float ms = 0.0f;
				status = time_it([&] {
					launch_wpost<VK, NPRE, NPOST>(st, blocks, (const float*)lv[l].rhs, (const float*)lv[l + 1].tmp,
					                              (float*)lv[l].tmp, lw, lv[l].xt, (int)g.rows, (int)g.cols, (int)cg.rows,
					                              (int)cg.cols, (int)ns, (int)nseg, sh, te, fx::red_scratch(be), nullptr);
				}, &ms);
				if (ms < bu) bu = ms, mg->wrows_post_epi = sh;
			}
			(void)hipStreamSynchronize(st);
			fx::dev_free(sc);
		}
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (ms_out) *ms_out = total;
	return status;
}
inline fx_status autotune(fx_multigrid* mg, int reps, double* ms_out) {
	const int a = mg->opt.n_pre, b = mg->opt.n_post;
	const bool vk = mg->lv[0].k != nullptr;
#define FMG_TUNE(A, B)                                                                       \
	if (a == A && b == B) return vk ? autotune_t<true, A, B>(mg, reps, ms_out)               \
	                                : autotune_t<false, A, B>(mg, reps, ms_out);
	FMG_TUNE(2, 2)
	FMG_TUNE(1, 1)
	FMG_TUNE(1, 2)
	FMG_TUNE(2, 1)
#undef FMG_TUNE
	return FX_ERR_UNSUPPORTED;
}

template <bool VK, class Epi, class Upd>
fx_status fast_cycle_v(fx_multigrid* mg, const int* skip, Epi epi0, Upd upd0) {
	const int a = mg->opt.n_pre, b = mg->opt.n_post;
	if (a == 2 && b == 2) return fast_cycle_t<VK, 2, 2>(mg, skip, epi0, upd0);
	if (a == 1 && b == 1) return fast_cycle_t<VK, 1, 1>(mg, skip, epi0, upd0);
	if (a == 1 && b == 2) return fast_cycle_t<VK, 1, 2>(mg, skip, epi0, upd0);
	return fast_cycle_t<VK, 2, 1>(mg, skip, epi0, upd0);
}

// One fused V-cycle with a zero initial guess.  The level-0 rhs must already be
// in lv[0].rhs (FP32), unless upd0 (Upd::ON, allowed when upd_ok) forms it
// from the FCG residual update in the level-0 restriction, followed by
// k_finish on upd0.fin(); the result is written to lv[0].tmp, with epi0's
// per-point terms reduced over level 0 (k_finish runs epi0.fin() on the sums).
template <class Epi, class Upd = NoUpd>
fx_status fast_cycle(fx_multigrid* mg, const int* skip, Epi epi0, Upd upd0 = Upd{}) {
	if (mg->lv[0].k) return fast_cycle_v<true>(mg, skip, epi0, upd0);
	return fast_cycle_v<false>(mg, skip, epi0, upd0);
}

}  // namespace fmg
}  // namespace
