// fx_mg.hpp — multigrid hierarchy shared between fx_mg.hip and the solvers.
#pragma once
#include "fx_internal.hpp"

// Linear same-domain transfer from one level to the next coarser one,
// tabulated per axis on the host so that kernels do no integer division.
// Fine index i sits at coarse coordinate i·(nc−1)/(nf−1) = i0[i] + w[i];
// coarse index I gathers the fine interior indices lo[I]..hi[I].
struct fx_xtab {
	const int32_t* i0r = nullptr;   // [fine rows]
	const double* wr = nullptr;     // [fine rows]
	const int32_t* i0c = nullptr;   // [fine cols]
	const double* wc = nullptr;     // [fine cols]
	const int32_t* lor = nullptr;   // [coarse rows]
	const int32_t* hir = nullptr;
	const int32_t* loc = nullptr;   // [coarse cols]
	const int32_t* hic = nullptr;
	float sy = 0.0f, sx = 0.0f;     // float((nc−1)/(nf−1)) per axis: the restriction scale (h/H)
	int ry = 0, rx = 0;             // LDS region (halo included) of the fused restriction's widest tile
	int gmax = 0;                   // most fine columns any coarse column gathers from
	int gmaxr = 0;                  // most fine rows any coarse row gathers from
	int wx = 0;                     // fine window (halo included) of the wave restriction's widest strip
	int adv = 0;                    // widest fine-column advance between the wave restriction's consecutive strips
};

struct fx_multigrid {
	struct Level {
		fx_grid g;
		void* u = nullptr;     // T[n]
		void* rhs = nullptr;   // T[n]
		void* tmp = nullptr;   // T[n]
		void* w0 = nullptr;    // Chebyshev work (T[n])
		void* w1 = nullptr;
		void* k = nullptr;     // nodal coefficient (T[n]) of a variable-coefficient hierarchy
		void* vx = nullptr;    // nodal velocity (T[n] each) of a convection-diffusion hierarchy
		void* vy = nullptr;
		void* kx = nullptr;    // FP32 fused hierarchies: harmonic-mean face coefficient east of each point
		void* ky = nullptr;    //   and south of each point (float[n]; 0 past the last column / row)
		void* wd = nullptr;    //   and ω/diagonal of the level operator (float[n]; 0 on the boundary)
		void* kp = nullptr;    //   the same three packed per point, {kx, ky, ω/d} (float[3n]): one load per row
		double w[5];           // sigma·I + scale·Δ_h weights, laplacian_2d_5pt order
		double inv_diag;
		fx_xtab xt;            // transfer to the next coarser level (linear transfers only)
		void* xbuf = nullptr;  // device storage behind xt
	};
	fx_backend* be = nullptr;
	fx_mg_options opt{};
	fx_boundaries bc{};
	bool f32 = false;
	// Fused-cycle level kernels: FX_MGK_WAVE (wave-streaming, default), FX_MGK_RING (LDS-ring
	// streaming) or FX_MGK_TILED; levels a variant cannot run fall back to the tiled kernels.
	int kernels = 2;
	std::vector<int> wrows_post, wrows_restrict;   // tuned wave segment heights per level (0: heuristic;
	                                               //   −1, levels ≥ 1: the LDS-tile kernel instead)
	int wrows_restrict_upd = 0;   // level 0's restriction carrying the FCG residual update (0: as wrows_restrict)
	int wrows_post_epi = 0;       // level 0's post kernel with a solver epilogue (0: as wrows_post)
	std::vector<Level> lv;
};

namespace fx {
// z = M·r with z's initial guess taken from z (the solvers zero it first).
// Kernels return early while *skip != 0 (skip may be null).
fx_status mg_apply(fx_multigrid* mg, double* z, const double* r, const int* skip);
}
