// fx_mg.hpp — multigrid hierarchy shared between fx_mg.hip and the solvers.
#pragma once
#include "fx_internal.hpp"

struct fx_multigrid {
	struct Level {
		fx_grid g;
		void* u = nullptr;     // T[n]
		void* rhs = nullptr;   // T[n]
		void* tmp = nullptr;   // T[n]
		void* w0 = nullptr;    // Chebyshev work (T[n])
		void* w1 = nullptr;
		void* k = nullptr;     // nodal coefficient (T[n]) of a variable-coefficient hierarchy
		double w[5];           // sigma·I + scale·Δ_h weights, laplacian_2d_5pt order
		double inv_diag;
	};
	fx_backend* be = nullptr;
	fx_mg_options opt{};
	fx_boundaries bc{};
	bool f32 = false;
	std::vector<Level> lv;
};

namespace fx {
// z = M·r with z's initial guess taken from z (the solvers zero it first).
// Kernels return early while *skip != 0 (skip may be null).
fx_status mg_apply(fx_multigrid* mg, double* z, const double* r, const int* skip);
}
