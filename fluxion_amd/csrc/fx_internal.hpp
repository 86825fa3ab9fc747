// fx_internal.hpp — host-side internals shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fluxion_hip.h"
#include "fx_device.hpp"

// HipBackend analogue (backend-hip/src/backend.rs:61-79): one device, one
// in-order stream, reduction scratch and a pinned mirror for host results.
struct fx_backend {
	int device = 0;
	hipStream_t stream = nullptr;
	int num_cu = 0;
	char name[96] = {0};
	fxd::Acc* partials = nullptr;   // reduction partials
	size_t partials_cap = 0;        // in Acc entries
	double* dres = nullptr;         // device result slots for host-returning reductions
	double* hres = nullptr;         // pinned host mirror of dres
	// Kernel probe (fx_probe_set): event pairs around launches of one family.
	int probe_kind = 0;
	std::vector<hipEvent_t> probe_a, probe_b;
	size_t probe_used = 0;
	std::vector<float> probe_ms;
	// This device's address of the static scratch behind fx_dot_noalloc (set up at creation).
	void* noalloc_scratch = nullptr;
};

// HipBuffer analogue (backend-hip/src/buffer.rs:16-123).
struct fx_buf {
	fx_backend* be = nullptr;
	void* ptr = nullptr;
	size_t len = 0;
	fx_dtype dtype = FX_F64;
};

namespace fx {

void set_error(const char* fmt, ...);

#define FX_HIP(expr)                                                                   \
	do {                                                                               \
		hipError_t e_ = (expr);                                                        \
		if (e_ != hipSuccess) {                                                        \
			fx::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
			return FX_ERR_HIP;                                                         \
		}                                                                              \
	} while (0)

#define FX_CHECK(expr)                  \
	do {                                \
		fx_status s_ = (expr);          \
		if (s_ != FX_OK) return s_;     \
	} while (0)

#define FX_REQUIRE(cond, code, ...)     \
	do {                                \
		if (!(cond)) {                  \
			fx::set_error(__VA_ARGS__); \
			return code;                \
		}                               \
	} while (0)

// Checks the launch that was just issued (the reference ignores this status:
// backend-hip/src/backend.rs:293-307).
fx_status check_launch(const char* what);

// Grows the reduction partials buffer to hold `n` Acc entries.
fx_status ensure_partials(fx_backend* be, size_t n);
// Event records around a launch of probe family `kind` (no-ops unless selected).
void probe_begin(fx_backend* be, int kind);
void probe_end(fx_backend* be, int kind);
fx_status probe_drain(fx_backend* be);
inline fxd::RedScratch red_scratch(fx_backend* be) { return fxd::RedScratch{be->partials}; }

// Device allocation on the backend's device; zero-filled on the stream.
fx_status dev_alloc(fx_backend* be, size_t bytes, void** out);
void dev_free(void* p);

// One explicit diffusion stage y = x + coeff·S(x) inside, x on the boundary, then
// y += dt·src (src may be null) and y = (y + avg)·0.5 (avg may be null), in one
// pass of the 5-point streaming kernel (or the generic one for other stencils).
fx_status explicit_stage(fx_backend* be, const double* x, double* y, double coeff, const double* src, double dt,
                         const double* avg, const fx_stencil* st, int64_t rows, int64_t cols);

// One classical RK4 step of ∂u/∂t = α·S(u) (+ source) in four stencil passes and
// three BC-perimeter passes, u in place; k holds five scratch vectors of u's size.
fx_status rk4_passes(fx_backend* be, double* u, double* const k[5], const double* src, double dt, double alpha,
                     const fx_stencil* st, const fx_boundaries* bc, const fx_grid* g);

// Validation helpers.
fx_status check_grid(const fx_grid* g);
fx_status check_buf(const fx_buf* b, size_t len, fx_dtype dt, const char* name);
fx_status check_bc(const fx_boundaries* bc);
fx_status stencil_args(const fx_stencil* st, fxd::W9* w9, fxd::Ent* ent);

// Stencil launch with the default streaming shape.
inline dim3 stream_grid(int64_t rows, int64_t cols, int rpb) {
	return dim3((unsigned)((cols + fxd::BLOCK - 1) / fxd::BLOCK), (unsigned)((rows + rpb - 1) / rpb));
}
inline unsigned elem_blocks(size_t n, int per_thread = 4) {
	size_t b = (n + (size_t)fxd::BLOCK * per_thread - 1) / ((size_t)fxd::BLOCK * per_thread);
	if (b < 1) b = 1;
	if (b > 65535u * 16u) b = 65535u * 16u;
	return (unsigned)b;
}
// Fixed reduction grid for a length (deterministic for a given n).
inline unsigned red_blocks(size_t n) {
	size_t b = (n + (size_t)fxd::BLOCK * 8 - 1) / ((size_t)fxd::BLOCK * 8);
	if (b < 1) b = 1;
	if (b > 2048) b = 2048;
	return (unsigned)b;
}

// ── internal kernels exported across translation units ────────────────
// Device-pointer level operations used by the solvers and the multigrid.
fx_status k_copy(fx_backend* be, const double* src, double* dst, size_t n);
fx_status k_fill(fx_backend* be, double* dst, double v, size_t n);
fx_status k_fill_f32(fx_backend* be, float* dst, float v, size_t n);
// y = A·x for an operator (fx_operator semantics); identity rows on the boundary.
fx_status k_apply_operator(fx_backend* be, const fx_operator* op, const double* x, double* y, const fx_grid* g,
                           const int* skip = nullptr);
// Finds this device's address of fx_dot_noalloc's static scratch (once per device, with the
// backend's device current) and stores it in be->noalloc_scratch.
fx_status noalloc_setup(fx_backend* be);
// Host-returning dot of device arrays (synchronises).
fx_status k_dot_host(fx_backend* be, const double* x, const double* y, size_t n, double* out);

// Network inference on a device FP32 input [B][H][W][cin]; returns the FP32 output
// ([B][H][W], channel 0 of the output layer) in *y.  All kernels return early while *skip != 0.
fx_status unet_infer(fx_unet* u, const float* x, int B, int H, int W, const int* skip, const float** y);

}  // namespace fx
