// fx_implicit.hip — callers either side of the solve: the Crank–Nicolson
// stepper (crates/runtime/src/implicit.rs:1-238), IMEX (imex.rs:1-274), BDF1–4 (bdf.rs:1-256), the explicit diffusion
// stepper (crates/runtime/src/solver.rs:1-494) and the raw field format
// (crates/io/src/raw.rs:1-77).
//
// The stepper is host code over the library's own entry points: one operator
// pass assembles the right-hand side (the reference's copy + copy +
// stencil_axpy, implicit.rs:187-194, in one read of u and one write), an axpy
// adds the source, and the solver of the step runs from u^n as its initial
// guess, so every step is warm-started as in the reference.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <new>
#include <sys/stat.h>

#include "fx_internal.hpp"

struct fx_cn {
	fx_backend* be = nullptr;
	fx_grid g{};
	double alpha = 0, dt = 0;
	uint64_t steps = 0;
	fx_boundaries bc{};
	fx_stencil_entry ent[5]{};
	fx_stencil st{};
	int32_t op_kind = FX_OP_STENCIL;
	int32_t solver_kind = FX_SOLVER_CG;
	const fx_buf* k = nullptr;
	fx_solver* solver = nullptr;
	fx_multigrid* mg = nullptr;
	fx_buf* rhs = nullptr;
	fx_precond user_pc{};      // fx_cn_set_preconditioner: used instead of the stepper's multigrid
	bool has_user_pc = false;
};

// DiffusionSolver (solver.rs:25-48).  Explicit steps write a second buffer
// instead of copying u first: Euler and the first Heun stage go u → s, the
// second Heun stage s → u, and step_n pairs Euler steps u → s → u, so no step
// copies data except a lone Euler step, which copies s back to u.  RK4 keeps
// u as u_old until its last pass and its stages in two of its five scratch
// vectors (the other three hold k1 … k3).  The graph
// of step_n holds many units (graph_units), replayed while whole graphs fit.
struct fx_diffusion {
	fx_backend* be = nullptr;
	fx_grid g{};
	double alpha = 0, dt = 0;
	int32_t scheme = FX_SCHEME_EULER;
	uint64_t steps = 0;
	fx_boundaries bc{};
	fx_stencil_entry ent[5]{};
	fx_stencil st{};
	std::vector<fx_buf*> scratch;
	bool use_graph = false;
	// Captured HIP graph of graph_units() step_n units, for the u it was captured on.
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	void* graph_u = nullptr;
};

// ImexSolver (imex.rs:10-45): a Crank–Nicolson step in κ whose source is the
// explicit upwind convection v·∇u of the current state (apply_conv_diff with a
// zero κ field: imex.rs:215-274), i.e. the reference's step sequence
// (imex.rs:147-197) as conv_diff, then fx_cn_step(u, conv).
struct fx_imex {
	fx_backend* be = nullptr;
	fx_grid g{};
	fx_boundaries bc{};
	fx_cn* cn = nullptr;
	fx_buf *conv = nullptr, *kzero = nullptr, *vx = nullptr, *vy = nullptr;
};

// BdfSolver (bdf.rs:9-37): BDF1–4 with CG or MG-PCG per step.  The history is a
// ring of buffers: the reference's shift (history[i] = history[i − 1], then
// history[0] = u) is a rotation of the ring plus one copy of u.
struct fx_bdf {
	fx_backend* be = nullptr;
	fx_grid g{};
	double alpha = 0, dt = 0;
	int order = 1;
	uint64_t steps = 0;
	fx_boundaries bc{};
	fx_stencil_entry ent[5]{};
	fx_stencil st{};
	fx_solver* solver = nullptr;
	fx_multigrid* mg = nullptr;
	fx_buf* rhs = nullptr;
	std::vector<fx_buf*> hist;   // hist[0] = u^n, hist[1] = u^{n−1}, …
};

namespace {

// BDF coefficients (bdf.rs:41-56): history weights of u^n, u^{n−1}, … and the lhs weight.
void bdf_coeffs(int order, double* h, double* lhs) {
	switch (order) {
	case 1: h[0] = -1.0; *lhs = 1.0; break;
	case 2: h[0] = -2.0; h[1] = 0.5; *lhs = 1.5; break;
	case 3: h[0] = -3.0; h[1] = 1.5; h[2] = -1.0 / 3.0; *lhs = 11.0 / 6.0; break;
	default: h[0] = -4.0; h[1] = 3.0; h[2] = -4.0 / 3.0; h[3] = 0.25; *lhs = 25.0 / 12.0; break;
	}
}

// rhs = (u·(−h0) − h1·u^{n−1} − h2·u^{n−2} − h3·u^{n−3})·(1/lhs): the reference's copy, scale, axpys and
// scale (bdf.rs:189-206) per point, in their order, in one pass.
__global__ void k_bdf_rhs(const double* __restrict__ u, const double* __restrict__ h1, const double* __restrict__ h2,
                          const double* __restrict__ h3, int nh, double m0, double m1, double m2, double m3, double inv,
                          double* __restrict__ rhs, int64_t n) {
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
		double t = u[i] * m0;
		if (nh > 0) t = fma(m1, h1[i], t);
		if (nh > 1) t = fma(m2, h2[i], t);
		if (nh > 2) t = fma(m3, h3[i], t);
		rhs[i] = t * inv;
	}
}

void bdf_free(fx_bdf* b) {
	if (!b) return;
	if (b->solver) fx_solver_destroy(b->solver);
	if (b->mg) fx_multigrid_destroy(b->mg);
	if (b->rhs) fx_buffer_free(b->rhs);
	for (fx_buf* h : b->hist) fx_buffer_free(h);
	delete b;
}

// Stencil::laplacian_2d_5pt (core/src/stencil.rs:56-72), entry order kept.
void laplacian_5pt(double dx, double dy, fx_stencil_entry* e) {
	const double inv_dx2 = 1.0 / (dx * dx);
	const double inv_dy2 = 1.0 / (dy * dy);
	e[0] = {0, 0, -2.0 * (inv_dx2 + inv_dy2)};
	e[1] = {-1, 0, inv_dy2};
	e[2] = {1, 0, inv_dy2};
	e[3] = {0, -1, inv_dx2};
	e[4] = {0, 1, inv_dx2};
}

// The step's two operators.  With h = (0.5·dt)·α (implicit.rs:183, evaluated
// left to right as the reference writes it):
//   stencil form: explicit u + h·S(u), implicit u − h·S(u)  (A·v = v + coeff·S(v),
//                 identity rows on the boundary: cg.rs:104-111)
//   variable k:   explicit u − h·(−∇·(k∇u)), implicit u + h·(−∇·(k∇u))
fx_operator cn_operator(const fx_cn* cn, bool implicit_side) {
	const double h = 0.5 * cn->dt * cn->alpha;
	fx_operator op{};
	op.kind = cn->op_kind;
	op.sigma = 1.0;
	if (cn->op_kind == FX_OP_STENCIL) {
		op.coeff = implicit_side ? -h : h;
		op.stencil = &cn->st;
		op.bc = &cn->bc;
	} else {
		op.coeff = implicit_side ? h : -h;
		op.k = cn->k;
	}
	return op;
}

void cn_free(fx_cn* cn) {
	if (!cn) return;
	if (cn->solver) fx_solver_destroy(cn->solver);
	if (cn->mg) fx_multigrid_destroy(cn->mg);
	if (cn->rhs) fx_buffer_free(cn->rhs);
	delete cn;
}

// ── DiffusionSolver steps (solver.rs:339-494) ────────────────────────
fx_status euler_stage(fx_diffusion* d, const fx_buf* x, fx_buf* y, const fx_buf* src) {
	// u += α·dt·Δ(u) + dt·source (solver.rs:344-367), written to y.
	return fx::explicit_stage(d->be, (const double*)x->ptr, (double*)y->ptr, d->alpha * d->dt,
	                          src ? (const double*)src->ptr : nullptr, d->dt, nullptr, &d->st, (int64_t)d->g.rows,
	                          (int64_t)d->g.cols);
}

// Heun (solver.rs:371-418): u* = u + dt·α·Δu (+ dt·s); u** = u* + dt·α·Δu* (+ dt·s); u = (u** + u)·0.5.
fx_status rk2_step(fx_diffusion* d, fx_buf* u, const fx_buf* src) {
	const double coeff = d->dt * d->alpha;
	const double* sp = src ? (const double*)src->ptr : nullptr;
	fx_buf* s = d->scratch[0];
	const int64_t rows = (int64_t)d->g.rows, cols = (int64_t)d->g.cols;
	FX_CHECK(fx::explicit_stage(d->be, (const double*)u->ptr, (double*)s->ptr, coeff, sp, d->dt, nullptr, &d->st,
	                            rows, cols));
	return fx::explicit_stage(d->be, (const double*)s->ptr, (double*)u->ptr, coeff, sp, d->dt,
	                          (const double*)u->ptr, &d->st, rows, cols);
}

// Classical RK4 (solver.rs:421-494): four fused stencil passes (fx::rk4_passes),
// each point computed by the reference's operations in its order.
fx_status rk4_step(fx_diffusion* d, fx_buf* u, const fx_buf* src) {
	double* k[5];
	for (int i = 0; i < 5; ++i) k[i] = (double*)d->scratch[i]->ptr;
	return fx::rk4_passes(d->be, (double*)u->ptr, k, src ? (const double*)src->ptr : nullptr, d->dt, d->alpha, &d->st,
	                      &d->bc, &d->g);
}

// One step from u back into u.
fx_status diff_step(fx_diffusion* d, fx_buf* u, const fx_buf* src) {
	switch (d->scheme) {
	case FX_SCHEME_EULER:
		FX_CHECK(euler_stage(d, u, d->scratch[0], src));
		return fx_copy(d->be, d->scratch[0], u);
	case FX_SCHEME_RK2: return rk2_step(d, u, src);
	default: return rk4_step(d, u, src);
	}
}

// step_n's unit: two Euler steps u → s → u, or one RK step.
fx_status diff_unit(fx_diffusion* d, fx_buf* u, const fx_buf* src) {
	if (d->scheme == FX_SCHEME_EULER) {
		FX_CHECK(euler_stage(d, u, d->scratch[0], src));
		return euler_stage(d, d->scratch[0], u, src);
	}
	return diff_step(d, u, src);
}
int unit_steps(const fx_diffusion* d) { return d->scheme == FX_SCHEME_EULER ? 2 : 1; }
// Units per captured graph: about 64 kernels, so that one graph launch (which
// costs more than a kernel launch on this runtime) is spread over many steps.
int graph_units(const fx_diffusion* d) {
	return d->scheme == FX_SCHEME_EULER ? 32 : d->scheme == FX_SCHEME_RK2 ? 32 : 8;
}

void drop_graph(fx_diffusion* d) {
	if (d->exec) (void)hipGraphExecDestroy(d->exec);
	if (d->graph) (void)hipGraphDestroy(d->graph);
	d->exec = nullptr;
	d->graph = nullptr;
	d->graph_u = nullptr;
}

// Captures step_n's unit on the backend's stream (the reference captures one
// step: solver.rs:259-303) for this u; a different u is captured anew.
fx_status capture_unit(fx_diffusion* d, fx_buf* u) {
	if (d->exec && d->graph_u == u->ptr) return FX_OK;
	drop_graph(d);
	hipStream_t st = d->be->stream;
	FX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
	fx_status s = FX_OK;
	for (int i = 0; i < graph_units(d) && s == FX_OK; ++i) s = diff_unit(d, u, nullptr);
	hipGraph_t g = nullptr;
	hipError_t e = hipStreamEndCapture(st, &g);
	if (s != FX_OK) {
		if (g) (void)hipGraphDestroy(g);
		return s;
	}
	if (e != hipSuccess) {
		fx::set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
		return FX_ERR_HIP;
	}
	d->graph = g;
	e = hipGraphInstantiate(&d->exec, g, nullptr, nullptr, 0);
	if (e != hipSuccess) {
		drop_graph(d);
		fx::set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
		return FX_ERR_HIP;
	}
	d->graph_u = u->ptr;
	return FX_OK;
}

void diff_free(fx_diffusion* d) {
	if (!d) return;
	drop_graph(d);
	for (fx_buf* b : d->scratch) fx_buffer_free(b);
	delete d;
}

// ── raw field format (io/src/raw.rs:7-77) ────────────────────────────
constexpr size_t RAW_HEADER = 32;

void put_le(unsigned char* p, uint64_t v) {
	for (int i = 0; i < 8; ++i) p[i] = (unsigned char)(v >> (8 * i));
}
uint64_t get_le(const unsigned char* p) {
	uint64_t v = 0;
	for (int i = 0; i < 8; ++i) v |= (uint64_t)p[i] << (8 * i);
	return v;
}
uint64_t f64_bits(double d) {
	uint64_t u;
	std::memcpy(&u, &d, 8);
	return u;
}
double bits_f64(uint64_t u) {
	double d;
	std::memcpy(&d, &u, 8);
	return d;
}

struct File {
	FILE* f = nullptr;
	~File() {
		if (f) std::fclose(f);
	}
};

fx_status raw_write(const char* path, const double* data, const fx_grid* g) {
	FX_REQUIRE(path && data && g, FX_ERR_ARGUMENT, "null argument");
	File f;
	f.f = std::fopen(path, "wb");
	FX_REQUIRE(f.f, FX_ERR_IO, "cannot create %s: %s", path, std::strerror(errno));
	unsigned char h[RAW_HEADER];
	put_le(h, g->rows);
	put_le(h + 8, g->cols);
	put_le(h + 16, f64_bits(g->dx));
	put_le(h + 24, f64_bits(g->dy));
	FX_REQUIRE(std::fwrite(h, 1, RAW_HEADER, f.f) == RAW_HEADER, FX_ERR_IO, "write to %s failed", path);
	const size_t n = (size_t)(g->rows * g->cols);
	std::vector<unsigned char> buf(8 * (size_t)65536);
	for (size_t i0 = 0; i0 < n; i0 += 65536) {
		const size_t m = std::min<size_t>(65536, n - i0);
		for (size_t i = 0; i < m; ++i) put_le(buf.data() + 8 * i, f64_bits(data[i0 + i]));
		FX_REQUIRE(std::fwrite(buf.data(), 8, m, f.f) == m, FX_ERR_IO, "write to %s failed", path);
	}
	FX_REQUIRE(std::fclose(f.f) == 0, FX_ERR_IO, "closing %s failed", path);
	f.f = nullptr;
	return FX_OK;
}

// Header, validated as Grid::new validates it (grid.rs:15-36); then, if data is
// not null, exactly rows·cols values (a short file is an error, as read_exact's is).
fx_status raw_read(const char* path, fx_grid* g, double* data, size_t len) {
	FX_REQUIRE(path && g, FX_ERR_ARGUMENT, "null argument");
	File f;
	f.f = std::fopen(path, "rb");
	FX_REQUIRE(f.f, FX_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
	unsigned char h[RAW_HEADER];
	FX_REQUIRE(std::fread(h, 1, RAW_HEADER, f.f) == RAW_HEADER, FX_ERR_IO, "%s: file shorter than its header",
	           path);
	fx_grid r{get_le(h), get_le(h + 8), bits_f64(get_le(h + 16)), bits_f64(get_le(h + 24))};
	FX_CHECK(fx::check_grid(&r));
	FX_REQUIRE(r.rows <= (uint64_t)1 << 31 && r.cols <= (uint64_t)1 << 31, FX_ERR_GRID,
	           "%s: grid %llux%llu is too large", path, (unsigned long long)r.rows, (unsigned long long)r.cols);
	// The file must hold all rows·cols values before anything is sized from the header, so that a
	// truncated or corrupt header is an I/O error rather than a huge allocation.
	const size_t n = (size_t)(r.rows * r.cols);
	struct stat sb;
	FX_REQUIRE(fstat(fileno(f.f), &sb) == 0, FX_ERR_IO, "cannot stat %s: %s", path, std::strerror(errno));
	const uint64_t held = (uint64_t)sb.st_size >= RAW_HEADER ? ((uint64_t)sb.st_size - RAW_HEADER) / 8 : 0;
	FX_REQUIRE(held >= (uint64_t)n, FX_ERR_IO, "%s: file holds fewer than %zu values", path, n);
	*g = r;
	if (!data) return FX_OK;
	FX_REQUIRE(len == n, FX_ERR_DIMENSION, "dimension mismatch: expected %zu, got %zu (data)", n, len);
	std::vector<unsigned char> buf(8 * (size_t)65536);
	for (size_t i0 = 0; i0 < n; i0 += 65536) {
		const size_t m = std::min<size_t>(65536, n - i0);
		FX_REQUIRE(std::fread(buf.data(), 8, m, f.f) == m, FX_ERR_IO, "%s: file holds fewer than %zu values", path,
		           n);
		for (size_t i = 0; i < m; ++i) data[i0 + i] = bits_f64(get_le(buf.data() + 8 * i));
	}
	return FX_OK;
}

}  // namespace

extern "C" {

fx_status fx_cn_create(fx_backend* be, const fx_grid* grid, double alpha, double dt, const fx_boundaries* bc,
                       double cg_tol, uint64_t cg_max_iter, const fx_cn_options* opt, fx_cn** out) {
	FX_REQUIRE(be && bc && out, FX_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	FX_CHECK(fx::check_grid(grid));
	FX_CHECK(fx::check_bc(bc));
	FX_REQUIRE(std::isfinite(alpha) && std::isfinite(dt), FX_ERR_ARGUMENT, "alpha and dt must be finite");
	fx_cn_options o{};
	o.op = FX_OP_STENCIL;
	o.solver = FX_SOLVER_CG;
	if (opt) o = *opt;
	FX_REQUIRE(o.op == FX_OP_STENCIL || o.op == FX_OP_VARK, FX_ERR_ARGUMENT,
	           "Crank-Nicolson operator must be FX_OP_STENCIL or FX_OP_VARK, got %d", (int)o.op);
	FX_REQUIRE(o.solver == FX_SOLVER_CG || o.solver == FX_SOLVER_FCG, FX_ERR_ARGUMENT,
	           "Crank-Nicolson solver must be FX_SOLVER_CG or FX_SOLVER_FCG, got %d", (int)o.solver);
	FX_REQUIRE(o.solver != FX_SOLVER_FCG || o.mg, FX_ERR_ARGUMENT, "flexible CG needs a multigrid preconditioner");
	const size_t n = (size_t)(grid->rows * grid->cols);
	if (o.op == FX_OP_VARK) FX_CHECK(fx::check_buf(o.k, n, FX_F64, "k"));

	fx_cn* cn = new fx_cn;
	cn->be = be;
	cn->g = *grid;
	cn->alpha = alpha;
	cn->dt = dt;
	cn->bc = *bc;
	laplacian_5pt(grid->dx, grid->dy, cn->ent);
	cn->st = {5, cn->ent};
	cn->op_kind = o.op;
	cn->solver_kind = o.solver;
	cn->k = o.op == FX_OP_VARK ? o.k : nullptr;
	fx_status s = fx_allocate(be, n, FX_F64, &cn->rhs);
	if (s == FX_OK) s = fx_solver_create(be, o.solver, grid, 0, cg_tol, cg_max_iter, &cn->solver);
	if (s == FX_OK && o.mg) s = fx_multigrid_create_ex(be, grid, bc, o.mg, &cn->mg);
	if (s != FX_OK) {
		cn_free(cn);
		return s;
	}
	*out = cn;
	return FX_OK;
}

void fx_cn_destroy(fx_cn* cn) { cn_free(cn); }

fx_status fx_cn_step(fx_cn* cn, fx_buf* u, const fx_buf* source, fx_solve_stats* stats) {
	FX_REQUIRE(cn, FX_ERR_ARGUMENT, "solver is null");
	const size_t n = (size_t)(cn->g.rows * cn->g.cols);
	FX_CHECK(fx::check_buf(u, n, FX_F64, "u"));
	if (source) FX_CHECK(fx::check_buf(source, n, FX_F64, "source"));
	// RHS = (I + (dt/2)·α·Δ) u^n + dt·source   (implicit.rs:185-197)
	const fx_operator ex = cn_operator(cn, false);
	FX_CHECK(fx_apply_operator(cn->be, &ex, u, cn->rhs, &cn->g));
	if (source) FX_CHECK(fx_axpy(cn->be, cn->dt, source, cn->rhs));
	// (I − (dt/2)·α·Δ) u^{n+1} = RHS from u^n   (implicit.rs:199-218)
	const fx_operator im = cn_operator(cn, true);
	fx_precond pc{FX_PC_MULTIGRID, 0, cn->mg};
	const fx_precond* use = cn->has_user_pc ? &cn->user_pc : cn->mg ? &pc : nullptr;
	fx_solve_stats st{};
	FX_CHECK(fx_solver_solve_op(cn->solver, u, cn->rhs, &im, use, &st));
	cn->steps += 1;
	if (stats) *stats = st;
	return FX_OK;
}

fx_status fx_cn_step_n(fx_cn* cn, fx_buf* u, uint64_t n) {
	for (uint64_t i = 0; i < n; ++i) FX_CHECK(fx_cn_step(cn, u, nullptr, nullptr));
	return FX_OK;
}

double fx_cn_dt(const fx_cn* cn) { return cn ? cn->dt : 0.0; }
uint64_t fx_cn_steps_done(const fx_cn* cn) { return cn ? cn->steps : 0; }
double fx_cn_sim_time(const fx_cn* cn) { return cn ? (double)cn->steps * cn->dt : 0.0; }

fx_status fx_cn_set_preconditioner(fx_cn* cn, const fx_precond* pc) {
	FX_REQUIRE(cn, FX_ERR_ARGUMENT, "solver is null");
	if (!pc) {
		cn->has_user_pc = false;
		return FX_OK;
	}
	FX_REQUIRE(pc->kind != FX_PC_LEARNED || cn->solver_kind == FX_SOLVER_FCG, FX_ERR_ARGUMENT,
	           "the learned preconditioner varies from step to step: it needs the flexible CG stepper");
	cn->user_pc = *pc;
	cn->has_user_pc = true;
	return FX_OK;
}

fx_solver* fx_cn_solver(fx_cn* cn) { return cn ? cn->solver : nullptr; }
fx_multigrid* fx_cn_multigrid(fx_cn* cn) { return cn ? cn->mg : nullptr; }

fx_status fx_diffusion_create(fx_backend* be, const fx_grid* grid, double alpha, const double* dt,
                              const fx_boundaries* bc, int32_t scheme, fx_diffusion** out) {
	FX_REQUIRE(be && bc && out, FX_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	FX_CHECK(fx::check_grid(grid));
	FX_CHECK(fx::check_bc(bc));
	FX_REQUIRE(scheme == FX_SCHEME_EULER || scheme == FX_SCHEME_RK2 || scheme == FX_SCHEME_RK4, FX_ERR_ARGUMENT,
	           "unknown time scheme %d", (int)scheme);
	FX_REQUIRE(std::isfinite(alpha) && (!dt || std::isfinite(*dt)), FX_ERR_ARGUMENT, "alpha and dt must be finite");
	fx_diffusion* d = new fx_diffusion;
	d->be = be;
	d->g = *grid;
	d->alpha = alpha;
	d->scheme = scheme;
	d->bc = *bc;
	// dt = safety·h_min²/(4α) unless given; safety 0.4, 0.8, 1.4 (solver.rs:103-113).
	const double h_min = std::min(grid->dx, grid->dy);
	const double safety = scheme == FX_SCHEME_EULER ? 0.4 : scheme == FX_SCHEME_RK2 ? 0.8 : 1.4;
	d->dt = dt ? *dt : safety * h_min * h_min / (4.0 * alpha);
	laplacian_5pt(grid->dx, grid->dy, d->ent);
	d->st = {5, d->ent};
	const int nscr = scheme == FX_SCHEME_RK4 ? 5 : 1;
	for (int i = 0; i < nscr; ++i) {
		fx_buf* b = nullptr;
		fx_status s = fx_allocate(be, (size_t)(grid->rows * grid->cols), FX_F64, &b);
		if (s != FX_OK) {
			diff_free(d);
			return s;
		}
		d->scratch.push_back(b);
	}
	*out = d;
	return FX_OK;
}

void fx_diffusion_destroy(fx_diffusion* d) { diff_free(d); }

fx_status fx_diffusion_enable_graph(fx_diffusion* d) {
	FX_REQUIRE(d, FX_ERR_ARGUMENT, "solver is null");
	d->use_graph = true;
	return FX_OK;
}

fx_status fx_diffusion_step(fx_diffusion* d, fx_buf* u, const fx_buf* source) {
	FX_REQUIRE(d, FX_ERR_ARGUMENT, "solver is null");
	const size_t n = (size_t)(d->g.rows * d->g.cols);
	FX_CHECK(fx::check_buf(u, n, FX_F64, "u"));
	if (source) FX_CHECK(fx::check_buf(source, n, FX_F64, "source"));
	FX_CHECK(diff_step(d, u, source));
	d->steps += 1;
	return FX_OK;
}

static fx_status step_many(fx_diffusion* d, fx_buf* u, uint64_t n, const fx_buf* source, bool graph) {
	FX_REQUIRE(d, FX_ERR_ARGUMENT, "solver is null");
	const size_t len = (size_t)(d->g.rows * d->g.cols);
	FX_CHECK(fx::check_buf(u, len, FX_F64, "u"));
	if (source) FX_CHECK(fx::check_buf(source, len, FX_F64, "source"));
	const uint64_t per = (uint64_t)unit_steps(d);
	uint64_t done = 0;
	const uint64_t gsteps = per * (uint64_t)graph_units(d);
	if (graph && n >= gsteps) {
		FX_CHECK(capture_unit(d, u));
		for (; done + gsteps <= n; done += gsteps) {
			FX_HIP(hipGraphLaunch(d->exec, d->be->stream));
			d->steps += gsteps;
		}
	}
	for (; done + per <= n; done += per) {
		FX_CHECK(diff_unit(d, u, source));
		d->steps += per;
	}
	for (uint64_t i = done; i < n; ++i) {
		FX_CHECK(diff_step(d, u, source));
		d->steps += 1;
	}
	return FX_OK;
}

fx_status fx_diffusion_step_n(fx_diffusion* d, fx_buf* u, uint64_t n) {
	return step_many(d, u, n, nullptr, d && d->use_graph);
}

fx_status fx_diffusion_step_n_with_source(fx_diffusion* d, fx_buf* u, uint64_t n, const fx_buf* source) {
	FX_REQUIRE(source, FX_ERR_ARGUMENT, "source is null");
	return step_many(d, u, n, source, false);
}

double fx_diffusion_dt(const fx_diffusion* d) { return d ? d->dt : 0.0; }
uint64_t fx_diffusion_steps_done(const fx_diffusion* d) { return d ? d->steps : 0; }
double fx_diffusion_sim_time(const fx_diffusion* d) { return d ? (double)d->steps * d->dt : 0.0; }
int32_t fx_diffusion_scheme(const fx_diffusion* d) { return d ? d->scheme : -1; }

fx_status fx_imex_create(fx_backend* be, const fx_grid* grid, double kappa, double vx, double vy, double dt,
                         const fx_boundaries* bc, double cg_tol, uint64_t cg_max_iter, const fx_mg_options* mg,
                         fx_imex** out) {
	FX_REQUIRE(be && bc && out, FX_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	FX_CHECK(fx::check_grid(grid));
	FX_REQUIRE(std::isfinite(vx) && std::isfinite(vy), FX_ERR_ARGUMENT, "velocity must be finite");
	fx_imex* m = new fx_imex;
	m->be = be;
	m->g = *grid;
	m->bc = *bc;
	auto done = [&](fx_status s) {
		if (s == FX_OK) {
			*out = m;
			return s;
		}
		fx_imex_destroy(m);
		return s;
	};
	fx_cn_options o{FX_OP_STENCIL, FX_SOLVER_CG, nullptr, mg};
	fx_status s = fx_cn_create(be, grid, kappa, dt, bc, cg_tol, cg_max_iter, &o, &m->cn);
	const size_t n = (size_t)(grid->rows * grid->cols);
	if (s == FX_OK) s = fx_allocate(be, n, FX_F64, &m->conv);
	if (s == FX_OK) s = fx_allocate(be, n, FX_F64, &m->kzero);
	if (s == FX_OK) s = fx_allocate(be, n, FX_F64, &m->vx);
	if (s == FX_OK) s = fx_allocate(be, n, FX_F64, &m->vy);
	if (s == FX_OK) s = fx_fill(be, m->vx, vx);
	if (s == FX_OK) s = fx_fill(be, m->vy, vy);
	return done(s);
}

void fx_imex_destroy(fx_imex* m) {
	if (!m) return;
	if (m->cn) fx_cn_destroy(m->cn);
	for (fx_buf* b : {m->conv, m->kzero, m->vx, m->vy})
		if (b) fx_buffer_free(b);
	delete m;
}

fx_status fx_imex_step(fx_imex* m, fx_buf* u, fx_solve_stats* stats) {
	FX_REQUIRE(m, FX_ERR_ARGUMENT, "solver is null");
	FX_CHECK(fx_apply_conv_diff(m->be, u, m->conv, m->kzero, m->vx, m->vy, &m->g, &m->bc));
	return fx_cn_step(m->cn, u, m->conv, stats);
}

fx_status fx_imex_step_n(fx_imex* m, fx_buf* u, uint64_t n) {
	for (uint64_t i = 0; i < n; ++i) FX_CHECK(fx_imex_step(m, u, nullptr));
	return FX_OK;
}

double fx_imex_dt(const fx_imex* m) { return m ? fx_cn_dt(m->cn) : 0.0; }
uint64_t fx_imex_steps_done(const fx_imex* m) { return m ? fx_cn_steps_done(m->cn) : 0; }
double fx_imex_sim_time(const fx_imex* m) { return m ? fx_cn_sim_time(m->cn) : 0.0; }

fx_status fx_bdf_create(fx_backend* be, const fx_grid* grid, double alpha, double dt, uint32_t order,
                        const fx_boundaries* bc, double cg_tol, uint64_t cg_max_iter, const fx_mg_options* mg,
                        fx_bdf** out) {
	FX_REQUIRE(be && bc && out, FX_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	FX_CHECK(fx::check_grid(grid));
	FX_CHECK(fx::check_bc(bc));
	FX_REQUIRE(order >= 1 && order <= 4, FX_ERR_ARGUMENT, "BDF order must be 1-4, got %u", order);
	FX_REQUIRE(std::isfinite(alpha) && std::isfinite(dt), FX_ERR_ARGUMENT, "alpha and dt must be finite");
	fx_bdf* b = new fx_bdf;
	b->be = be;
	b->g = *grid;
	b->alpha = alpha;
	b->dt = dt;
	b->order = (int)order;
	b->bc = *bc;
	laplacian_5pt(grid->dx, grid->dy, b->ent);
	b->st = {5, b->ent};
	const size_t n = (size_t)(grid->rows * grid->cols);
	fx_status s = fx_allocate(be, n, FX_F64, &b->rhs);
	for (uint32_t i = 0; i < order && s == FX_OK; ++i) {
		fx_buf* h = nullptr;
		s = fx_allocate(be, n, FX_F64, &h);
		if (s == FX_OK) b->hist.push_back(h);
	}
	if (s == FX_OK) s = fx_solver_create(be, FX_SOLVER_CG, grid, 0, cg_tol, cg_max_iter, &b->solver);
	if (s == FX_OK && mg) s = fx_multigrid_create_ex(be, grid, bc, mg, &b->mg);
	if (s != FX_OK) {
		bdf_free(b);
		return s;
	}
	*out = b;
	return FX_OK;
}

void fx_bdf_destroy(fx_bdf* b) { bdf_free(b); }

fx_status fx_bdf_step(fx_bdf* b, fx_buf* u, fx_solve_stats* stats) {
	FX_REQUIRE(b, FX_ERR_ARGUMENT, "solver is null");
	const size_t n = (size_t)(b->g.rows * b->g.cols);
	FX_CHECK(fx::check_buf(u, n, FX_F64, "u"));
	// Effective order min(order, steps_done + 1) bootstraps the history (bdf.rs:176-179).
	const int eff = (int)std::min<uint64_t>((uint64_t)b->order, b->steps + 1);
	double h[4] = {0, 0, 0, 0}, lhs = 1.0;
	bdf_coeffs(eff, h, &lhs);
	const double* hp[3] = {nullptr, nullptr, nullptr};
	for (int i = 1; i < eff; ++i) hp[i - 1] = (const double*)b->hist[i - 1]->ptr;
	const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
	hipLaunchKernelGGL(k_bdf_rhs, dim3(blocks), dim3(256), 0, b->be->stream, (const double*)u->ptr, hp[0], hp[1], hp[2],
	                   eff - 1, -h[0], -h[1], -h[2], -h[3], 1.0 / lhs, (double*)b->rhs->ptr, (int64_t)n);
	FX_CHECK(fx::check_launch("bdf rhs"));
	// (I + coeff·Δ) u = rhs with coeff = −dt·α/lhs (bdf.rs:196-205), after the history shift (207-217).
	const double coeff = -b->dt * b->alpha / lhs;
	std::rotate(b->hist.rbegin(), b->hist.rbegin() + 1, b->hist.rend());
	FX_CHECK(fx_copy(b->be, u, b->hist[0]));
	fx_operator op{};
	op.kind = FX_OP_STENCIL;
	op.coeff = coeff;
	op.sigma = 1.0;
	op.stencil = &b->st;
	op.bc = &b->bc;
	fx_precond pc{FX_PC_MULTIGRID, 0, b->mg};
	fx_solve_stats st{};
	FX_CHECK(fx_solver_solve_op(b->solver, u, b->rhs, &op, b->mg ? &pc : nullptr, &st));
	b->steps += 1;
	if (stats) *stats = st;
	return FX_OK;
}

fx_status fx_bdf_step_n(fx_bdf* b, fx_buf* u, uint64_t n) {
	for (uint64_t i = 0; i < n; ++i) FX_CHECK(fx_bdf_step(b, u, nullptr));
	return FX_OK;
}

double fx_bdf_dt(const fx_bdf* b) { return b ? b->dt : 0.0; }
uint64_t fx_bdf_steps_done(const fx_bdf* b) { return b ? b->steps : 0; }
double fx_bdf_sim_time(const fx_bdf* b) { return b ? (double)b->steps * b->dt : 0.0; }
uint32_t fx_bdf_order(const fx_bdf* b) { return b ? (uint32_t)b->order : 0; }

fx_status fx_write_raw_host(const char* path, const double* data, const fx_grid* grid) {
	FX_CHECK(fx::check_grid(grid));
	return raw_write(path, data, grid);
}

fx_status fx_read_raw_host(const char* path, fx_grid* grid, double* data, size_t len) {
	return raw_read(path, grid, data, len);
}

fx_status fx_write_raw(fx_backend* be, const char* path, const fx_buf* field, const fx_grid* grid) {
	FX_REQUIRE(be && field, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(fx::check_grid(grid));
	const size_t n = (size_t)(grid->rows * grid->cols);
	FX_REQUIRE(field->len == n, FX_ERR_DIMENSION, "dimension mismatch: expected %zu, got %zu (field)", n, field->len);
	try {
		std::vector<double> host(n);
		if (field->dtype == FX_F64) {
			FX_CHECK(fx_copy_to_host_f64(be, field, host.data(), n));
		} else {
			// Field::to_vec_f64 (runtime/src/field.rs:134-138) returns an F32 field widened to f64.
			std::vector<float> h32(n);
			FX_CHECK(fx_copy_to_host_f32(be, field, h32.data(), n));
			for (size_t i = 0; i < n; ++i) host[i] = (double)h32[i];
		}
		return raw_write(path, host.data(), grid);
	} catch (const std::bad_alloc&) {
		fx::set_error("%s: no host memory for %zu values", path, n);
		return FX_ERR_ALLOC;
	}
}

fx_status fx_read_raw(fx_backend* be, const char* path, fx_grid* grid, fx_buf** out) {
	FX_REQUIRE(be && grid && out, FX_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	fx_grid g{};
	FX_CHECK(raw_read(path, &g, nullptr, 0));
	const size_t n = (size_t)(g.rows * g.cols);
	try {
		std::vector<double> host(n);
		FX_CHECK(raw_read(path, &g, host.data(), n));
		FX_CHECK(fx_upload_f64(be, host.data(), n, out));
	} catch (const std::bad_alloc&) {
		fx::set_error("%s: no host memory for %zu values", path, n);
		return FX_ERR_ALLOC;
	}
	*grid = g;
	return FX_OK;
}

}  // extern "C"
