/*
 * fluxion_hip.h — C ABI of libfluxion_hip.so, a gfx950 (MI355X) backend for
 * fluxion's matrix-free structured-grid linear solve path.
 *
 * Every entry point replaces one method of the reference's backend plugin
 * contract `fluxion_core::Backend` (crates/core/src/backend.rs:43-406) or one
 * solver / preconditioner type of `fluxion_runtime` (crates/runtime/src/).
 * The reference file:line each declaration stands in for is given beside it.
 *
 * Conventions
 *   - Plain C types only: opaque handles, pointers, sizes.  No C++ or torch types.
 *   - Every call returns an fx_status (0 = FX_OK).  On failure a message is
 *     available from fx_last_error() (thread-local).  Every kernel launch is
 *     checked, unlike the reference (backend-hip/src/backend.rs:293-307).
 *   - All operations are ordered on the backend's single HIP stream.  Calls
 *     that return a host value (reductions, downloads) synchronise the stream,
 *     as the reference's reductions do (backend.rs:356-413).
 *   - Grids are row-major, index = row*cols + col (core/src/grid.rs:59-61).
 *   - Buffers are owned by the library; callers release them with
 *     fx_buffer_free.  Input and output buffers of one call may alias only
 *     where stated (pointwise ops with z == x or z == y).
 *   - One backend per host thread, or an external lock (the reference relies
 *     on a process-wide singleton: runtime/src/storage.rs:120-138).
 */
#ifndef FLUXION_HIP_H
#define FLUXION_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ── status codes (core/src/error.rs:3-20) ───────────────────────────── */
typedef enum {
	FX_OK = 0,
	FX_ERR_PERIODIC = 1,    /* ops.rs:11-23 reject_periodic */
	FX_ERR_DIMENSION = 2,   /* CoreError::DimensionMismatch */
	FX_ERR_GRID = 3,        /* CoreError::InvalidGrid (grid.rs:15-36) */
	FX_ERR_STENCIL = 4,     /* CoreError::InvalidStencil (stencil.rs:23-30); also |offset| > 1 */
	FX_ERR_ARGUMENT = 5,    /* null handle, wrong dtype, bad option */
	FX_ERR_HIP = 6,         /* CoreError::BackendError from the HIP runtime */
	FX_ERR_ALLOC = 7,       /* device allocation failed */
	FX_ERR_NO_DEVICE = 8,   /* no HIP device / not gfx950 */
	FX_ERR_UNSUPPORTED = 9,
	FX_ERR_IO = 10          /* std::io::Error of the raw field format (io/src/raw.rs:13-77) */
} fx_status;

typedef enum { FX_F32 = 0, FX_F64 = 1 } fx_dtype;            /* core/src/dtype.rs:1-16 */

typedef struct fx_backend fx_backend;  /* HipBackend (backend-hip/src/backend.rs:61-79) */
typedef struct fx_buf fx_buf;          /* HipBuffer (backend-hip/src/buffer.rs:16-123) */

typedef struct { uint64_t rows, cols; double dx, dy; } fx_grid;            /* grid.rs:4-10 */
typedef struct { int32_t dr, dc; double weight; } fx_stencil_entry;       /* stencil.rs:3-9 */
typedef struct { uint32_t n; const fx_stencil_entry* entries; } fx_stencil; /* stencil.rs:16-18 */

typedef enum { FX_BC_DIRICHLET = 0, FX_BC_NEUMANN = 1, FX_BC_ROBIN = 2, FX_BC_PERIODIC = 3 } fx_bc_kind;
/* boundary.rs:1-15: Dirichlet(a) | Neumann(a) | Robin{alpha=a, beta=b, g} | Periodic */
typedef struct { int32_t kind; double a, b, g; } fx_bc;
typedef struct { fx_bc top, bottom, left, right; } fx_boundaries;          /* boundary.rs:18-24 */

/* ── backend lifetime (backend.rs:103-264, storage.rs:120-131) ───────── */
fx_status fx_backend_create(int device, fx_backend** out);
void fx_backend_destroy(fx_backend* be);
const char* fx_backend_name(const fx_backend* be);                        /* Backend::name backend.rs:47 */
fx_status fx_synchronize(fx_backend* be);
size_t fx_last_error(char* buf, size_t len);
/* Raw HIP stream of the backend (hipStream_t), for callers that time or order work around it. */
void* fx_backend_stream(fx_backend* be);

/* ── kernel probe (measurement extension; no reference counterpart) ─────
 * HIP events recorded on the backend's stream immediately before and after
 * every launch of one kernel family.  fx_probe_set selects the family (OFF
 * disables) and clears the accumulator; fx_probe_read synchronises and returns
 * the summed duration and count of the probed launches.  Launches that exited
 * at once on a solver's convergence flag (shorter than a tenth of the median)
 * are left out of both figures. */
typedef enum {
	FX_PROBE_OFF = 0,
	FX_PROBE_FCG_UPDATE = 1,      /* FCG x/r update + FP32 residual copy + r·r        */
	FX_PROBE_FCG_STEP = 2,        /* FCG p update + A·p + p·Ap                        */
	FX_PROBE_VC_POST0 = 3,        /* fused V-cycle, finest-level post kernel          */
	FX_PROBE_VC_RESTRICT0 = 4,    /* fused V-cycle, finest-level residual+restriction */
	FX_PROBE_STENCIL_AXPY = 5,    /* Backend::stencil_axpy, 5-point streaming kernel  */
	FX_PROBE_REDUCE = 6           /* dot / dot2 / norm / sum block-reduction kernel   */
} fx_probe_kind;
fx_status fx_probe_set(fx_backend* be, int32_t kind);
fx_status fx_probe_read(fx_backend* be, double* total_ms, uint64_t* launches);
/* Attainable HBM bandwidth of plain streaming kernels (measurement extension; no reference
 * counterpart), the ceilings the solver kernels are read against.  Each kind streams n points once
 * per launch, 16-byte accesses, grid-stride, over four rotating copies of its arrays (4·n·bytes
 * beyond the 256 MB last-level cache for n = 2^24) allocated and freed by the call; out_us[kind]
 * receives the mean microseconds per launch over `reps` launches (HIP events, after 3 warm-ups):
 *   FX_HBM_READ   8 B/pt read (sum of an FP64 array)       FX_HBM_WRITE 8 B/pt written (fill)
 *   FX_HBM_COPY  16 B/pt (FP64 copy)
 *   FX_HBM_MIX52 52 B/pt: FP32 + 3 FP64 arrays read, 3 FP64 arrays written (the byte mix of the
 *                FCG step kernel), non-temporal stores. */
enum { FX_HBM_READ = 0, FX_HBM_WRITE = 1, FX_HBM_COPY = 2, FX_HBM_MIX52 = 3, FX_HBM_NKINDS = 4 };
fx_status fx_hbm_ceilings(fx_backend* be, size_t n, int32_t reps, double* out_us);

/* ── memory (backend.rs:50-79, BackendBuffer 13-36) ──────────────────── */
fx_status fx_allocate(fx_backend* be, size_t len, fx_dtype dtype, fx_buf** out);   /* zero-filled */
fx_status fx_upload_f64(fx_backend* be, const double* data, size_t len, fx_buf** out);
fx_status fx_upload_f32(fx_backend* be, const float* data, size_t len, fx_buf** out);
void fx_buffer_free(fx_buf* buf);                                          /* Drop for HipBuffer */
size_t fx_buffer_len(const fx_buf* buf);
fx_dtype fx_buffer_dtype(const fx_buf* buf);
void* fx_buffer_device_ptr(fx_buf* buf);
fx_status fx_copy_to_host_f64(fx_backend* be, const fx_buf* buf, double* dst, size_t len);
fx_status fx_copy_to_host_f32(fx_backend* be, const fx_buf* buf, float* dst, size_t len);
/* Overwrite an existing buffer from host memory (upload without allocation). */
fx_status fx_write_f64(fx_backend* be, fx_buf* buf, const double* src, size_t len);

/* ── stencil operations (backend.rs:88-174) ──────────────────────────── */
fx_status fx_apply_stencil(fx_backend* be, const fx_buf* input, fx_buf* output, const fx_grid* grid,
                           const fx_stencil* st, const fx_boundaries* bc);            /* backend.rs:88-95 */
fx_status fx_apply_stencil_var(fx_backend* be, const fx_buf* input, fx_buf* output, const fx_buf* coeff,
                               const fx_grid* grid, const fx_stencil* st,
                               const fx_boundaries* bc);                              /* backend.rs:148-156 */
fx_status fx_stencil_axpy(fx_backend* be, double alpha, const fx_buf* x, fx_buf* y, const fx_grid* grid,
                          const fx_stencil* st, const fx_boundaries* bc);             /* backend.rs:166-174 */

/* ── BLAS-1 (backend.rs:101-136, 285-361) ────────────────────────────── */
fx_status fx_copy(fx_backend* be, const fx_buf* src, fx_buf* dst);                   /* backend.rs:101-105 */
fx_status fx_fill(fx_backend* be, fx_buf* buf, double value);                        /* backend.rs:111-115 */
fx_status fx_axpy(fx_backend* be, double alpha, const fx_buf* x, fx_buf* y);         /* backend.rs:121-126 */
fx_status fx_scale(fx_backend* be, fx_buf* buf, double alpha);                       /* backend.rs:132-136 */
fx_status fx_waxpy(fx_backend* be, double alpha, const fx_buf* x, double beta, const fx_buf* y,
                   fx_buf* w);                                                        /* backend.rs:309-316 */
fx_status fx_aypx(fx_backend* be, double alpha, const fx_buf* x, fx_buf* y);         /* backend.rs:322-327 */
fx_status fx_pointwise_mult(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* z); /* backend.rs:285-290 */
fx_status fx_pointwise_div(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* z);  /* backend.rs:296-301 */
fx_status fx_pointwise_max(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* z);  /* backend.rs:348-354 */
fx_status fx_pointwise_min(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* z);  /* backend.rs:356-361 */
fx_status fx_reciprocal(fx_backend* be, fx_buf* buf);                                /* backend.rs:333-336 */
fx_status fx_abs_val(fx_backend* be, fx_buf* buf);                                   /* backend.rs:342-345 */

/* ── reductions (backend.rs:180-230); results are copied to the host ── */
fx_status fx_norm_l2(fx_backend* be, const fx_buf* buf, double* out);                /* backend.rs:180 */
fx_status fx_reduce_sum(fx_backend* be, const fx_buf* buf, double* out);             /* backend.rs:188 */
fx_status fx_reduce_max(fx_backend* be, const fx_buf* buf, double* out);             /* backend.rs:194 */
fx_status fx_reduce_min(fx_backend* be, const fx_buf* buf, double* out);             /* backend.rs:200 */
fx_status fx_dot(fx_backend* be, const fx_buf* x, const fx_buf* y, double* out);     /* backend.rs:206-210 */
/* No counterpart in the reference: fx_dot for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It allocates nothing and
 * does not synchronise: the block partials live in a static device array of 2048 partials (32 KiB)
 * that each GPU holds until the process ends, and x·y is stored into out[0] (FP64, len >= 1) on the
 * backend's stream — the same value as fx_dot, bit for bit.  Every input length fits (the reduction
 * never launches more than 2048 workgroups).  The array is shared by every call on one GPU: calls
 * that overlap on one GPU (from two backends or threads) race. */
fx_status fx_dot_noalloc(fx_backend* be, const fx_buf* x, const fx_buf* y, fx_buf* out);
/* out[0] = a·b, out[1] = c·d in one pass (backend.rs:219-230; HIP dot2 backend-hip/src/backend.rs:1143-1210) */
fx_status fx_dot2(fx_backend* be, const fx_buf* a, const fx_buf* b, const fx_buf* c, const fx_buf* d,
                  double* out);

/* ── multigrid primitives (backend.rs:240-277) ───────────────────────── */
fx_status fx_restrict(fx_backend* be, const fx_buf* fine, fx_buf* coarse, const fx_grid* fine_grid,
                      const fx_grid* coarse_grid);                                    /* backend.rs:240-246 */
fx_status fx_prolong(fx_backend* be, const fx_buf* coarse, fx_buf* fine, const fx_grid* coarse_grid,
                     const fx_grid* fine_grid);                                       /* backend.rs:254-260 */
/* One damped sweep x += ω·D⁻¹·(b − S·x) on interior points (backend.rs:269-277).  The reference CPU
 * sweep is in place and lexicographic (ops.rs:447-471); this one is a two-buffer Jacobi sweep: every
 * point reads the previous iterate only. */
fx_status fx_weighted_jacobi(fx_backend* be, fx_buf* x, const fx_buf* b, double omega, const fx_grid* grid,
                             const fx_stencil* st, const fx_boundaries* bc);

/* ── convection-diffusion (backend.rs:375-405) ───────────────────────── */
fx_status fx_apply_conv_diff(fx_backend* be, const fx_buf* u, fx_buf* output, const fx_buf* kappa,
                             const fx_buf* vx, const fx_buf* vy, const fx_grid* grid,
                             const fx_boundaries* bc);                                /* backend.rs:375-384 */
fx_status fx_conv_diff_axpy(fx_backend* be, double alpha, const fx_buf* u, fx_buf* output, const fx_buf* kappa,
                            const fx_buf* vx, const fx_buf* vy, const fx_grid* grid,
                            const fx_boundaries* bc);                                 /* backend.rs:395-405 */

/* ── HIP-only extras (backend-hip/src/backend.rs:463-531) ────────────── */
fx_status fx_zero_boundaries(fx_backend* be, fx_buf* buf, uint64_t rows, uint64_t cols); /* backend.rs:510-531 */
/* z=q+βz; t=w+βt; p=r+βp; x+=αp; r−=αt; w−=αz (backend.rs:469-504, pipelined_cg.rs:172-200) */
fx_status fx_pipelined_cg_fused(fx_backend* be, double alpha, double beta, fx_buf* z, fx_buf* t, fx_buf* p,
                                fx_buf* x, fx_buf* r, fx_buf* w, const fx_buf* q);

/* ════════════════════════════════════════════════════════════════════
 * Solvers (crates/runtime/src/{cg,pipelined_cg,bicgstab,gmres}.rs)
 * ════════════════════════════════════════════════════════════════════ */

typedef enum { FX_CONVERGED = 0, FX_MAX_ITERATIONS = 1, FX_BREAKDOWN = 2 } fx_reason;  /* cg.rs:11-18 */
typedef struct {                                                                       /* cg.rs:22-29 */
	uint64_t iterations;
	double residual;
	int32_t reason;
	int32_t pad;
	/* Extension: ‖b − A·x‖₂ recomputed in FP64 after the solve (the reference never recomputes it). */
	double true_residual;
} fx_solve_stats;

/* Operator A.  FX_OP_STENCIL is the reference operator A·v = v + coeff·S(v) with identity rows on the
 * boundary (cg.rs:104-111, copy + stencil_axpy).  The other kinds are extensions with no reference
 * solver counterpart: conv-diff sigma·v + coeff·(κΔv + w·∇⁻v) (ops.rs:674-791 with an identity weight)
 * and variable coefficient sigma·v + coeff·(−∇·(k∇v)) with harmonic face averages of nodal k. */
typedef enum { FX_OP_STENCIL = 0, FX_OP_CONV_DIFF = 1, FX_OP_VARK = 2 } fx_op_kind;
typedef struct {
	int32_t kind;
	int32_t pad;
	double coeff;
	double sigma;                 /* identity weight (kinds 1, 2) */
	const fx_stencil* stencil;    /* kind 0 */
	const fx_boundaries* bc;      /* kind 0 */
	const fx_buf* kappa;          /* kind 1: diffusivity field, or NULL for kappa_scalar */
	double kappa_scalar;
	const fx_buf* vx;             /* kind 1 */
	const fx_buf* vy;             /* kind 1 */
	const fx_buf* k;              /* kind 2: nodal coefficient */
} fx_operator;

/* Multigrid V-cycle preconditioner (multigrid.rs:12-608). */
typedef struct fx_multigrid fx_multigrid;
typedef enum { FX_SMOOTH_JACOBI = 0, FX_SMOOTH_CHEBYSHEV = 1 } fx_smoother;          /* multigrid.rs:12-20 */
typedef enum { FX_TRANSFER_REFERENCE = 0, FX_TRANSFER_FW_BILINEAR = 1 } fx_transfer;
typedef struct {
	int32_t n_pre, n_post;
	double omega;
	int32_t smoother;        /* fx_smoother */
	int32_t transfer;        /* FX_TRANSFER_REFERENCE: 2x2 average + injection, halving sizes (ops.rs:369-423,
	                          * multigrid.rs:121-135).  FX_TRANSFER_FW_BILINEAR: every level spans the fine domain
	                          * (sizes (n+1)/2 for odd n, n/2 for even n); linear interpolation and its transpose
	                          * scaled by (h/H)², i.e. textbook full weighting + bilinear on nested 2^k+1 grids. */
	int32_t coarsest;        /* stop coarsening once rows or cols <= coarsest (reference: 4, multigrid.rs:125) */
	int32_t extra_coarse;    /* extra sweeps on the coarsest level (reference: 10, multigrid.rs:215-216) */
	int32_t precision;       /* FX_F64 (reference) or FX_F32 (mixed-precision V-cycle) */
	int32_t fused;           /* 1: fused FP32 kernels (smoother+residual+restriction); needs FX_F32, Jacobi, FW */
	double sigma, scale;     /* level operator sigma·I + scale·Δ_h (reference: 0, 1 — multigrid.rs:105-106) */
	const fx_buf* k;         /* optional nodal coefficient: level operator sigma·I + scale·(−∇·k∇) */
	/* Optional nodal velocity (FP64, both or neither; not with k, Jacobi only, unfused): level operator
	 * sigma·I + scale·(κΔ + v·∇⁻v), the upwind convection-diffusion of ops.rs:707-731 rediscretised on
	 * every level with the velocity interpolated like k, κ = kappa. */
	const fx_buf* vx;
	const fx_buf* vy;
	double kappa;
} fx_mg_options;

/* Multigrid::new(grid, bc, device, n_pre, n_post, omega) / Multigrid::build(.., smoother)
 * (multigrid.rs:69-146): reference options. */
fx_status fx_multigrid_create(fx_backend* be, const fx_grid* grid, const fx_boundaries* bc, uint32_t n_pre,
                              uint32_t n_post, double omega, int32_t smoother, fx_multigrid** out);
fx_status fx_multigrid_create_ex(fx_backend* be, const fx_grid* grid, const fx_boundaries* bc,
                                 const fx_mg_options* opt, fx_multigrid** out);
void fx_multigrid_destroy(fx_multigrid* mg);
uint32_t fx_multigrid_depth(const fx_multigrid* mg);                                  /* multigrid.rs:172-175 */
/* Multigrid::v_cycle(x, b) (multigrid.rs:188-230): x is the initial guess, overwritten. */
fx_status fx_multigrid_vcycle(fx_multigrid* mg, fx_buf* x, const fx_buf* b);
/* Level kernels of the fused FP32 V-cycle (extension; no reference counterpart).  FX_MGK_WAVE (the
 * default) runs each level as independent wave-wide row streams, FX_MGK_RING as LDS-ring row streams and
 * FX_MGK_TILED as LDS tiles; all three give bit-identical results.  fx_multigrid_set_wave_rows sets
 * level `level`'s wave segment heights (fine rows per wave of its post kernel, coarse rows per wave of
 * its restriction; 0 restores the built-in choice; −1 on levels ≥ 1 runs that kernel of the level as
 * LDS tiles instead). */
typedef enum { FX_MGK_TILED = 0, FX_MGK_RING = 1, FX_MGK_WAVE = 2 } fx_mg_kernels;
fx_status fx_multigrid_set_kernels(fx_multigrid* mg, int32_t kernels);
fx_status fx_multigrid_set_wave_rows(fx_multigrid* mg, uint32_t level, int32_t post_rows, int32_t restrict_rows);
/* Autotuner (extension): times every level's wave kernels over candidate segment heights with HIP
 * events (reps launches each, on the hierarchy's own buffers, whose contents it overwrites) and keeps
 * the fastest per level, the LDS-tile kernels competing on levels ≥ 1; *ms_out (may be NULL) gets the
 * summed per-launch time of the chosen shapes.
 * Level 0 is also timed as the fused FCG loop runs it (restriction with the residual update, post
 * kernel with the solver's reductions), on scratch vectors it allocates and frees; those heights are
 * kept apart.  fx_multigrid_set_wave_rows on level 0 sets both. */
fx_status fx_multigrid_autotune(fx_multigrid* mg, int32_t reps, double* ms_out);

/* Preconditioner handle (preconditioner.rs:10-37). */
typedef enum { FX_PC_IDENTITY = 0, FX_PC_MULTIGRID = 1, FX_PC_LEARNED = 2 } fx_pc_kind;
typedef struct { int32_t kind; int32_t pad; void* handle; } fx_precond;

typedef struct fx_solver fx_solver;
typedef enum {
	FX_SOLVER_CG = 0,            /* CgSolver (cg.rs:52-256) */
	FX_SOLVER_PIPELINED_CG = 1,  /* PipelinedCgSolver (pipelined_cg.rs:25-238) */
	FX_SOLVER_BICGSTAB = 2,      /* BiCgStabSolver (bicgstab.rs:19-346) */
	FX_SOLVER_GMRES = 3,         /* GmresSolver (gmres.rs:24-490) */
	FX_SOLVER_FCG = 4,           /* extension: flexible CG with FP64 true-residual refinement: after the
	                              * recurrence meets tol, restarts from x while ‖b − A·x‖ ≥ tol and each
	                              * restart halves it; FX_CONVERGED if the true residual ends below tol or
	                              * at that FP64 floor (stats.true_residual says which) */
	FX_SOLVER_FGMRES = 5         /* extension: flexible right-preconditioned GMRES(m) (Saad 1993): each step
	                              * stores z_k = M⁻¹v_k and applies A to it, x += Σ y_i·z_i, so the Givens
	                              * estimate is ‖b − A·x‖ and the preconditioner may vary (FP32 V-cycle);
	                              * without a preconditioner it is GmresSolver (gmres.rs:116-283) */
} fx_solver_kind;

/* XxxSolver::new(grid, device, [restart,] tol, max_iter): all work vectors allocated once
 * (cg.rs:72-86, pipelined_cg.rs:50-70, bicgstab.rs:61-82, gmres.rs:56-81). restart is used by GMRES and
 * FGMRES only. */
fx_status fx_solver_create(fx_backend* be, int32_t kind, const fx_grid* grid, uint32_t restart, double tol,
                           uint64_t max_iter, fx_solver** out);
void fx_solver_destroy(fx_solver* s);
/* LinearSolver::name (linear_solver.rs:30): "CG", "PipelinedCG", "BiCGSTAB", "GMRES", "FCG", "FGMRES". */
const char* fx_solver_name(const fx_solver* s);
/* LinearSolver::solve(x, b, coeff, stencil, bc) (linear_solver.rs:13-27). */
fx_status fx_solver_solve(fx_solver* s, fx_buf* x, const fx_buf* b, double coeff, const fx_stencil* st,
                          const fx_boundaries* bc, fx_solve_stats* stats);
/* solve_preconditioned (cg.rs:178-256, bicgstab.rs:218-346, gmres.rs:294-454).  PipelinedCgSolver has no
 * preconditioned form in the reference (pipelined_cg.rs:96-167); here an exact preconditioner runs the
 * Ghysels–Vanroose recurrence and an FP32 multigrid runs the flexible single-reduction pipelined form
 * (M applied to the current residual each iteration) inside FP64 true-residual refinement passes. */
fx_status fx_solver_solve_preconditioned(fx_solver* s, fx_buf* x, const fx_buf* b, double coeff,
                                         const fx_stencil* st, const fx_boundaries* bc, const fx_precond* pc,
                                         fx_solve_stats* stats);
/* Extension: any operator kind, optional preconditioner (NULL = unpreconditioned). */
fx_status fx_solver_solve_op(fx_solver* s, fx_buf* x, const fx_buf* b, const fx_operator* op,
                             const fx_precond* pc, fx_solve_stats* stats);
/* Extension: per-iteration residual history of the last solve (entries copied: min(len, its+1)). */
size_t fx_solver_history(const fx_solver* s, double* out, size_t len);

/* ════════════════════════════════════════════════════════════════════
 * Learned preconditioner network (extension; no reference counterpart).
 * A 4-level U-Net (3×3 convs, 16/32/48/64 channels, BF16 weights and
 * activations, FP32 accumulation on MFMA).  Tensors are FX_F32 buffers in
 * NHWC order: x has B·H·W·cin values, y and t have B·H·W; H, W divisible by 8.
 * Parameters: conv weights W_l[tap][cout][cin] in layer order, then biases
 * (the layout of oracle/unet.cpp).
 * ════════════════════════════════════════════════════════════════════ */
typedef struct fx_unet fx_unet;
fx_status fx_unet_create(fx_backend* be, int32_t cin, fx_unet** out);
void fx_unet_destroy(fx_unet* u);
size_t fx_unet_nparams(const fx_unet* u);
fx_status fx_unet_set_params(fx_unet* u, const float* host, size_t n);   /* also resets the Adam state */
fx_status fx_unet_get_params(fx_unet* u, float* host, size_t n);
fx_status fx_unet_get_grad(fx_unet* u, float* host, size_t n);
fx_status fx_unet_forward(fx_unet* u, const fx_buf* x, fx_buf* y, int32_t B, int32_t H, int32_t W);
/* Σ_b ‖y_b − t_b‖²/‖t_b‖² and its gradient (kept on the device). */
fx_status fx_unet_loss_grad(fx_unet* u, const fx_buf* x, const fx_buf* t, int32_t B, int32_t H, int32_t W,
                            double* loss);
/* Relative energy-norm loss Σ_b dᵀA_b d / tᵀA_b t (d = y − t, A_b the unit-spacing 5-point
 * −∇·(k∇·) on interior pixels with k = exp of input channel 1) and its gradient. */
fx_status fx_unet_loss_grad_energy(fx_unet* u, const fx_buf* x, const fx_buf* t, int32_t B, int32_t H, int32_t W,
                                   double* loss);
fx_status fx_unet_adam_step(fx_unet* u, float lr, float beta1, float beta2, float eps);

/* Learned preconditioner (FX_PC_LEARNED): z = V(r) + α*·c with c = s·h²·N([r1/s, log k]), r1 = r − A·V(r),
 * s = rms(r1), V the multigrid V-cycle, N the network (h² = dx·dy, logk may be NULL for k ≡ 1) and
 * α* = (c·r1)/(c·A·c) clipped at 0, the step that minimises the A-norm of the remaining error.  Used by
 * flexible CG.  While the network is in use the solve recomputes the FP64 true residual ‖b − A·x‖ after
 * every iteration, and falls back to z = V(r) for the rest of the solve once an iteration whose
 * direction came from a network application fails to cut it by the factor `stall`. */
typedef struct fx_learned fx_learned;
fx_status fx_learned_create(fx_backend* be, fx_unet* net, fx_multigrid* mg, const fx_buf* logk, const fx_grid* grid,
                            double stall, fx_learned** out);
void fx_learned_destroy(fx_learned* l);
/* 1 if a solve since the previous call switched to the plain V-cycle (and clears the flag). */
int32_t fx_learned_fell_back(fx_learned* l);
/* Schedule of the network inside a solve: applied at the preconditioner applications made after
 * first, first + every, first + 2·every, … iterations, at most max_apps times per solve pass (0: no
 * limit); the other applications are the V-cycle alone.  Default every = 1, max_apps = 0, first = 0. */
fx_status fx_learned_set_schedule(fx_learned* l, int32_t every, int32_t max_apps, int32_t first);

/* Extension: A·x for an operator (same arithmetic as the solvers use). */
fx_status fx_apply_operator(fx_backend* be, const fx_operator* op, const fx_buf* x, fx_buf* y,
                            const fx_grid* grid);

/* ── batched multi-problem solver (extension; no reference counterpart) ───────
 * nprob independent problems −∇·(k_i∇u_i) = f_i (zero Dirichlet) on grids of uneven sizes, each
 * solved by FP64 flexible CG around its own fused FP32 coefficient-aware V-cycle (options `opt`:
 * FX_F32, fused, Jacobi, FX_TRANSFER_FW_BILINEAR), all problems advancing together: every kernel of
 * an iteration is one launch over a work list of (problem, level, tile) items and the iterations are
 * replayed from a captured HIP graph until every problem has stopped.  Problem i's points occupy
 * [off_i, off_i + rows_i·cols_i) of every concatenated vector (problems in order, row-major).  `k`
 * (FP64, concatenated) must stay allocated while the batch is used.  Launch shapes that leave the
 * per-point results unchanged are picked by timing candidates at creation (fx_batch_info). */
typedef struct fx_batch fx_batch;
fx_status fx_batch_create(fx_backend* be, uint32_t nprob, const fx_grid* grids, const fx_buf* k,
                          const fx_mg_options* opt, fx_batch** out);
/* x: initial guesses in, solutions out; tol[i] is problem i's absolute ‖r‖ tolerance; stats[nprob]. */
fx_status fx_batch_solve(fx_batch* bt, fx_buf* x, const fx_buf* b, const double* tol, uint64_t max_iter,
                         fx_solve_stats* stats);
void fx_batch_destroy(fx_batch* bt);
fx_status fx_batch_info(const fx_batch* bt, int32_t* rows_per_block, int32_t* points_per_block, uint64_t* points,
                        int32_t* graph_iterations);
/* Level kernels of the batch's V-cycles: wave = 1 when every problem's levels run on the
 * wave-streaming kernels (one wave per strip and row segment; post_rows fine rows and restrict_rows
 * coarse rows per segment, tuned at creation), 0 for the LDS-tiled kernels (both heights 0). */
fx_status fx_batch_level_kernels(const fx_batch* bt, int32_t* wave, int32_t* post_rows, int32_t* restrict_rows);

/* ── Crank–Nicolson stepper (CrankNicolsonSolver, runtime/src/implicit.rs:1-238) ──────────
 * One step solves (I − (dt/2)·α·Δ) u^{n+1} = (I + (dt/2)·α·Δ) u^n + dt·source from u^n as the initial
 * guess (implicit.rs:178-222), with Δ = Stencil::laplacian_2d_5pt(dx, dy) and the step's boundary
 * conditions.  opt == NULL is CrankNicolsonSolver::new (unpreconditioned CG, implicit.rs:41-75).
 * With opt->mg the solver is preconditioned by the multigrid those options build: the reference
 * options of fx_multigrid_create are CrankNicolsonSolver::with_multigrid / with_multigrid_smoother
 * (implicit.rs:77-154).  Extensions: opt->op = FX_OP_VARK steps ∂u/∂t = α·∇·(k∇u) + source with
 * the variable-coefficient operator (nodal k, harmonic face means, zero-Dirichlet identity rows;
 * opt->k must stay allocated), and opt->solver = FX_SOLVER_FCG runs flexible CG (needs opt->mg;
 * for the fused FP32 cycle on this operator give the level operator sigma = 1, scale = (dt/2)·α,
 * k = opt->k). */
typedef struct fx_cn fx_cn;
typedef struct {
	int32_t op;                 /* FX_OP_STENCIL (reference) or FX_OP_VARK */
	int32_t solver;             /* FX_SOLVER_CG (reference) or FX_SOLVER_FCG */
	const fx_buf* k;            /* FX_OP_VARK: nodal coefficient, FP64, rows·cols values */
	const fx_mg_options* mg;    /* NULL: unpreconditioned */
} fx_cn_options;
fx_status fx_cn_create(fx_backend* be, const fx_grid* grid, double alpha, double dt, const fx_boundaries* bc,
                       double cg_tol, uint64_t cg_max_iter, const fx_cn_options* opt, fx_cn** out);
void fx_cn_destroy(fx_cn* cn);
/* CrankNicolsonSolver::step(u, source) (implicit.rs:178-222); source may be NULL. */
fx_status fx_cn_step(fx_cn* cn, fx_buf* u, const fx_buf* source, fx_solve_stats* stats);
/* CrankNicolsonSolver::step_n(u, n) (implicit.rs:224-237): n steps without a source (stats dropped). */
fx_status fx_cn_step_n(fx_cn* cn, fx_buf* u, uint64_t n);
double fx_cn_dt(const fx_cn* cn);                /* implicit.rs:156-160 */
uint64_t fx_cn_steps_done(const fx_cn* cn);      /* implicit.rs:162-166 */
double fx_cn_sim_time(const fx_cn* cn);          /* implicit.rs:168-172: steps_done · dt */
/* Extension: the step's solver (residual history of the last step: fx_solver_history) and
 * multigrid (NULL when unpreconditioned; fx_multigrid_autotune applies), owned by the stepper. */
fx_solver* fx_cn_solver(fx_cn* cn);
fx_multigrid* fx_cn_multigrid(fx_cn* cn);
/* Extension: solve every step with the caller's preconditioner instead of the stepper's multigrid (the
 * caller keeps it alive; NULL restores the multigrid).  FX_PC_LEARNED needs opt->solver = FX_SOLVER_FCG;
 * build it over fx_cn_multigrid(cn), whose level operator is the step's. */
fx_status fx_cn_set_preconditioner(fx_cn* cn, const fx_precond* pc);

/* ── IMEX convection–diffusion stepper (ImexSolver, runtime/src/imex.rs:1-274) ──────────────
 * Uniform velocity (vx, vy) and diffusivity κ: conv = v·∇u (upwind, apply_conv_diff with κ = 0), then
 * (I − (dt/2)·κ·Δ) u^{n+1} = u^n + (dt/2)·κ·Δu^n + dt·conv by CG, or PCG with the multigrid of `mg`
 * (ImexSolver::new / with_multigrid, imex.rs:46-137; step imex.rs:147-197). */
typedef struct fx_imex fx_imex;
fx_status fx_imex_create(fx_backend* be, const fx_grid* grid, double kappa, double vx, double vy, double dt,
                         const fx_boundaries* bc, double cg_tol, uint64_t cg_max_iter, const fx_mg_options* mg,
                         fx_imex** out);
void fx_imex_destroy(fx_imex* m);
fx_status fx_imex_step(fx_imex* m, fx_buf* u, fx_solve_stats* stats);
fx_status fx_imex_step_n(fx_imex* m, fx_buf* u, uint64_t n);
double fx_imex_dt(const fx_imex* m);
uint64_t fx_imex_steps_done(const fx_imex* m);
double fx_imex_sim_time(const fx_imex* m);

/* ── BDF stepper (BdfSolver, runtime/src/bdf.rs:1-256) ─────────────────────────────────────
 * ∂u/∂t = α·Δu by BDF of order 1–4, the first steps at order steps_done + 1 (bdf.rs:163-240):
 * (lhs·I − dt·α·Δ) u^{n+1} = −Σ a_i·u^{n+1−i}, solved as (I + coeff·Δ) u = rhs/lhs with coeff = −dt·α/lhs
 * by CG from u^n, or by PCG with the multigrid of `mg` (Multigrid::new options for with_multigrid,
 * bdf.rs:110-137; NULL: BdfSolver::new, bdf.rs:58-108). */
typedef struct fx_bdf fx_bdf;
fx_status fx_bdf_create(fx_backend* be, const fx_grid* grid, double alpha, double dt, uint32_t order,
                        const fx_boundaries* bc, double cg_tol, uint64_t cg_max_iter, const fx_mg_options* mg,
                        fx_bdf** out);
void fx_bdf_destroy(fx_bdf* b);
fx_status fx_bdf_step(fx_bdf* b, fx_buf* u, fx_solve_stats* stats);                  /* bdf.rs:163-240 */
fx_status fx_bdf_step_n(fx_bdf* b, fx_buf* u, uint64_t n);                            /* bdf.rs:242-255 */
double fx_bdf_dt(const fx_bdf* b);
uint64_t fx_bdf_steps_done(const fx_bdf* b);
double fx_bdf_sim_time(const fx_bdf* b);
uint32_t fx_bdf_order(const fx_bdf* b);

/* ── explicit diffusion stepper (DiffusionSolver, runtime/src/solver.rs:1-494) ─────────────
 * ∂u/∂t = α·Δu + source with Δ = Stencil::laplacian_2d_5pt(dx, dy) and the given boundary conditions.
 * Schemes (TimeScheme, solver.rs:14-23): forward Euler, Heun (RK2, Shu–Osher form), classical RK4.
 * dt == NULL takes safety·min(dx, dy)²/(4α) with safety 0.4 / 0.8 / 1.4 (solver.rs:96-113).  Each step
 * gives the reference's values (solver.rs:344-494) while writing a second buffer instead of copying
 * u first (RK4: four fused stencil passes instead of 17 operations).  fx_diffusion_enable_graph
 * (enable_hip_graph, solver.rs:173-184) makes step_n replay a captured HIP graph of about 64 kernels
 * (64 Euler, 32 Heun or 8 RK4 steps) on the backend's stream, the remaining steps launched directly;
 * the graph is recaptured when step_n is given a different u. */
typedef struct fx_diffusion fx_diffusion;
typedef enum { FX_SCHEME_EULER = 0, FX_SCHEME_RK2 = 1, FX_SCHEME_RK4 = 2 } fx_time_scheme;
/* DiffusionSolver::build(grid, alpha, dt, boundaries, scheme, device) (solver.rs:96-148);
 * new / with_boundaries are the Euler forms (solver.rs:51-93). */
fx_status fx_diffusion_create(fx_backend* be, const fx_grid* grid, double alpha, const double* dt,
                              const fx_boundaries* bc, int32_t scheme, fx_diffusion** out);
void fx_diffusion_destroy(fx_diffusion* d);
fx_status fx_diffusion_enable_graph(fx_diffusion* d);                                  /* solver.rs:182-184 */
fx_status fx_diffusion_step(fx_diffusion* d, fx_buf* u, const fx_buf* source);        /* solver.rs:187-222 */
fx_status fx_diffusion_step_n(fx_diffusion* d, fx_buf* u, uint64_t n);                /* solver.rs:224-320 */
fx_status fx_diffusion_step_n_with_source(fx_diffusion* d, fx_buf* u, uint64_t n,
                                          const fx_buf* source);                      /* solver.rs:322-337 */
double fx_diffusion_dt(const fx_diffusion* d);                                         /* solver.rs:149-153 */
uint64_t fx_diffusion_steps_done(const fx_diffusion* d);                               /* solver.rs:155-159 */
double fx_diffusion_sim_time(const fx_diffusion* d);                                   /* solver.rs:161-165 */
int32_t fx_diffusion_scheme(const fx_diffusion* d);                                    /* solver.rs:167-171 */

/* ── raw field format (io/src/raw.rs:1-77) ───────────────────────────────────────────────
 * [rows:u64][cols:u64][dx:f64][dy:f64][data:f64 × rows·cols], all little-endian, row-major.
 * Reading validates the header as Grid::new does (FX_ERR_GRID) and fails on a short file
 * (FX_ERR_IO). */
/* write_raw(path, field) (raw.rs:13-32); an FX_F32 field is written widened to f64. */
fx_status fx_write_raw(fx_backend* be, const char* path, const fx_buf* field, const fx_grid* grid);
/* read_raw(path, device) (raw.rs:38-77): *out is a new FX_F64 buffer, *grid its grid. */
fx_status fx_read_raw(fx_backend* be, const char* path, fx_grid* grid, fx_buf** out);
/* Host-memory forms (no device): data has rows·cols values; fx_read_raw_host with data == NULL
 * reads and validates the header only. */
fx_status fx_write_raw_host(const char* path, const double* data, const fx_grid* grid);
fx_status fx_read_raw_host(const char* path, fx_grid* grid, double* data, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* FLUXION_HIP_H */
