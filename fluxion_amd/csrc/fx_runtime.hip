// fx_runtime.hip — backend lifetime, buffers, errors and argument checks.
//
// Replaces HipBackend::new / HipBuffer (crates/backend-hip/src/backend.rs:103-264,
// buffer.rs:16-123) with a C ABI.  Unlike the reference there is no embedded
// code-object loading: kernels are linked into this library for gfx950.
#include <algorithm>
#include "fx_internal.hpp"

#include <cmath>
#include <cstring>
#include <mutex>

namespace fx {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	g_last_error = buf;
}

fx_status check_launch(const char* what) {
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		set_error("kernel launch %s failed: %s", what, hipGetErrorString(e));
		return FX_ERR_HIP;
	}
	return FX_OK;
}

fx_status dev_alloc(fx_backend* be, size_t bytes, void** out) {
	*out = nullptr;
	if (bytes == 0) bytes = 8;
	hipError_t e = hipMalloc(out, bytes);
	if (e != hipSuccess) {
		set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
		*out = nullptr;
		return FX_ERR_ALLOC;
	}
	FX_HIP(hipMemsetAsync(*out, 0, bytes, be->stream));
	return FX_OK;
}

void dev_free(void* p) {
	if (p) (void)hipFree(p);
}

fx_status ensure_partials(fx_backend* be, size_t n) {
	if (n <= be->partials_cap) return FX_OK;
	size_t cap = n < 8192 ? 8192 : n;
	FX_HIP(hipStreamSynchronize(be->stream));
	dev_free(be->partials);
	be->partials = nullptr;
	void* p = nullptr;
	FX_CHECK(dev_alloc(be, cap * sizeof(fxd::Acc), &p));
	be->partials = (fxd::Acc*)p;
	be->partials_cap = cap;
	return FX_OK;
}

fx_status probe_drain(fx_backend* be) {
	if (be->probe_used == 0) return FX_OK;
	FX_HIP(hipEventSynchronize(be->probe_b[be->probe_used - 1]));
	for (size_t i = 0; i < be->probe_used; ++i) {
		float ms = 0.0f;
		FX_HIP(hipEventElapsedTime(&ms, be->probe_a[i], be->probe_b[i]));
		be->probe_ms.push_back(ms);
	}
	be->probe_used = 0;
	return FX_OK;
}
void probe_begin(fx_backend* be, int kind) {
	if (be->probe_kind == 0 || be->probe_kind != kind) return;
	if (be->probe_used == be->probe_a.size()) (void)probe_drain(be);
	(void)hipEventRecord(be->probe_a[be->probe_used], be->stream);
}
void probe_end(fx_backend* be, int kind) {
	if (be->probe_kind == 0 || be->probe_kind != kind) return;
	(void)hipEventRecord(be->probe_b[be->probe_used], be->stream);
	++be->probe_used;
}

// grid.rs:15-36.
fx_status check_grid(const fx_grid* g) {
	FX_REQUIRE(g, FX_ERR_ARGUMENT, "grid is null");
	FX_REQUIRE(g->rows >= 3 && g->cols >= 3, FX_ERR_GRID, "minimum 3x3, got %llux%llu",
	           (unsigned long long)g->rows, (unsigned long long)g->cols);
	FX_REQUIRE(std::isfinite(g->dx) && std::isfinite(g->dy) && g->dx > 0 && g->dy > 0, FX_ERR_GRID,
	           "spacing must be finite and positive, got dx=%g, dy=%g", g->dx, g->dy);
	return FX_OK;
}

fx_status check_buf(const fx_buf* b, size_t len, fx_dtype dt, const char* name) {
	FX_REQUIRE(b && b->ptr, FX_ERR_ARGUMENT, "buffer %s is null", name);
	FX_REQUIRE(b->dtype == dt, FX_ERR_ARGUMENT, "buffer %s has dtype %d, expected %d", name, (int)b->dtype, (int)dt);
	FX_REQUIRE(b->len == len, FX_ERR_DIMENSION, "dimension mismatch: expected %zu, got %zu (%s)", len, b->len, name);
	return FX_OK;
}

fx_status check_bc(const fx_boundaries* bc) {
	FX_REQUIRE(bc, FX_ERR_ARGUMENT, "boundaries are null");
	const fx_bc* s[4] = {&bc->top, &bc->bottom, &bc->left, &bc->right};
	for (int i = 0; i < 4; ++i) {
		// ops.rs:11-23.
		FX_REQUIRE(s[i]->kind != FX_BC_PERIODIC, FX_ERR_PERIODIC, "periodic boundaries are not yet implemented");
		FX_REQUIRE(s[i]->kind >= 0 && s[i]->kind <= FX_BC_ROBIN, FX_ERR_ARGUMENT, "unknown boundary kind %d",
		           s[i]->kind);
	}
	return FX_OK;
}

// Weight extraction follows ops.rs:908-920 (last matching entry wins; the
// diagonal weights are read only for 9-entry stencils, ops.rs:954-969).
fx_status stencil_args(const fx_stencil* st, fxd::W9* w9, fxd::Ent* ent) {
	FX_REQUIRE(st && st->n > 0 && st->entries, FX_ERR_STENCIL, "stencil must have at least one entry");
	FX_REQUIRE(st->n <= 32, FX_ERR_STENCIL, "stencil has %u entries; at most 32 are supported", st->n);
	fxd::W9 k{};
	for (uint32_t i = 0; i < st->n; ++i) {
		const fx_stencil_entry& e = st->entries[i];
		FX_REQUIRE(e.dr >= -1 && e.dr <= 1 && e.dc >= -1 && e.dc <= 1, FX_ERR_STENCIL,
		           "stencil offset (%d,%d) exceeds radius 1", e.dr, e.dc);
		bool diag = st->n == 9;
		if (e.dr == 0 && e.dc == 0) k.c = e.weight;
		else if (e.dr == -1 && e.dc == 0) k.n = e.weight;
		else if (e.dr == 1 && e.dc == 0) k.s = e.weight;
		else if (e.dr == 0 && e.dc == -1) k.w = e.weight;
		else if (e.dr == 0 && e.dc == 1) k.e = e.weight;
		else if (diag && e.dr == -1 && e.dc == -1) k.nw = e.weight;
		else if (diag && e.dr == -1 && e.dc == 1) k.ne = e.weight;
		else if (diag && e.dr == 1 && e.dc == -1) k.sw = e.weight;
		else if (diag && e.dr == 1 && e.dc == 1) k.se = e.weight;
	}
	if (w9) *w9 = k;
	if (ent) {
		ent->n = (int)st->n;
		for (uint32_t i = 0; i < st->n; ++i) {
			ent->dr[i] = st->entries[i].dr;
			ent->dc[i] = st->entries[i].dc;
			ent->w[i] = st->entries[i].weight;
		}
	}
	return FX_OK;
}

}  // namespace fx

using namespace fx;

// ── attainable HBM bandwidth (fx_hbm_ceilings) ────────────────────────
namespace {
typedef double hbm_d2v __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(fxd::BLOCK) void k_hbm_read(const double2* __restrict__ a, double* out, size_t n2) {
	double s = 0.0;
	for (size_t i = (size_t)blockIdx.x * fxd::BLOCK + threadIdx.x; i < n2; i += (size_t)gridDim.x * fxd::BLOCK) {
		const double2 v = a[i];
		s += v.x + v.y;
	}
	if (s == -1.0) out[0] = s;   // keeps the loads; never true for the zero-filled arrays
}
__global__ __launch_bounds__(fxd::BLOCK) void k_hbm_write(double2* __restrict__ b, size_t n2) {
	for (size_t i = (size_t)blockIdx.x * fxd::BLOCK + threadIdx.x; i < n2; i += (size_t)gridDim.x * fxd::BLOCK)
		b[i] = double2{1.0, 2.0};
}
__global__ __launch_bounds__(fxd::BLOCK) void k_hbm_copy(const double2* __restrict__ a, double2* __restrict__ b,
                                                         size_t n2) {
	for (size_t i = (size_t)blockIdx.x * fxd::BLOCK + threadIdx.x; i < n2; i += (size_t)gridDim.x * fxd::BLOCK) b[i] = a[i];
}
__global__ __launch_bounds__(fxd::BLOCK) void k_hbm_mix52(const float2* __restrict__ z, const double2* __restrict__ po,
                                                          const double2* __restrict__ k, double2* __restrict__ x,
                                                          double2* __restrict__ p, double2* __restrict__ ap, size_t n2) {
	for (size_t i = (size_t)blockIdx.x * fxd::BLOCK + threadIdx.x; i < n2; i += (size_t)gridDim.x * fxd::BLOCK) {
		const float2 zz = z[i];
		const double2 pp = po[i], kk = k[i], xx = x[i];
		const double2 pn{zz.x + 0.5 * pp.x, zz.y + 0.5 * pp.y};
		__builtin_nontemporal_store(hbm_d2v{pn.x, pn.y}, (hbm_d2v*)&p[i]);
		__builtin_nontemporal_store(hbm_d2v{pn.x * kk.x, pn.y * kk.y}, (hbm_d2v*)&ap[i]);
		__builtin_nontemporal_store(hbm_d2v{xx.x + 0.25 * pp.x, xx.y + 0.25 * pp.y}, (hbm_d2v*)&x[i]);
	}
}
}  // namespace

extern "C" {

size_t fx_last_error(char* buf, size_t len) {
	const std::string& s = g_last_error;
	if (buf && len) {
		size_t n = s.size() < len - 1 ? s.size() : len - 1;
		std::memcpy(buf, s.data(), n);
		buf[n] = 0;
	}
	return s.size();
}

fx_status fx_backend_create(int device, fx_backend** out) {
	FX_REQUIRE(out, FX_ERR_ARGUMENT, "out is null");
	*out = nullptr;
	int count = 0;
	hipError_t e = hipGetDeviceCount(&count);
	FX_REQUIRE(e == hipSuccess && count > 0, FX_ERR_NO_DEVICE, "no HIP device available (%s)",
	           e == hipSuccess ? "count = 0" : hipGetErrorString(e));
	FX_REQUIRE(device >= 0 && device < count, FX_ERR_NO_DEVICE, "device %d out of range (%d devices)", device, count);
	FX_HIP(hipSetDevice(device));
	hipDeviceProp_t prop;
	FX_HIP(hipGetDeviceProperties(&prop, device));
	FX_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, FX_ERR_NO_DEVICE,
	           "this library is built for gfx950 only; device %d is %s", device, prop.gcnArchName);
	fx_backend* be = new fx_backend();
	be->device = device;
	be->num_cu = prop.multiProcessorCount;
	std::snprintf(be->name, sizeof be->name, "hip-%s", prop.gcnArchName);
	// Default-priority stream (stream priorities are left untouched).
	if (hipStreamCreateWithFlags(&be->stream, hipStreamNonBlocking) != hipSuccess) {
		delete be;
		set_error("hipStreamCreate failed");
		return FX_ERR_HIP;
	}
	void* p = nullptr;
	fx_status s;
	if ((s = dev_alloc(be, sizeof(double) * 64, &p)) != FX_OK) { fx_backend_destroy(be); return s; }
	be->dres = (double*)p;
	if (hipHostMalloc((void**)&be->hres, sizeof(double) * 64, hipHostMallocDefault) != hipSuccess) {
		fx_backend_destroy(be);
		set_error("hipHostMalloc failed");
		return FX_ERR_ALLOC;
	}
	if ((s = ensure_partials(be, 8192)) != FX_OK) { fx_backend_destroy(be); return s; }
	if ((s = noalloc_setup(be)) != FX_OK) { fx_backend_destroy(be); return s; }
	if (hipStreamSynchronize(be->stream) != hipSuccess) { fx_backend_destroy(be); set_error("sync failed"); return FX_ERR_HIP; }
	*out = be;
	return FX_OK;
}

void fx_backend_destroy(fx_backend* be) {
	if (!be) return;
	if (be->stream) (void)hipStreamSynchronize(be->stream);
	for (hipEvent_t e : be->probe_a) (void)hipEventDestroy(e);
	for (hipEvent_t e : be->probe_b) (void)hipEventDestroy(e);
	dev_free(be->partials);
	dev_free(be->dres);
	if (be->hres) (void)hipHostFree(be->hres);
	if (be->stream) (void)hipStreamDestroy(be->stream);
	delete be;
}

const char* fx_backend_name(const fx_backend* be) { return be ? be->name : "hip"; }
void* fx_backend_stream(fx_backend* be) { return be ? (void*)be->stream : nullptr; }

fx_status fx_probe_set(fx_backend* be, int32_t kind) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_REQUIRE(kind >= FX_PROBE_OFF && kind <= FX_PROBE_REDUCE, FX_ERR_ARGUMENT, "bad probe kind %d", kind);
	FX_HIP(hipStreamSynchronize(be->stream));
	if (kind != FX_PROBE_OFF && be->probe_a.empty()) {
		be->probe_a.resize(256);
		be->probe_b.resize(256);
		for (size_t i = 0; i < be->probe_a.size(); ++i) {
			FX_HIP(hipEventCreate(&be->probe_a[i]));
			FX_HIP(hipEventCreate(&be->probe_b[i]));
		}
	}
	be->probe_kind = kind;
	be->probe_used = 0;
	be->probe_ms.clear();
	return FX_OK;
}

fx_status fx_probe_read(fx_backend* be, double* total_ms, uint64_t* launches) {
	FX_REQUIRE(be && total_ms && launches, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(fx::probe_drain(be));
	std::vector<float> v = be->probe_ms;
	double tot = 0.0;
	uint64_t cnt = 0;
	if (!v.empty()) {
		std::sort(v.begin(), v.end());
		const float cut = 0.1f * v[v.size() / 2];
		for (float m : v)
			if (m >= cut) {
				tot += m;
				++cnt;
			}
	}
	*total_ms = tot;
	*launches = cnt;
	return FX_OK;
}

fx_status fx_synchronize(fx_backend* be) {
	FX_REQUIRE(be, FX_ERR_ARGUMENT, "backend is null");
	FX_HIP(hipStreamSynchronize(be->stream));
	return FX_OK;
}

// Backend::allocate (backend.rs:55-59): zero-initialised.
fx_status fx_allocate(fx_backend* be, size_t len, fx_dtype dtype, fx_buf** out) {
	FX_REQUIRE(be && out, FX_ERR_ARGUMENT, "null argument");
	FX_REQUIRE(dtype == FX_F32 || dtype == FX_F64, FX_ERR_ARGUMENT, "unknown dtype %d", (int)dtype);
	*out = nullptr;
	size_t es = dtype == FX_F64 ? 8 : 4;
	void* p = nullptr;
	FX_CHECK(dev_alloc(be, len * es, &p));
	fx_buf* b = new fx_buf();
	b->be = be;
	b->ptr = p;
	b->len = len;
	b->dtype = dtype;
	*out = b;
	return FX_OK;
}

fx_status fx_upload_f64(fx_backend* be, const double* data, size_t len, fx_buf** out) {
	FX_CHECK(fx_allocate(be, len, FX_F64, out));
	if (len) {
		FX_HIP(hipMemcpyAsync((*out)->ptr, data, len * 8, hipMemcpyHostToDevice, be->stream));
		FX_HIP(hipStreamSynchronize(be->stream));
	}
	return FX_OK;
}

fx_status fx_upload_f32(fx_backend* be, const float* data, size_t len, fx_buf** out) {
	FX_CHECK(fx_allocate(be, len, FX_F32, out));
	if (len) {
		FX_HIP(hipMemcpyAsync((*out)->ptr, data, len * 4, hipMemcpyHostToDevice, be->stream));
		FX_HIP(hipStreamSynchronize(be->stream));
	}
	return FX_OK;
}

fx_status fx_write_f64(fx_backend* be, fx_buf* buf, const double* src, size_t len) {
	FX_REQUIRE(be && buf && src, FX_ERR_ARGUMENT, "null argument");
	FX_CHECK(check_buf(buf, len, FX_F64, "dst"));
	FX_HIP(hipMemcpyAsync(buf->ptr, src, len * 8, hipMemcpyHostToDevice, be->stream));
	FX_HIP(hipStreamSynchronize(be->stream));
	return FX_OK;
}

void fx_buffer_free(fx_buf* buf) {
	if (!buf) return;
	if (buf->be && buf->be->stream) (void)hipStreamSynchronize(buf->be->stream);
	dev_free(buf->ptr);
	delete buf;
}

size_t fx_buffer_len(const fx_buf* buf) { return buf ? buf->len : 0; }
fx_dtype fx_buffer_dtype(const fx_buf* buf) { return buf ? buf->dtype : FX_F64; }
void* fx_buffer_device_ptr(fx_buf* buf) { return buf ? buf->ptr : nullptr; }

// BackendBuffer::copy_to_host_f64 (backend.rs:27-28); a length mismatch is an
// error here instead of a panic.
fx_status fx_copy_to_host_f64(fx_backend* be, const fx_buf* buf, double* dst, size_t len) {
	FX_REQUIRE(be && buf && dst, FX_ERR_ARGUMENT, "null argument");
	FX_REQUIRE(buf->len == len, FX_ERR_DIMENSION, "dimension mismatch: expected %zu, got %zu", buf->len, len);
	if (buf->dtype == FX_F64) {
		FX_HIP(hipMemcpyAsync(dst, buf->ptr, len * 8, hipMemcpyDeviceToHost, be->stream));
		FX_HIP(hipStreamSynchronize(be->stream));
	} else {
		std::vector<float> tmp(len);
		FX_HIP(hipMemcpyAsync(tmp.data(), buf->ptr, len * 4, hipMemcpyDeviceToHost, be->stream));
		FX_HIP(hipStreamSynchronize(be->stream));
		for (size_t i = 0; i < len; ++i) dst[i] = tmp[i];
	}
	return FX_OK;
}

fx_status fx_copy_to_host_f32(fx_backend* be, const fx_buf* buf, float* dst, size_t len) {
	FX_REQUIRE(be && buf && dst, FX_ERR_ARGUMENT, "null argument");
	FX_REQUIRE(buf->len == len, FX_ERR_DIMENSION, "dimension mismatch: expected %zu, got %zu", buf->len, len);
	if (buf->dtype == FX_F32) {
		FX_HIP(hipMemcpyAsync(dst, buf->ptr, len * 4, hipMemcpyDeviceToHost, be->stream));
		FX_HIP(hipStreamSynchronize(be->stream));
	} else {
		std::vector<double> tmp(len);
		FX_HIP(hipMemcpyAsync(tmp.data(), buf->ptr, len * 8, hipMemcpyDeviceToHost, be->stream));
		FX_HIP(hipStreamSynchronize(be->stream));
		for (size_t i = 0; i < len; ++i) dst[i] = (float)tmp[i];
	}
	return FX_OK;
}


fx_status fx_hbm_ceilings(fx_backend* be, size_t n, int32_t reps, double* out_us) {
	FX_REQUIRE(be && out_us && reps > 0 && n >= 2 && n % 2 == 0, FX_ERR_ARGUMENT,
	           "backend, output, reps > 0 and an even n >= 2 required");
	constexpr int NC = 4;   // rotating copies, so that consecutive launches do not hit in the last-level cache
	// Per copy: z (FP32) and six FP64 arrays; the other kinds reuse them.
	const size_t per = n * 4 + 6 * n * 8;
	void* mem = nullptr;
	FX_CHECK(dev_alloc(be, NC * per, &mem));
	hipStream_t st = be->stream;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	fx_status status = FX_OK;
	auto run = [&]() -> fx_status {
		FX_HIP(hipMemsetAsync(mem, 0, NC * per, st));
		FX_HIP(hipEventCreate(&e0));
		FX_HIP(hipEventCreate(&e1));
		const size_t n2 = n / 2;
		// Enough workgroups to fill the chip several times over, each thread looping over a few pairs.
		const unsigned blocks = (unsigned)std::min<size_t>(16384, (n2 + fxd::BLOCK - 1) / fxd::BLOCK);
		auto arr = [&](int c, int j) { return (char*)mem + c * per + n * 4 + (size_t)j * n * 8; };
		auto zf = [&](int c) { return (const float2*)((char*)mem + c * per); };
		double* sink = (double*)arr(0, 5);
		for (int kind = 0; kind < FX_HBM_NKINDS; ++kind) {
			auto launch = [&](int c) {
				switch (kind) {
				case FX_HBM_READ:
					hipLaunchKernelGGL(k_hbm_read, dim3(blocks), dim3(fxd::BLOCK), 0, st, (const double2*)arr(c, 0), sink, n2);
					break;
				case FX_HBM_WRITE:
					hipLaunchKernelGGL(k_hbm_write, dim3(blocks), dim3(fxd::BLOCK), 0, st, (double2*)arr(c, 1), n2);
					break;
				case FX_HBM_COPY:
					hipLaunchKernelGGL(k_hbm_copy, dim3(blocks), dim3(fxd::BLOCK), 0, st, (const double2*)arr(c, 0),
					                   (double2*)arr(c, 1), n2);
					break;
				default:
					hipLaunchKernelGGL(k_hbm_mix52, dim3(blocks), dim3(fxd::BLOCK), 0, st, zf(c), (const double2*)arr(c, 0),
					                   (const double2*)arr(c, 1), (double2*)arr(c, 2), (double2*)arr(c, 3),
					                   (double2*)arr(c, 4), n2);
				}
			};
			for (int w = 0; w < 3; ++w) launch(w % NC);
			FX_HIP(hipEventRecord(e0, st));
			for (int r = 0; r < reps; ++r) launch(r % NC);
			FX_HIP(hipEventRecord(e1, st));
			FX_HIP(hipEventSynchronize(e1));
			FX_CHECK(check_launch("hbm ceilings"));
			float ms = 0.0f;
			FX_HIP(hipEventElapsedTime(&ms, e0, e1));
			out_us[kind] = 1e3 * (double)ms / reps;
		}
		return FX_OK;
	};
	status = run();
	(void)hipStreamSynchronize(st);
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	dev_free(mem);
	return status;
}

}  // extern "C"
