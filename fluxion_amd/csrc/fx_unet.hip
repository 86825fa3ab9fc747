// fx_unet.hip — learned preconditioner network on gfx950 (no reference counterpart).
//
// A 4-level U-Net (3×3 convolutions, 16/32/48/64 channels) with BF16 weights
// and activations and FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16_bf16).  Layer order, parameter layout and loss are
// those of oracle/unet.cpp (the FP64 checker):
//   enc l=0..3: relu(conv3) ×2, skip, avgpool2 (l<3)
//   dec l=2..0: up2(nearest) ‖ skip_l, relu(conv3) ×2;   out: conv3 → 1 channel
// Activations are NHWC with the channel count padded to a multiple of 16.
//
// 3×3 convolution as an implicit GEMM: one wave computes 16 pixels of one
// image row × all output channels; K runs over (tap, 16-channel block).
// MFMA 16x16x16 operand layout (lane l, element j): A[m = l%16][k = 4(l/16)+j],
// B[k = 4(l/16)+j][n = l%16], D[m = 4(l/16)+j][n = l%16].
// The same kernel computes the data gradient with the transposed, flipped
// weights.  Weight gradients reduce over pixels with the pixel index as K.
#include "fx_internal.hpp"

#include <cmath>
#include <cstring>
#include <vector>

using namespace fxd;

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

__device__ __forceinline__ bf16 f2bf(float f) {
	uint32_t u = __float_as_uint(f);
	if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);   // inf / nan
	u += 0x7fffu + ((u >> 16) & 1u);                                    // round to nearest even
	return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 b) { return __uint_as_float((uint32_t)b << 16); }

constexpr int NLAYER = 15;
constexpr int CH[4] = {16, 32, 48, 64};

inline int pad16(int c) { return (c + 15) / 16 * 16; }

// Input kind of the conv kernel: BF16 activations, or FP32 gradients masked by
// the ReLU of the forward output (mask != nullptr) and rounded to BF16.
struct ConvArgs {
	const void* in;        // [B][H][W][cin_p]
	const bf16* mask;      // [B][H][W][cin_p] or null
	const bf16* w;         // [9][cout_p][cin_p]
	const float* bias;     // [cout] or null
	void* out;             // [B][H][W][cout_p] (bf16 or f32)
	int B, H, W, cin_p, cout_p, cout;
	int in_f32, out_f32, relu, flip;
	const int* skip;       // kernels return early while *skip != 0 (may be null)
};

__global__ __launch_bounds__(BLOCK) void k_conv3(ConvArgs a) {
	if (a.skip && *a.skip) return;
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int b = blockIdx.z, row = blockIdx.y;
	const int col0 = blockIdx.x * 64 + wv * 16;
	const int m = lane & 15, kq = lane >> 4;
	const int nt = a.cout_p / 16, cb = a.cin_p / 16;   // nt <= 8 (cout_p <= 128, checked on the host)
	f4 acc[8];
	for (int t = 0; t < 8; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
	for (int tap = 0; tap < 9; ++tap) {
		const int rr = row + tap / 3 - 1, cc = col0 + m + tap % 3 - 1;
		const bool ok = rr >= 0 && rr < a.H && cc >= 0 && cc < a.W;
		const size_t pix = ((size_t)b * a.H + (ok ? rr : 0)) * a.W + (ok ? cc : 0);
		const int wt = a.flip ? 8 - tap : tap;
		for (int c = 0; c < cb; ++c) {
			const int ch = c * 16 + kq * 4;
			s4 av = {0, 0, 0, 0};
			if (ok) {
				const size_t q = pix * a.cin_p + ch;
				if (a.in_f32) {
					const float* src = (const float*)a.in + q;
					float v0 = src[0], v1 = src[1], v2 = src[2], v3 = src[3];
					if (a.mask) {
						const bf16* mk = a.mask + q;
						v0 = bf2f(mk[0]) > 0.f ? v0 : 0.f;
						v1 = bf2f(mk[1]) > 0.f ? v1 : 0.f;
						v2 = bf2f(mk[2]) > 0.f ? v2 : 0.f;
						v3 = bf2f(mk[3]) > 0.f ? v3 : 0.f;
					}
					av = s4{(short)f2bf(v0), (short)f2bf(v1), (short)f2bf(v2), (short)f2bf(v3)};
				} else {
					av = *(const s4*)((const bf16*)a.in + q);
				}
			}
			for (int t = 0; t < 8; ++t) {
				if (t >= nt) break;
				const s4 bv = *(const s4*)(a.w + ((size_t)wt * a.cout_p + t * 16 + m) * a.cin_p + ch);
				acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
			}
		}
	}
	for (int t = 0; t < 8; ++t) {
		if (t >= nt) break;
		const int co = t * 16 + m;
		for (int j = 0; j < 4; ++j) {
			const int col = col0 + kq * 4 + j;
			if (col >= a.W) continue;
			float v = acc[t][j];
			if (co < a.cout && a.bias) v += a.bias[co];
			if (co >= a.cout) v = 0.f;
			if (a.relu && v < 0.f) v = 0.f;
			const size_t q = (((size_t)b * a.H + row) * a.W + col) * a.cout_p + co;
			if (a.out_f32) ((float*)a.out)[q] = v;
			else ((bf16*)a.out)[q] = f2bf(v);
		}
	}
}

// dW[tap][co][ci] += Σ_p g[p][co]·x[p+off][ci], g = dy masked by the ReLU of y.
// One wave: a 16(co) × 16(ci) tile of one tap over a chunk of pixels.
__global__ __launch_bounds__(BLOCK) void k_conv3_wgrad(const float* dy, const bf16* y, int relu, const bf16* x,
                                                       float* dW, int B, int H, int W, int cin_p, int cout_p,
                                                       int cin, int cout, int chunk) {
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int tap = blockIdx.y, tile = blockIdx.z;
	const int nco = cout_p / 16;
	const int cot = tile % nco, cit = tile / nco;
	const int m = lane & 15, kq = lane >> 4;
	const int dyo = tap / 3 - 1, dxo = tap % 3 - 1;
	const long total = (long)B * H * W;
	const long p0 = ((long)blockIdx.x * 4 + wv) * chunk;
	f4 acc = {0.f, 0.f, 0.f, 0.f};
	for (long pb = p0; pb < p0 + chunk && pb < total; pb += 16) {
		s4 av, bv;
		short ae[4], be_[4];
		for (int j = 0; j < 4; ++j) {
			long p = pb + kq * 4 + j;
			float g = 0.f;
			float xv = 0.f;
			if (p < total) {
				const int co = cot * 16 + m;
				g = dy[p * cout_p + co];
				if (relu && !(bf2f(y[p * cout_p + co]) > 0.f)) g = 0.f;
				long bb = p / ((long)H * W);
				int rem = (int)(p - bb * H * W);
				int rr = rem / W + dyo, cc = rem % W + dxo;
				if (rr >= 0 && rr < H && cc >= 0 && cc < W)
					xv = bf2f(x[(((size_t)bb * H + rr) * W + cc) * cin_p + cit * 16 + m]);
			}
			ae[j] = (short)f2bf(g);
			be_[j] = (short)f2bf(xv);
		}
		av = s4{ae[0], ae[1], ae[2], ae[3]};
		bv = s4{be_[0], be_[1], be_[2], be_[3]};
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
	}
	for (int j = 0; j < 4; ++j) {
		const int co = cot * 16 + kq * 4 + j, ci = cit * 16 + m;
		if (co < cout && ci < cin) atomicAdd(&dW[((size_t)tap * cout + co) * cin + ci], acc[j]);
	}
}

// db[co] += Σ_p g[p][co].
__global__ void k_bgrad(const float* dy, const bf16* y, int relu, float* db, long total, int cout_p, int cout) {
	const int co = threadIdx.x & 63;
	if (co >= cout) return;
	float s = 0.f;
	for (long p = (long)blockIdx.x * 4 + (threadIdx.x >> 6); p < total; p += (long)gridDim.x * 4) {
		float g = dy[p * cout_p + co];
		if (relu && !(bf2f(y[p * cout_p + co]) > 0.f)) g = 0.f;
		s += g;
	}
	atomicAdd(&db[co], s);
}

__global__ void k_pool2(const bf16* x, bf16* y, int B, int H, int W, int C, const int* skip) {
	if (skip && *skip) return;
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	int h = H / 2, w = W / 2;
	long n = (long)B * h * w * C;
	if (q >= n) return;
	int c = q % C;
	long p = q / C;
	int cc = p % w, rr = (p / w) % h;
	long b = p / ((long)w * h);
	auto X = [&](int r, int s) { return bf2f(x[(((size_t)b * H + r) * W + s) * C + c]); };
	y[q] = f2bf(0.25f * (X(2 * rr, 2 * cc) + X(2 * rr, 2 * cc + 1) + X(2 * rr + 1, 2 * cc) + X(2 * rr + 1, 2 * cc + 1)));
}
// dx (f32, [B][H][W][C]) += 0.25·dy[B][H/2][W/2][C]
__global__ void k_pool2_bwd(const float* dy, float* dx, int B, int H, int W, int C) {
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	long n = (long)B * H * W * C;
	if (q >= n) return;
	int c = q % C;
	long p = q / C;
	int cc = p % W, rr = (p / W) % H;
	long b = p / ((long)W * H);
	dx[q] += 0.25f * dy[(((size_t)b * (H / 2) + rr / 2) * (W / 2) + cc / 2) * C + c];
}
// out[B][H][W][ca+cs (padded co)] = up2(a) ‖ s
__global__ void k_upcat(const bf16* a, int ca, const bf16* s, int cs, bf16* out, int co_p, int B, int H, int W,
                        const int* skip) {
	if (skip && *skip) return;
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	long n = (long)B * H * W * co_p;
	if (q >= n) return;
	int c = q % co_p;
	long p = q / co_p;
	int cc = p % W, rr = (p / W) % H;
	long b = p / ((long)W * H);
	bf16 v = 0;
	if (c < ca) v = a[(((size_t)b * (H / 2) + rr / 2) * (W / 2) + cc / 2) * ca + c];
	else if (c < ca + cs) v = s[(((size_t)b * H + rr) * W + cc) * cs + (c - ca)];
	out[q] = v;
}
// Splits d(up2(a) ‖ s): da[B][H/2][W/2][ca] = Σ 2x2, ds[B][H][W][cs] (+=).
__global__ void k_upcat_bwd_a(const float* d, int co_p, int ca, float* da, int B, int H, int W) {
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	int h = H / 2, w = W / 2;
	long n = (long)B * h * w * ca;
	if (q >= n) return;
	int c = q % ca;
	long p = q / ca;
	int cc = p % w, rr = (p / w) % h;
	long b = p / ((long)w * h);
	auto D = [&](int r, int s) { return d[(((size_t)b * H + r) * W + s) * co_p + c]; };
	da[q] = D(2 * rr, 2 * cc) + D(2 * rr, 2 * cc + 1) + D(2 * rr + 1, 2 * cc) + D(2 * rr + 1, 2 * cc + 1);
}
__global__ void k_upcat_bwd_s(const float* d, int co_p, int ca, int cs, float* ds, long npix) {
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= npix * cs) return;
	int c = q % cs;
	long p = q / cs;
	ds[q] += d[p * co_p + ca + c];
}
// Relative squared error per sample: partial sums of (y−t)² and t².
__global__ void k_loss(const float* y, int ystride, const float* t, float* sums, int B, long hw) {
	const int b = blockIdx.y;
	float e = 0.f, tt = 0.f;
	for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < hw; p += (long)gridDim.x * blockDim.x) {
		float yv = y[((size_t)b * hw + p) * ystride], tv = t[(size_t)b * hw + p];
		e += (yv - tv) * (yv - tv);
		tt += tv * tv;
	}
	for (int off = 32; off > 0; off >>= 1) {
		e += __shfl_xor(e, off);
		tt += __shfl_xor(tt, off);
	}
	if ((threadIdx.x & 63) == 0) {
		atomicAdd(&sums[2 * b], e);
		atomicAdd(&sums[2 * b + 1], tt);
	}
}
__global__ void k_loss_grad(const float* y, int ystride, const float* t, const float* sums, float* dy, int B, long hw) {
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= (long)B * hw) return;
	int b = q / hw;
	float tt = sums[2 * b + 1];
	dy[q * ystride] = 2.f * (y[q * ystride] - t[q]) / tt;
}
// Relative energy-norm loss per sample: dᵀA d / tᵀA t with d = y − t and A the unit-spacing
// 5-point −∇·(k∇·) on interior pixels (k = exp of input channel 1, harmonic face means,
// zero outside the interior).  Writes A·d to dy (k_eloss_scale turns it into the gradient).
__device__ __forceinline__ float hmk(float a, float b) { return 2.f * a * b / (a + b); }
__global__ void k_eloss(const float* x, int cin, const float* y, int ystride, const float* t, float* sums, float* dy,
                        int H, int W) {
	const int b = blockIdx.y;
	const long hw = (long)H * W;
	const size_t base = (size_t)b * hw;
	auto inner = [&](int r, int c) { return r > 0 && c > 0 && r < H - 1 && c < W - 1; };
	auto K = [&](long p) { return expf(x[(base + p) * cin + 1]); };
	auto D = [&](int r, int c) { return inner(r, c) ? y[(base + (long)r * W + c) * ystride] - t[base + (long)r * W + c] : 0.f; };
	auto T = [&](int r, int c) { return inner(r, c) ? t[base + (long)r * W + c] : 0.f; };
	float e = 0.f, tt = 0.f;
	for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < hw; p += (long)gridDim.x * blockDim.x) {
		const int r = (int)(p / W), c = (int)(p % W);
		if (!inner(r, c)) {
			dy[(base + p) * ystride] = 0.f;
			continue;
		}
		const float kc = K(p);
		const float hw_ = hmk(kc, K(p - 1)), he = hmk(kc, K(p + 1)), hn = hmk(kc, K(p - W)), hs = hmk(kc, K(p + W));
		const float dc = D(r, c), tc = T(r, c);
		const float ad = hw_ * (dc - D(r, c - 1)) + he * (dc - D(r, c + 1)) + hn * (dc - D(r - 1, c)) + hs * (dc - D(r + 1, c));
		const float at = hw_ * (tc - T(r, c - 1)) + he * (tc - T(r, c + 1)) + hn * (tc - T(r - 1, c)) + hs * (tc - T(r + 1, c));
		dy[(base + p) * ystride] = ad;
		e += dc * ad;
		tt += tc * at;
	}
	for (int off = 32; off > 0; off >>= 1) {
		e += __shfl_xor(e, off);
		tt += __shfl_xor(tt, off);
	}
	if ((threadIdx.x & 63) == 0) {
		atomicAdd(&sums[2 * b], e);
		atomicAdd(&sums[2 * b + 1], tt);
	}
}
__global__ void k_eloss_scale(float* dy, int ystride, const float* sums, int B, long hw) {
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= (long)B * hw) return;
	int b = q / hw;
	dy[q * ystride] *= 2.f / sums[2 * b + 1];
}
// FP32 master weights [tap][cout][cin] -> BF16 padded [tap][cout_p][cin_p] and the
// transposed copy [tap][cin_p][cout_p] used for the data gradient.
__global__ void k_pack_w(const float* w, bf16* wp, bf16* wt, int cin, int cout, int cin_p, int cout_p) {
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	long n = 9L * cout_p * cin_p;
	if (q >= n) return;
	int ci = q % cin_p;
	int co = (q / cin_p) % cout_p;
	int t = q / ((long)cin_p * cout_p);
	float v = (ci < cin && co < cout) ? w[((size_t)t * cout + co) * cin + ci] : 0.f;
	bf16 b = f2bf(v);
	wp[q] = b;
	wt[((size_t)t * cin_p + ci) * cout_p + co] = b;
}
__global__ void k_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n) return;
	float gg = g[q];
	float mm = b1 * m[q] + (1.f - b1) * gg;
	float vv = b2 * v[q] + (1.f - b2) * gg * gg;
	m[q] = mm;
	v[q] = vv;
	p[q] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}
// NHWC f32 input [B][H][W][cin] -> BF16 [B][H][W][cin_p].
__global__ void k_pack_in(const float* x, bf16* out, long npix, int cin, int cin_p, const int* skip) {
	if (skip && *skip) return;
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= npix * cin_p) return;
	int c = q % cin_p;
	long p = q / cin_p;
	out[q] = c < cin ? f2bf(x[p * cin + c]) : (bf16)0;
}

// y[p] = src[p·stride] (channel 0 of the padded network output).
__global__ void k_extract(const float* src, int stride, float* y, long n) {
	long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (q < n) y[q] = src[q * stride];
}

unsigned nb(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// ════════════════════════════════════════════════════════════════════════
struct fx_unet {
	fx_backend* be = nullptr;
	int cin = 2;
	int lcin[NLAYER], lcout[NLAYER];
	size_t woff[NLAYER], boff[NLAYER], nparams = 0;
	float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;   // FP32 master, grad, Adam moments
	bf16 *Wp[NLAYER] = {}, *Wt[NLAYER] = {};
	long step = 0;
	// Activation workspace for the current (B, H, W).
	int B = 0, H = 0, W = 0;
	const int* skip = nullptr;   // set for inference inside a solver
	std::vector<void*> ws;
	bf16 *xin = nullptr, *act[NLAYER] = {}, *cat[3] = {}, *pool[3] = {};
	float* yout = nullptr;   // layer-14 output, FP32 [B][H][W][16]
	float *g1 = nullptr, *g2 = nullptr;   // gradient ping-pong, sized for the largest activation
	float *gskip[3] = {};
	float* sums = nullptr;
	float* dy = nullptr;
};

namespace {

fx_status alloc(fx_unet* u, size_t bytes, void** p) {
	FX_CHECK(fx::dev_alloc(u->be, bytes, p));
	u->ws.push_back(*p);
	return FX_OK;
}

void free_ws(fx_unet* u) {
	if (u->be) (void)hipStreamSynchronize(u->be->stream);
	for (void* p : u->ws) fx::dev_free(p);
	u->ws.clear();
	u->B = u->H = u->W = 0;
}

int level_of(int li) { return li < 8 ? li / 2 : (li < 14 ? 2 - (li - 8) / 2 : 0); }

// Buffers for batch B of H×W images.
fx_status ensure_ws(fx_unet* u, int B, int H, int W) {
	if (u->B == B && u->H == H && u->W == W) return FX_OK;
	free_ws(u);
	u->B = B; u->H = H; u->W = W;
	auto npix = [&](int l) { return (size_t)B * (H >> l) * (W >> l); };
	void* p;
	FX_CHECK(alloc(u, npix(0) * pad16(u->cin) * 2, &p)); u->xin = (bf16*)p;
	for (int li = 0; li < NLAYER; ++li) {
		int l = level_of(li);
		FX_CHECK(alloc(u, npix(l) * pad16(u->lcout[li]) * 2, &p));
		u->act[li] = (bf16*)p;
	}
	for (int l = 0; l < 3; ++l) {
		FX_CHECK(alloc(u, npix(l) * pad16(CH[l + 1] + CH[l]) * 2, &p)); u->cat[l] = (bf16*)p;
		FX_CHECK(alloc(u, npix(l + 1) * CH[l] * 2, &p)); u->pool[l] = (bf16*)p;
		FX_CHECK(alloc(u, npix(l) * CH[l] * 4, &p)); u->gskip[l] = (float*)p;
	}
	FX_CHECK(alloc(u, npix(0) * 16 * 4, &p)); u->yout = (float*)p;
	size_t gmax = npix(0) * 128;
	FX_CHECK(alloc(u, gmax * 4, &p)); u->g1 = (float*)p;
	FX_CHECK(alloc(u, gmax * 4, &p)); u->g2 = (float*)p;
	FX_CHECK(alloc(u, (size_t)B * 2 * 4, &p)); u->sums = (float*)p;
	FX_CHECK(alloc(u, npix(0) * 16 * 4, &p)); u->dy = (float*)p;
	return FX_OK;
}

fx_status pack_weights(fx_unet* u) {
	for (int li = 0; li < NLAYER; ++li) {
		long n = 9L * pad16(u->lcout[li]) * pad16(u->lcin[li]);
		hipLaunchKernelGGL(k_pack_w, dim3(nb(n)), dim3(256), 0, u->be->stream, u->P + u->woff[li], u->Wp[li], u->Wt[li],
		                   u->lcin[li], u->lcout[li], pad16(u->lcin[li]), pad16(u->lcout[li]));
	}
	return fx::check_launch("unet pack weights");
}

fx_status conv(fx_unet* u, int li, const void* in, int in_f32, const bf16* mask, void* out, int out_f32, int l,
               bool data_grad) {
	ConvArgs a;
	a.in = in;
	a.mask = mask;
	a.B = u->B;
	a.H = u->H >> l;
	a.W = u->W >> l;
	if (!data_grad) {
		a.w = u->Wp[li];
		a.bias = u->P + u->boff[li];
		a.cin_p = pad16(u->lcin[li]);
		a.cout_p = pad16(u->lcout[li]);
		a.cout = u->lcout[li];
		a.relu = li < NLAYER - 1;
		a.flip = 0;
	} else {
		a.w = u->Wt[li];
		a.bias = nullptr;
		a.cin_p = pad16(u->lcout[li]);
		a.cout_p = pad16(u->lcin[li]);
		a.cout = u->lcin[li];
		a.relu = 0;
		a.flip = 1;
	}
	FX_REQUIRE(a.cout_p <= 128 && a.cin_p <= 128, FX_ERR_ARGUMENT, "conv channel count out of range");
	a.out = out;
	a.skip = data_grad ? nullptr : u->skip;
	a.in_f32 = in_f32;
	a.out_f32 = out_f32;
	dim3 grid((unsigned)((a.W + 63) / 64), (unsigned)a.H, (unsigned)a.B);
	hipLaunchKernelGGL(k_conv3, grid, dim3(BLOCK), 0, u->be->stream, a);
	return fx::check_launch("unet conv");
}

// Forward over the packed input u->xin; result in u->yout (channel 0 of 16).
fx_status forward(fx_unet* u) {
	hipStream_t st = u->be->stream;
	const bf16* h = u->xin;
	int li = 0;
	for (int l = 0; l < 4; ++l) {
		for (int k = 0; k < 2; ++k, ++li) {
			FX_CHECK(conv(u, li, h, 0, nullptr, u->act[li], 0, l, false));
			h = u->act[li];
		}
		if (l < 3) {
			long n = (long)u->B * (u->H >> (l + 1)) * (u->W >> (l + 1)) * CH[l];
			hipLaunchKernelGGL(k_pool2, dim3(nb(n)), dim3(256), 0, st, h, u->pool[l], u->B, u->H >> l, u->W >> l, CH[l],
			                   u->skip);
			h = u->pool[l];
		}
	}
	for (int l = 2; l >= 0; --l) {
		int ca = CH[l + 1], cs = CH[l], cop = pad16(ca + cs);
		long n = (long)u->B * (u->H >> l) * (u->W >> l) * cop;
		hipLaunchKernelGGL(k_upcat, dim3(nb(n)), dim3(256), 0, st, h, ca, (const bf16*)u->act[2 * l + 1], cs, u->cat[l],
		                   cop, u->B, u->H >> l, u->W >> l, u->skip);
		h = u->cat[l];
		for (int k = 0; k < 2; ++k, ++li) {
			FX_CHECK(conv(u, li, h, 0, nullptr, u->act[li], 0, l, false));
			h = u->act[li];
		}
	}
	FX_CHECK(conv(u, li, h, 0, nullptr, u->yout, 1, 0, false));
	return fx::check_launch("unet forward");
}

// Backward from u->dy ([B][H][W][16] f32, channel 0) into u->G.
fx_status backward(fx_unet* u) {
	hipStream_t st = u->be->stream;
	FX_HIP(hipMemsetAsync(u->G, 0, u->nparams * 4, st));
	for (int l = 0; l < 3; ++l)
		FX_HIP(hipMemsetAsync(u->gskip[l], 0, (size_t)u->B * (u->H >> l) * (u->W >> l) * CH[l] * 4, st));
	auto wgrad = [&](int li, const float* g, const bf16* yact, int relu, const bf16* xin, int l) -> fx_status {
		int Hl = u->H >> l, Wl = u->W >> l;
		long total = (long)u->B * Hl * Wl;
		int chunk = 256;
		int cip = pad16(u->lcin[li]), cop = pad16(u->lcout[li]);
		dim3 grid((unsigned)((total + 4L * chunk - 1) / (4L * chunk)), 9, (unsigned)((cop / 16) * (cip / 16)));
		hipLaunchKernelGGL(k_conv3_wgrad, grid, dim3(BLOCK), 0, st, g, yact, relu, xin, u->G + u->woff[li], u->B, Hl, Wl,
		                   cip, cop, u->lcin[li], u->lcout[li], chunk);
		hipLaunchKernelGGL(k_bgrad, dim3(256), dim3(256), 0, st, g, yact, relu, u->G + u->boff[li], total, cop,
		                   u->lcout[li]);
		return fx::check_launch("unet wgrad");
	};
	// Inputs of each layer in forward order.
	const bf16* lin[NLAYER];
	{
		const bf16* h = u->xin;
		int li = 0;
		for (int l = 0; l < 4; ++l) {
			for (int k = 0; k < 2; ++k, ++li) { lin[li] = h; h = u->act[li]; }
			if (l < 3) h = u->pool[l];
		}
		for (int l = 2; l >= 0; --l) {
			h = u->cat[l];
			for (int k = 0; k < 2; ++k, ++li) { lin[li] = h; h = u->act[li]; }
		}
		lin[li] = h;
	}
	// Output layer (no ReLU): gradient wrt its pre-activation is dy itself.
	int li = NLAYER - 1;
	FX_CHECK(wgrad(li, u->dy, nullptr, 0, lin[li], 0));
	float* g = u->g1;
	float* gn = u->g2;
	FX_CHECK(conv(u, li, u->dy, 1, nullptr, g, 1, 0, true));   // d(act13) (16 ch)
	--li;
	for (int l = 0; l <= 2; ++l) {
		for (int k = 0; k < 2; ++k, --li) {
			FX_CHECK(wgrad(li, g, u->act[li], 1, lin[li], l));
			FX_CHECK(conv(u, li, g, 1, u->act[li], gn, 1, l, true));
			std::swap(g, gn);
		}
		// g = d(cat_l) [B][H_l][W_l][pad16(ca+cs)] -> d(a) at level l+1, d(skip_l) accumulated.
		int ca = CH[l + 1], cs = CH[l], cop = pad16(ca + cs);
		long npix = (long)u->B * (u->H >> l) * (u->W >> l);
		hipLaunchKernelGGL(k_upcat_bwd_s, dim3(nb(npix * cs)), dim3(256), 0, st, g, cop, ca, cs, u->gskip[l], npix);
		long na = (long)u->B * (u->H >> (l + 1)) * (u->W >> (l + 1)) * ca;
		hipLaunchKernelGGL(k_upcat_bwd_a, dim3(nb(na)), dim3(256), 0, st, g, cop, ca, gn, u->B, u->H >> l, u->W >> l);
		std::swap(g, gn);
	}
	for (int l = 3; l >= 0; --l) {
		if (l < 3) {
			// d(act at level l) = pool_bwd(g) + d(skip_l)
			long n = (long)u->B * (u->H >> l) * (u->W >> l) * CH[l];
			FX_HIP(hipMemcpyAsync(gn, u->gskip[l], n * 4, hipMemcpyDeviceToDevice, st));
			hipLaunchKernelGGL(k_pool2_bwd, dim3(nb(n)), dim3(256), 0, st, g, gn, u->B, u->H >> l, u->W >> l, CH[l]);
			std::swap(g, gn);
		}
		for (int k = 0; k < 2; ++k, --li) {
			FX_CHECK(wgrad(li, g, u->act[li], 1, lin[li], l));
			if (li > 0) {
				FX_CHECK(conv(u, li, g, 1, u->act[li], gn, 1, l, true));
				std::swap(g, gn);
			}
		}
	}
	return fx::check_launch("unet backward");
}

}  // namespace

using namespace fx;

extern "C" {

fx_status fx_unet_create(fx_backend* be, int32_t cin, fx_unet** out) {
	FX_REQUIRE(be && out && cin >= 1 && cin <= 16, FX_ERR_ARGUMENT, "bad arguments");
	fx_unet* u = new fx_unet();
	u->be = be;
	u->cin = cin;
	int c = cin, i = 0;
	for (int l = 0; l < 4; ++l) {
		u->lcin[i] = c; u->lcout[i++] = CH[l];
		u->lcin[i] = CH[l]; u->lcout[i++] = CH[l];
		c = CH[l];
	}
	for (int l = 2; l >= 0; --l) {
		u->lcin[i] = c + CH[l]; u->lcout[i++] = CH[l];
		u->lcin[i] = CH[l]; u->lcout[i++] = CH[l];
		c = CH[l];
	}
	u->lcin[i] = c; u->lcout[i++] = 1;
	size_t o = 0;
	for (int l = 0; l < NLAYER; ++l) { u->woff[l] = o; o += (size_t)9 * u->lcin[l] * u->lcout[l]; }
	for (int l = 0; l < NLAYER; ++l) { u->boff[l] = o; o += u->lcout[l]; }
	u->nparams = o;
	void* p;
	fx_status s;
	auto fail = [&](fx_status st) { fx_unet_destroy(u); return st; };
	if ((s = dev_alloc(be, o * 4, &p)) != FX_OK) return fail(s);
	u->P = (float*)p;
	if ((s = dev_alloc(be, o * 4, &p)) != FX_OK) return fail(s);
	u->G = (float*)p;
	if ((s = dev_alloc(be, o * 4, &p)) != FX_OK) return fail(s);
	u->M = (float*)p;
	if ((s = dev_alloc(be, o * 4, &p)) != FX_OK) return fail(s);
	u->V = (float*)p;
	for (int l = 0; l < NLAYER; ++l) {
		size_t n = (size_t)9 * pad16(u->lcin[l]) * pad16(u->lcout[l]);
		if ((s = dev_alloc(be, n * 2, &p)) != FX_OK) return fail(s);
		u->Wp[l] = (bf16*)p;
		if ((s = dev_alloc(be, n * 2, &p)) != FX_OK) return fail(s);
		u->Wt[l] = (bf16*)p;
	}
	*out = u;
	return FX_OK;
}

void fx_unet_destroy(fx_unet* u) {
	if (!u) return;
	free_ws(u);
	dev_free(u->P); dev_free(u->G); dev_free(u->M); dev_free(u->V);
	for (int l = 0; l < NLAYER; ++l) { dev_free(u->Wp[l]); dev_free(u->Wt[l]); }
	delete u;
}

size_t fx_unet_nparams(const fx_unet* u) { return u ? u->nparams : 0; }

fx_status fx_unet_set_params(fx_unet* u, const float* host, size_t n) {
	FX_REQUIRE(u && host && n == u->nparams, FX_ERR_DIMENSION, "expected %zu parameters", u ? u->nparams : 0);
	FX_HIP(hipMemcpyAsync(u->P, host, n * 4, hipMemcpyHostToDevice, u->be->stream));
	FX_HIP(hipMemsetAsync(u->M, 0, n * 4, u->be->stream));
	FX_HIP(hipMemsetAsync(u->V, 0, n * 4, u->be->stream));
	u->step = 0;
	FX_CHECK(pack_weights(u));
	FX_HIP(hipStreamSynchronize(u->be->stream));
	return FX_OK;
}

fx_status fx_unet_get_params(fx_unet* u, float* host, size_t n) {
	FX_REQUIRE(u && host && n == u->nparams, FX_ERR_DIMENSION, "expected %zu parameters", u ? u->nparams : 0);
	FX_HIP(hipMemcpyAsync(host, u->P, n * 4, hipMemcpyDeviceToHost, u->be->stream));
	FX_HIP(hipStreamSynchronize(u->be->stream));
	return FX_OK;
}

fx_status fx_unet_get_grad(fx_unet* u, float* host, size_t n) {
	FX_REQUIRE(u && host && n == u->nparams, FX_ERR_DIMENSION, "expected %zu parameters", u ? u->nparams : 0);
	FX_HIP(hipMemcpyAsync(host, u->G, n * 4, hipMemcpyDeviceToHost, u->be->stream));
	FX_HIP(hipStreamSynchronize(u->be->stream));
	return FX_OK;
}

static fx_status run_forward(fx_unet* u, const float* x, int32_t B, int32_t H, int32_t W) {
	FX_REQUIRE(u && x && B > 0 && H % 8 == 0 && W % 8 == 0 && H > 0 && W > 0, FX_ERR_ARGUMENT,
	           "batch > 0 and H, W divisible by 8 required");
	FX_CHECK(ensure_ws(u, B, H, W));
	long npix = (long)B * H * W;
	hipLaunchKernelGGL(k_pack_in, dim3(nb(npix * pad16(u->cin))), dim3(256), 0, u->be->stream, x, u->xin, npix, u->cin,
	                   pad16(u->cin), u->skip);
	return forward(u);
}


static fx_status check_tensors(fx_unet* u, const fx_buf* x, const fx_buf* y, int32_t B, int32_t H, int32_t W) {
	FX_REQUIRE(u && x && y && B > 0 && H > 0 && W > 0, FX_ERR_ARGUMENT, "bad arguments");
	FX_REQUIRE(H % 8 == 0 && W % 8 == 0, FX_ERR_DIMENSION,
	           "the U-Net pools three times: image sides must be multiples of 8, got %dx%d", H, W);
	size_t npix = (size_t)B * H * W;
	FX_CHECK(check_buf(x, npix * u->cin, FX_F32, "x"));
	FX_CHECK(check_buf(y, npix, FX_F32, "y/t"));
	return FX_OK;
}

// x: FP32 [B][H][W][cin]; y: FP32 [B][H][W].
fx_status fx_unet_forward(fx_unet* u, const fx_buf* xb, fx_buf* yb, int32_t B, int32_t H, int32_t W) {
	FX_CHECK(check_tensors(u, xb, yb, B, H, W));
	FX_CHECK(run_forward(u, (const float*)xb->ptr, B, H, W));
	long npix = (long)B * H * W;
	hipLaunchKernelGGL(k_extract, dim3(nb(npix)), dim3(256), 0, u->be->stream, (const float*)u->yout, 16,
	                   (float*)yb->ptr, npix);
	return check_launch("unet forward");
}

// Loss Σ_b ‖y_b − t_b‖²/‖t_b‖² and its gradient (kept on the device; fx_unet_get_grad reads it).
fx_status fx_unet_loss_grad(fx_unet* u, const fx_buf* xb, const fx_buf* tb, int32_t B, int32_t H, int32_t W,
                            double* loss) {
	FX_CHECK(check_tensors(u, xb, tb, B, H, W));
	const float* t = (const float*)tb->ptr;
	FX_CHECK(run_forward(u, (const float*)xb->ptr, B, H, W));
	hipStream_t st = u->be->stream;
	long hw = (long)H * W;
	FX_HIP(hipMemsetAsync(u->sums, 0, (size_t)B * 2 * 4, st));
	hipLaunchKernelGGL(k_loss, dim3(64, B), dim3(256), 0, st, u->yout, 16, t, u->sums, B, hw);
	FX_HIP(hipMemsetAsync(u->dy, 0, (size_t)B * hw * 16 * 4, st));
	hipLaunchKernelGGL(k_loss_grad, dim3(nb(B * hw)), dim3(256), 0, st, u->yout, 16, t, u->sums, u->dy, B, hw);
	FX_CHECK(check_launch("unet loss"));
	FX_CHECK(backward(u));
	std::vector<float> sums((size_t)B * 2);
	FX_HIP(hipMemcpyAsync(sums.data(), u->sums, sums.size() * 4, hipMemcpyDeviceToHost, st));
	FX_HIP(hipStreamSynchronize(st));
	double L = 0;
	for (int b = 0; b < B; ++b) L += sums[2 * b] / sums[2 * b + 1];
	if (loss) *loss = L;
	return FX_OK;
}

// Relative energy-norm loss (k_eloss) and its gradient (kept on the device).
fx_status fx_unet_loss_grad_energy(fx_unet* u, const fx_buf* xb, const fx_buf* tb, int32_t B, int32_t H, int32_t W,
                                   double* loss) {
	FX_CHECK(check_tensors(u, xb, tb, B, H, W));
	FX_REQUIRE(u->cin >= 2, FX_ERR_ARGUMENT, "the energy loss reads log k from input channel 1");
	const float* t = (const float*)tb->ptr;
	FX_CHECK(run_forward(u, (const float*)xb->ptr, B, H, W));
	hipStream_t st = u->be->stream;
	long hw = (long)H * W;
	FX_HIP(hipMemsetAsync(u->sums, 0, (size_t)B * 2 * 4, st));
	FX_HIP(hipMemsetAsync(u->dy, 0, (size_t)B * hw * 16 * 4, st));
	hipLaunchKernelGGL(k_eloss, dim3(64, B), dim3(256), 0, st, (const float*)xb->ptr, u->cin, u->yout, 16, t, u->sums,
	                   u->dy, H, W);
	hipLaunchKernelGGL(k_eloss_scale, dim3(nb(B * hw)), dim3(256), 0, st, u->dy, 16, u->sums, B, hw);
	FX_CHECK(check_launch("unet energy loss"));
	FX_CHECK(backward(u));
	std::vector<float> sums((size_t)B * 2);
	FX_HIP(hipMemcpyAsync(sums.data(), u->sums, sums.size() * 4, hipMemcpyDeviceToHost, st));
	FX_HIP(hipStreamSynchronize(st));
	double L = 0;
	for (int b = 0; b < B; ++b) L += sums[2 * b] / sums[2 * b + 1];
	if (loss) *loss = L;
	return FX_OK;
}

fx_status fx_unet_adam_step(fx_unet* u, float lr, float b1, float b2, float eps) {
	FX_REQUIRE(u, FX_ERR_ARGUMENT, "null");
	u->step += 1;
	float c1 = 1.f - powf(b1, (float)u->step), c2 = 1.f - powf(b2, (float)u->step);
	hipLaunchKernelGGL(k_adam, dim3(nb((long)u->nparams)), dim3(256), 0, u->be->stream, u->P, u->G, u->M, u->V,
	                   (long)u->nparams, lr, b1, b2, eps, c1, c2);
	FX_CHECK(check_launch("adam"));
	return pack_weights(u);
}

}  // extern "C"


namespace fx {
fx_status unet_infer(fx_unet* u, const float* x, int B, int H, int W, const int* skip, const float** y) {
	FX_REQUIRE(H % 8 == 0 && W % 8 == 0, FX_ERR_DIMENSION, "U-Net image sides must be multiples of 8, got %dx%d", H, W);
	u->skip = skip;
	fx_status s = run_forward(u, x, B, H, W);
	u->skip = nullptr;
	*y = u->yout;
	return s;
}
}  // namespace fx
