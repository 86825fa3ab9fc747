// fx_unet.hip — learned preconditioner network on gfx950 (no reference counterpart).
//
// A 4-level U-Net (3×3 convolutions, 16/32/48/64 channels) with BF16 weights
// and activations and FP32 accumulation on the matrix cores.  Layer order,
// parameter layout and loss are those of oracle/unet.cpp (the FP64 checker):
//   enc l=0..3: relu(conv3) ×2, skip, avgpool2 (l<3)
//   dec l=2..0: up2(nearest) ‖ skip_l, relu(conv3) ×2;   out: conv3 → 1 channel
// Activations are NHWC with the channel count padded to a multiple of 16.
//
// 3×3 convolution (k_conv3m) as an implicit GEMM on gfx950's
// v_mfma_f32_16x16x32_bf16: M = pixels, N = output channels, K = (tap, input
// channel) flattened, k = tap·cin_p + c.  A workgroup stages the input tile of
// its 4 output rows × 64 columns (plus the one-pixel halo, zero outside the
// image) in LDS once, with 16-byte loads; each wave then computes one row of
// 64 pixels × all output channels, so every A fragment feeds all N tiles and
// every B fragment (weights, read through the cache) all four M tiles.
// MFMA 16x16x32 operand layout (lane l, element j): A[m = l%16][k = 8(l/16)+j],
// B[k = 8(l/16)+j][n = l%16], D[m = 4(l/16)+j][n = l%16].
// The same kernel computes the data gradient with the transposed, flipped
// weights (its staging masks the FP32 gradient by the forward ReLU and rounds
// it to BF16).  Inference fuses the decoder's upsample + concatenation into
// the staging and writes only output channel 0.  Weight gradients reduce over
// pixels with the pixel index as K.
#include "fx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace fxd;

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

// FP32 → BF16, round to nearest even (gfx950's v_cvt_pk_bf16_f32).
__device__ __forceinline__ bf16 f2bf(float f) { return __builtin_bit_cast(bf16, (__bf16)f); }
__device__ __forceinline__ float bf2f(bf16 b) { return __uint_as_float((uint32_t)b << 16); }

constexpr int NLAYER = 15;
constexpr int CH[4] = {16, 32, 48, 64};

inline int pad16(int c) { return (c + 15) / 16 * 16; }

// Source of a convolution's staged input tile, and the form of its output.  SRC_F32IN is the
// network's raw FP32 input [B][H][W][ca], rounded to BF16 and zero-padded to cin_p as it is staged
// (the bits k_pack_in would store), so inference needs no packed copy of it; SRC_F32IN_K is the same
// input (two channels) read by k_conv3p, whose K packs the nine taps' channels densely.
enum { SRC_BF16 = 0, SRC_UPCAT = 1, SRC_F32 = 2, SRC_F32IN = 3, SRC_F32IN_K = 4 };
enum { OUT_BF16 = 0, OUT_F32 = 1, OUT_F32_CH0 = 2 };

struct ConvArgs {
	const void* in;        // SRC_BF16 bf16 [B][H][W][cin_p]; SRC_F32 f32 [B][H][W][cin_p];
	                       // SRC_UPCAT bf16 [B][H/2][W/2][ca] (the coarser level, upsampled by r/2, c/2)
	const bf16* in2;       // SRC_UPCAT: the skip, bf16 [B][H][W][cin_p − ca], channels ca.. of the tile
	const bf16* mask;      // SRC_F32: forward output whose ReLU masks the gradient (bf16 [B][H][W][cin_p]), or null
	const bf16* w;         // [9][cout_p][cin_p]
	const float* bias;     // [cout] or null
	void* out;             // OUT_BF16 bf16 / OUT_F32 f32 [B][H][W][cout_p]; OUT_F32_CH0 f32 [B][H][W]
	bf16* pool;            // OUT_BF16: also the 2×2 average of the output, bf16 [B][H/2][W/2][cout_p], or null
	int B, H, W, cin_p, cout_p, cout, ca;
	int relu, flip;
	const int* skip;       // kernels return early while *skip != 0 (may be null)
};

constexpr int CTW = 64, CTH = 4;   // output tile of a workgroup: CTH rows (one per wave) × CTW columns
// SRC_UPCAT stages its upsampled operand at the coarse resolution: the coarse pixels under the
// (rows + 2) × (CTW + 2) fine input pixels of a tile of `rows` rows, rows / 2 + 2 rows × CUW columns
// (rows and CTW even, tiles start on even rows and columns).
constexpr int CUW = CTW / 2 + 2, CUR = CTH / 2 + 2, CUPX = CUR * CUW;
constexpr int coarse_px(int rows) { return (rows / 2 + 2) * CUW; }
// LDS of a workgroup whose output tile has `rows` rows: the staged input tile, reused afterwards for
// the BF16 output transpose (rows × CTW pixels × cout_p).  ca > 0: the up+cat tile (coarse operand,
// ca channels per pixel, then the fine skip, cin_p − ca).  Pixels are stored unpadded with their
// 16-byte channel chunks swizzled (chunk_sw).
inline size_t conv_lds_bytes(int cin_p, int cout_p, int ca = 0, int rows = CTH) {
	const size_t tpx = (size_t)(rows + 2) * (CTW + 2);
	const size_t in = ca > 0 ? ((size_t)coarse_px(rows) * ca + tpx * (cin_p - ca)) * sizeof(bf16)
	                         : tpx * cin_p * sizeof(bf16);
	const size_t out = (size_t)rows * CTW * cout_p * sizeof(bf16);
	return in > out ? in : out;
}

// LDS swizzle of the staged tiles: 8-channel chunk c of staged pixel P, with n (even) chunks per
// pixel, sits at chunk c ^ chunk_sw(P, n) of the pixel.  An A fragment's 16-byte reads by 8
// consecutive pixels (one LDS pass) then fall on 8 different groups of four banks without padding the
// pixels: with d = gcd(n, 8), pixels P and P + 8/d share a bank group unswizzled, and the XOR with
// (P / (8/d)) mod d separates them (it permutes chunks within aligned blocks of d, so c ^ sw < n).
// The value is the same for P and P + 16 (P + 8 on the coarse tile), i.e. for all four M tiles.
__device__ __forceinline__ int chunk_sw(int P, int n) {
	const int d = (n & -n) < 8 ? (n & -n) : 8;
	const int sh = d == 8 ? 0 : (d == 4 ? 1 : 2);
	return (P >> sh) & (d - 1);
}

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

// This lane's bias values, channel 16·t + (lane % 16) for t < NT (0 past cout or without a bias).
template <int NT>
__device__ __forceinline__ void load_bias(const ConvArgs& a, float (&bv)[NT]) {
	const int m = threadIdx.x & 15;
#pragma unroll
	for (int t = 0; t < NT; ++t) bv[t] = a.bias && 16 * t + m < a.cout ? a.bias[16 * t + m] : 0.f;
}

// Bias, ReLU and the output form of a workgroup's (CTH·RPW) × CTW tile from the MFMA accumulators
// (wave wv holds rows row0 + RPW·wv + r, r < RPW; acc[4r + mt][t][j] is row r's pixel 16·mt + 4·g + j,
// channel 16·t + m of lane 16·g + m); bv from load_bias.  Every thread of the workgroup calls it (it
// holds workgroup barriers, and returns nowhere early); the input tile in LDS is reused for the output
// transpose, tile row lr at element lr·CTW·cout_p.
template <int OUTK, int NT, int RPW = 1>
__device__ __forceinline__ void conv_epilogue(const ConvArgs& a, f4 (&acc)[4 * RPW][NT], const float (&bv)[NT],
                                              bf16* tile, int b, int row0, int col0) {
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int m = lane & 15, g = lane >> 4;
	// Output value of an accumulator: + bias, ReLU, and 0 for channels past cout, as one clamp per
	// value, v_med3(acc + bias, lo, hi) with per-lane bounds: [0, ∞) under ReLU, (−∞, ∞) without, and
	// [0, 0] past cout.  The outputs are cout_p = 16·NT channels (CP) in every form.
	constexpr int CP = 16 * NT;
	float badd[NT], lo[NT], hi[NT];
#pragma unroll
	for (int t = 0; t < NT; ++t) {
		const bool keep = 16 * t + m < a.cout;
		badd[t] = keep && a.bias ? bv[t] : 0.f;
		lo[t] = keep ? (a.relu ? 0.f : -__builtin_inff()) : 0.f;
		hi[t] = keep ? __builtin_inff() : 0.f;
	}
	auto value = [&](int i, int t, int j) {
		return __builtin_amdgcn_fmed3f(acc[i][t][j] + badd[t], lo[t], hi[t]);
	};
	if constexpr (OUTK == OUT_F32) {
		// Training's data gradients: FP32 stores straight from the fragments (16 lanes write 64
		// contiguous bytes).
#pragma unroll
		for (int r = 0; r < RPW; ++r) {
			const int orow = row0 + RPW * wv + r;
#pragma unroll
			for (int mt = 0; mt < 4; ++mt)
#pragma unroll
				for (int t = 0; t < NT; ++t)
#pragma unroll
					for (int j = 0; j < 4; ++j) {
						const int col = col0 + 16 * mt + 4 * g + j;
						if (orow < a.H && col < a.W)
							((float*)a.out)[(((size_t)b * a.H + orow) * a.W + col) * CP + 16 * t + m] = value(4 * r + mt, t, j);
					}
		}
	} else {
		// A row's 64 pixels × cout_p outputs are contiguous in memory: transpose them through LDS (the
		// input tile is no longer read once every wave has passed the barrier) and store 16 bytes per
		// lane.
		__syncthreads();
		const int nvalid = min(CTW, a.W - col0);
		if constexpr (OUTK == OUT_BF16) {
			bf16* st = tile + (size_t)wv * RPW * CTW * CP;
#pragma unroll
			for (int r = 0; r < RPW; ++r)
#pragma unroll
				for (int mt = 0; mt < 4; ++mt)
#pragma unroll
					for (int t = 0; t < NT; ++t)
#pragma unroll
						for (int j = 0; j < 4; ++j)
							st[(r * CTW + 16 * mt + 4 * g + j) * CP + 16 * t + m] = f2bf(value(4 * r + mt, t, j));
			__syncthreads();
#pragma unroll
			for (int r = 0; r < RPW; ++r) {
				const int lr = RPW * wv + r, orow = row0 + lr;
				const bool rowok = orow < a.H;
				const bf16* s0 = st + (size_t)r * CTW * CP;
				bf16* dst = (bf16*)a.out + (((size_t)b * a.H + orow) * a.W + col0) * CP;
				const int nchunk = rowok ? nvalid * CP / 8 : 0;
				for (int q = lane; q < nchunk; q += 64) *(u4v*)(dst + 8 * q) = *(const u4v*)(s0 + 8 * q);
				if (a.pool && (lr & 1) == 0 && rowok) {
					// The encoder's 2×2 average pooling of tile rows lr and lr + 1 (both in LDS; the next
					// wave's first row when RPW = 1): y[r][c] = bf16(0.25·(x[2r][2c] + x[2r][2c+1] +
					// x[2r+1][2c] + x[2r+1][2c+1])), summed in FP32 left to right (the oracle's pool2 in
					// BF16 storage).
					const bf16* s1 = s0 + CTW * CP;
					const int nck = CP >> 3, npc = nvalid >> 1;
					bf16* pdst = a.pool + (((size_t)b * (a.H >> 1) + (orow >> 1)) * (a.W >> 1) + (col0 >> 1)) * CP;
					for (int q = lane; q < npc * nck; q += 64) {
						const int pc = q / nck, ch = (q - pc * nck) << 3;
						const int e0 = 2 * pc * CP + ch;
						const u4v a0 = *(const u4v*)(s0 + e0), a1 = *(const u4v*)(s0 + e0 + CP);
						const u4v a2 = *(const u4v*)(s1 + e0), a3 = *(const u4v*)(s1 + e0 + CP);
						u4v o;
#pragma unroll
						for (int e = 0; e < 4; ++e) {
							uint32_t r2 = 0;
#pragma unroll
							for (int half = 0; half < 2; ++half) {
								const int sh = 16 * half;
								const float v = 0.25f * (bf2f((bf16)(a0[e] >> sh)) + bf2f((bf16)(a1[e] >> sh)) +
								                         bf2f((bf16)(a2[e] >> sh)) + bf2f((bf16)(a3[e] >> sh)));
								r2 |= (uint32_t)f2bf(v) << sh;
							}
							o[e] = r2;
						}
						*(u4v*)(pdst + (size_t)pc * CP + ch) = o;
					}
				}
			}
		} else {
			float* st = (float*)tile + (size_t)wv * RPW * CTW;
			if (m == 0)
#pragma unroll
				for (int r = 0; r < RPW; ++r)
#pragma unroll
					for (int mt = 0; mt < 4; ++mt)
#pragma unroll
						for (int j = 0; j < 4; ++j) st[r * CTW + 16 * mt + 4 * g + j] = value(4 * r + mt, 0, j);
			__syncthreads();
			const int col = col0 + lane;
#pragma unroll
			for (int r = 0; r < RPW; ++r) {
				const int orow = row0 + RPW * wv + r;
				if (orow < a.H && col < a.W) ((float*)a.out)[((size_t)b * a.H + orow) * a.W + col] = st[r * CTW + lane];
			}
		}
	}
}

// The K loop of a CTH × CTW output tile over the input tile staged in LDS (k_conv3m's layout: cin_p
// elements per pixel, or for SRC_UPCAT the coarse operand then the skip, chunks swizzled by
// chunk_sw): acc = Σ_k A·B over k = (tap, channel), weights read through the cache one K step ahead.
template <int SRC, int NT>
__device__ __forceinline__ void conv_kloop(const ConvArgs& a, const bf16* tile, f4 (&acc)[4][NT]) {
	constexpr int TW = CTW + 2;
	const int pitch = a.cin_p;
	const int cpitch = a.ca, spitch = a.cin_p - a.ca, soff = CUPX * cpitch;
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int m = lane & 15, g = lane >> 4;
#pragma unroll
	for (int mt = 0; mt < 4; ++mt)
#pragma unroll
		for (int t = 0; t < NT; ++t) acc[mt][t] = f4{0.f, 0.f, 0.f, 0.f};
	// This lane's k = 8g + j (j = 0..7) within each 32-wide K step, as (tap, first channel c0).
	int tap = 0, c0 = 8 * g;
	while (c0 >= a.cin_p) { c0 -= a.cin_p; ++tap; }
	const int nsteps = (9 * a.cin_p + 31) >> 5;
	const u4v zero = {0u, 0u, 0u, 0u};
	// A fragment rows: tile pixel P = (wv + dr)·TW + 16·mt + m + dc, its chunk c0/8 swizzled; B fragment
	// rows: output channel 16·t + m of tap wt at wrow[t] + wt·cout_p·cin_p.  SRC_UPCAT: a lane whose 8
	// channels lie in the upsampled operand (c0 < ca) reads the coarse pixel under that fine pixel,
	// coarse tile row (wv + dr + 1)/2 and column 8·mt + (m + dc + 1)/2; otherwise the skip tile at
	// channel c0 − ca.
	const bf16* wrow[NT];
#pragma unroll
	for (int t = 0; t < NT; ++t) wrow[t] = a.w + (size_t)(16 * t + m) * a.cin_p;
	const size_t wtap = (size_t)a.cout_p * a.cin_p;
	// Weight fragments of the K step at (tap, c0); a step past 9·cin_p (only in the last step) reads
	// tap 0, in bounds, and its fragments are replaced by zeros where they are used.
	auto wload = [&](int tp, int c, u4v* bw) {
		const int t9 = tp < 9 ? tp : 0;
		const int wt = a.flip ? 8 - t9 : t9;
#pragma unroll
		for (int t = 0; t < NT; ++t) bw[t] = *(const u4v*)(wrow[t] + wt * wtap + c);
	};
	auto adv = [&](int& tp, int& c) {
		c += 32;
		while (c >= a.cin_p) { c -= a.cin_p; ++tp; }
	};
	// One K step; LAST (the final step, the only one whose k can pass 9·cin_p) masks the lanes past it.
	auto kstep = [&](auto last, const u4v* bc) {
		constexpr bool LAST = decltype(last)::value;
		const bool ok = !LAST || tap < 9;
		const int t9 = LAST ? (ok ? tap : 0) : tap;
		const int dr = (t9 * 11) >> 5, dc = t9 - 3 * dr;   // t9 / 3 and t9 % 3 for t9 < 9
		const bf16* ap;
		int mstride;
		if constexpr (SRC == SRC_UPCAT) {
			if (c0 < a.ca) {
				const int Q = ((wv + dr + 1) >> 1) * CUW + ((m + dc + 1) >> 1);
				ap = tile + Q * cpitch + (((c0 >> 3) ^ chunk_sw(Q, cpitch >> 3)) << 3);
				mstride = 8 * cpitch;
			} else {
				const int P = (wv + dr) * TW + m + dc;
				ap = tile + soff + P * spitch + ((((c0 - a.ca) >> 3) ^ chunk_sw(P, spitch >> 3)) << 3);
				mstride = 16 * spitch;
			}
		} else {
			const int P = (wv + dr) * TW + m + dc;
			ap = tile + P * pitch + (((c0 >> 3) ^ chunk_sw(P, pitch >> 3)) << 3);
			mstride = 16 * pitch;
		}
		u4v x[4];
#pragma unroll
		for (int mt = 0; mt < 4; ++mt) x[mt] = *(const u4v*)(ap + mt * mstride);
#pragma unroll
		for (int mt = 0; mt < 4; ++mt) {
			const bf8 av = __builtin_bit_cast(bf8, LAST && !ok ? zero : x[mt]);
#pragma unroll
			for (int t = 0; t < NT; ++t) {
				const u4v bv = LAST && !ok ? zero : bc[t];
				acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf8, bv), acc[mt][t], 0, 0, 0);
			}
		}
	};
	u4v bcur[NT];
	wload(tap, c0, bcur);
	for (int s1 = 1; s1 < nsteps; ++s1) {
		// The next step's weights are fetched while this step computes.
		int ntap = tap, nc0 = c0;
		adv(ntap, nc0);
		u4v bnext[NT];
		wload(ntap, nc0, bnext);
		kstep(std::false_type{}, bcur);
#pragma unroll
		for (int t = 0; t < NT; ++t) bcur[t] = bnext[t];
		tap = ntap;
		c0 = nc0;
	}
	kstep(std::true_type{}, bcur);
}

// One K segment of the compile-time K loop: the NCH 8-channel chunks of every tap of one staged
// operand, K step s covering chunks 4s .. 4s + 3 of the segment's (tap, chunk) order (lane g takes
// chunk 4s + g).  A step spans at most two taps: lanes g ≥ NCH − (4s mod NCH) wrap to the next one,
// so a lane's tap, chunk and LDS address are a compare and selects between constants.  COARSE reads
// the up+cat tile's coarse operand (pixel under the fine pixel, rows and columns halved), otherwise a
// fine tile at element BASE; CH0 is the segment's first channel in the weights.  Weights are fetched
// two steps ahead.  A wave computes RPW output rows (tile rows RPW·wv + r), each weight fragment
// feeding all of them.
template <int NT, int CINP, int NCH, bool COARSE, int BASE, int CH0, int RPW>
__device__ __forceinline__ void conv_kseg(const ConvArgs& a, const bf16* tile, f4 (&acc)[4 * RPW][NT]) {
	constexpr int TW = CTW + 2, PITCH = 8 * NCH, NSTEP = (9 * NCH + 3) / 4, PD = 2;
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int m = lane & 15, g = lane >> 4;
	const bf16* wl = a.w + (size_t)m * CINP + CH0;   // + (tap·cout_p + 16t)·CINP + 8·chunk
	const int wtap = a.cout_p * CINP;
	auto step = [&](int s, int& tap, int& c8) {
		const int q0 = 4 * s, t0 = q0 / NCH, r0 = q0 - t0 * NCH;
		const bool wrap = g >= NCH - r0;
		tap = t0 + (wrap ? 1 : 0);
		c8 = r0 + g - (wrap ? NCH : 0);
	};
	u4v wb[NSTEP][NT];
	auto wload = [&](int s) {
		int tap, c8;
		step(s, tap, c8);
		const int t9 = tap < 9 ? tap : 0;
		const int wt = a.flip ? 8 - t9 : t9;
#pragma unroll
		for (int t = 0; t < NT; ++t) wb[s][t] = *(const u4v*)(wl + (size_t)(wt * wtap + 16 * t * CINP + 8 * c8));
	};
#pragma unroll
	for (int s = 0; s < PD && s < NSTEP; ++s) wload(s);
	const u4v zero = {0u, 0u, 0u, 0u};
#pragma unroll
	for (int s = 0; s < NSTEP; ++s) {
		if (s + PD < NSTEP) wload(s + PD);
		int tap, c8;
		step(s, tap, c8);
		const bool ok = tap < 9;   // false only past the last tap, in the last step
		const int t9 = ok ? tap : 0;
		const int dr = t9 / 3, dc = t9 - 3 * dr;
		u4v x[4 * RPW];
#pragma unroll
		for (int r = 0; r < RPW; ++r) {
			const int lr = RPW * wv + r;
			int addr, mstride;
			if constexpr (COARSE) {
				const int Q = ((lr + dr + 1) >> 1) * CUW + ((m + dc + 1) >> 1);
				addr = BASE + Q * PITCH + ((c8 ^ chunk_sw(Q, NCH)) << 3);
				mstride = 8 * PITCH;
			} else {
				const int P = (lr + dr) * TW + m + dc;
				addr = BASE + P * PITCH + ((c8 ^ chunk_sw(P, NCH)) << 3);
				mstride = 16 * PITCH;
			}
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) x[4 * r + mt] = *(const u4v*)(tile + addr + mt * mstride);
		}
#pragma unroll
		for (int i = 0; i < 4 * RPW; ++i) {
			const bf8 av = __builtin_bit_cast(bf8, ok ? x[i] : zero);
#pragma unroll
			for (int t = 0; t < NT; ++t)
				acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf8, ok ? wb[s][t] : zero),
				                                                    acc[i][t], 0, 0, 0);
		}
	}
}

// conv_kloop with the channel counts known at compile time (CINP input channels, of which CA
// upsampled for SRC_UPCAT): the K loop fully unrolled over conv_kseg segments — for up+cat first every
// tap's coarse chunks, then every tap's skip chunks, so no step mixes the two tiles.  The tile has
// CTH·RPW rows.
template <int SRC, int NT, int CINP, int CA, int RPW>
__device__ __forceinline__ void conv_kloop_ct(const ConvArgs& a, const bf16* tile, f4 (&acc)[4 * RPW][NT]) {
#pragma unroll
	for (int i = 0; i < 4 * RPW; ++i)
#pragma unroll
		for (int t = 0; t < NT; ++t) acc[i][t] = f4{0.f, 0.f, 0.f, 0.f};
	if constexpr (SRC == SRC_UPCAT) {
		conv_kseg<NT, CINP, CA / 8, true, 0, 0, RPW>(a, tile, acc);
		conv_kseg<NT, CINP, (CINP - CA) / 8, false, coarse_px(CTH * RPW) * CA, CA, RPW>(a, tile, acc);
	} else {
		conv_kseg<NT, CINP, CINP / 8, false, 0, 0, RPW>(a, tile, acc);
	}
}

// One (CTH·RPW) × CTW output tile per workgroup, wave wv computing rows RPW·wv .. RPW·wv + RPW − 1
// (RPW > 1 only with the compile-time K loop, CINP > 0: each weight fragment then feeds RPW rows, and
// the staged halo is 2 rows in CTH·RPW + 2 instead of in CTH + 2).
template <int SRC, int OUTK, int NT, int CINP = 0, int CA = 0, int RPW = 1>
__global__ __launch_bounds__(BLOCK, RPW > 1 ? 2 : 1) void k_conv3m(ConvArgs a) {
	static_assert(RPW == 1 || CINP > 0, "several rows per wave need the compile-time K loop");
	if (a.skip && *a.skip) return;
	extern __shared__ __align__(16) bf16 tile[];
	constexpr int TR = CTH * RPW;
	const int b = blockIdx.z, row0 = blockIdx.y * TR, col0 = blockIdx.x * CTW;
	float bv[NT];
	load_bias<NT>(a, bv);
	// Pixels are stored unpadded, their chunks swizzled by chunk_sw.
	const int pitch = a.cin_p;
	constexpr int TW = CTW + 2, TPX = (TR + 2) * TW, CPX = coarse_px(TR);
	// Staging, chunk-major: a thread owns tile pixels threadIdx.x + k·BLOCK (k < PPT), whose image
	// position, bounds test and source pointers it forms once; the loop over 8-channel chunks is then
	// wave-uniform, and CG chunks × PPT pixels of 16-byte loads are in flight together.  Pixels
	// outside the image stage zeros.
	constexpr int PPT = (TPX + BLOCK - 1) / BLOCK, CG = 2;
	// SRC_UPCAT: the upsampled operand's CPX coarse pixels (TR / 2 + 2 rows × CUW) at element 0 with
	// pitch cpitch, the skip's fine pixels after them (element soff) with pitch spitch.  A fine pixel
	// outside the image is exactly a coarse pixel outside the coarse image (H and W are even), so both
	// stage zeros there.
	const int cpitch = a.ca, spitch = a.cin_p - a.ca, soff = CPX * cpitch;
	const int nch = CINP > 0 ? (SRC == SRC_UPCAT ? (CINP - CA) >> 3 : CINP >> 3)
	                         : (SRC == SRC_UPCAT ? (a.cin_p - a.ca) >> 3 : a.cin_p >> 3);
	const char* src0[PPT];   // the pixel's first byte in `in` (SRC_UPCAT: in the skip `in2`)
	const char* src1[PPT];   // SRC_F32: the pixel's mask row
	bool pin[PPT], pst[PPT];
	int lofs[PPT], swz[PPT];
#pragma unroll
	for (int k = 0; k < PPT; ++k) {
		const int p = threadIdx.x + k * BLOCK;
		const int pr = p / TW, pc = p - pr * TW;
		const int rr = row0 - 1 + pr, cc = col0 - 1 + pc;
		pst[k] = p < TPX;
		pin[k] = pst[k] && rr >= 0 && rr < a.H && cc >= 0 && cc < a.W;
		const size_t pix = pin[k] ? ((size_t)b * a.H + rr) * a.W + cc : 0;
		lofs[k] = SRC == SRC_UPCAT ? soff + p * spitch : p * pitch;
		swz[k] = chunk_sw(p, nch);
		if constexpr (SRC == SRC_BF16) {
			src0[k] = (const char*)((const bf16*)a.in + pix * a.cin_p);
			src1[k] = nullptr;
		} else if constexpr (SRC == SRC_UPCAT) {
			src0[k] = (const char*)(a.in2 + pix * (a.cin_p - a.ca));
			src1[k] = nullptr;
		} else if constexpr (SRC == SRC_F32IN) {
			src0[k] = (const char*)((const float*)a.in + pix * a.ca);
			src1[k] = nullptr;
		} else {
			src0[k] = (const char*)((const float*)a.in + pix * a.cin_p);
			src1[k] = a.mask ? (const char*)(a.mask + pix * a.cin_p) : nullptr;
		}
	}
	if constexpr (SRC == SRC_UPCAT) {
		// The coarse operand: thread t stages coarse tile pixels t, t + BLOCK, …, four 16-byte loads in
		// flight.
		for (int t = threadIdx.x; t < CPX; t += BLOCK) {
			const int pr = t / CUW, pc = t - pr * CUW;
			const int Hc = a.H >> 1, Wc = a.W >> 1;
			const int rr = (row0 >> 1) - 1 + pr, cc = (col0 >> 1) - 1 + pc;
			const bool in = rr >= 0 && rr < Hc && cc >= 0 && cc < Wc;
			const bf16* src = (const bf16*)a.in + (in ? (((size_t)b * Hc + rr) * Wc + cc) * a.ca : 0);
			bf16* dst = tile + t * cpitch;
			const int nc = (CINP > 0 ? CA : a.ca) >> 3, sw = chunk_sw(t, nc);
			for (int c8 = 0; c8 < nc; c8 += 4) {
				u4v v[4];
#pragma unroll
				for (int q = 0; q < 4; ++q) {
					v[q] = u4v{0u, 0u, 0u, 0u};
					if (in && c8 + q < nc) v[q] = *(const u4v*)(src + ((c8 + q) << 3));
				}
#pragma unroll
				for (int q = 0; q < 4; ++q)
					if (c8 + q < nc) *(u4v*)(dst + (((c8 + q) ^ sw) << 3)) = v[q];
			}
		}
	}
	auto fetch = [&](int k, int ch) -> u4v {
		u4v v = {0u, 0u, 0u, 0u};
		if (!pin[k]) return v;
		if constexpr (SRC == SRC_BF16 || SRC == SRC_UPCAT) {
			v = *(const u4v*)((const bf16*)src0[k] + ch);
		} else if constexpr (SRC == SRC_F32IN) {
			const float* src = (const float*)src0[k];
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				const int c = ch + 2 * e;
				const bf16 lo = c < a.ca ? f2bf(src[c]) : (bf16)0;
				const bf16 hi = c + 1 < a.ca ? f2bf(src[c + 1]) : (bf16)0;
				v[e] = (uint32_t)lo | ((uint32_t)hi << 16);
			}
		} else {
			const float* src = (const float*)src0[k] + ch;
			const float4 lo = *(const float4*)src, hi = *(const float4*)(src + 4);
			float f[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
			if (src1[k]) {
				const u4v mk = *(const u4v*)((const bf16*)src1[k] + ch);
#pragma unroll
				for (int e = 0; e < 8; ++e) {
					const bf16 mb = (bf16)(mk[e >> 1] >> ((e & 1) * 16));
					if (!(bf2f(mb) > 0.f)) f[e] = 0.f;
				}
			}
#pragma unroll
			for (int e = 0; e < 4; ++e) v[e] = (uint32_t)f2bf(f[2 * e]) | ((uint32_t)f2bf(f[2 * e + 1]) << 16);
		}
		return v;
	};
	for (int c8 = 0; c8 < nch; c8 += CG) {
		u4v v[CG][PPT];
#pragma unroll
		for (int q = 0; q < CG; ++q)
#pragma unroll
			for (int k = 0; k < PPT; ++k) v[q][k] = fetch(k, (c8 + q < nch ? c8 + q : c8) << 3);
#pragma unroll
		for (int q = 0; q < CG; ++q)
#pragma unroll
			for (int k = 0; k < PPT; ++k)
				if (c8 + q < nch && pst[k]) *(u4v*)(tile + lofs[k] + (((c8 + q) ^ swz[k]) << 3)) = v[q][k];
	}
	__syncthreads();

	f4 acc[4 * RPW][NT];
	if constexpr (CINP > 0) conv_kloop_ct<SRC, NT, CINP, CA, RPW>(a, tile, acc);
	else conv_kloop<SRC, NT>(a, tile, acc);
	conv_epilogue<OUTK, NT, RPW>(a, acc, bv, tile, b, row0, col0);
}

// Persistent convolutions with 16 output channels at full resolution, for two inputs: a 16-channel
// BF16 map (SRC_BF16) and the network's raw 2-channel FP32 input (SRC_F32IN_K, rounded to BF16 as it
// is staged, the bits k_pack_in stores), whose K packs the nine taps' two channels densely as
// k = 2·tap + c: one 32-wide K step instead of five over channels padded to 16.  A workgroup loops over
// output tiles t = blockIdx.x, blockIdx.x + gridDim.x, … with its weights and bias in registers, and
// loads its next tile's input into registers while it computes and stores the current one, so the
// staging latency of one tile overlaps the work of the previous one (the non-persistent k_conv3m runs
// each workgroup's load, compute and store one after another).  Output tile, LDS layout, MFMA
// fragments and epilogue are k_conv3m's.
template <int SRC, int OUTK>
__global__ __launch_bounds__(BLOCK, 4) void k_conv3p(ConvArgs a) {
	if (a.skip && *a.skip) return;
	extern __shared__ __align__(16) bf16 tile[];
	constexpr bool PK = SRC == SRC_F32IN_K;
	static_assert(PK || SRC == SRC_BF16, "16-channel BF16 or packed 2-channel FP32 input");
	constexpr int TW = CTW + 2, TPX = (CTH + 2) * TW, PPT = (TPX + BLOCK - 1) / BLOCK;
	constexpr int PITCH = PK ? 2 : 16 + 8;   // LDS elements per staged pixel
	constexpr int NCH = PK ? 1 : 2;          // loads per staged pixel (PK: one 8-byte pair of floats)
	constexpr int NSTEP = PK ? 1 : 5;        // 32-wide K steps (9·16 = 144 → 5, the last half zero)
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int m = lane & 15, g = lane >> 4;
	const int ntx = (a.W + CTW - 1) / CTW, nty = (a.H + CTH - 1) / CTH, ntiles = ntx * nty * a.B;
	float bv[1];
	load_bias<1>(a, bv);
	// B fragments of every K step (weights w[tap][16][16], output channel m), and this lane's A
	// offsets from tile pixel (wv, m): SRC_BF16 step s covers taps 2s and 2s + 1, the lane's 8
	// channels starting at c0 = 8(g % 2) of tap 2s + g/2; PK: the lane's k = 8g..8g+7 are taps
	// 4g..4g+3, channels 0 and 1 of each (one dword of the staged pixel).
	u4v wf[NSTEP];
	int aoff[PK ? 4 : NSTEP];
	bool aok[PK ? 4 : NSTEP];
	if constexpr (PK) {
		wf[0] = u4v{0u, 0u, 0u, 0u};
#pragma unroll
		for (int i = 0; i < 4; ++i) {
			const int tap = 4 * g + i;
			aok[i] = tap < 9;
			const int t9 = aok[i] ? tap : 0, dr = t9 / 3, dc = t9 - 3 * dr;
			aoff[i] = (dr * TW + dc) * PITCH;
			const bf16 w0 = aok[i] ? a.w[(size_t)(t9 * 16 + m) * a.cin_p] : (bf16)0;
			const bf16 w1 = aok[i] ? a.w[(size_t)(t9 * 16 + m) * a.cin_p + 1] : (bf16)0;
			wf[0][i] = (uint32_t)w0 | ((uint32_t)w1 << 16);
		}
	} else {
#pragma unroll
		for (int s = 0; s < NSTEP; ++s) {
			const int tap = 2 * s + (g >> 1), c0 = 8 * (g & 1);
			aok[s] = tap < 9;
			const int t9 = aok[s] ? tap : 0, dr = t9 / 3, dc = t9 - 3 * dr;
			aoff[s] = (dr * TW + dc) * PITCH + c0;
			wf[s] = aok[s] ? *(const u4v*)(a.w + (size_t)(t9 * 16 + m) * 16 + c0) : u4v{0u, 0u, 0u, 0u};
		}
	}
	const int lbase = (wv * TW + m) * PITCH;
	// Staging registers of the next tile: pixel threadIdx.x + k·BLOCK of its (CTH + 2) × (CTW + 2)
	// input window (PK: the raw FP32 pair in sv[k][0][0..1]).  The loads are unconditional, from pixel
	// 0 of the image where the window pixel lies outside it (pin false: zeros are staged instead), so
	// nothing waits for them before the next tile's staging stores.
	u4v sv[PPT][NCH];
	bool pin[PPT];
	auto issue = [&](int t) {
		const int b = t / (ntx * nty), rem = t - b * ntx * nty, ty = rem / ntx, tx = rem - ty * ntx;
		const int row0 = ty * CTH, col0 = tx * CTW;
#pragma unroll
		for (int k = 0; k < PPT; ++k) {
			const int p = threadIdx.x + k * BLOCK;
			const int pr = p / TW, pc = p - pr * TW;
			const int rr = row0 - 1 + pr, cc = col0 - 1 + pc;
			pin[k] = p < TPX && rr >= 0 && rr < a.H && cc >= 0 && cc < a.W;
			const size_t pix = pin[k] ? ((size_t)b * a.H + rr) * a.W + cc : 0;
			if constexpr (PK) {
				const float2 f = *(const float2*)((const float*)a.in + pix * 2);
				sv[k][0] = u4v{__float_as_uint(f.x), __float_as_uint(f.y), 0u, 0u};
			} else {
#pragma unroll
				for (int q = 0; q < NCH; ++q) sv[k][q] = *(const u4v*)((const bf16*)a.in + pix * 16 + 8 * q);
			}
		}
	};
	int t = blockIdx.x;
	if (t < ntiles) issue(t);
	for (; t < ntiles; t += gridDim.x) {
		const int b = t / (ntx * nty), rem = t - b * ntx * nty, ty = rem / ntx, tx = rem - ty * ntx;
		const int row0 = ty * CTH, col0 = tx * CTW;
#pragma unroll
		for (int k = 0; k < PPT; ++k) {
			const int p = threadIdx.x + k * BLOCK;
			if (p < TPX) {
				if constexpr (PK) {
					const uint32_t v = (uint32_t)f2bf(__uint_as_float(sv[k][0][0])) |
					                   ((uint32_t)f2bf(__uint_as_float(sv[k][0][1])) << 16);
					*(uint32_t*)(tile + p * PITCH) = pin[k] ? v : 0u;
				} else {
#pragma unroll
					for (int q = 0; q < NCH; ++q)
						*(u4v*)(tile + p * PITCH + 8 * q) = pin[k] ? sv[k][q] : u4v{0u, 0u, 0u, 0u};
				}
			}
		}
		__syncthreads();
		if (t + (int)gridDim.x < ntiles) issue(t + gridDim.x);
		f4 acc[4][1];
#pragma unroll
		for (int mt = 0; mt < 4; ++mt) acc[mt][0] = f4{0.f, 0.f, 0.f, 0.f};
		if constexpr (PK) {
#pragma unroll
			for (int mt = 0; mt < 4; ++mt) {
				u4v x;
#pragma unroll
				for (int i = 0; i < 4; ++i) x[i] = aok[i] ? *(const uint32_t*)(tile + lbase + 16 * mt * PITCH + aoff[i]) : 0u;
				acc[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, x),
				                                                      __builtin_bit_cast(bf8, wf[0]), acc[mt][0], 0, 0, 0);
			}
		} else {
#pragma unroll
			for (int s = 0; s < NSTEP; ++s) {
				u4v x[4];
#pragma unroll
				for (int mt = 0; mt < 4; ++mt)
					x[mt] = aok[s] ? *(const u4v*)(tile + lbase + 16 * mt * PITCH + aoff[s]) : u4v{0u, 0u, 0u, 0u};
#pragma unroll
				for (int mt = 0; mt < 4; ++mt)
					acc[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, x[mt]),
					                                                      __builtin_bit_cast(bf8, wf[s]), acc[mt][0], 0, 0, 0);
			}
		}
		conv_epilogue<OUTK, 1>(a, acc, bv, tile, b, row0, col0);
		__syncthreads();   // the epilogue's LDS reads end before the next tile is staged
	}
}

// Grid of a persistent convolution: as many workgroups as fit on the device at once (at most one per
// tile); `resident` caches the count for the kernel.
inline fx_status persistent_grid(const void* kern, size_t lds, const ConvArgs& a, int* resident, unsigned* grid) {
	if (*resident == 0) {
		int dev = 0, ncu = 0, per_cu = 0;
		FX_HIP(hipGetDevice(&dev));
		FX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
		FX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BLOCK, lds));
		*resident = std::max(1, ncu * std::max(1, per_cu));
	}
	const long ntiles = (long)((a.W + CTW - 1) / CTW) * ((a.H + CTH - 1) / CTH) * a.B;
	FX_REQUIRE(ntiles < (1L << 31), FX_ERR_DIMENSION, "too many convolution tiles");
	*grid = (unsigned)std::min<long>(ntiles, *resident);
	return FX_OK;
}

inline size_t conv3p_lds_bytes(int src) {
	const size_t in = (size_t)(CTH + 2) * (CTW + 2) * (src == SRC_F32IN_K ? 2 : 24) * sizeof(bf16);
	const size_t out = (size_t)CTH * CTW * 16 * sizeof(bf16);
	return in > out ? in : out;
}

template <int SRC, int OUTK>
fx_status launch_conv3p(const ConvArgs& a, hipStream_t st) {
	static int resident = 0;
	const void* kern = (const void*)k_conv3p<SRC, OUTK>;
	const size_t lds = conv3p_lds_bytes(SRC);
	unsigned grid = 0;
	FX_CHECK(persistent_grid(kern, lds, a, &resident, &grid));
	hipLaunchKernelGGL((k_conv3p<SRC, OUTK>), dim3(grid), dim3(BLOCK), lds, st, a);
	return fx::check_launch("unet persistent conv");
}

template <int SRC, int OUTK, int NT>
fx_status launch_conv_nt(const ConvArgs& a, hipStream_t st) {
	static bool attr = false;   // raise the dynamic LDS limit once per instantiation
	if (!attr) {
		FX_HIP(hipFuncSetAttribute((const void*)k_conv3m<SRC, OUTK, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                           (int)conv_lds_bytes(128, 128)));
		attr = true;
	}
	dim3 grid((unsigned)((a.W + CTW - 1) / CTW), (unsigned)((a.H + CTH - 1) / CTH), (unsigned)a.B);
	hipLaunchKernelGGL((k_conv3m<SRC, OUTK, NT>), grid, dim3(BLOCK),
	                   conv_lds_bytes(a.cin_p, a.cout_p, SRC == SRC_UPCAT ? a.ca : 0), st, a);
	return fx::check_launch("unet conv");
}
// The forward convolutions of the network with their channel counts fixed at compile time
// (conv_kloop_ct); launch_ct reports whether this shape has one.
template <int SRC, int NT, int CINP, int CA, int RPW = 1>
fx_status launch_ct1(const ConvArgs& a, hipStream_t st) {
	static bool attr = false;
	const void* kern = (const void*)k_conv3m<SRC, OUT_BF16, NT, CINP, CA, RPW>;
	constexpr int TR = CTH * RPW;
	const size_t lds = conv_lds_bytes(CINP, 16 * NT, SRC == SRC_UPCAT ? CA : 0, TR);
	if (!attr) {
		FX_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
		attr = true;
	}
	dim3 grid((unsigned)((a.W + CTW - 1) / CTW), (unsigned)((a.H + TR - 1) / TR), (unsigned)a.B);
	hipLaunchKernelGGL((k_conv3m<SRC, OUT_BF16, NT, CINP, CA, RPW>), grid, dim3(BLOCK), lds, st, a);
	return fx::check_launch("unet conv");
}
// Workgroups of a launch with tiles of `rows` rows.
inline long conv_tiles(const ConvArgs& a, int rows) {
	return (long)((a.W + CTW - 1) / CTW) * ((a.H + rows - 1) / rows) * a.B;
}
fx_status launch_ct(int src, int outk, const ConvArgs& a, hipStream_t st, bool* done) {
	*done = true;
	// Two rows per wave where the launch keeps at least 8 such tiles per CU (four rows per wave on the
	// 16-channel up+cat layer measured the same as two, 386–407 against 393 µs at 4096²).
	const bool tall2 = conv_tiles(a, 2 * CTH) >= 2048;
	if (outk == OUT_BF16 && src == SRC_BF16) {
		const int nt = a.cout_p / 16, ci = a.cin_p;
		if (nt == 2 && ci == 16) return tall2 ? launch_ct1<SRC_BF16, 2, 16, 0, 2>(a, st) : launch_ct1<SRC_BF16, 2, 16, 0>(a, st);
		if (nt == 2 && ci == 32) return tall2 ? launch_ct1<SRC_BF16, 2, 32, 0, 2>(a, st) : launch_ct1<SRC_BF16, 2, 32, 0>(a, st);
		// (The 48-channel layers were slower with two rows per wave: 62 and 75 against 60 and 71 µs at
		// 1024².)
		if (nt == 3 && ci == 32) return launch_ct1<SRC_BF16, 3, 32, 0>(a, st);
		if (nt == 3 && ci == 48) return launch_ct1<SRC_BF16, 3, 48, 0>(a, st);
		if (nt == 4 && ci == 48) return launch_ct1<SRC_BF16, 4, 48, 0>(a, st);
		if (nt == 4 && ci == 64) return launch_ct1<SRC_BF16, 4, 64, 0>(a, st);
	}
	if (outk == OUT_BF16 && src == SRC_UPCAT) {
		const int nt = a.cout_p / 16, ci = a.cin_p, ca = a.ca;
		if (nt == 1 && ci == 48 && ca == 32)
			return tall2 ? launch_ct1<SRC_UPCAT, 1, 48, 32, 2>(a, st) : launch_ct1<SRC_UPCAT, 1, 48, 32>(a, st);
		if (nt == 2 && ci == 80 && ca == 48)
			return tall2 ? launch_ct1<SRC_UPCAT, 2, 80, 48, 2>(a, st) : launch_ct1<SRC_UPCAT, 2, 80, 48>(a, st);
		if (nt == 3 && ci == 112 && ca == 64) return launch_ct1<SRC_UPCAT, 3, 112, 64>(a, st);
	}
	*done = false;
	return FX_OK;
}

template <int SRC, int OUTK>
fx_status launch_conv(const ConvArgs& a, hipStream_t st) {
	switch (a.cout_p / 16) {
	case 1: return launch_conv_nt<SRC, OUTK, 1>(a, st);
	case 2: return launch_conv_nt<SRC, OUTK, 2>(a, st);
	case 3: return launch_conv_nt<SRC, OUTK, 3>(a, st);
	case 4: return launch_conv_nt<SRC, OUTK, 4>(a, st);
	case 5: return launch_conv_nt<SRC, OUTK, 5>(a, st);
	case 6: return launch_conv_nt<SRC, OUTK, 6>(a, st);
	case 7: return launch_conv_nt<SRC, OUTK, 7>(a, st);
	case 8: return launch_conv_nt<SRC, OUTK, 8>(a, st);
	default: fx::set_error("conv with %d output channels is not supported", a.cout_p); return FX_ERR_ARGUMENT;
	}
}

// dW[tap][co][ci] += Σ_p g[p][co]·x[p+off][ci], g = dy masked by the ReLU of y.
// One wave: a 16(co) × 16(ci) tile of all nine taps over `rpw` image rows (rows counted over the batch),
// 32 pixels of a row per v_mfma_f32_16x16x32_bf16 (K = pixels): lane (m, q) feeds pixels 8q..8q+7 of
// the chunk, g rounded to BF16 as the A row co = m (loaded once for the nine taps) and x as the B
// column ci = m.  For each of the three input rows the lane loads the 10 pixels its three column
// shifts read, so a chunk costs 8 + 8 + 30 loads for 9 MFMAs.
__global__ __launch_bounds__(BLOCK) void k_conv3_wgrad(const float* dy, const bf16* y, int relu, const bf16* x,
                                                       float* dW, int B, int H, int W, int cin_p, int cout_p,
                                                       int cin, int cout, int rpw) {
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int tile = blockIdx.z;
	const int nco = cout_p / 16;
	const int cot = tile % nco, cit = tile / nco;
	const int m = lane & 15, q = lane >> 4;
	const int co = cot * 16 + m, ci = cit * 16 + m;
	const long rows = (long)B * H;
	const long rw0 = ((long)blockIdx.x * NWAVES + wv) * rpw;
	f4 acc[9];
#pragma unroll
	for (int t = 0; t < 9; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
	for (long rw = rw0; rw < rw0 + rpw && rw < rows; ++rw) {
		const long b = rw / H;
		const int r = (int)(rw - b * H);
		const size_t prow = (size_t)rw * W;   // first pixel of this output row
		for (int cb = 0; cb < W; cb += 32) {
			const int c0 = cb + 8 * q;
			u4v av;
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				uint32_t pk = 0;
#pragma unroll
				for (int h = 0; h < 2; ++h) {
					const int c = c0 + 2 * e + h;
					float gv = 0.f;
					if (c < W) {
						gv = dy[(prow + c) * cout_p + co];
						if (relu && !(bf2f(y[(prow + c) * cout_p + co]) > 0.f)) gv = 0.f;
					}
					pk |= (uint32_t)f2bf(gv) << (16 * h);
				}
				av[e] = pk;
			}
#pragma unroll
			for (int dr = 0; dr < 3; ++dr) {
				const int rs = r + dr - 1;
				const bool row_ok = rs >= 0 && rs < H;
				const size_t srow = ((size_t)b * H + (row_ok ? rs : 0)) * W;
				uint16_t xv[10];   // pixels c0 − 1 .. c0 + 8 of input row rs, channel ci
#pragma unroll
				for (int j = 0; j < 10; ++j) {
					const int cs = c0 - 1 + j;
					xv[j] = row_ok && cs >= 0 && cs < W ? x[(srow + cs) * cin_p + ci] : (uint16_t)0;
				}
#pragma unroll
				for (int dc = 0; dc < 3; ++dc) {
					u4v bv;
#pragma unroll
					for (int e = 0; e < 4; ++e) {
						// output pixels c0 + 2e, c0 + 2e + 1 read input columns + dc − 1; past the row's end
						// the g values are zero, whatever x holds
						bv[e] = (uint32_t)xv[2 * e + dc] | ((uint32_t)xv[2 * e + 1 + dc] << 16);
					}
					acc[3 * dr + dc] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
					    __builtin_bit_cast(bf8, av), __builtin_bit_cast(bf8, bv), acc[3 * dr + dc], 0, 0, 0);
				}
			}
		}
	}
#pragma unroll
	for (int t = 0; t < 9; ++t)
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const int o = cot * 16 + 4 * q + j;
			if (o < cout && ci < cin) atomicAdd(&dW[((size_t)t * cout + o) * cin + ci], acc[t][j]);
		}
}

// db[co] += Σ_p g[p][co]: a block's threads cover `slots` pixels × cout_p channels of consecutive memory
// per step (coalesced), sum their pixels, and merge the slots through LDS into one atomic per channel.
__global__ __launch_bounds__(BLOCK) void k_bgrad(const float* dy, const bf16* y, int relu, float* db, long total,
                                                 int cout_p, int cout) {
	const int slots = BLOCK / cout_p;
	const int c = threadIdx.x % cout_p, slot = threadIdx.x / cout_p;
	float s = 0.f;
	if (slot < slots)
		for (long p = (long)blockIdx.x * slots + slot; p < total; p += (long)gridDim.x * slots) {
			float g = dy[p * cout_p + c];
			if (relu && !(bf2f(y[p * cout_p + c]) > 0.f)) g = 0.f;
			s += g;
		}
	__shared__ float sh[BLOCK];
	sh[threadIdx.x] = slot < slots ? s : 0.f;
	__syncthreads();
	if (threadIdx.x < cout_p && threadIdx.x < cout) {
		float t = 0.f;
		for (int k = 0; k < slots; ++k) t += sh[k * cout_p + threadIdx.x];
		atomicAdd(&db[threadIdx.x], t);
	}
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
// NHWC f32 input [B][H][W][cin] -> BF16 [B][H][W][cin_p], one pixel per thread (16-byte stores).
__global__ void k_pack_in(const float* x, bf16* out, long npix, int cin, int cin_p, const int* skip) {
	if (skip && *skip) return;
	for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (long)gridDim.x * blockDim.x) {
		for (int c8 = 0; c8 < cin_p; c8 += 8) {
			u4v v;
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				const int c = c8 + 2 * e;
				const bf16 lo = c < cin ? f2bf(x[p * cin + c]) : (bf16)0;
				const bf16 hi = c + 1 < cin ? f2bf(x[p * cin + c + 1]) : (bf16)0;
				v[e] = (uint32_t)lo | ((uint32_t)hi << 16);
			}
			*(u4v*)(out + p * cin_p + c8) = v;
		}
	}
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
	float* yout = nullptr;   // layer-14 output while training, FP32 [B][H][W][16]
	float* ych0 = nullptr;   // channel 0 of the layer-14 output in inference, FP32 [B][H][W]
	bool train_ws = false;   // the backward's buffers are allocated for (B, H, W)
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
	u->train_ws = false;
}

int level_of(int li) { return li < 8 ? li / 2 : (li < 14 ? 2 - (li - 8) / 2 : 0); }

// Buffers for batch B of H×W images.
// Buffers for batch B of H×W images: the forward's always, the backward's (concatenations, FP32
// gradients, the output layer's 16 channels) only once a training call asks for them.
fx_status ensure_ws(fx_unet* u, int B, int H, int W, bool train) {
	if (!(u->B == B && u->H == H && u->W == W)) {
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
			FX_CHECK(alloc(u, npix(l + 1) * CH[l] * 2, &p)); u->pool[l] = (bf16*)p;
		}
		FX_CHECK(alloc(u, npix(0) * 4, &p)); u->ych0 = (float*)p;
		u->train_ws = false;
	}
	if (train && !u->train_ws) {
		auto npix = [&](int l) { return (size_t)B * (H >> l) * (W >> l); };
		void* p;
		for (int l = 0; l < 3; ++l) {
			FX_CHECK(alloc(u, npix(l) * pad16(CH[l + 1] + CH[l]) * 2, &p)); u->cat[l] = (bf16*)p;
			FX_CHECK(alloc(u, npix(l) * CH[l] * 4, &p)); u->gskip[l] = (float*)p;
		}
		FX_CHECK(alloc(u, npix(0) * 16 * 4, &p)); u->yout = (float*)p;
		size_t gmax = npix(0) * 128;
		FX_CHECK(alloc(u, gmax * 4, &p)); u->g1 = (float*)p;
		FX_CHECK(alloc(u, gmax * 4, &p)); u->g2 = (float*)p;
		FX_CHECK(alloc(u, (size_t)B * 2 * 4, &p)); u->sums = (float*)p;
		FX_CHECK(alloc(u, npix(0) * 16 * 4, &p)); u->dy = (float*)p;
		u->train_ws = true;
	}
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

// One convolution of layer li at level l: the forward (weights Wp, bias, ReLU on all but the output
// layer) or, with data_grad, the data gradient (the transposed, flipped weights Wt).
fx_status conv(fx_unet* u, int li, int src, const void* in, const bf16* in2, int ca, const bf16* mask, void* out,
               int outk, int l, bool data_grad, bf16* pool = nullptr) {
	ConvArgs a;
	a.in = in;
	a.pool = pool;
	a.in2 = in2;
	a.ca = ca;
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
	hipStream_t st = u->be->stream;
	// 16 → 16 channel forward convolutions (and the packed 2-channel input) run persistent.
	if (src == SRC_F32IN_K && outk == OUT_BF16) {
		FX_REQUIRE(a.ca == 2 && a.cin_p == 16 && a.cout_p == 16, FX_ERR_ARGUMENT, "packed-K first conv: 2 -> 16 channels");
		return launch_conv3p<SRC_F32IN_K, OUT_BF16>(a, st);
	}
	const bool p16 = src == SRC_BF16 && a.cin_p == 16 && a.cout_p == 16 && !a.flip;
	if (p16 && outk == OUT_BF16) return launch_conv3p<SRC_BF16, OUT_BF16>(a, st);
	if (p16 && outk == OUT_F32_CH0) return launch_conv3p<SRC_BF16, OUT_F32_CH0>(a, st);
	bool done = false;
	const fx_status cs = launch_ct(src, outk, a, st, &done);
	if (done) return cs;
	if (src == SRC_BF16 && outk == OUT_BF16) return launch_conv<SRC_BF16, OUT_BF16>(a, st);
	if (src == SRC_F32IN && outk == OUT_BF16) return launch_conv<SRC_F32IN, OUT_BF16>(a, st);
	if (src == SRC_UPCAT && outk == OUT_BF16) return launch_conv<SRC_UPCAT, OUT_BF16>(a, st);
	if (src == SRC_BF16 && outk == OUT_F32) return launch_conv<SRC_BF16, OUT_F32>(a, st);
	if (src == SRC_BF16 && outk == OUT_F32_CH0) return launch_conv<SRC_BF16, OUT_F32_CH0>(a, st);
	if (src == SRC_F32 && outk == OUT_F32) return launch_conv<SRC_F32, OUT_F32>(a, st);
	fx::set_error("unsupported conv form (source %d, output %d)", src, outk);
	return FX_ERR_UNSUPPORTED;
}

// Forward over the packed input u->xin.  Training (infer false) keeps what the backward needs: the
// decoder's concatenations materialised in u->cat and the output layer's 16 FP32 channels in u->yout.
// Inference upsamples and concatenates inside the decoder convs' staging and writes output channel 0
// only, to y ([B][H][W] FP32).
fx_status forward(fx_unet* u, bool infer, float* y, const float* xraw = nullptr) {
	hipStream_t st = u->be->stream;
	const bf16* h = u->xin;
	int li = 0;
	for (int l = 0; l < 4; ++l) {
		for (int k = 0; k < 2; ++k, ++li) {
			// The level's second convolution also writes the pooled map for the next level (l < 3).
			bf16* pool = k == 1 && l < 3 ? u->pool[l] : nullptr;
			if (li == 0 && xraw)   // the raw FP32 input, rounded (and padded) while staged
				FX_CHECK(conv(u, li, u->cin == 2 ? SRC_F32IN_K : SRC_F32IN, xraw, nullptr, u->cin, nullptr,
				              u->act[li], OUT_BF16, l, false));
			else
				FX_CHECK(conv(u, li, SRC_BF16, h, nullptr, 0, nullptr, u->act[li], OUT_BF16, l, false, pool));
			h = u->act[li];
		}
		if (l < 3) h = u->pool[l];
	}
	for (int l = 2; l >= 0; --l) {
		int ca = CH[l + 1], cs = CH[l], cop = pad16(ca + cs);
		const bf16* skipv = u->act[2 * l + 1];
		if (infer) {
			FX_CHECK(conv(u, li, SRC_UPCAT, h, skipv, ca, nullptr, u->act[li], OUT_BF16, l, false));
		} else {
			long n = (long)u->B * (u->H >> l) * (u->W >> l) * cop;
			hipLaunchKernelGGL(k_upcat, dim3(nb(n)), dim3(256), 0, st, h, ca, skipv, cs, u->cat[l], cop, u->B, u->H >> l,
			                   u->W >> l, u->skip);
			FX_CHECK(conv(u, li, SRC_BF16, u->cat[l], nullptr, 0, nullptr, u->act[li], OUT_BF16, l, false));
		}
		h = u->act[li];
		++li;
		FX_CHECK(conv(u, li, SRC_BF16, h, nullptr, 0, nullptr, u->act[li], OUT_BF16, l, false));
		h = u->act[li];
		++li;
	}
	if (infer) FX_CHECK(conv(u, li, SRC_BF16, h, nullptr, 0, nullptr, y, OUT_F32_CH0, 0, false));
	else FX_CHECK(conv(u, li, SRC_BF16, h, nullptr, 0, nullptr, u->yout, OUT_F32, 0, false));
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
		int cip = pad16(u->lcin[li]), cop = pad16(u->lcout[li]);
		// Rows per wave: about 256 pixels (all nine taps) of work per wave, at least one row.
		const long rows = (long)u->B * Hl;
		const int rpw = (int)std::max<long>(1, 256 / Wl);
		dim3 grid((unsigned)((rows + (long)NWAVES * rpw - 1) / ((long)NWAVES * rpw)), 1, (unsigned)((cop / 16) * (cip / 16)));
		hipLaunchKernelGGL(k_conv3_wgrad, grid, dim3(BLOCK), 0, st, g, yact, relu, xin, u->G + u->woff[li], u->B, Hl, Wl,
		                   cip, cop, u->lcin[li], u->lcout[li], rpw);
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
	FX_CHECK(conv(u, li, SRC_F32, u->dy, nullptr, 0, nullptr, g, OUT_F32, 0, true));   // d(act13) (16 ch)
	--li;
	for (int l = 0; l <= 2; ++l) {
		for (int k = 0; k < 2; ++k, --li) {
			FX_CHECK(wgrad(li, g, u->act[li], 1, lin[li], l));
			FX_CHECK(conv(u, li, SRC_F32, g, nullptr, 0, u->act[li], gn, OUT_F32, l, true));
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
				FX_CHECK(conv(u, li, SRC_F32, g, nullptr, 0, u->act[li], gn, OUT_F32, l, true));
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

// Runs the forward on x (FP32 [B][H][W][cin]): inference reads x in its first convolution and writes
// channel 0 to y; training first packs x to BF16 (the backward reads the packed copy).
static fx_status run_forward(fx_unet* u, const float* x, int32_t B, int32_t H, int32_t W, bool infer, float* y) {
	FX_REQUIRE(u && x && B > 0 && H % 8 == 0 && W % 8 == 0 && H > 0 && W > 0, FX_ERR_ARGUMENT,
	           "batch > 0 and H, W divisible by 8 required");
	FX_CHECK(ensure_ws(u, B, H, W, !infer));
	// Inference reads x in the first convolution; training keeps the packed copy for the backward.
	if (infer) return forward(u, infer, y ? y : u->ych0, x);
	long npix = (long)B * H * W;
	hipLaunchKernelGGL(k_pack_in, dim3((unsigned)std::min<long>(nb(npix), 65536)), dim3(256), 0, u->be->stream, x, u->xin,
	                   npix, u->cin, pad16(u->cin), u->skip);
	return forward(u, infer, y ? y : u->ych0);
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
	return run_forward(u, (const float*)xb->ptr, B, H, W, true, (float*)yb->ptr);
}

// Loss Σ_b ‖y_b − t_b‖²/‖t_b‖² and its gradient (kept on the device; fx_unet_get_grad reads it).
fx_status fx_unet_loss_grad(fx_unet* u, const fx_buf* xb, const fx_buf* tb, int32_t B, int32_t H, int32_t W,
                            double* loss) {
	FX_CHECK(check_tensors(u, xb, tb, B, H, W));
	const float* t = (const float*)tb->ptr;
	FX_CHECK(run_forward(u, (const float*)xb->ptr, B, H, W, false, nullptr));
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
	FX_CHECK(run_forward(u, (const float*)xb->ptr, B, H, W, false, nullptr));
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
	fx_status s = run_forward(u, x, B, H, W, true, nullptr);
	u->skip = nullptr;
	*y = u->ych0;
	return s;
}
}  // namespace fx
