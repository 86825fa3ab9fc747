// unet.cpp — FP64 CPU restatement of the learned preconditioner network.
//
// TEST INFRASTRUCTURE ONLY (see oracle.cpp): the checker for the gfx950
// BF16/MFMA kernels in fluxion_amd/csrc/fx_unet.hip.  No reference
// counterpart exists; the network is specified here and in the HIP file:
//
//   input  x: H×W×CIN (NHWC), H and W divisible by 8
//   enc l=0..3:  h = relu(conv3(h)); h = relu(conv3(h)); skip_l = h; h = avgpool2(h) (l<3)
//   dec l=2..0:  h = up2(h) (nearest); h = concat(h, skip_l); h = relu(conv3(h)); h = relu(conv3(h))
//   out:         y = conv3(h) (1 channel, no activation)
//   channels c = {16, 32, 48, 64}; conv3 = 3×3, zero padding, weights W[tap][cout][cin], bias[cout].
//   conv3 is a correlation: y[r][c][o] = b[o] + Σ_t Σ_i W[t][o][i]·x[r + t/3 − 1][c + t%3 − 1][i].
//
// BF16 mode (the `bf` flag of the entry points) rounds every value to the
// precision the GPU stores it in, so the GPU can be held to a tight bar:
// input, weights and the activations of layers 0-13 and of the pools to BF16
// (through FP32, round to nearest even), biases to FP32; in the backward pass
// the ReLU-masked gradient that multiplies weights and activations is rounded
// to BF16 (the bias gradient sums it unrounded).  Sums stay FP64.
//
// Layer order (15 convolutions):
//   0: CIN->16  1: 16->16  2: 16->32  3: 32->32  4: 32->48  5: 48->48  6: 48->64  7: 64->64
//   8: 64+48->48  9: 48->48  10: 48+32->32  11: 32->32  12: 32+16->16  13: 16->16  14: 16->1
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// FP64 -> FP32 -> BF16 (round to nearest even) -> FP64, as the GPU's f2bf of an FP32 value.
double bf16r(double v) {
	float f = (float)v;
	uint32_t u;
	std::memcpy(&u, &f, 4);
	if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
	u &= 0xffff0000u;
	std::memcpy(&f, &u, 4);
	return (double)f;
}
double f32r(double v) { return (double)(float)v; }

struct T4 {
	int h, w, c;
	std::vector<double> v;
	T4() : h(0), w(0), c(0) {}
	T4(int h_, int w_, int c_) : h(h_), w(w_), c(c_), v((size_t)h_ * w_ * c_, 0.0) {}
	double& at(int y, int x, int k) { return v[((size_t)y * w + x) * c + k]; }
	double at(int y, int x, int k) const { return v[((size_t)y * w + x) * c + k]; }
};

const int CH[4] = {16, 32, 48, 64};

struct Layer { int cin, cout; bool relu; };

void layers(int cin0, Layer* L) {
	int i = 0;
	int c = cin0;
	for (int l = 0; l < 4; ++l) {
		L[i++] = {c, CH[l], true};
		L[i++] = {CH[l], CH[l], true};
		c = CH[l];
	}
	for (int l = 2; l >= 0; --l) {
		L[i++] = {c + CH[l], CH[l], true};
		L[i++] = {CH[l], CH[l], true};
		c = CH[l];
	}
	L[i++] = {c, 1, false};
}

void conv3(const T4& x, const double* W, const double* b, int cout, bool relu, T4& y, T4* pre, bool bf) {
	y = T4(x.h, x.w, cout);
	if (pre) *pre = T4(x.h, x.w, cout);
	for (int r = 0; r < x.h; ++r)
		for (int c = 0; c < x.w; ++c)
			for (int o = 0; o < cout; ++o) {
				double acc = b[o];
				for (int t = 0; t < 9; ++t) {
					int rr = r + t / 3 - 1, cc = c + t % 3 - 1;
					if (rr < 0 || cc < 0 || rr >= x.h || cc >= x.w) continue;
					const double* wp = W + ((size_t)t * cout + o) * x.c;
					for (int i = 0; i < x.c; ++i) acc += wp[i] * x.at(rr, cc, i);
				}
				if (pre) pre->at(r, c, o) = acc;
				double v = relu && acc < 0 ? 0.0 : acc;
				// Only the ReLU layers' outputs are stored in BF16; the output layer stays FP32-exact here.
				if (bf && relu) v = bf16r(v);
				y.at(r, c, o) = v;
			}
}

// Backward of y = act(conv3(x)): given dy, accumulate dW, db and return dx.
void conv3_bwd(const T4& x, const double* W, const T4& pre, bool relu, const T4& dy, double* dW, double* db, T4& dx,
               bool bf) {
	int cout = dy.c;
	dx = T4(x.h, x.w, x.c);
	for (int r = 0; r < x.h; ++r)
		for (int c = 0; c < x.w; ++c)
			for (int o = 0; o < cout; ++o) {
				double g = dy.at(r, c, o);
				if (relu && pre.at(r, c, o) <= 0.0) g = 0.0;
				if (g == 0.0) continue;
				db[o] += g;
				if (bf) g = bf16r(g);
				for (int t = 0; t < 9; ++t) {
					int rr = r + t / 3 - 1, cc = c + t % 3 - 1;
					if (rr < 0 || cc < 0 || rr >= x.h || cc >= x.w) continue;
					const double* wp = W + ((size_t)t * cout + o) * x.c;
					double* dwp = dW + ((size_t)t * cout + o) * x.c;
					for (int i = 0; i < x.c; ++i) {
						dwp[i] += g * x.at(rr, cc, i);
						dx.at(rr, cc, i) += g * wp[i];
					}
				}
			}
}

T4 pool2(const T4& x, bool bf) {
	T4 y(x.h / 2, x.w / 2, x.c);
	for (int r = 0; r < y.h; ++r)
		for (int c = 0; c < y.w; ++c)
			for (int k = 0; k < x.c; ++k)
				y.at(r, c, k) = 0.25 * (x.at(2 * r, 2 * c, k) + x.at(2 * r, 2 * c + 1, k) + x.at(2 * r + 1, 2 * c, k) +
				                        x.at(2 * r + 1, 2 * c + 1, k));
	if (bf)
		for (double& v : y.v) v = bf16r(v);
	return y;
}
T4 pool2_bwd(const T4& dy, int h, int w) {
	T4 dx(h, w, dy.c);
	for (int r = 0; r < h; ++r)
		for (int c = 0; c < w; ++c)
			for (int k = 0; k < dy.c; ++k) dx.at(r, c, k) = 0.25 * dy.at(r / 2, c / 2, k);
	return dx;
}
// concat(up2(a), s) along channels.
T4 up_cat(const T4& a, const T4& s) {
	T4 y(s.h, s.w, a.c + s.c);
	for (int r = 0; r < s.h; ++r)
		for (int c = 0; c < s.w; ++c) {
			for (int k = 0; k < a.c; ++k) y.at(r, c, k) = a.at(r / 2, c / 2, k);
			for (int k = 0; k < s.c; ++k) y.at(r, c, a.c + k) = s.at(r, c, k);
		}
	return y;
}
void up_cat_bwd(const T4& dy, int ac, T4& da, T4& ds) {
	da = T4(dy.h / 2, dy.w / 2, ac);
	ds = T4(dy.h, dy.w, dy.c - ac);
	for (int r = 0; r < dy.h; ++r)
		for (int c = 0; c < dy.w; ++c) {
			for (int k = 0; k < ac; ++k) da.at(r / 2, c / 2, k) += dy.at(r, c, k);
			for (int k = ac; k < dy.c; ++k) ds.at(r, c, k - ac) = dy.at(r, c, k);
		}
}

struct Net {
	int cin0;
	bool bf = false;   // BF16 storage emulation
	Layer L[15];
	std::vector<size_t> woff, boff;
	explicit Net(int cin) : cin0(cin) {
		layers(cin, L);
		size_t o = 0;
		for (int i = 0; i < 15; ++i) { woff.push_back(o); o += (size_t)9 * L[i].cin * L[i].cout; }
		for (int i = 0; i < 15; ++i) { boff.push_back(o); o += L[i].cout; }
	}
	size_t nparams() const { return boff[14] + L[14].cout; }
};

struct Tape {
	T4 in[15], pre[15], out[15];
	T4 skip[4];
	int skip_h[4], skip_w[4];
};

void forward(const Net& N, const double* P, const T4& x, T4& y, Tape& tp) {
	T4 h = x;
	int li = 0;
	for (int l = 0; l < 4; ++l) {
		for (int k = 0; k < 2; ++k, ++li) {
			tp.in[li] = h;
			conv3(h, P + N.woff[li], P + N.boff[li], N.L[li].cout, N.L[li].relu, tp.out[li], &tp.pre[li], N.bf);
			h = tp.out[li];
		}
		tp.skip[l] = h;
		if (l < 3) h = pool2(h, N.bf);
	}
	for (int l = 2; l >= 0; --l) {
		h = up_cat(h, tp.skip[l]);
		for (int k = 0; k < 2; ++k, ++li) {
			tp.in[li] = h;
			conv3(h, P + N.woff[li], P + N.boff[li], N.L[li].cout, N.L[li].relu, tp.out[li], &tp.pre[li], N.bf);
			h = tp.out[li];
		}
	}
	tp.in[li] = h;
	conv3(h, P + N.woff[li], P + N.boff[li], 1, false, tp.out[li], &tp.pre[li], N.bf);
	y = tp.out[li];
}

// Backward of the network given dL/dy; accumulates into G (same layout as P).
void backward(const Net& N, const double* P, const Tape& tp, const T4& dy, double* G, T4* dxin) {
	T4 g = dy, dx;
	int li = 14;
	conv3_bwd(tp.in[li], P + N.woff[li], tp.pre[li], false, g, G + N.woff[li], G + N.boff[li], dx, N.bf);
	g = dx;
	--li;
	T4 dskip[4];
	for (int l = 0; l < 3; ++l) dskip[l] = T4(tp.skip[l].h, tp.skip[l].w, tp.skip[l].c);
	for (int l = 0; l <= 2; ++l) {
		// decoder level l in reverse order of the forward (forward went l = 2, 1, 0)
		int lev = l;
		for (int k = 0; k < 2; ++k, --li) {
			conv3_bwd(tp.in[li], P + N.woff[li], tp.pre[li], N.L[li].relu, g, G + N.woff[li], G + N.boff[li], dx, N.bf);
			g = dx;
		}
		T4 da, ds;
		up_cat_bwd(g, g.c - tp.skip[lev].c, da, ds);
		for (size_t q = 0; q < ds.v.size(); ++q) dskip[lev].v[q] += ds.v[q];
		g = da;
	}
	for (int l = 3; l >= 0; --l) {
		if (l < 3) {
			T4 up = pool2_bwd(g, tp.skip[l].h, tp.skip[l].w);
			for (size_t q = 0; q < up.v.size(); ++q) up.v[q] += dskip[l].v[q];
			g = up;
		}
		for (int k = 0; k < 2; ++k, --li) {
			conv3_bwd(tp.in[li], P + N.woff[li], tp.pre[li], N.L[li].relu, g, G + N.woff[li], G + N.boff[li], dx, N.bf);
			g = dx;
		}
	}
	if (dxin) *dxin = g;
}

// The parameters the network computes with: P itself, or (BF16 mode) weights rounded to BF16 and
// biases to FP32.
const double* net_params(const Net& N, const double* P, std::vector<double>& store) {
	if (!N.bf) return P;
	store.assign(P, P + N.nparams());
	for (size_t q = 0; q < N.boff[0]; ++q) store[q] = bf16r(store[q]);
	for (size_t q = N.boff[0]; q < store.size(); ++q) store[q] = f32r(store[q]);
	return store.data();
}
// One sample's input (BF16 mode: rounded to BF16).
T4 net_input(const Net& N, const double* x, int h, int w, int cin) {
	T4 xi(h, w, cin);
	std::memcpy(xi.v.data(), x, sizeof(double) * h * w * cin);
	if (N.bf)
		for (double& v : xi.v) v = bf16r(v);
	return xi;
}

}  // namespace

extern "C" {

size_t orc_unet_nparams(int cin) { return Net(cin).nparams(); }

// Parameter layout: all conv weights W_l[tap][cout][cin] in layer order, then all biases.
void orc_unet_layout(int cin, int64_t* woff, int64_t* boff, int* lcin, int* lcout) {
	Net N(cin);
	for (int i = 0; i < 15; ++i) {
		woff[i] = (int64_t)N.woff[i];
		boff[i] = (int64_t)N.boff[i];
		lcin[i] = N.L[i].cin;
		lcout[i] = N.L[i].cout;
	}
}

// y[b][h][w] = net(x[b][h][w][cin]); bf selects the BF16 storage emulation.
void orc_unet_forward(const double* P, const double* x, double* y, int batch, int h, int w, int cin, int bf) {
	Net N(cin);
	N.bf = bf != 0;
	std::vector<double> pst;
	P = net_params(N, P, pst);
	for (int b = 0; b < batch; ++b) {
		T4 xi = net_input(N, x + (size_t)b * h * w * cin, h, w, cin), yo;
		Tape tp;
		forward(N, P, xi, yo, tp);
		std::memcpy(y + (size_t)b * h * w, yo.v.data(), sizeof(double) * h * w);
	}
}

// Loss = Σ_b Σ_p (y − t)² / Σ_p t²  (relative squared error per sample, summed over the batch);
// returns the loss and writes dLoss/dP into G (zeroed here).
double orc_unet_loss_grad(const double* P, const double* x, const double* t, double* G, int batch, int h, int w,
                          int cin, int bf) {
	Net N(cin);
	N.bf = bf != 0;
	std::vector<double> pst;
	P = net_params(N, P, pst);
	std::memset(G, 0, sizeof(double) * N.nparams());
	double loss = 0.0;
	for (int b = 0; b < batch; ++b) {
		T4 xi = net_input(N, x + (size_t)b * h * w * cin, h, w, cin), yo;
		Tape tp;
		forward(N, P, xi, yo, tp);
		const double* tb = t + (size_t)b * h * w;
		double tt = 0.0, ee = 0.0;
		for (int q = 0; q < h * w; ++q) tt += tb[q] * tb[q];
		T4 dy(h, w, 1);
		for (int q = 0; q < h * w; ++q) {
			double d = yo.v[q] - tb[q];
			ee += d * d;
			dy.v[q] = 2.0 * d / tt;
		}
		loss += ee / tt;
		backward(N, P, tp, dy, G, nullptr);
	}
	return loss;
}

// Relative energy-norm loss: Σ_b dᵀA_b d / tᵀA_b t with d = y − t and A_b the unit-spacing
// 5-point −∇·(k∇·) on interior pixels (k = exp of input channel 1, harmonic face means,
// values outside the interior taken as zero); dLoss/dy = 2·A_b d / tᵀA_b t on interior pixels.
double orc_unet_loss_grad_energy(const double* P, const double* x, const double* t, double* G, int batch, int h,
                                 int w, int cin, int bf) {
	Net N(cin);
	N.bf = bf != 0;
	std::vector<double> pst;
	P = net_params(N, P, pst);
	std::memset(G, 0, sizeof(double) * N.nparams());
	double loss = 0.0;
	for (int b = 0; b < batch; ++b) {
		const double* xb = x + (size_t)b * h * w * cin;
		T4 xi = net_input(N, xb, h, w, cin), yo;
		Tape tp;
		forward(N, P, xi, yo, tp);
		const double* tb = t + (size_t)b * h * w;
		auto inner = [&](int r, int c) { return r > 0 && c > 0 && r < h - 1 && c < w - 1; };
		// k from the unrounded input, as the loss reads it.
		auto kk = [&](int r, int c) { return std::exp(xb[((size_t)r * w + c) * cin + 1]); };
		auto hm = [](double a, double c) { return 2.0 * a * c / (a + c); };
		auto apply = [&](const std::vector<double>& v, std::vector<double>& out) {
			for (int r = 0; r < h; ++r)
				for (int c = 0; c < w; ++c) {
					double o = 0.0;
					if (inner(r, c)) {
						const double kc = kk(r, c), vc = v[(size_t)r * w + c];
						const int dr[4] = {0, 0, -1, 1}, dc[4] = {-1, 1, 0, 0};
						for (int e = 0; e < 4; ++e) {
							const int rr = r + dr[e], cc = c + dc[e];
							const double vn = inner(rr, cc) ? v[(size_t)rr * w + cc] : 0.0;
							o += hm(kc, kk(rr, cc)) * (vc - vn);
						}
					}
					out[(size_t)r * w + c] = o;
				}
		};
		std::vector<double> d((size_t)h * w), tv((size_t)h * w), ad((size_t)h * w), at((size_t)h * w);
		for (int q = 0; q < h * w; ++q) {
			const int r = q / w, c = q % w;
			d[q] = inner(r, c) ? yo.v[q] - tb[q] : 0.0;
			tv[q] = inner(r, c) ? tb[q] : 0.0;
		}
		apply(d, ad);
		apply(tv, at);
		double ee = 0.0, tt = 0.0;
		for (int q = 0; q < h * w; ++q) {
			ee += d[q] * ad[q];
			tt += tv[q] * at[q];
		}
		T4 dy(h, w, 1);
		for (int q = 0; q < h * w; ++q) dy.v[q] = 2.0 * ad[q] / tt;
		loss += ee / tt;
		backward(N, P, tp, dy, G, nullptr);
	}
	return loss;
}

}  // extern "C"
