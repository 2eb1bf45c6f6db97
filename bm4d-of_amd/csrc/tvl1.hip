// TV-L1 optical flow on gfx950: Zach-Pock-Bischof duality scheme as the
// reference implements it (lib/tvl1flow/tvl1flow_lib.c, mask.c,
// bicubic_interpolation.c, zoom.c).
//
// All frame pairs of a sequence are solved together: every kernel takes a
// pair index in blockIdx.y and skips pairs that have converged.  Per-pixel
// arithmetic follows the reference operation by operation (this file is
// compiled with -ffp-contract=off; bicubic and Gaussian sums in double as in
// the reference), so a pixel's update is bit-identical given identical
// inputs.  The per-iteration mean squared update is reduced in double in a
// fixed order; the reference sums it in float in pixel order, so the stopping
// iteration can differ by one where the error lands next to epsilon^2.
#include "ctx.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include <algorithm>
#include <array>
#include <mutex>
#include <stdexcept>

namespace bm4dof {

#define TV_MAX_ITERATIONS 300      // tvl1flow_lib.c:24
#define TV_PRESMOOTHING_SIGMA 0.8  // tvl1flow_lib.c:25
#define TV_GRAD_IS_ZERO 1E-10      // tvl1flow_lib.c:26
#define TV_ZOOM_SIGMA_ZERO 0.6     // zoom.c:17
#define TV_THREADS 256

// --------------------------------------------------------------- bicubic

__device__ __forceinline__ int nbc(int x, int n, bool &out) // bicubic_interpolation.c:25-39
{
	if (x < 0) { out = true; return 0; }
	if (x >= n) { out = true; return n - 1; }
	return x;
}

__device__ __forceinline__ double cubic_cell(double v0, double v1, double v2, double v3, double x)
{ // bicubic_interpolation.c:102-112
	return v1 + 0.5 * x * (v2 - v0 +
		x * (2.0 * v0 - 5.0 * v1 + 4.0 * v2 - v3 +
		x * (3.0 * (v1 - v2) + v3 - v0)));
}

// bicubic_interpolation_at (bicubic_interpolation.c:138-234), Neumann case;
// the y-neighbour above uses the x sign, as in the reference.
__device__ float bicubic_at(const float *__restrict__ in, float uu, float vv, int nx, int ny,
                            bool border_out)
{
	const int sx = (uu < 0) ? -1 : 1;
	const int sy = (vv < 0) ? -1 : 1;
	bool out = false;
	const int x = nbc((int)uu, nx, out);
	const int y = nbc((int)vv, ny, out);
	const int mx = nbc((int)uu - sx, nx, out);
	const int my = nbc((int)vv - sx, ny, out);
	const int dx = nbc((int)uu + sx, nx, out);
	const int dy = nbc((int)vv + sy, ny, out);
	const int ddx = nbc((int)uu + 2 * sx, nx, out);
	const int ddy = nbc((int)vv + 2 * sy, ny, out);
	if (out && border_out) return 0.0f;
	const int cols[4] = {mx, x, dx, ddx};
	const double fy = (double)(vv - y);
	double v[4];
#pragma unroll
	for (int c = 0; c < 4; c++) {
		const double a0 = in[cols[c] + nx * my];
		const double a1 = in[cols[c] + nx * y];
		const double a2 = in[cols[c] + nx * dy];
		const double a3 = in[cols[c] + nx * ddy];
		v[c] = cubic_cell(a0, a1, a2, a3, fy);
	}
	return (float)cubic_cell(v[0], v[1], v[2], v[3], (double)(uu - x));
}

// ------------------------------------------------------------ pyramid

// min/max of I0 and I1 per pair (tvl1flow_lib.c:286-320): partials per block.
__global__ void tv_minmax_kernel(const float *__restrict__ I, int size, const int *__restrict__ i0,
                                 const int *__restrict__ i1, float *__restrict__ part)
{
	const int pr = blockIdx.y;
	const float *a = I + (size_t)i0[pr] * size;
	const float *b = I + (size_t)i1[pr] * size;
	float mn0 = INFINITY, mx0 = -INFINITY, mn1 = INFINITY, mx1 = -INFINITY;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += gridDim.x * blockDim.x) {
		mn0 = fminf(mn0, a[i]); mx0 = fmaxf(mx0, a[i]);
		mn1 = fminf(mn1, b[i]); mx1 = fmaxf(mx1, b[i]);
	}
	__shared__ float s[4][TV_THREADS];
	s[0][threadIdx.x] = mn0; s[1][threadIdx.x] = mx0; s[2][threadIdx.x] = mn1; s[3][threadIdx.x] = mx1;
	__syncthreads();
	for (int st = TV_THREADS / 2; st > 0; st >>= 1) {
		if (threadIdx.x < st) {
			s[0][threadIdx.x] = fminf(s[0][threadIdx.x], s[0][threadIdx.x + st]);
			s[1][threadIdx.x] = fmaxf(s[1][threadIdx.x], s[1][threadIdx.x + st]);
			s[2][threadIdx.x] = fminf(s[2][threadIdx.x], s[2][threadIdx.x + st]);
			s[3][threadIdx.x] = fmaxf(s[3][threadIdx.x], s[3][threadIdx.x + st]);
		}
		__syncthreads();
	}
	if (threadIdx.x < 4) part[((size_t)pr * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = s[threadIdx.x][0];
}

// image_normalization (tvl1flow_lib.c:303-337) into the level-0 buffers.
__global__ void tv_normalize_kernel(const float *__restrict__ I, int size, const int *__restrict__ i0,
                                    const int *__restrict__ i1, const float *__restrict__ part,
                                    int nparts, float *__restrict__ I0n, float *__restrict__ I1n)
{
	const int pr = blockIdx.y;
	float mn0 = INFINITY, mx0 = -INFINITY, mn1 = INFINITY, mx1 = -INFINITY;
	for (int k = 0; k < nparts; k++) {
		const float *q = part + ((size_t)pr * nparts + k) * 4;
		mn0 = fminf(mn0, q[0]); mx0 = fmaxf(mx0, q[1]); mn1 = fminf(mn1, q[2]); mx1 = fmaxf(mx1, q[3]);
	}
	const float max = (mx0 > mx1) ? mx0 : mx1;
	const float min = (mn0 < mn1) ? mn0 : mn1;
	const float den = max - min;
	const float *a = I + (size_t)i0[pr] * size;
	const float *b = I + (size_t)i1[pr] * size;
	float *oa = I0n + (size_t)pr * size, *ob = I1n + (size_t)pr * size;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += gridDim.x * blockDim.x) {
		if (den > 0) {
			oa[i] = 255.0 * (a[i] - min) / den;
			ob[i] = 255.0 * (b[i] - min) / den;
		} else {
			oa[i] = a[i];
			ob[i] = b[i];
		}
	}
}

// Gaussian smoothing with reflecting boundaries (mask.c:217-333), one pass.
// rows: convolve along x; else along y.  Taps in constant-size array B.
__global__ void tv_gauss_kernel(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                int stride_pair, const double *__restrict__ B, int size, int rows)
{
	const int pr = blockIdx.y;
	const float *I = in + (size_t)pr * stride_pair;
	float *O = out + (size_t)pr * stride_pair;
	const int n = nx * ny;
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
		const int i = p / nx, j = p - i * nx;
		// extended coordinate e = (rows ? j : i) + size; R[e] with reflection
		const int len = rows ? nx : ny;
		const int pos = rows ? j : i;
		auto R = [&](int e) -> double {
			int q;
			if (e < size) q = size - e;                  // R[i] = I[size - i]
			else if (e >= len + size) q = len - 1 - (e - len - size); // R[bd+i] = I[len-1-i]
			else q = e - size;
			return rows ? (double)I[i * nx + q] : (double)I[q * nx + j];
		};
		const int e = pos + size;
		double sum = B[0] * R(e);
		for (int t = 1; t < size; t++) sum += B[t] * (R(e - t) + R(e + t));
		O[p] = (float)sum;
	}
}

// zoom_out resampling (zoom.c:66-76) of the pre-smoothed image.
__global__ void tv_zoom_out_kernel(const float *__restrict__ Is, int nx, int ny, int in_stride,
                                   float *__restrict__ out, int nxx, int nyy, int out_stride,
                                   float factor)
{
	const int pr = blockIdx.y;
	const float *I = Is + (size_t)pr * in_stride;
	float *O = out + (size_t)pr * out_stride;
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nxx * nyy; p += gridDim.x * blockDim.x) {
		const int i1 = p / nxx, j1 = p - i1 * nxx;
		const float i2 = (float)i1 / factor;
		const float j2 = (float)j1 / factor;
		O[p] = bicubic_at(I, j2, i2, nx, ny, false);
	}
}

// zoom_in (zoom.c:87-111) followed by the 1/zfactor scaling
// (tvl1flow_lib.c:426-435); `scale` is (float)1.0/zfactor.
__global__ void tv_zoom_in_kernel(const float *__restrict__ in, int nx, int ny, int in_stride,
                                  float *__restrict__ out, int nxx, int nyy, int out_stride,
                                  float scale)
{
	const int pr = blockIdx.y;
	const float *I = in + (size_t)pr * in_stride;
	float *O = out + (size_t)pr * out_stride;
	const float factorx = ((float)nxx / nx);
	const float factory = ((float)nyy / ny);
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nxx * nyy; p += gridDim.x * blockDim.x) {
		const int i1 = p / nxx, j1 = p - i1 * nxx;
		const float i2 = (float)i1 / factory;
		const float j2 = (float)j1 / factorx;
		O[p] = bicubic_at(I, j2, i2, nx, ny, false) * scale;
	}
}

// ------------------------------------------------------------ one scale

// centered_gradient (mask.c:152-209): clamped central differences, 0.5*()
// in double.
__global__ void tv_cgrad_kernel(const float *__restrict__ I1, int nx, int ny, int stride,
                                float *__restrict__ Ix, float *__restrict__ Iy)
{
	const int pr = blockIdx.y;
	const float *in = I1 + (size_t)pr * stride;
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nx * ny; p += gridDim.x * blockDim.x) {
		const int i = p / nx, j = p - i * nx;
		const int jl = j > 0 ? j - 1 : 0, jr = j < nx - 1 ? j + 1 : nx - 1;
		const int iu = i > 0 ? i - 1 : 0, id = i < ny - 1 ? i + 1 : ny - 1;
		Ix[(size_t)pr * stride + p] = 0.5 * (in[i * nx + jr] - in[i * nx + jl]);
		Iy[(size_t)pr * stride + p] = 0.5 * (in[id * nx + j] - in[iu * nx + j]);
	}
}

// Warping of I1, I1x, I1y by u and the constant part of rho
// (tvl1flow_lib.c:142-161).
__global__ void tv_warp_kernel(const float *__restrict__ I0, const float *__restrict__ I1,
                               const float *__restrict__ I1x, const float *__restrict__ I1y,
                               const float *__restrict__ u1a, const float *__restrict__ u2a,
                               const float *__restrict__ u1b, const float *__restrict__ u2b,
                               const int *__restrict__ cur, int nx, int ny, int stride,
                               float *__restrict__ I1wx, float *__restrict__ I1wy, float *__restrict__ grad,
                               float *__restrict__ rho_c)
{
	const int pr = blockIdx.y;
	const size_t off = (size_t)pr * stride;
	const bool flip = cur[pr] != 0;
	const float *__restrict__ u1 = flip ? u1b : u1a;
	const float *__restrict__ u2 = flip ? u2b : u2a;
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nx * ny; p += gridDim.x * blockDim.x) {
		const int i = p / nx, j = p - i * nx;
		const float uu = (float)(j + u1[off + p]);
		const float vv = (float)(i + u2[off + p]);
		const float w = bicubic_at(I1 + off, uu, vv, nx, ny, true);
		const float wx = bicubic_at(I1x + off, uu, vv, nx, ny, true);
		const float wy = bicubic_at(I1y + off, uu, vv, nx, ny, true);
		const float Ix2 = wx * wx;
		const float Iy2 = wy * wy;
		I1wx[off + p] = wx;
		I1wy[off + p] = wy;
		grad[off + p] = (Ix2 + Iy2);
		rho_c[off + p] = (w - wx * u1[off + p] - wy * u2[off + p] - I0[off + p]);
	}
}

// ------------------------------------------------- primal-dual iterations
// K iterations of the TV-L1 primal-dual scheme (tvl1flow_lib.c:165-246) per
// launch, on LDS tiles: a block loads its TX x TY tile with a K-pixel halo
// (clipped to the image) and iterates there; after iteration j the pixels at
// least j pixels inside a cut halo edge are exact (an iteration reaches one
// pixel in each direction: the divergence backwards, the gradient forwards),
// so after K iterations the tile itself is, and only it is written back.
// Per iteration: the TH step, divergence of p and u update for every pixel
// (phase A), then the forward gradient of the new u and the p update
// (phase B); each pixel's arithmetic is the reference's, operation for
// operation, as in the one-iteration form.  The squared update of every
// tile pixel and iteration is summed in double (pixels in thread order,
// lanes by xor tree, waves in order) into part[pair][j][tile].
//
// The state (u1, u2, p11, p12, p21, p22) is double-buffered: a pair reads
// buffer cur[pair] and writes the other.  ctl[pair] is the number of
// iterations to run this round (0: idle).  tv_round_kernel then sums the
// tiles' partials and applies the stopping test to each iteration of the
// round in order (tv_round_decide); if it fires before the round's last
// iteration, the pair runs that many iterations again from the same source
// buffer (an exact re-run), so the state and the iteration count are those
// of the iteration the test fired on.
// ---------------------------------------------------------------------------
struct TVPlanes {
	float *u1, *u2, *p11, *p12, *p21, *p22;
};

#define TV_TX 64
#define TV_TY 32
#define TV_K 4
#define TV_NT 512
enum { TV_RUN = 0, TV_REDO = 1 };

// The stopping test of every iteration of a round, in order
// (tvl1flow_lib.c:165, 225), given errs[j] = the mean squared update of
// iteration j: the loop ends when error <= epsilon^2 or the count reaches
// MAX_ITERATIONS.  state[pair] receives the final count.  A round that
// stops before its last iteration is run again from the same source buffer
// with that many iterations (mode TV_REDO), so the state and the count are
// those of the iteration the test fired on.
__device__ void tv_round_decide(int pr, int kk, const float *errs, float eps2, int K, int *__restrict__ ctl,
                                int *__restrict__ cur, int *__restrict__ iters, int *__restrict__ mode,
                                int *__restrict__ state, int *__restrict__ n_active)
{
	const int done = iters[pr];
	if (mode[pr] == TV_REDO) {   // the exact re-run of the iterations up to the test
		iters[pr] = done + kk;
		cur[pr] ^= 1;
		state[pr] = done + kk;
		ctl[pr] = 0;
		atomicSub(n_active, 1);
		return;
	}
	int stop = -1;
	for (int j = 0; j < kk && stop < 0; j++)
		if (!(errs[j] > eps2) || done + j + 1 >= TV_MAX_ITERATIONS) stop = j;
	if (stop < 0) {
		iters[pr] = done + kk;
		cur[pr] ^= 1;
		ctl[pr] = min(K, TV_MAX_ITERATIONS - (done + kk));
	} else if (stop == kk - 1) {
		iters[pr] = done + kk;
		cur[pr] ^= 1;
		state[pr] = done + kk;
		ctl[pr] = 0;
		atomicSub(n_active, 1);
	} else {
		mode[pr] = TV_REDO;
		ctl[pr] = stop + 1;
	}
}


// Per-slot flags of a tile pixel (the LDS index is in the low 16 bits).
enum : unsigned {
	TVF_VALID = 1u << 16,  // the slot holds a pixel
	TVF_TILE = 1u << 17,   // inside the tile itself (written back, counted in the error)
	TVF_TOP = 1u << 18,    // image row 0
	TVF_BOT = 1u << 19,    // image row ny - 1
	TVF_LEFT = 1u << 20,   // image column 0
	TVF_RIGHT = 1u << 21,  // image column nx - 1
	TVF_BORDER = TVF_TOP | TVF_BOT | TVF_LEFT | TVF_RIGHT,
};

// sqrt of a double that is +0 or at least 2^-767, as the compiler's
// correctly rounded sqrt computes it (rsq, then the Newton steps below) with
// its range scaling left out: that scaling applies only below 2^-767, and
// x^2 + y^2 of floats is +0 or at least 2^-298.  Same bits, five
// instructions fewer.
__device__ __forceinline__ double sqrt_sum_sq(double s)
{
	const double y = __builtin_amdgcn_rsq(s);
	double g = s * y, h = y * 0.5;
	const double r = __builtin_fma(-h, g, 0.5);
	g = __builtin_fma(g, r, g);
	h = __builtin_fma(h, r, h);
	double d = __builtin_fma(-g, g, s);
	g = __builtin_fma(d, h, g);
	d = __builtin_fma(-g, g, s);
	g = __builtin_fma(d, h, g);
	return __builtin_amdgcn_class(s, 0x260) ? s : g;   // +-0 and +inf give themselves
}

// divergence of (v1, v2) with backward differences at an image-border pixel
// (mask.c:43-92); v1m / v2u are the left / upper neighbours.
__device__ __forceinline__ float div_border(float v1, float v1m, float v2, float v2u, unsigned f)
{
	const bool top = f & TVF_TOP, bot = f & TVF_BOT, left = f & TVF_LEFT, right = f & TVF_RIGHT;
	if (top && !left && !right) return v1 - v1m + v2;
	if (bot && !left && !right) return v1 - v1m - v2u;
	if (left && !top && !bot) return v1 + v2 - v2u;
	if (right && !top && !bot) return -v1m + v2 - v2u;
	if (top && left) return v1 + v2;
	if (top && right) return -v1m + v2;
	if (bot && left) return v1 - v2u;
	return -v1m - v2u;
}

// kk iterations on the LDS tile (planes at L + c * PS, rows of RW floats);
// BORDER: the tile touches the image border, so the boundary cases of the
// divergence and gradient apply to some of its pixels.
template <int RW, int PS, int NPT, int K, bool BORDER>
__device__ __forceinline__ void tv_iterate(float *L, unsigned (&sl)[NPT], const float (&cwx)[NPT],
                                           const float (&cwy)[NPT], const float (&cg)[NPT],
                                           const float (&crc)[NPT], int kk, float l_t, float theta,
                                           float taut, double (&err)[K])
{
	constexpr int U1 = 0, U2 = PS, P11 = 2 * PS, P12 = 3 * PS, P21 = 4 * PS, P22 = 5 * PS;
	for (int j = 0; j < kk; j++) {
		// the flags are re-read every iteration (not held as masks across it)
#pragma unroll
		for (int k = 0; k < NPT; k++) asm volatile("" : "+v"(sl[k]));
		// phase A: TH, divergence, u (tvl1flow_lib.c:170-225)
		double ej = 0.0;
#pragma unroll
		for (int k = 0; k < NPT; k++) {
			const unsigned f = sl[k];
			if (!(f & TVF_VALID)) continue;
			const int e = f & 0xffff;
			const float u1k = L[U1 + e], u2k = L[U2 + e];
			const float wx = cwx[k], wy = cwy[k], g = cg[k];
			const float rho = crc[k] + (wx * u1k + wy * u2k);
			float d1, d2;
			if (rho < -l_t * g) { d1 = l_t * wx; d2 = l_t * wy; }
			else if (rho > l_t * g) { d1 = -l_t * wx; d2 = -l_t * wy; }
			else if (g < TV_GRAD_IS_ZERO) d1 = d2 = 0;
			else { const float fi = -rho / g; d1 = fi * wx; d2 = fi * wy; }
			const float v1 = u1k + d1, v2 = u2k + d2;
			const float a11 = L[P11 + e], a11m = L[P11 + e - 1], a12 = L[P12 + e], a12u = L[P12 + e - RW];
			const float a21 = L[P21 + e], a21m = L[P21 + e - 1], a22 = L[P22 + e], a22u = L[P22 + e - RW];
			float dp1, dp2;
			if (!BORDER || !(f & TVF_BORDER)) {
				dp1 = (a11 - a11m) + (a12 - a12u);
				dp2 = (a21 - a21m) + (a22 - a22u);
			} else {
				dp1 = div_border(a11, a11m, a12, a12u, f);
				dp2 = div_border(a21, a21m, a22, a22u, f);
			}
			const float n1 = v1 + theta * dp1;
			const float n2 = v2 + theta * dp2;
			L[U1 + e] = n1;
			L[U2 + e] = n2;
			// squared update of the tile's own pixels (an and-mask, not a branch)
			const double q = (double)((n1 - u1k) * (n1 - u1k) + (n2 - u2k) * (n2 - u2k));
			const long long tm = (f & TVF_TILE) ? -1ll : 0ll;
			ej += __longlong_as_double(__double_as_longlong(q) & tm);
		}
#pragma unroll
		for (int q = 0; q < K; q++) if (q == j) err[q] = ej;
		__syncthreads();
		// phase B: forward gradient of the new u, p (tvl1flow_lib.c:228-245,
		// mask.c:101-144)
#pragma unroll
		for (int k = 0; k < NPT; k++) {
			const unsigned f = sl[k];
			if (!(f & TVF_VALID)) continue;
			const int e = f & 0xffff;
			const float a1 = L[U1 + e], a2 = L[U2 + e];
			const float r1 = L[U1 + e + 1], d1 = L[U1 + e + RW], r2 = L[U2 + e + 1], d2 = L[U2 + e + RW];
			float u1x = r1 - a1, u1y = d1 - a1, u2x = r2 - a2, u2y = d2 - a2;
			if (BORDER) {
				if (f & TVF_RIGHT) u1x = u2x = 0.f;
				if (f & TVF_BOT) u1y = u2y = 0.f;
			}
			const float g1 = (float)sqrt_sum_sq((double)u1x * u1x + (double)u1y * u1y);
			const float g2 = (float)sqrt_sum_sq((double)u2x * u2x + (double)u2y * u2y);
			// the reference adds 1.0 in double and stores to float; for a
			// float t >= 0 that is the float sum 1 + t: the double sum is
			// exact for 2^-29 <= t < 2^53, rounds to a value that rounds to 1
			// below, and to t above, as the float sum does
			const float ng1 = 1.0f + taut * g1;
			const float ng2 = 1.0f + taut * g2;
			L[P11 + e] = (L[P11 + e] + taut * u1x) / ng1;
			L[P12 + e] = (L[P12 + e] + taut * u1y) / ng1;
			L[P21 + e] = (L[P21 + e] + taut * u2x) / ng2;
			L[P22 + e] = (L[P22 + e] + taut * u2y) / ng2;
		}
		__syncthreads();
	}
}

template <int TX, int TY, int K, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
void tv_iter_kernel(const float *__restrict__ I1wx, const float *__restrict__ I1wy,
                    const float *__restrict__ grad, const float *__restrict__ rho_c, TVPlanes A, TVPlanes B,
                    int nx, int ny, int stride, int tiles_x, float l_t, float theta, float taut,
                    const int *__restrict__ ctl, const int *__restrict__ cur, double *__restrict__ part)
{
	// LDS: six planes u1, u2, p11, p12, p21, p22 of the tile with its halo,
	// rows of RW floats, one spare row before and after so that the
	// neighbours of every slot are inside the plane
	constexpr int RW = TX + 2 * K, RH = TY + 2 * K, EN = RW * RH, PS = EN + 2 * RW;
	constexpr int NPT = (EN + NT - 1) / NT, NW = NT / 64;
	constexpr int U1 = 0, U2 = PS, P11 = 2 * PS, P12 = 3 * PS, P21 = 4 * PS, P22 = 5 * PS;
	__shared__ float L[6 * PS];
	__shared__ double red[NW][K];
	const int pr = blockIdx.y;
	const int kk = ctl[pr];
	if (kk <= 0) return;
	const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
	const int tile = blockIdx.x, ntiles = gridDim.x;
	const int tx0 = (tile % tiles_x) * TX, ty0 = (tile / tiles_x) * TY;
	const int xs = max(0, tx0 - K), ys = max(0, ty0 - K);
	const int EW = min(nx, tx0 + TX + K) - xs, EH = min(ny, ty0 + TY + K) - ys, en = EW * EH;
	const bool flip = cur[pr] != 0;
	const TVPlanes S = flip ? B : A;
	const size_t off = (size_t)pr * stride;
	float cwx[NPT], cwy[NPT], cg[NPT], crc[NPT];
	unsigned sl[NPT];
#pragma unroll
	for (int k = 0; k < NPT; k++) {
		const int q = tid + k * NT;
		sl[k] = 0;
		cwx[k] = cwy[k] = cg[k] = crc[k] = 0.f;
		if (q < en) {
			const int ey = q / EW, ex = q - ey * EW;
			const int gx = xs + ex, gy = ys + ey;
			const int e = RW + ey * RW + ex;
			const size_t gi = off + (size_t)gy * nx + gx;
			cwx[k] = I1wx[gi]; cwy[k] = I1wy[gi]; cg[k] = grad[gi]; crc[k] = rho_c[gi];
			L[U1 + e] = S.u1[gi]; L[U2 + e] = S.u2[gi];
			L[P11 + e] = S.p11[gi]; L[P12 + e] = S.p12[gi]; L[P21 + e] = S.p21[gi]; L[P22 + e] = S.p22[gi];
			unsigned f = (unsigned)e | TVF_VALID;
			if (gx >= tx0 && gx < tx0 + TX && gy >= ty0 && gy < ty0 + TY) f |= TVF_TILE;
			if (gy == 0) f |= TVF_TOP;
			if (gy == ny - 1) f |= TVF_BOT;
			if (gx == 0) f |= TVF_LEFT;
			if (gx == nx - 1) f |= TVF_RIGHT;
			sl[k] = f;
		}
	}
	double err[K];
#pragma unroll
	for (int j = 0; j < K; j++) err[j] = 0.0;
	__syncthreads();
	// block-uniform: does the tile with its halo reach the image border
	if (xs == 0 || ys == 0 || xs + EW == nx || ys + EH == ny)
		tv_iterate<RW, PS, NPT, K, true>(L, sl, cwx, cwy, cg, crc, kk, l_t, theta, taut, err);
	else
		tv_iterate<RW, PS, NPT, K, false>(L, sl, cwx, cwy, cg, crc, kk, l_t, theta, taut, err);
	// the tile itself back to the other buffer
	const TVPlanes D = flip ? A : B;
#pragma unroll
	for (int k = 0; k < NPT; k++) {
		const unsigned f = sl[k];
		if (!(f & TVF_TILE)) continue;
		const int q = tid + k * NT;
		const int ey = q / EW, ex = q - ey * EW;
		const int e = f & 0xffff;
		const size_t gi = off + (size_t)(ys + ey) * nx + (xs + ex);
		D.u1[gi] = L[U1 + e]; D.u2[gi] = L[U2 + e];
		D.p11[gi] = L[P11 + e]; D.p12[gi] = L[P12 + e]; D.p21[gi] = L[P21 + e]; D.p22[gi] = L[P22 + e];
	}
	// squared updates of the tile, per iteration
#pragma unroll
	for (int j = 0; j < K; j++) {
		double v = err[j];
#pragma unroll
		for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
		if (lane == 0) red[wv][j] = v;
	}
	__syncthreads();
	if (tid < kk) {
		double v = 0.0;
		for (int w = 0; w < NW; w++) v += red[w][tid];
		part[((size_t)pr * K + tid) * ntiles + tile] = v;
	}
}

// The stopping test of a round: error of iteration j = (sum over the
// tiles' partials) / size, then tv_round_decide.  One wave per pair: 256
// strided sums (sum v: tiles v, v + 256, ...) held four to a lane, then a
// tree over them (sum v += sum v + s for s = 128, 64, ..., 1).
__global__ __launch_bounds__(64)
void tv_round_kernel(const double *__restrict__ part, int K, int ntiles, int size, float eps2,
                     int *__restrict__ ctl, int *__restrict__ cur, int *__restrict__ iters,
                     int *__restrict__ mode, int *__restrict__ state, int *__restrict__ n_active)
{
	const int pr = blockIdx.x, lane = threadIdx.x;
	const int kk = ctl[pr];
	if (kk <= 0) return;
	__shared__ float errs[TV_K];
#pragma unroll
	for (int j = 0; j < TV_K; j++) {
		if (j >= kk) continue;
		const double *pj = part + ((size_t)pr * K + j) * ntiles;
		double v[4];
#pragma unroll
		for (int i = 0; i < 4; i++) {
			double e = 0.0;
			for (int t = lane + 64 * i; t < ntiles; t += 256) e += pj[t];
			v[i] = e;
		}
		v[0] = v[0] + v[2];   // s = 128
		v[1] = v[1] + v[3];
		double x = v[0] + v[1];   // s = 64
#pragma unroll
		for (int st = 32; st > 0; st >>= 1) {
			const double o = __shfl_down(x, st, 64);
			if (lane < st) x = x + o;
		}
		if (lane == 0) errs[j] = (float)x / (float)size;
	}
	if (lane == 0) tv_round_decide(pr, kk, errs, eps2, K, ctl, cur, iters, mode, state, n_active);
}

// Per pair: run state of the solve of one warp.
// The state stays in the buffer the last round wrote (cur) from one warp to
// the next; reset_cur at the start of a scale, where it is in the first.
__global__ void tv_round_init_kernel(int P, int K, int *__restrict__ ctl, int *__restrict__ cur,
                                     int *__restrict__ iters, int *__restrict__ mode, int *__restrict__ state,
                                     int *__restrict__ n_active, int reset_cur)
{
	const int pr = blockIdx.x * blockDim.x + threadIdx.x;
	if (pr == 0) *n_active = P;
	if (pr >= P) return;
	ctl[pr] = min(K, TV_MAX_ITERATIONS);
	if (reset_cur) cur[pr] = 0;
	iters[pr] = 0;
	mode[pr] = TV_RUN;
	state[pr] = -1;
}

// End of a scale: pairs whose flow is in the second buffer copy it to the
// first (the pyramid level the zoom and the output read).
__global__ void tv_settle_kernel(TVPlanes A, TVPlanes B, int n, int stride, const int *__restrict__ cur)
{
	const int pr = blockIdx.y;
	if (!cur[pr]) return;
	const size_t off = (size_t)pr * stride;
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
		A.u1[off + p] = B.u1[off + p];
		A.u2[off + p] = B.u2[off + p];
	}
}

__global__ void tv_fill_kernel(float *__restrict__ a, size_t n, float v)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		a[i] = v;
}

__global__ void tv_fill_int_kernel(int *__restrict__ a, int n, int v)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

// Writes pair results into the sequence flow layout [F][2][H][W]:
// dst frame `frame` gets (u, v) of pair pr.
__global__ void tv_store_kernel(const float *__restrict__ u1, const float *__restrict__ u2, int size,
                                const int *__restrict__ dst_frame, float *__restrict__ flow)
{
	const int pr = blockIdx.y;
	const int f = dst_frame[pr];
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < size; p += gridDim.x * blockDim.x) {
		flow[(size_t)f * 2 * size + p] = u1[(size_t)pr * size + p];
		flow[(size_t)f * 2 * size + size + p] = u2[(size_t)pr * size + p];
	}
}

__global__ void tv_copy_frame_kernel(float *__restrict__ flow, int size, int from, int to)
{
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < 2 * size; p += gridDim.x * blockDim.x)
		flow[(size_t)to * 2 * size + p] = flow[(size_t)from * 2 * size + p];
}

} // namespace bm4dof

// ====================================================================== host

using namespace bm4dof;

// Scratch of bm4dof_tvl1_noalloc: one fixed region per GPU, part of the
// loaded code object, held until the process ends.  Host code never uses
// this name as an address: tv_scratch_init records, per device, the address
// hipGetSymbolAddress gives with that device current, and the entry point
// passes that address to the kernels.
static __device__ __align__(256) unsigned char scratch[BM4DOF_TVL1_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
} // namespace

// Called by bm4dof_ctx_create with `device` current.
void bm4dof::tv_scratch_init(int device)
{
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (device < 0 || device >= kMaxDevices) throw ArgError("device index above the no-alloc scratch table");
	if (g_scratch_ptr[device]) return;
	void *ptr = nullptr;
	HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
	g_scratch_ptr[device] = ptr;
}

namespace {

inline int tv_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + TV_THREADS - 1) / TV_THREADS, 4096)); }
inline int nblk_tv(int64_t n) { return (int)((n + TV_THREADS - 1) / TV_THREADS); }

// Normalised 1-D Gaussian taps (mask.c:227-249), computed on the host with
// the same libm calls as the reference.
int gauss_taps(double sigma, std::vector<double> &B)
{
	const double den = 2 * sigma * sigma;
	const int size = (int)(5 * sigma) + 1;
	B.assign(size, 0.0);
	for (int i = 0; i < size; i++) B[i] = 1 / (sigma * sqrt(2.0 * 3.1415926)) * exp(-i * i / den);
	double norm = 0;
	for (int i = 0; i < size; i++) norm += B[i];
	norm *= 2;
	norm -= B[0];
	for (int i = 0; i < size; i++) B[i] /= norm;
	return size;
}

struct TVPlan {
	int P = 0;                       // number of pairs
	int nscales = 0;
	std::vector<int> nx, ny;         // per scale
	std::vector<size_t> off;         // per-scale offset (floats) in a pyramid of P images
	size_t pyr = 0;                  // floats per pyramid
};

TVPlan tv_plan(int P, int nxx, int nyy, int nscales, float zfactor)
{
	TVPlan pl;
	pl.P = P;
	pl.nscales = nscales;
	pl.nx.resize(nscales);
	pl.ny.resize(nscales);
	pl.off.resize(nscales);
	pl.nx[0] = nxx;
	pl.ny[0] = nyy;
	for (int s = 1; s < nscales; s++) { // zoom_size, zoom.c:24-36
		pl.nx[s] = (int)((float)pl.nx[s - 1] * zfactor + 0.5);
		pl.ny[s] = (int)((float)pl.ny[s - 1] * zfactor + 0.5);
	}
	size_t o = 0;
	for (int s = 0; s < nscales; s++) {
		pl.off[s] = o;
		o += (size_t)P * pl.nx[s] * pl.ny[s];
	}
	pl.pyr = o;
	return pl;
}

// Gaussian smoothing of P images (in -> out, via tmp).
void tv_gauss(bm4dof_ctx *ctx, double *dB, const float *in, float *tmp, float *out, int nx, int ny, int P,
              double sigma)
{
	std::vector<double> B;
	const int size = gauss_taps(sigma, B);
	if (size > nx) throw ArgError("GaussianSmooth: sigma too large");
	if (size >= ny) throw ArgError("GaussianSmooth: sigma too large for the image height");
	if (size > 64) throw ArgError("GaussianSmooth: sigma too large");
	HIPCHK(hipMemcpyAsync(dB, B.data(), size * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
	const int n = nx * ny;
	hipLaunchKernelGGL(tv_gauss_kernel, dim3(tv_grid(n), P), dim3(TV_THREADS), 0, ctx->stream, in, tmp,
	                   nx, ny, n, dB, size, 1);
	hipLaunchKernelGGL(tv_gauss_kernel, dim3(tv_grid(n), P), dim3(TV_THREADS), 0, ctx->stream, tmp, out,
	                   nx, ny, n, dB, size, 0);
	HIPCHK(hipGetLastError());
	// the taps buffer is reused by the next call: keep ordering on the stream
	HIPCHK(hipStreamSynchronize(ctx->stream));
}

// Device memory of one multiscale solve of P pairs: pair indices, the four
// pyramids (I0, I1, u1, u2), 16 working planes at the finest size (12 of
// them the scheme's arrays, the last four with the two Gaussian temporaries
// the second state buffer of the iterations), min/max partials, the
// iterations' per-tile partial sums, the per-pair round control (ctl, cur,
// iters, mode, state, then the active count) and Gaussian taps.
struct TVScratch {
	int *pairs = nullptr;
	float *pyr = nullptr, *work = nullptr, *mm = nullptr;
	double *part = nullptr, *taps = nullptr;
	int *state = nullptr;
};
constexpr int TV_MMB = 64;      // blocks of the min/max reduction per pair
constexpr int TV_WORK = 16;     // working planes at the finest size

inline int tv_tiles(int nx, int ny) { return ((nx + TV_TX - 1) / TV_TX) * ((ny + TV_TY - 1) / TV_TY); }

// Scales whose 64 x 32 tiles give fewer than TV_SMALL_BLOCKS blocks over all
// pairs run 32 x 16 tiles in blocks of three waves: more blocks to spread over
// the CUs where the launch is too small to fill them.  A pixel's arithmetic
// does not depend on the tile it is in.
#define TV_STX 32
#define TV_STY 16
#define TV_SNT 192
#define TV_SMALL_BLOCKS 1024
inline bool tv_small(int nx, int ny, int P) { return (int64_t)tv_tiles(nx, ny) * P < TV_SMALL_BLOCKS; }
inline int tv_tiles_x(int nx, int ny, int P) { return (nx + (tv_small(nx, ny, P) ? TV_STX : TV_TX) - 1) / (tv_small(nx, ny, P) ? TV_STX : TV_TX); }
inline int tv_tiles_at(int nx, int ny, int P)
{
	return tv_small(nx, ny, P) ? ((nx + TV_STX - 1) / TV_STX) * ((ny + TV_STY - 1) / TV_STY) : tv_tiles(nx, ny);
}
// most tiles of any scale: the per-iteration partial sums per pair
inline int tv_part_tiles(const TVPlan &pl)
{
	int m = 0;
	for (int s = 0; s < pl.nscales; s++) m = std::max(m, tv_tiles_at(pl.nx[s], pl.ny[s], pl.P));
	return m;
}

// Byte offsets of the parts of a TVScratch carved from one region, and the
// region's size (the last entry), each part 256-byte aligned.
std::array<size_t, 8> tv_scratch_layout(const TVPlan &pl, size_t size0)
{
	auto al = [](size_t b) { return (b + 255) / 256 * 256; };
	const size_t P = (size_t)pl.P;
	std::array<size_t, 8> o{};
	o[0] = 0;                                                          // pairs
	o[1] = o[0] + al(2 * P * sizeof(int));                             // pyramids
	o[2] = o[1] + al(pl.pyr * 4 * sizeof(float));                      // working planes
	o[3] = o[2] + al(P * size0 * TV_WORK * sizeof(float));             // min/max
	o[4] = o[3] + al(P * TV_MMB * 4 * sizeof(float));                  // iteration partial sums
	o[5] = o[4] + al(P * TV_K * (size_t)tv_part_tiles(pl) * sizeof(double));
	o[6] = o[5] + al((5 * P + 1) * sizeof(int));                       // taps
	o[7] = o[6] + al(64 * sizeof(double));                             // total
	return o;
}

TVScratch tv_scratch_carve(char *base, const std::array<size_t, 8> &o)
{
	TVScratch t;
	t.pairs = (int *)(base + o[0]);
	t.pyr = (float *)(base + o[1]);
	t.work = (float *)(base + o[2]);
	t.mm = (float *)(base + o[3]);
	t.part = (double *)(base + o[4]);
	t.state = (int *)(base + o[5]);
	t.taps = (double *)(base + o[6]);
	return t;
}

TVScratch tv_scratch_ctx(bm4dof_ctx *ctx, const TVPlan &pl, size_t size0)
{
	const size_t P = (size_t)pl.P;
	TVScratch t;
	t.pairs = ctx->tv_pairs.as<int>(2 * P);
	t.pyr = ctx->tv_pyr.as<float>(pl.pyr * 4);
	t.work = ctx->tv_work.as<float>(P * size0 * TV_WORK);
	t.mm = ctx->tv_mm.as<float>(P * TV_MMB * 4);
	t.part = ctx->tv_part.as<double>(P * TV_K * (size_t)tv_part_tiles(pl));
	t.state = ctx->tv_state.as<int>(5 * P + 1);
	t.taps = ctx->tv_taps.as<double>(64);
	return t;
}

void tv_check_params(const bm4dof_tvl1_params &prm)
{
	if (prm.nscales < 1 || prm.warps < 0 || prm.fscale < 0 || !(prm.zfactor > 0 && prm.zfactor < 1))
		throw ArgError("invalid TV-L1 parameters");
}

// Multiscale TV-L1 of P pairs.  frames: device, nfr frames of nxx*nyy;
// pair pr uses frames i0[pr] -> i1[pr].  Result in u1/u2 (P*nxx*nyy each,
// device).  iters (host, may be NULL): P*nscales*warps ints.  Device
// scratch from `fixed` when given (nothing is allocated then), else from the
// context's buffers.
void tv_flow_pairs(bm4dof_ctx *ctx, const bm4dof_tvl1_params &prm, const float *frames, int nxx,
                   int nyy, const std::vector<int> &i0, const std::vector<int> &i1, float *u1out,
                   float *u2out, int *iters, const TVScratch *fixed = nullptr)
{
	const int P = (int)i0.size();
	if (P == 0) return;
	tv_check_params(prm);
	const int nscales = prm.nscales;
	TVPlan pl = tv_plan(P, nxx, nyy, nscales, prm.zfactor);
	ctx->tv_pixel_iters = 0;
	for (int s = 0; s < nscales; s++)
		if (pl.nx[s] < 3 || pl.ny[s] < 3)
			throw ArgError("TV-L1 pyramid level " + std::to_string(s) + " is smaller than 3x3; lower nscales");
	const size_t size0 = (size_t)nxx * nyy;
	hipStream_t st = ctx->stream;

	const TVScratch sc = fixed ? *fixed : tv_scratch_ctx(ctx, pl, size0);
	int *d_pairs = sc.pairs;
	HIPCHK(hipMemcpyAsync(d_pairs, i0.data(), P * sizeof(int), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(d_pairs + P, i1.data(), P * sizeof(int), hipMemcpyHostToDevice, st));
	// pyramids: I0s, I1s, u1s, u2s
	float *pyr = sc.pyr;
	float *I0s = pyr, *I1s = pyr + pl.pyr, *U1 = pyr + 2 * pl.pyr, *U2 = pyr + 3 * pl.pyr;
	// working arrays at the finest size: tmp, tmp2, I1x, I1y, I1wx, I1wy, grad, rho_c, p11..p22,
	// and four planes that with tmp and tmp2 hold the iterations' second state buffer
	const size_t W = (size_t)P * size0;
	float *work = sc.work;
	float *tmp = work, *tmp2 = work + W, *I1x = work + 2 * W, *I1y = work + 3 * W,
	      *I1wx = work + 4 * W, *I1wy = work + 5 * W, *grad = work + 6 * W, *rho_c = work + 7 * W,
	      *p11 = work + 8 * W, *p12 = work + 9 * W, *p21 = work + 10 * W, *p22 = work + 11 * W;
	const int mmb = TV_MMB;
	float *d_mm = sc.mm;
	double *d_part = sc.part;
	int *d_ctl = sc.state, *d_cur = d_ctl + P, *d_iters = d_cur + P, *d_mode = d_iters + P, *d_state = d_mode + P;
	int *d_nact = d_state + P;

	// normalisation + presmoothing (tvl1flow_lib.c:381-386)
	hipLaunchKernelGGL(tv_minmax_kernel, dim3(mmb, P), dim3(TV_THREADS), 0, st, frames, (int)size0,
	                   d_pairs, d_pairs + P, d_mm);
	hipLaunchKernelGGL(tv_normalize_kernel, dim3(tv_grid(size0), P), dim3(TV_THREADS), 0, st, frames,
	                   (int)size0, d_pairs, d_pairs + P, d_mm, mmb, tmp2, tmp2 + W);
	HIPCHK(hipGetLastError());
	tv_gauss(ctx, sc.taps, tmp2, tmp, I0s + pl.off[0], nxx, nyy, P, TV_PRESMOOTHING_SIGMA);
	tv_gauss(ctx, sc.taps, tmp2 + W, tmp, I1s + pl.off[0], nxx, nyy, P, TV_PRESMOOTHING_SIGMA);
	// pyramid (tvl1flow_lib.c:389-403; zoom_out, zoom.c:43-79)
	const float zsig = TV_ZOOM_SIGMA_ZERO * sqrt(1.0 / (prm.zfactor * prm.zfactor) - 1.0);
	for (int s = 1; s < nscales; s++) {
		const int nx = pl.nx[s - 1], ny = pl.ny[s - 1], n = nx * ny;
		for (int im = 0; im < 2; im++) {
			float *src = (im ? I1s : I0s) + pl.off[s - 1];
			float *dst = (im ? I1s : I0s) + pl.off[s];
			tv_gauss(ctx, sc.taps, src, tmp, tmp2, nx, ny, P, zsig);
			hipLaunchKernelGGL(tv_zoom_out_kernel, dim3(tv_grid(pl.nx[s] * pl.ny[s]), P), dim3(TV_THREADS), 0,
			                   st, tmp2, nx, ny, n, dst, pl.nx[s], pl.ny[s], pl.nx[s] * pl.ny[s], prm.zfactor);
			HIPCHK(hipGetLastError());
		}
	}
	{
		const int s = nscales - 1;
		const size_t n = (size_t)P * pl.nx[s] * pl.ny[s];
		hipLaunchKernelGGL(tv_fill_kernel, dim3(tv_grid(n)), dim3(TV_THREADS), 0, st, U1 + pl.off[s], n, 0.f);
		hipLaunchKernelGGL(tv_fill_kernel, dim3(tv_grid(n)), dim3(TV_THREADS), 0, st, U2 + pl.off[s], n, 0.f);
	}
	if (iters) std::fill(iters, iters + (size_t)P * nscales * prm.warps, -1);
	const float l_t = prm.lambda * prm.theta;
	const float taut = prm.tau / prm.theta;
	const float eps2 = prm.epsilon * prm.epsilon;
	const float scale = (float)1.0 / prm.zfactor;
	std::vector<int> h_state(P);
	// Rounds of TV_K iterations run until the pair's test stops them; the
	// host reads the active count after a batch of rounds: the first batch
	// long enough for the previous warp's iteration count, then two rounds
	// at a time (a round of a finished pair is a launch whose blocks exit).
	// With the pinned words, a poll copies the count and the GPU runs one
	// more round while the host waits for the copy, and a warp's iteration
	// counts are copied at its end and read at the next poll.
	const int max_rounds = (TV_MAX_ITERATIONS + TV_K - 1) / TV_K + 1;
	int guess = 0;   // most iterations of a pair in the last warp read back
	const bool piped = ctx->tv_pin && P <= ctx->tv_pin_pairs;
	int pend_s = -1, pend_w = -1, pend_slot = 0, nwarp = 0;
	size_t pend_n = 0;
	auto take_counts = [&](const int *hs, int s, int w, size_t n) {
		guess = 0;
		for (int pr = 0; pr < P; pr++) {
			ctx->tv_pixel_iters += (uint64_t)std::max(hs[pr], 0) * (uint64_t)n;
			guess = std::max(guess, hs[pr]);
		}
		if (iters)
			for (int pr = 0; pr < P; pr++) iters[((size_t)pr * nscales + s) * prm.warps + w] = hs[pr];
		if (prm.verbose)
			for (int pr = 0; pr < P; pr++)
				fprintf(stderr, "Pair %d scale %d warping %d: iterations %d\n", pr, s, w, hs[pr]);
	};
	// the counts of a warp whose copy a completed poll has passed
	auto take_pending = [&]() {
		if (pend_w < 0) return;
		take_counts(ctx->tv_pin + 1 + (size_t)pend_slot * P, pend_s, pend_w, pend_n);
		pend_w = -1;
	};

	for (int s = nscales - 1; s >= prm.fscale; s--) {
		// Dual_TVL1_optic_flow, tvl1flow_lib.c:93-275
		const int nx = pl.nx[s], ny = pl.ny[s], n = nx * ny;
		float *u1 = U1 + pl.off[s], *u2 = U2 + pl.off[s];
		const float *I0 = I0s + pl.off[s], *I1 = I1s + pl.off[s];
		const int g = tv_grid(n);
		const TVPlanes SA = {u1, u2, p11, p12, p21, p22};
		const TVPlanes SB = {tmp, tmp2, work + 12 * W, work + 13 * W, work + 14 * W, work + 15 * W};
		const bool small = tv_small(nx, ny, P);
		const int tiles_x = tv_tiles_x(nx, ny, P), ntiles = tv_tiles_at(nx, ny, P);
		auto launch_round = [&]() {
			if (small)
				hipLaunchKernelGGL((tv_iter_kernel<TV_STX, TV_STY, TV_K, TV_SNT>), dim3(ntiles, P), dim3(TV_SNT), 0, st,
				                   I1wx, I1wy, grad, rho_c, SA, SB, nx, ny, n, tiles_x, l_t, prm.theta, taut, d_ctl,
				                   d_cur, d_part);
			else
				hipLaunchKernelGGL((tv_iter_kernel<TV_TX, TV_TY, TV_K, TV_NT>), dim3(ntiles, P), dim3(TV_NT), 0, st,
				                   I1wx, I1wy, grad, rho_c, SA, SB, nx, ny, n, tiles_x, l_t, prm.theta, taut, d_ctl,
				                   d_cur, d_part);
			hipLaunchKernelGGL(tv_round_kernel, dim3(P), dim3(64), 0, st, d_part, TV_K, ntiles, n, eps2,
			                   d_ctl, d_cur, d_iters, d_mode, d_state, d_nact);
		};
		hipLaunchKernelGGL(tv_cgrad_kernel, dim3(g, P), dim3(TV_THREADS), 0, st, I1, nx, ny, n, I1x, I1y);
		const size_t nP = (size_t)P * n;
		for (float *q : {p11, p12, p21, p22})
			hipLaunchKernelGGL(tv_fill_kernel, dim3(tv_grid(nP)), dim3(TV_THREADS), 0, st, q, nP, 0.f);
		for (int w = 0; w < prm.warps; w++, nwarp++) {
			hipLaunchKernelGGL(tv_round_init_kernel, dim3(nblk_tv(P)), dim3(TV_THREADS), 0, st, P, TV_K, d_ctl,
			                   d_cur, d_iters, d_mode, d_state, d_nact, w == 0 ? 1 : 0);
			hipLaunchKernelGGL(tv_warp_kernel, dim3(g, P), dim3(TV_THREADS), 0, st, I0, I1, I1x, I1y, SA.u1, SA.u2,
			                   SB.u1, SB.u2, d_cur, nx, ny, n, I1wx, I1wy, grad, rho_c);
			HIPCHK(hipGetLastError());
			int batch = guess > 0 ? (guess + TV_K - 1) / TV_K : 3;
			for (int r = 0; r < max_rounds;) {
				const int rb = std::min(batch, max_rounds - r);
				for (int i = 0; i < rb; i++) launch_round();
				r += rb;
				HIPCHK(hipGetLastError());
				int nact = 0;
				if (piped) {
					HIPCHK(hipMemcpyAsync(ctx->tv_pin, d_nact, sizeof(int), hipMemcpyDeviceToHost, st));
					HIPCHK(hipEventRecord(ctx->ev_tvpoll, st));
					if (r < max_rounds) {   // runs while the host waits; a no-op once every pair stopped
						launch_round();
						r++;
					}
					HIPCHK(hipGetLastError());
					HIPCHK(hipEventSynchronize(ctx->ev_tvpoll));
					nact = ctx->tv_pin[0];
					take_pending();
				} else {
					HIPCHK(hipMemcpyAsync(&nact, d_nact, sizeof(int), hipMemcpyDeviceToHost, st));
					HIPCHK(hipStreamSynchronize(st));
				}
				if (nact == 0) break;
				batch = 2;
			}
			if (piped) {
				pend_slot = nwarp & 1;
				HIPCHK(hipMemcpyAsync(ctx->tv_pin + 1 + (size_t)pend_slot * P, d_state, P * sizeof(int),
				                      hipMemcpyDeviceToHost, st));
				pend_s = s;
				pend_w = w;
				pend_n = (size_t)n;
			} else {
				HIPCHK(hipMemcpyAsync(h_state.data(), d_state, P * sizeof(int), hipMemcpyDeviceToHost, st));
				HIPCHK(hipStreamSynchronize(st));
				take_counts(h_state.data(), s, w, (size_t)n);
			}
		}
		// the flow of every pair into the pyramid level
		if (prm.warps > 0)
			hipLaunchKernelGGL(tv_settle_kernel, dim3(g, P), dim3(TV_THREADS), 0, st, SA, SB, n, n, d_cur);
		HIPCHK(hipGetLastError());
		if (!s) break;
		hipLaunchKernelGGL(tv_zoom_in_kernel, dim3(tv_grid(pl.nx[s - 1] * pl.ny[s - 1]), P), dim3(TV_THREADS), 0, st,
		                   u1, nx, ny, n, U1 + pl.off[s - 1], pl.nx[s - 1], pl.ny[s - 1], pl.nx[s - 1] * pl.ny[s - 1], scale);
		hipLaunchKernelGGL(tv_zoom_in_kernel, dim3(tv_grid(pl.nx[s - 1] * pl.ny[s - 1]), P), dim3(TV_THREADS), 0, st,
		                   u2, nx, ny, n, U2 + pl.off[s - 1], pl.nx[s - 1], pl.ny[s - 1], pl.nx[s - 1] * pl.ny[s - 1], scale);
		HIPCHK(hipGetLastError());
	}
	for (int s = prm.fscale - 1; s >= 0; s--) { // tvl1flow_lib.c:438-455
		if (!s) break;
		const int n = pl.nx[s] * pl.ny[s], nn = pl.nx[s - 1] * pl.ny[s - 1];
		hipLaunchKernelGGL(tv_zoom_in_kernel, dim3(tv_grid(nn), P), dim3(TV_THREADS), 0, st, U1 + pl.off[s],
		                   pl.nx[s], pl.ny[s], n, U1 + pl.off[s - 1], pl.nx[s - 1], pl.ny[s - 1], nn, scale);
		hipLaunchKernelGGL(tv_zoom_in_kernel, dim3(tv_grid(nn), P), dim3(TV_THREADS), 0, st, U2 + pl.off[s],
		                   pl.nx[s], pl.ny[s], n, U2 + pl.off[s - 1], pl.nx[s - 1], pl.ny[s - 1], nn, scale);
		HIPCHK(hipGetLastError());
	}
	HIPCHK(hipMemcpyAsync(u1out, U1 + pl.off[0], (size_t)P * size0 * sizeof(float), hipMemcpyDeviceToDevice, st));
	HIPCHK(hipMemcpyAsync(u2out, U2 + pl.off[0], (size_t)P * size0 * sizeof(float), hipMemcpyDeviceToDevice, st));
	HIPCHK(hipStreamSynchronize(st));
	take_pending();
}

bm4dof_ctx *g_tv_ctx = nullptr;

} // namespace

extern "C" {

__attribute__((visibility("default")))
int bm4dof_tvl1(bm4dof_ctx *ctx, const bm4dof_tvl1_params *prm, const float *I0, const float *I1,
                float *u1, float *u2, int nx, int ny, int on_device, int *iters)
{
	if (!ctx || !prm || !I0 || !I1 || !u1 || !u2) { set_error(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		HIPCHK(hipSetDevice(ctx->device));
		if (nx < 3 || ny < 3) throw ArgError("image smaller than 3x3");
		const size_t n = (size_t)nx * ny;
		float *fr = ctx->tv_frames.as<float>(4 * n);
		float *du1 = fr + 2 * n, *du2 = fr + 3 * n;
		const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
		HIPCHK(hipMemcpyAsync(fr, I0, n * 4, k, ctx->stream));
		HIPCHK(hipMemcpyAsync(fr + n, I1, n * 4, k, ctx->stream));
		tv_flow_pairs(ctx, *prm, fr, nx, ny, {0}, {1}, du1, du2, iters);
		const hipMemcpyKind kb = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
		HIPCHK(hipMemcpyAsync(u1, du1, n * 4, kb, ctx->stream));
		HIPCHK(hipMemcpyAsync(u2, du2, n * 4, kb, ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
	});
}

__attribute__((visibility("default")))
int bm4dof_tvl1_noalloc(bm4dof_ctx *ctx, const bm4dof_tvl1_params *prm, const float *I0, const float *I1,
                        float *u1, float *u2, int nx, int ny, int *iters)
{
	if (!ctx || !prm || !I0 || !I1 || !u1 || !u2) { set_error(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		HIPCHK(hipSetDevice(ctx->device));
		if (nx < 3 || ny < 3) throw ArgError("image smaller than 3x3");
		tv_check_params(*prm);
		void *base = nullptr;
		{
			std::lock_guard<std::mutex> lk(g_scratch_mu);
			if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch_ptr[ctx->device];
		}
		if (!base) throw ArgError("bm4dof_tvl1_noalloc: no scratch address recorded for this device");
		const size_t n = (size_t)nx * ny;
		const TVPlan pl = tv_plan(1, nx, ny, prm->nscales, prm->zfactor);
		const std::array<size_t, 8> lay = tv_scratch_layout(pl, n);
		const size_t need = lay[7] + (2 * n * sizeof(float) + 255) / 256 * 256;
		if (need > (size_t)BM4DOF_TVL1_NOALLOC_BYTES)
			throw ArgError("bm4dof_tvl1_noalloc: a " + std::to_string(nx) + "x" + std::to_string(ny) +
			               " pair with these parameters needs " + std::to_string(need) +
			               " bytes of scratch; the fixed scratch holds " +
			               std::to_string((size_t)BM4DOF_TVL1_NOALLOC_BYTES));
		char *b = (char *)base;
		const TVScratch t = tv_scratch_carve(b, lay);
		float *fr = (float *)(b + lay[7]);
		HIPCHK(hipMemcpyAsync(fr, I0, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
		HIPCHK(hipMemcpyAsync(fr + n, I1, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
		tv_flow_pairs(ctx, *prm, fr, nx, ny, {0}, {1}, u1, u2, iters, &t);
	});
}

__attribute__((visibility("default")))
int bm4dof_tvl1_sequence(bm4dof_ctx *ctx, const bm4dof_tvl1_params *prm, const float *frames, int nx,
                         int ny, int F, float *fflow, float *bflow, int on_device)
{
	if (!ctx || !prm || !frames || !fflow || !bflow) { set_error(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		HIPCHK(hipSetDevice(ctx->device));
		if (F < 1 || nx < 3 || ny < 3) throw ArgError("empty sequence or frames smaller than 3x3");
		const size_t n = (size_t)nx * ny;
		float *fr = ctx->tv_frames.as<float>(n * F + (size_t)4 * n * F);
		float *ff = fr + n * F, *bf = ff + 2 * n * F;
		hipStream_t st = ctx->stream;
		HIPCHK(hipMemcpyAsync(fr, frames, n * F * 4, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
		HIPCHK(hipMemsetAsync(ff, 0, 4 * n * F * 4, st));
		if (F > 1) {
			// scripts/tvl1flow-seq.sh:31-61: bwd(i) = I_i -> I_{i-1}, fwd(i) = I_i -> I_{i+1}
			std::vector<int> i0, i1, dst;
			for (int i = 1; i < F; i++) { i0.push_back(i); i1.push_back(i - 1); dst.push_back(i); }
			for (int i = 0; i < F - 1; i++) { i0.push_back(i); i1.push_back(i + 1); dst.push_back(i); }
			const int P = (int)i0.size();
			DevBuf ub;
			float *u = ub.as<float>((size_t)2 * P * n);
			tv_flow_pairs(ctx, *prm, fr, nx, ny, i0, i1, u, u + (size_t)P * n, nullptr);
			int *d_dst = ctx->tv_pairs.as<int>(P);
			HIPCHK(hipMemcpyAsync(d_dst, dst.data(), P * sizeof(int), hipMemcpyHostToDevice, st));
			const int half = F - 1;
			hipLaunchKernelGGL(tv_store_kernel, dim3(tv_grid(n), half), dim3(TV_THREADS), 0, st, u,
			                   u + (size_t)P * n, (int)n, d_dst, bf);
			hipLaunchKernelGGL(tv_store_kernel, dim3(tv_grid(n), half), dim3(TV_THREADS), 0, st, u + half * n,
			                   u + (size_t)P * n + half * n, (int)n, d_dst + half, ff);
			hipLaunchKernelGGL(tv_copy_frame_kernel, dim3(tv_grid(2 * n)), dim3(TV_THREADS), 0, st, bf, (int)n, 1, 0);
			hipLaunchKernelGGL(tv_copy_frame_kernel, dim3(tv_grid(2 * n)), dim3(TV_THREADS), 0, st, ff, (int)n, F - 2, F - 1);
			HIPCHK(hipGetLastError());
			HIPCHK(hipStreamSynchronize(st));
			ub.release();
		}
		const hipMemcpyKind kb = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
		HIPCHK(hipMemcpyAsync(fflow, ff, 2 * n * F * 4, kb, st));
		HIPCHK(hipMemcpyAsync(bflow, bf, 2 * n * F * 4, kb, st));
		HIPCHK(hipStreamSynchronize(st));
	});
}

__attribute__((visibility("default")))
void Dual_TVL1_optic_flow_multiscale(float *I0, float *I1, float *u1, float *u2, int nxx, int nyy,
                                     float tau, float lambda, float theta, int nscales, int fscale,
                                     float zfactor, int warps, float epsilon, bool verbose)
{
	if (!g_tv_ctx && bm4dof_ctx_create(&g_tv_ctx, 0) != BM4DOF_OK) {
		fprintf(stderr, "Dual_TVL1_optic_flow_multiscale: %s\n", bm4dof_global_error());
		abort();
	}
	bm4dof_tvl1_params prm = {tau, lambda, theta, nscales, fscale, zfactor, warps, epsilon, verbose ? 1 : 0};
	if (bm4dof_tvl1(g_tv_ctx, &prm, I0, I1, u1, u2, nxx, nyy, 0, nullptr) != BM4DOF_OK) {
		fprintf(stderr, "Dual_TVL1_optic_flow_multiscale: %s\n", bm4dof_last_error(g_tv_ctx));
		abort();
	}
}

} // extern "C"
