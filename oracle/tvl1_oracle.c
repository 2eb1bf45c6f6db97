/*
 * tvl1_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's Zach-Pock-Bischof dual TV-L1 optical
 * flow (lib/tvl1flow/{tvl1flow_lib,mask,bicubic_interpolation,zoom}.c), used
 * as the checker for the GPU flow kernels and pinned against the reference's
 * own compiled library (oracle/_ref).  Single-threaded: the reference's
 * per-iteration error is an OpenMP float reduction whose value depends on
 * the thread count (tvl1flow_lib.c:213); one thread makes it a plain
 * sequential sum.  Build with -ffp-contract=off.
 *
 * It also reports the iteration count of every (scale, warp) solve.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define OR_EXPORT __attribute__((visibility("default")))

#define TV_MAX_ITERATIONS 300      /* tvl1flow_lib.c:24 */
#define TV_PRESMOOTHING_SIGMA 0.8  /* tvl1flow_lib.c:25 */
#define TV_GRAD_IS_ZERO 1E-10      /* tvl1flow_lib.c:26 */
#define TV_ZOOM_SIGMA_ZERO 0.6     /* zoom.c:17 */

/* divergence with backward differences, mask.c:43-92 */
OR_EXPORT void or_tv_divergence(const float *v1, const float *v2, float *div, int nx, int ny)
{
	for (int i = 1; i < ny - 1; i++)
		for (int j = 1; j < nx - 1; j++) {
			const int p = i * nx + j;
			div[p] = (v1[p] - v1[p - 1]) + (v2[p] - v2[p - nx]);
		}
	for (int j = 1; j < nx - 1; j++) {
		const int p = (ny - 1) * nx + j;
		div[j] = v1[j] - v1[j - 1] + v2[j];
		div[p] = v1[p] - v1[p - 1] - v2[p - nx];
	}
	for (int i = 1; i < ny - 1; i++) {
		const int p1 = i * nx, p2 = (i + 1) * nx - 1;
		div[p1] = v1[p1] + v2[p1] - v2[p1 - nx];
		div[p2] = -v1[p2 - 1] + v2[p2] - v2[p2 - nx];
	}
	div[0] = v1[0] + v2[0];
	div[nx - 1] = -v1[nx - 2] + v2[nx - 1];
	div[(ny - 1) * nx] = v1[(ny - 1) * nx] - v2[(ny - 2) * nx];
	div[ny * nx - 1] = -v1[ny * nx - 2] - v2[(ny - 1) * nx - 1];
}

/* forward-difference gradient, mask.c:101-144 */
OR_EXPORT void or_tv_forward_gradient(const float *f, float *fx, float *fy, int nx, int ny)
{
	for (int i = 0; i < ny - 1; i++)
		for (int j = 0; j < nx - 1; j++) {
			const int p = i * nx + j;
			fx[p] = f[p + 1] - f[p];
			fy[p] = f[p + nx] - f[p];
		}
	for (int j = 0; j < nx - 1; j++) {
		const int p = (ny - 1) * nx + j;
		fx[p] = f[p + 1] - f[p];
		fy[p] = 0;
	}
	for (int i = 1; i < ny; i++) {
		const int p = i * nx - 1;
		fx[p] = 0;
		fy[p] = f[p + nx] - f[p];
	}
	fx[ny * nx - 1] = 0;
	fy[ny * nx - 1] = 0;
}

/* centred-difference gradient, mask.c:152-209 (0.5*(...) in double) */
OR_EXPORT void or_tv_centered_gradient(const float *in, float *dx, float *dy, int nx, int ny)
{
	for (int i = 1; i < ny - 1; i++)
		for (int j = 1; j < nx - 1; j++) {
			const int k = i * nx + j;
			dx[k] = 0.5 * (in[k + 1] - in[k - 1]);
			dy[k] = 0.5 * (in[k + nx] - in[k - nx]);
		}
	for (int j = 1; j < nx - 1; j++) {
		dx[j] = 0.5 * (in[j + 1] - in[j - 1]);
		dy[j] = 0.5 * (in[j + nx] - in[j]);
		const int k = (ny - 1) * nx + j;
		dx[k] = 0.5 * (in[k + 1] - in[k - 1]);
		dy[k] = 0.5 * (in[k] - in[k - nx]);
	}
	for (int i = 1; i < ny - 1; i++) {
		const int p = i * nx;
		dx[p] = 0.5 * (in[p + 1] - in[p]);
		dy[p] = 0.5 * (in[p + nx] - in[p - nx]);
		const int k = (i + 1) * nx - 1;
		dx[k] = 0.5 * (in[k] - in[k - 1]);
		dy[k] = 0.5 * (in[k + nx] - in[k - nx]);
	}
	dx[0] = 0.5 * (in[1] - in[0]);
	dy[0] = 0.5 * (in[nx] - in[0]);
	dx[nx - 1] = 0.5 * (in[nx - 1] - in[nx - 2]);
	dy[nx - 1] = 0.5 * (in[2 * nx - 1] - in[nx - 1]);
	dx[(ny - 1) * nx] = 0.5 * (in[(ny - 1) * nx + 1] - in[(ny - 1) * nx]);
	dy[(ny - 1) * nx] = 0.5 * (in[(ny - 1) * nx] - in[(ny - 2) * nx]);
	dx[ny * nx - 1] = 0.5 * (in[ny * nx - 1] - in[ny * nx - 2]);
	dy[ny * nx - 1] = 0.5 * (in[ny * nx - 1] - in[(ny - 1) * nx - 1]);
}

/* 1-D normalised Gaussian taps, mask.c:227-249.  Returns the tap count. */
OR_EXPORT int or_tv_gauss_taps(double sigma, double *B, int maxn)
{
	const double den = 2 * sigma * sigma;
	const int size = (int)(5 * sigma) + 1;
	if (size > maxn) return -1;
	for (int i = 0; i < size; i++)
		B[i] = 1 / (sigma * sqrt(2.0 * 3.1415926)) * exp(-i * i / den);
	double norm = 0;
	for (int i = 0; i < size; i++) norm += B[i];
	norm *= 2;
	norm -= B[0];
	for (int i = 0; i < size; i++) B[i] /= norm;
	return size;
}

/* In-place separable Gaussian with reflecting boundaries, mask.c:217-333 */
OR_EXPORT void or_tv_gaussian(float *I, int xdim, int ydim, double sigma)
{
	double B[256];
	const int size = or_tv_gauss_taps(sigma, B, 256);
	if (size < 0 || size > xdim) { fprintf(stderr, "GaussianSmooth: sigma too large\n"); abort(); }
	const int bdx = xdim + size, bdy = ydim + size;
	double *R = (double *)malloc((size_t)(size + xdim + size) * sizeof *R);
	for (int k = 0; k < ydim; k++) {
		for (int i = size; i < bdx; i++) R[i] = I[k * xdim + i - size];
		for (int i = 0, j = bdx; i < size; i++, j++) {
			R[i] = I[k * xdim + size - i];
			R[j] = I[k * xdim + xdim - i - 1];
		}
		for (int i = size; i < bdx; i++) {
			double sum = B[0] * R[i];
			for (int j = 1; j < size; j++) sum += B[j] * (R[i - j] + R[i + j]);
			I[k * xdim + i - size] = sum;
		}
	}
	double *T = (double *)malloc((size_t)(size + ydim + size) * sizeof *T);
	for (int k = 0; k < xdim; k++) {
		for (int i = size; i < bdy; i++) T[i] = I[(i - size) * xdim + k];
		for (int i = 0, j = bdy; i < size; i++, j++) {
			T[i] = I[(size - i) * xdim + k];
			T[j] = I[(ydim - i - 1) * xdim + k];
		}
		for (int i = size; i < bdy; i++) {
			double sum = B[0] * T[i];
			for (int j = 1; j < size; j++) sum += B[j] * (T[i - j] + T[i + j]);
			I[(i - size) * xdim + k] = sum;
		}
	}
	free(R); free(T);
}

/* Neumann clamp, bicubic_interpolation.c:25-39 */
static int nbc(int x, int n, int *out)
{
	if (x < 0) { *out = 1; return 0; }
	if (x >= n) { *out = 1; return n - 1; }
	return x;
}

static double cubic_cell(const double v[4], double x) /* bicubic_interpolation.c:102-112 */
{
	return v[1] + 0.5 * x * (v[2] - v[0] +
		x * (2.0 * v[0] - 5.0 * v[1] + 4.0 * v[2] - v[3] +
		x * (3.0 * (v[1] - v[2]) + v[3] - v[0])));
}

/* bicubic_interpolation_at, bicubic_interpolation.c:138-234 (Neumann case;
 * the y-neighbour `my` is offset by the x sign `sx`, as in the reference). */
OR_EXPORT float or_tv_bicubic_at(const float *in, float uu, float vv, int nx, int ny, int border_out)
{
	const int sx = (uu < 0) ? -1 : 1;
	const int sy = (vv < 0) ? -1 : 1;
	int out = 0;
	const int x = nbc((int)uu, nx, &out);
	const int y = nbc((int)vv, ny, &out);
	const int mx = nbc((int)uu - sx, nx, &out);
	const int my = nbc((int)vv - sx, ny, &out);
	const int dx = nbc((int)uu + sx, nx, &out);
	const int dy = nbc((int)vv + sy, ny, &out);
	const int ddx = nbc((int)uu + 2 * sx, nx, &out);
	const int ddy = nbc((int)vv + 2 * sy, ny, &out);
	if (out && border_out) return 0.0;
	const int cols[4] = { mx, x, dx, ddx }, rows[4] = { my, y, dy, ddy };
	double v[4];
	for (int c = 0; c < 4; c++) {
		double col[4];
		for (int r = 0; r < 4; r++) col[r] = in[cols[c] + nx * rows[r]];
		v[c] = cubic_cell(col, vv - y);
	}
	return cubic_cell(v, uu - x);
}

static void warp(const float *in, const float *u, const float *v, float *out, int nx, int ny)
{
	for (int i = 0; i < ny; i++)
		for (int j = 0; j < nx; j++) {
			const int p = i * nx + j;
			const float uu = (float)(j + u[p]);
			const float vv = (float)(i + v[p]);
			out[p] = or_tv_bicubic_at(in, uu, vv, nx, ny, 1);
		}
}

/* Dual_TVL1_optic_flow, tvl1flow_lib.c:93-275 (one scale).  iters receives
 * the iteration count of every warp (may be NULL). */
OR_EXPORT void or_tv_solve(const float *I0, const float *I1, float *u1, float *u2, int nx, int ny,
                           float tau, float lambda, float theta, int warps, float epsilon, int *iters)
{
	const int size = nx * ny;
	const float l_t = lambda * theta;
	float *buf = (float *)calloc((size_t)size * 19, sizeof(float));
	float *I1x = buf, *I1y = buf + size, *I1w = buf + 2 * size, *I1wx = buf + 3 * size,
	      *I1wy = buf + 4 * size, *rho_c = buf + 5 * size, *v1 = buf + 6 * size, *v2 = buf + 7 * size,
	      *p11 = buf + 8 * size, *p12 = buf + 9 * size, *p21 = buf + 10 * size, *p22 = buf + 11 * size,
	      *grad = buf + 12 * size, *div_p1 = buf + 13 * size, *div_p2 = buf + 14 * size,
	      *u1x = buf + 15 * size, *u1y = buf + 16 * size, *u2x = buf + 17 * size, *u2y = buf + 18 * size;
	or_tv_centered_gradient(I1, I1x, I1y, nx, ny);
	for (int w = 0; w < warps; w++) {
		warp(I1, u1, u2, I1w, nx, ny);
		warp(I1x, u1, u2, I1wx, nx, ny);
		warp(I1y, u1, u2, I1wy, nx, ny);
		for (int i = 0; i < size; i++) {
			const float Ix2 = I1wx[i] * I1wx[i];
			const float Iy2 = I1wy[i] * I1wy[i];
			grad[i] = (Ix2 + Iy2);
			rho_c[i] = (I1w[i] - I1wx[i] * u1[i] - I1wy[i] * u2[i] - I0[i]);
		}
		int n = 0;
		float error = INFINITY;
		while (error > epsilon * epsilon && n < TV_MAX_ITERATIONS) {
			n++;
			for (int i = 0; i < size; i++) {
				const float rho = rho_c[i] + (I1wx[i] * u1[i] + I1wy[i] * u2[i]);
				float d1, d2;
				if (rho < -l_t * grad[i]) { d1 = l_t * I1wx[i]; d2 = l_t * I1wy[i]; }
				else if (rho > l_t * grad[i]) { d1 = -l_t * I1wx[i]; d2 = -l_t * I1wy[i]; }
				else if (grad[i] < TV_GRAD_IS_ZERO) d1 = d2 = 0;
				else { float fi = -rho / grad[i]; d1 = fi * I1wx[i]; d2 = fi * I1wy[i]; }
				v1[i] = u1[i] + d1;
				v2[i] = u2[i] + d2;
			}
			or_tv_divergence(p11, p12, div_p1, nx, ny);
			or_tv_divergence(p21, p22, div_p2, nx, ny);
			error = 0.0;
			for (int i = 0; i < size; i++) {
				const float u1k = u1[i], u2k = u2[i];
				u1[i] = v1[i] + theta * div_p1[i];
				u2[i] = v2[i] + theta * div_p2[i];
				error += (u1[i] - u1k) * (u1[i] - u1k) + (u2[i] - u2k) * (u2[i] - u2k);
			}
			error /= size;
			or_tv_forward_gradient(u1, u1x, u1y, nx, ny);
			or_tv_forward_gradient(u2, u2x, u2y, nx, ny);
			for (int i = 0; i < size; i++) {
				const float taut = tau / theta;
				const float g1 = hypot(u1x[i], u1y[i]);
				const float g2 = hypot(u2x[i], u2y[i]);
				const float ng1 = 1.0 + taut * g1;
				const float ng2 = 1.0 + taut * g2;
				p11[i] = (p11[i] + taut * u1x[i]) / ng1;
				p12[i] = (p12[i] + taut * u1y[i]) / ng1;
				p21[i] = (p21[i] + taut * u2x[i]) / ng2;
				p22[i] = (p22[i] + taut * u2y[i]) / ng2;
			}
		}
		if (iters) iters[w] = n;
	}
	free(buf);
}

/* zoom_size, zoom.c:24-36 */
OR_EXPORT void or_tv_zoom_size(int nx, int ny, int *nxx, int *nyy, float factor)
{
	*nxx = (int)((float)nx * factor + 0.5);
	*nyy = (int)((float)ny * factor + 0.5);
}

/* zoom_out, zoom.c:43-79 */
OR_EXPORT void or_tv_zoom_out(const float *I, float *Iout, int nx, int ny, float factor)
{
	float *Is = (float *)malloc((size_t)nx * ny * sizeof(float));
	memcpy(Is, I, (size_t)nx * ny * sizeof(float));
	int nxx, nyy;
	or_tv_zoom_size(nx, ny, &nxx, &nyy, factor);
	const float sigma = TV_ZOOM_SIGMA_ZERO * sqrt(1.0 / (factor * factor) - 1.0);
	or_tv_gaussian(Is, nx, ny, sigma);
	for (int i1 = 0; i1 < nyy; i1++)
		for (int j1 = 0; j1 < nxx; j1++) {
			const float i2 = (float)i1 / factor;
			const float j2 = (float)j1 / factor;
			Iout[i1 * nxx + j1] = or_tv_bicubic_at(Is, j2, i2, nx, ny, 0);
		}
	free(Is);
}

/* zoom_in, zoom.c:87-111 */
OR_EXPORT void or_tv_zoom_in(const float *I, float *Iout, int nx, int ny, int nxx, int nyy)
{
	const float factorx = ((float)nxx / nx);
	const float factory = ((float)nyy / ny);
	for (int i1 = 0; i1 < nyy; i1++)
		for (int j1 = 0; j1 < nxx; j1++) {
			float i2 = (float)i1 / factory;
			float j2 = (float)j1 / factorx;
			Iout[i1 * nxx + j1] = or_tv_bicubic_at(I, j2, i2, nx, ny, 0);
		}
}

/* image_normalization, tvl1flow_lib.c:303-337 */
static void normalize_pair(const float *I0, const float *I1, float *I0n, float *I1n, int size)
{
	float max0 = I0[0], min0 = I0[0], max1 = I1[0], min1 = I1[0];
	for (int i = 1; i < size; i++) {
		if (I0[i] < min0) min0 = I0[i];
		if (I0[i] > max0) max0 = I0[i];
	}
	for (int i = 1; i < size; i++) {
		if (I1[i] < min1) min1 = I1[i];
		if (I1[i] > max1) max1 = I1[i];
	}
	const float max = (max0 > max1) ? max0 : max1;
	const float min = (min0 < min1) ? min0 : min1;
	const float den = max - min;
	if (den > 0)
		for (int i = 0; i < size; i++) {
			I0n[i] = 255.0 * (I0[i] - min) / den;
			I1n[i] = 255.0 * (I1[i] - min) / den;
		}
	else
		for (int i = 0; i < size; i++) { I0n[i] = I0[i]; I1n[i] = I1[i]; }
}

/* Dual_TVL1_optic_flow_multiscale, tvl1flow_lib.c:345-474.  iters: nscales*warps
 * ints (scale-major, finest first), -1 where a scale is not solved. */
OR_EXPORT void or_tv_flow(const float *I0, const float *I1, float *u1, float *u2, int nxx, int nyy,
                          float tau, float lambda, float theta, int nscales, int fscale,
                          float zfactor, int warps, float epsilon, int *iters)
{
	float **I0s = (float **)calloc(nscales, sizeof(float *));
	float **I1s = (float **)calloc(nscales, sizeof(float *));
	float **u1s = (float **)calloc(nscales, sizeof(float *));
	float **u2s = (float **)calloc(nscales, sizeof(float *));
	int *nx = (int *)calloc(nscales, sizeof(int)), *ny = (int *)calloc(nscales, sizeof(int));
	const int size = nxx * nyy;
	if (iters) for (int k = 0; k < nscales * warps; k++) iters[k] = -1;
	I0s[0] = (float *)malloc((size_t)size * sizeof(float));
	I1s[0] = (float *)malloc((size_t)size * sizeof(float));
	u1s[0] = u1; u2s[0] = u2; nx[0] = nxx; ny[0] = nyy;
	normalize_pair(I0, I1, I0s[0], I1s[0], size);
	or_tv_gaussian(I0s[0], nx[0], ny[0], TV_PRESMOOTHING_SIGMA);
	or_tv_gaussian(I1s[0], nx[0], ny[0], TV_PRESMOOTHING_SIGMA);
	for (int s = 1; s < nscales; s++) {
		or_tv_zoom_size(nx[s - 1], ny[s - 1], &nx[s], &ny[s], zfactor);
		const int sz = nx[s] * ny[s];
		I0s[s] = (float *)malloc((size_t)sz * sizeof(float));
		I1s[s] = (float *)malloc((size_t)sz * sizeof(float));
		u1s[s] = (float *)malloc((size_t)sz * sizeof(float));
		u2s[s] = (float *)malloc((size_t)sz * sizeof(float));
		or_tv_zoom_out(I0s[s - 1], I0s[s], nx[s - 1], ny[s - 1], zfactor);
		or_tv_zoom_out(I1s[s - 1], I1s[s], nx[s - 1], ny[s - 1], zfactor);
	}
	for (int i = 0; i < nx[nscales - 1] * ny[nscales - 1]; i++)
		u1s[nscales - 1][i] = u2s[nscales - 1][i] = 0.0;
	for (int s = nscales - 1; s >= fscale; s--) {
		or_tv_solve(I0s[s], I1s[s], u1s[s], u2s[s], nx[s], ny[s], tau, lambda, theta, warps,
		            epsilon, iters ? iters + s * warps : NULL);
		if (!s) break;
		or_tv_zoom_in(u1s[s], u1s[s - 1], nx[s], ny[s], nx[s - 1], ny[s - 1]);
		or_tv_zoom_in(u2s[s], u2s[s - 1], nx[s], ny[s], nx[s - 1], ny[s - 1]);
		for (int i = 0; i < nx[s - 1] * ny[s - 1]; i++) {
			u1s[s - 1][i] *= (float)1.0 / zfactor;
			u2s[s - 1][i] *= (float)1.0 / zfactor;
		}
	}
	for (int s = fscale - 1; s >= 0; s--) {
		if (!s) break;
		or_tv_zoom_in(u1s[s], u1s[s - 1], nx[s], ny[s], nx[s - 1], ny[s - 1]);
		or_tv_zoom_in(u2s[s], u2s[s - 1], nx[s], ny[s], nx[s - 1], ny[s - 1]);
		for (int i = 0; i < nx[s - 1] * ny[s - 1]; i++) {
			u1s[s - 1][i] *= (float)1.0 / zfactor;
			u2s[s - 1][i] *= (float)1.0 / zfactor;
		}
	}
	for (int s = 1; s < nscales; s++) { free(I0s[s]); free(I1s[s]); free(u1s[s]); free(u2s[s]); }
	free(I0s[0]); free(I1s[0]);
	free(I0s); free(I1s); free(u1s); free(u2s); free(nx); free(ny);
}
