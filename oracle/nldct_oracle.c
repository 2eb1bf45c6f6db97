/*
 * nldct_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's two-step BM4D-OF / NL-DCT denoiser
 * (pariasm/bm4d-of, src/core/nldct.cpp and the helpers it calls).  It is the
 * checker the GPU path is compared with: only tests/, the smoke check and the
 * CPU-baseline leg of bench.py may load it.  It is never part of the product.
 *
 * Every function names the reference lines it follows.  Build with
 * -ffp-contract=off (oracle/Makefile): the reference's float arithmetic is
 * reproduced operation by operation, and products must not be fused.
 *
 * Matrix products follow a sequential-K sgemm with a fused multiply-add per
 * term: the reference calls an external CBLAS whose summation order it does
 * not fix (src/core/matrix_funs.cpp:152, cmake/modules/FindCBLAS.cmake);
 * this oracle fixes it as acc = 0; for p in 0..K-1: acc = fmaf(a, b, acc),
 * the accumulation of an FMA BLAS micro-kernel over K = 64 (one K block,
 * alpha = 1, beta = 0).  Every other product is separately rounded.
 *
 * Neighbour ties: the reference orders candidates with std::partial_sort
 * (libstdc++ heap select + sort_heap) under a strict `<` on the distance
 * (src/core/nldct.cpp:1280-1282, src/utils/utilities.cpp:35-40).
 * tie_mode 0 reproduces that heap order exactly; tie_mode 1 orders by
 * (distance, reference patch first, scan position), the total order the GPU
 * path uses.  The two differ only when two candidate distances are equal.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/bm4dof.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define OR_EXPORT __attribute__((visibility("default")))

static char g_err[512];
OR_EXPORT const char *or_last_error(void) { return g_err; }
static int fail(const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); return -1; }

/* ------------------------------------------------------------------------ */
/* Parameters: src/core/nldct.cpp:81-210                                     */
/* ------------------------------------------------------------------------ */

/* initializeNlbParameters, src/core/nldct.cpp:81-148.  betaMean is left
 * uninitialised by the reference; it is set to beta here (it is only read
 * with orderInvariance, src/core/nldct.cpp:1810). */
OR_EXPORT void or_init_params(bm4dof_params *p, unsigned step, float sigma,
                              unsigned frames, int verbose)
{
	const int s1 = (step == 1);
	memset(p, 0, sizeof *p);
	if (frames == 1) {
		if (sigma <= 40) {
			p->sizePatch = 8;
			p->nSimilarPatches = s1 ? 16 : 32;
			p->offSet = 3;
			p->beta = s1 ? 2.7f : 1.f;
			p->tau = s1 ? 54.8f : 20.f;
		} else {
			p->sizePatch = s1 ? 8 : 11;
			p->nSimilarPatches = 32;
			p->offSet = s1 ? 4 : 6;
			p->beta = s1 ? 2.8f : 1.f;
			p->tau = s1 ? 158.f : 59.2f;
		}
		p->sizePatchTime = 1;
		p->sizeSearchTimeRangeFwd = 0;
		p->sizeSearchTimeRangeBwd = 0;
		p->sizeSearchWindowPred = 0;
		p->offSetTime = 1;
		p->dsub = 0;
		p->nSimilarPatchesPred = p->nSimilarPatches;
		p->sizeSearchWindow = 39;
	} else {
		p->sizePatch = (s1 || sigma >= 30) ? 8 : 7;
		p->sizePatchTime = 1;
		p->nSimilarPatches = 8;
		p->sizeSearchWindow = 7;
		p->sizeSearchWindowPred = 5;
		p->sizeSearchTimeRangeFwd = 4;
		p->sizeSearchTimeRangeBwd = 4;
		p->nSimilarPatchesPred = 2;
		p->offSet = s1 ? 6 : 4;
		p->offSetTime = 1;
		p->beta = s1 ? 2.7f : 1.f;
		p->tau = s1 ? 67.1f : 54.8f;
		p->dsub = s1 ? 7 : 3;
	}
	p->betaMean = p->beta;
	p->orderInvariance = 0;
	p->isFirstStep = s1;
	p->sigma = sigma;
	p->verbose = verbose;
	p->colorSpace = BM4DOF_YUV;
	p->doPasteBoost = 0;
	p->agg_window = 1;
	p->boundary = 2 * (p->sizeSearchWindow / 2) + (p->sizePatch - 1);
	p->transform = s1 ? BM4DOF_BIOR1_5 : BM4DOF_DCT;
}

static unsigned umin(unsigned a, unsigned b) { return a < b ? a : b; }

/* setSizeSearchWindow, src/core/nldct.cpp:160-168 */
OR_EXPORT void or_set_search_window(bm4dof_params *p, unsigned sw)
{
	p->sizeSearchWindow = sw;
	p->boundary = 2 * (sw / 2) + (p->sizePatch - 1);
	p->nSimilarPatches = umin(p->nSimilarPatches, sw * sw *
		(p->sizeSearchTimeRangeFwd + p->sizeSearchTimeRangeBwd + 1));
}

/* setSizePatch, src/core/nldct.cpp:179-188.  For video the step-1 offset is
 * (float)px/4.*3. truncated to unsigned. */
OR_EXPORT void or_set_patch_size(bm4dof_params *p, unsigned frames, unsigned px)
{
	p->sizePatch = px;
	p->boundary = 2 * (p->sizeSearchWindow / 2) + (p->sizePatch - 1);
	if (p->sizePatch != 8) p->transform = BM4DOF_DCT;
	if (frames > 1)
		p->offSet = p->isFirstStep ? (unsigned)((double)(float)px / 4. * 3.) : px / 2;
}

/* setNSimilarPatches, src/core/nldct.cpp:199-210 (tests Fwd+Fwd, as the
 * reference does) */
OR_EXPORT void or_set_n_similar(bm4dof_params *p, unsigned np)
{
	p->nSimilarPatches = umin(np, p->sizeSearchWindow * p->sizeSearchWindow *
		(p->sizeSearchTimeRangeFwd + p->sizeSearchTimeRangeBwd + 1));
	if (p->sizeSearchTimeRangeFwd + p->sizeSearchTimeRangeFwd == 0)
		p->nSimilarPatchesPred = p->nSimilarPatches;
}

/* ------------------------------------------------------------------------ */
/* Video container: src/utils/lib_videot.hpp:40-121                          */
/* ------------------------------------------------------------------------ */

typedef struct {
	unsigned w, h, f, c;
	size_t wh, whc, whcf, whf;
	float *d;
} vid_t;

static void vid_init(vid_t *v, unsigned w, unsigned h, unsigned f, unsigned c)
{
	v->w = w; v->h = h; v->f = f; v->c = c;
	v->wh = (size_t)w * h; v->whc = v->wh * c; v->whcf = v->whc * f; v->whf = v->wh * f;
	v->d = (float *)calloc(v->whcf ? v->whcf : 1, sizeof(float));
}
static void vid_free(vid_t *v) { free(v->d); v->d = NULL; }
/* VideoSize::index(x,y,t,c), lib_videot.hpp:88-92 */
static inline size_t vidx(const vid_t *v, unsigned x, unsigned y, unsigned t, unsigned c)
{ return t * v->whc + c * v->wh + (size_t)y * v->w + x; }

/* ------------------------------------------------------------------------ */
/* Colour transform: VideoUtils::transformColorSpace, lib_videot.hpp:728-838 */
/* ------------------------------------------------------------------------ */
OR_EXPORT void or_color_transform(float *d, unsigned w, unsigned h, unsigned f,
                                  unsigned c, int forward)
{
	if (c == 1) return;
	const size_t wh = (size_t)w * h, whc = wh * c;
	for (unsigned t = 0; t < f; t++) {
		float *b0 = d + t * whc;
		if (forward && c == 3) {
			const float a = 1.f / sqrtf(3.f), b = 1.f / sqrtf(2.f), cc = 2.f * a * sqrtf(2.f);
			for (size_t k = 0; k < wh; k++) {
				const float r = b0[k], g = b0[k + wh], bl = b0[k + 2 * wh];
				const float y0 = a * (r + g + bl);
				const float y1 = b * (r - bl);
				const float y2 = cc * (0.25f * r - 0.5f * g + 0.25f * bl);
				b0[k] = y0; b0[k + wh] = y1; b0[k + 2 * wh] = y2;
			}
		} else if (forward) { /* 4 channels */
			const float a = 0.5f, b = 1.f / sqrtf(2.f);
			for (size_t k = 0; k < wh; k++) {
				const float Gr = b0[k], R = b0[k + wh], B = b0[k + 2 * wh], Gb = b0[k + 3 * wh];
				const float w0 = a * (Gr + R + B + Gb);
				const float w1 = b * (R - B);
				const float w2 = a * (-Gr + R + B - Gb);
				const float w3 = b * (-Gr + Gb);
				b0[k] = w0; b0[k + wh] = w1; b0[k + 2 * wh] = w2; b0[k + 3 * wh] = w3;
			}
		} else if (c == 3) {
			const float a = 1.f / sqrtf(3.f), b = 1.f / sqrtf(2.f), cc = a / b;
			for (size_t k = 0; k < wh; k++) {
				const float y0 = b0[k], y1 = b0[k + wh], y2 = b0[k + 2 * wh];
				const float r = a * y0 + b * y1 + cc * 0.5f * y2;
				const float g = a * y0 - cc * y2;
				const float bl = a * y0 - b * y1 + cc * 0.5f * y2;
				b0[k] = r; b0[k + wh] = g; b0[k + 2 * wh] = bl;
			}
		} else {
			const float a = 0.5f, b = 1.f / sqrtf(2.f);
			for (size_t k = 0; k < wh; k++) {
				const float Gr = b0[k], R = b0[k + wh], B = b0[k + 2 * wh], Gb = b0[k + 3 * wh];
				const float w0 = a * Gr - a * B - b * Gb;
				const float w1 = a * Gr + b * R + a * B;
				const float w2 = a * Gr - b * R + a * B;
				const float w3 = a * Gr - a * B + b * Gb;
				b0[k] = w0; b0[k + wh] = w1; b0[k + 2 * wh] = w2; b0[k + 3 * wh] = w3;
			}
		}
	}
}

/* PSNR/RMSE: VideoUtils::computePSNR, lib_videot.hpp:517-535 */
OR_EXPORT void or_psnr(const float *a, const float *b, size_t n, float *psnr, float *rmse)
{
	double sum = 0.f;
	for (size_t k = 0; k < n; k++)
		sum += ((double)a[k] - (double)b[k]) * ((double)a[k] - (double)b[k]);
	*rmse = sqrtf((float)(sum / (double)n));
	*psnr = 20.f * log10f(255.f / *rmse);
}

/* ------------------------------------------------------------------------ */
/* Tiling: determineFactor utilities.cpp:137-158, subDivideTight            */
/* lib_videot.hpp:854-916, subBuildTight lib_videot.hpp:1102-1163           */
/* ------------------------------------------------------------------------ */

OR_EXPORT void or_determine_factor(unsigned n, unsigned *a, unsigned *b)
{
	if (n == 1) { *a = 1; *b = 1; return; }
	*b = 2;
	while (n % *b > 0) (*b)++;
	*a = n / *b;
	if (*b > *a) { *a = *b; *b = n / *a; }
}

typedef struct {
	int ox, oy, ex, ey;  /* crop [ox,ex) x [oy,ey), all frames */
	int ix0, iy0, ix1, iy1; /* interior written back by subBuildTight */
} crop_t;

/* Fills crops[nparts]; returns 0, or -1 when the tiling is degenerate (a
 * crop of non-positive size: the reference then reads out of bounds). */
OR_EXPORT int or_tiles(unsigned W, unsigned H, int bdry, int nparts, int *out /* 8 ints per tile */)
{
	unsigned unW, unH;
	or_determine_factor((unsigned)nparts, &unW, &unH);
	const int nW = (int)unW, nH = (int)unH;
	const int wTmp = (int)ceilf((float)W / (float)nW);
	const int hTmp = (int)ceilf((float)H / (float)nH);
	for (int p = 0, n = 0; p < nH; p++)
		for (int q = 0; q < nW; q++, n++) {
			int *o = out + 8 * n;
			o[0] = q * wTmp - bdry > 0 ? q * wTmp - bdry : 0;
			o[1] = p * hTmp - bdry > 0 ? p * hTmp - bdry : 0;
			o[2] = (q + 1) * wTmp + bdry < (int)W ? (q + 1) * wTmp + bdry : (int)W;
			o[3] = (p + 1) * hTmp + bdry < (int)H ? (p + 1) * hTmp + bdry : (int)H;
			o[4] = q * wTmp;
			o[5] = p * hTmp;
			o[6] = (q + 1) * wTmp < (int)W ? (q + 1) * wTmp : (int)W;
			o[7] = (p + 1) * hTmp < (int)H ? (p + 1) * hTmp : (int)H;
			if (o[2] <= o[0] || o[3] <= o[1]) return fail("degenerate tiling: too many parts for the image size");
		}
	return 0;
}

/* crop with symmetric access, lib_videot.hpp:612-683 (crops lie inside the
 * image, so the symmetric extension is never taken) */
static void crop_video(const vid_t *src, vid_t *dst, const crop_t *cr)
{
	vid_init(dst, cr->ex - cr->ox, cr->ey - cr->oy, src->f, src->c);
	for (unsigned t = 0; t < dst->f; t++)
		for (unsigned c = 0; c < dst->c; c++)
			for (unsigned y = 0; y < dst->h; y++)
				memcpy(dst->d + vidx(dst, 0, y, t, c),
				       src->d + vidx(src, cr->ox, y + cr->oy, t, c), dst->w * sizeof(float));
}

/* ------------------------------------------------------------------------ */
/* libstdc++ std::partial_sort (heap select + sort_heap) on (dist, index)   */
/* pairs with comparator comparaisonFirst (utilities.cpp:35-40).            */
/* ------------------------------------------------------------------------ */

typedef struct { float d; uint32_t i; int32_t rank; } dpair;

static void push_heap_(dpair *first, long hole, long top, dpair value)
{
	long parent = (hole - 1) / 2;
	while (hole > top && first[parent].d < value.d) {
		first[hole] = first[parent];
		hole = parent;
		parent = (hole - 1) / 2;
	}
	first[hole] = value;
}

static void adjust_heap_(dpair *first, long hole, long len, dpair value)
{
	const long top = hole;
	long second = hole;
	while (second < (len - 1) / 2) {
		second = 2 * (second + 1);
		if (first[second].d < first[second - 1].d) second--;
		first[hole] = first[second];
		hole = second;
	}
	if ((len & 1) == 0 && second == (len - 2) / 2) {
		second = 2 * (second + 1);
		first[hole] = first[second - 1];
		hole = second - 1;
	}
	push_heap_(first, hole, top, value);
}

static void make_heap_(dpair *first, long len)
{
	if (len < 2) return;
	long parent = (len - 2) / 2;
	for (;;) {
		dpair v = first[parent];
		adjust_heap_(first, parent, len, v);
		if (parent == 0) return;
		parent--;
	}
}

static void pop_heap_(dpair *first, long len, dpair *result)
{
	dpair v = *result;
	*result = *first;
	adjust_heap_(first, 0, len, v);
}

static void partial_sort_libstdcxx(dpair *first, long middle, long last)
{
	make_heap_(first, middle);
	for (long i = middle; i < last; ++i)
		if (first[i].d < first[0].d) pop_heap_(first, middle, &first[i]);
	for (long l = middle; l > 1;) { --l; pop_heap_(first, l, &first[l]); }
}

static int cmp_total(const void *a, const void *b)
{
	const dpair *x = (const dpair *)a, *y = (const dpair *)b;
	if (x->d < y->d) return -1;
	if (y->d < x->d) return 1;
	return (x->rank > y->rank) - (x->rank < y->rank);
}

/* Exported for the library pin test (compared with std::partial_sort). */
OR_EXPORT void or_partial_sort(float *d, uint32_t *idx, long middle, long n, int tie_mode)
{
	dpair *a = (dpair *)malloc((n ? n : 1) * sizeof *a);
	for (long k = 0; k < n; k++) { a[k].d = d[k]; a[k].i = idx[k]; a[k].rank = (int32_t)k; }
	if (tie_mode == 0) partial_sort_libstdcxx(a, middle, n);
	else qsort(a, n, sizeof *a, cmp_total);
	for (long k = 0; k < n; k++) { d[k] = a[k].d; idx[k] = a[k].i; }
	free(a);
}

/* ------------------------------------------------------------------------ */
/* Patch bases and aggregation window: processNlBayes set-up,               */
/* src/core/nldct.cpp:571-737                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
	int P;         /* px*px*pt */
	float *basis;  /* P*P, basis[k*P + pix]     (nldct.cpp:597-607) */
	float *binv;   /* P*P, binv[k*P + pix]      (nldct.cpp:610, 685-695) */
	float *win;    /* P                          (nldct.cpp:698-737) */
} bases_t;

static const double bior_fwd[64] = { /* nldct.cpp:640-647 */
	 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274,
	 0.219417649252501, 0.449283757993216, 0.449283757993216, 0.219417649252501,-0.219417649252501,-0.449283757993216,-0.449283757993216,-0.219417649252501,
	 0.569359398342846, 0.402347308162278,-0.402347308162278,-0.569359398342846,-0.083506045090284, 0.083506045090284,-0.083506045090284, 0.083506045090284,
	-0.083506045090284, 0.083506045090284,-0.083506045090284, 0.083506045090284, 0.569359398342846, 0.402347308162278,-0.402347308162278,-0.569359398342846,
	 0.707106781186547,-0.707106781186547, 0, 0, 0, 0, 0, 0,
	 0, 0, 0.707106781186547,-0.707106781186547, 0, 0, 0, 0,
	 0, 0, 0, 0, 0.707106781186547,-0.707106781186547, 0, 0,
	 0, 0, 0, 0, 0, 0, 0.707106781186547,-0.707106781186547 };
static const double bior_inv[64] = { /* nldct.cpp:670-677 */
	0.353553390593274, 0.353553390593274, 0.5, 0.0, 0.707106781186547,-0.121533978016438,-0.000000000000000, 0.121533978016438,
	0.353553390593274, 0.353553390593274, 0.5,-0.0,-0.707106781186547,-0.121533978016438, 0.000000000000000, 0.121533978016438,
	0.353553390593274, 0.353553390593274,-0.5,-0.0, 0.121533978016438, 0.707106781186548,-0.121533978016438, 0.000000000000000,
	0.353553390593274, 0.353553390593274,-0.5, 0.0, 0.121533978016438,-0.707106781186548,-0.121533978016438, 0.000000000000000,
	0.353553390593274,-0.353553390593274, 0.0, 0.5, 0.000000000000000, 0.121533978016438, 0.707106781186547,-0.121533978016438,
	0.353553390593274,-0.353553390593274, 0.0, 0.5, 0.000000000000000, 0.121533978016438,-0.707106781186547,-0.121533978016438,
	0.353553390593274,-0.353553390593274, 0.0,-0.5,-0.121533978016438,-0.000000000000000, 0.121533978016438, 0.707106781186547,
	0.353553390593274,-0.353553390593274, 0.0,-0.5,-0.121533978016438,-0.000000000000000, 0.121533978016438,-0.707106781186547 };
static const float kaiser8[64] = { /* nldct.cpp:706-713 */
	0.1924f, 0.2989f, 0.3846f, 0.4325f, 0.4325f, 0.3846f, 0.2989f, 0.1924f,
	0.2989f, 0.4642f, 0.5974f, 0.6717f, 0.6717f, 0.5974f, 0.4642f, 0.2989f,
	0.3846f, 0.5974f, 0.7688f, 0.8644f, 0.8644f, 0.7688f, 0.5974f, 0.3846f,
	0.4325f, 0.6717f, 0.8644f, 0.9718f, 0.9718f, 0.8644f, 0.6717f, 0.4325f,
	0.4325f, 0.6717f, 0.8644f, 0.9718f, 0.9718f, 0.8644f, 0.6717f, 0.4325f,
	0.3846f, 0.5974f, 0.7688f, 0.8644f, 0.8644f, 0.7688f, 0.5974f, 0.3846f,
	0.2989f, 0.4642f, 0.5974f, 0.6717f, 0.6717f, 0.5974f, 0.4642f, 0.2989f,
	0.1924f, 0.2989f, 0.3846f, 0.4325f, 0.4325f, 0.3846f, 0.2989f, 0.1924f };
static const float kaiser10[100] = { /* nldct.cpp:720-729 */
	0.1924f, 0.2763f, 0.3503f, 0.4055f, 0.4349f, 0.4349f, 0.4055f, 0.3503f, 0.2763f, 0.1924f,
	0.2763f, 0.3967f, 0.5030f, 0.5822f, 0.6245f, 0.6245f, 0.5822f, 0.5030f, 0.3967f, 0.2763f,
	0.3503f, 0.5030f, 0.6377f, 0.7381f, 0.7917f, 0.7917f, 0.7381f, 0.6377f, 0.5030f, 0.3503f,
	0.4055f, 0.5822f, 0.7381f, 0.8544f, 0.9164f, 0.9164f, 0.8544f, 0.7381f, 0.5822f, 0.4055f,
	0.4349f, 0.6245f, 0.7917f, 0.9164f, 0.9829f, 0.9829f, 0.9164f, 0.7917f, 0.6245f, 0.4349f,
	0.4349f, 0.6245f, 0.7917f, 0.9164f, 0.9829f, 0.9829f, 0.9164f, 0.7917f, 0.6245f, 0.4349f,
	0.4055f, 0.5822f, 0.7381f, 0.8544f, 0.9164f, 0.9164f, 0.8544f, 0.7381f, 0.5822f, 0.4055f,
	0.3503f, 0.5030f, 0.6377f, 0.7381f, 0.7917f, 0.7917f, 0.7381f, 0.6377f, 0.5030f, 0.3503f,
	0.2763f, 0.3967f, 0.5030f, 0.5822f, 0.6245f, 0.6245f, 0.5822f, 0.5030f, 0.3967f, 0.2763f,
	0.1924f, 0.2763f, 0.3503f, 0.4055f, 0.4349f, 0.4349f, 0.4055f, 0.3503f, 0.2763f, 0.1924f };

/* Builds the bases of one step.  Returns -1 for bior with px != 8
 * (nldct.cpp:633-635). */
static int build_bases(const bm4dof_params *p, bases_t *b)
{
	const unsigned sPx = p->sizePatch, sPt = p->sizePatchTime;
	const int P = (int)(sPx * sPx * sPt);
	b->P = P;
	b->basis = (float *)calloc((size_t)P * P, sizeof(float));
	b->binv = (float *)calloc((size_t)P * P, sizeof(float));
	b->win = (float *)calloc(P, sizeof(float));
	if (p->transform == BM4DOF_DCT) {
		float *cosx = (float *)malloc(sPx * sPx * sizeof(float));
		float *cost = (float *)malloc(sPt * sPt * sizeof(float));
		for (unsigned kx = 0, i = 0; kx < sPx; kx++)
			for (unsigned nx = 0; nx < sPx; nx++, i++)
				cosx[i] = (float)(cos((0.5 + (float)nx) * (float)kx * M_PI / (float)sPx)
				                  * sqrt((kx == 0 ? 1. : 2.) / (float)sPx));
		for (unsigned kt = 0, i = 0; kt < sPt; kt++)
			for (unsigned nt = 0; nt < sPt; nt++, i++)
				cost[i] = (float)(cos((0.5 + (float)nt) * (float)kt * M_PI / (float)sPt)
				                  * sqrt((kt == 0 ? 1. : 2.) / (float)sPt));
		size_t i = 0;
		for (unsigned kt = 0; kt < sPt; kt++)
		for (unsigned ky = 0; ky < sPx; ky++)
		for (unsigned kx = 0; kx < sPx; kx++)
			for (unsigned nt = 0; nt < sPt; nt++)
			for (unsigned ny = 0; ny < sPx; ny++)
			for (unsigned nx = 0; nx < sPx; nx++, i++) {
				float v = cosx[kx * sPx + nx] * cosx[ky * sPx + ny];
				b->basis[i] = v * cost[kt * sPt + nt];
			}
		memcpy(b->binv, b->basis, (size_t)P * P * sizeof(float));
		free(cosx); free(cost);
	} else {
		if (sPx != 8) {
			free(b->basis); free(b->binv); free(b->win);
			return fail("VideoNLB::processNlBayes: bior1_5 transform is only implemented for patches of size 8 x 8 x pt.");
		}
		float bx[64], bxi[64];
		for (int k = 0; k < 64; k++) { bx[k] = (float)bior_fwd[k]; bxi[k] = (float)bior_inv[k]; }
		size_t i = 0;
		for (unsigned kt = 0; kt < sPt; kt++)
		for (unsigned ky = 0; ky < 8; ky++)
		for (unsigned kx = 0; kx < 8; kx++)
			for (unsigned nt = 0; nt < sPt; nt++)
			for (unsigned ny = 0; ny < 8; ny++)
			for (unsigned nx = 0; nx < 8; nx++, i++) {
				float v = bx[kx * 8 + nx] * bx[ky * 8 + ny];
				b->basis[i] = v * (nt == kt ? 1.f : 0.f);
			}
		i = 0;
		for (unsigned nt = 0; nt < sPt; nt++)
		for (unsigned ny = 0; ny < 8; ny++)
		for (unsigned nx = 0; nx < 8; nx++)
			for (unsigned kt = 0; kt < sPt; kt++)
			for (unsigned ky = 0; ky < 8; ky++)
			for (unsigned kx = 0; kx < 8; kx++, i++) {
				float v = bxi[kx * 8 + nx] * bxi[ky * 8 + ny];
				b->binv[i] = v * (nt == kt ? 1.f : 0.f);
			}
	}
	if (p->agg_window && sPx == 8)
		for (unsigned t = 0; t < sPt; t++) memcpy(b->win + 64 * t, kaiser8, sizeof kaiser8);
	else if (p->agg_window && sPx == 10)
		for (unsigned t = 0; t < sPt; t++) memcpy(b->win + 100 * t, kaiser10, sizeof kaiser10);
	else
		for (int k = 0; k < P; k++) b->win[k] = 1.f;
	return 0;
}

static void free_bases(bases_t *b) { free(b->basis); free(b->binv); free(b->win); }

/* Exported for pin tests: writes basis and inverse basis (P*P each) and the
 * window (P).  Returns P or -1. */
OR_EXPORT int or_bases(const bm4dof_params *p, float *basis, float *binv, float *win)
{
	bases_t b;
	if (build_bases(p, &b)) return -1;
	memcpy(basis, b.basis, (size_t)b.P * b.P * sizeof(float));
	memcpy(binv, b.binv, (size_t)b.P * b.P * sizeof(float));
	memcpy(win, b.win, (size_t)b.P * sizeof(float));
	free_bases(&b);
	return b.P;
}

/* ------------------------------------------------------------------------ */
/* productMatrix (matrix_funs.cpp:136-164) with a sequential-K FMA sgemm.   */
/* Column-major: C(n x l) = A(n x m) * op(B).                               */
/* ------------------------------------------------------------------------ */
static void sgemm_nn(float *C, const float *A, const float *B, int n, int l, int m)
{
	/* j, p, i: every C(i, j) is accumulated in p order; the i loop runs over
	 * independent sums (the CPU overlaps them) */
	for (int j = 0; j < l; j++) {
		float *c = C + (size_t)j * n;
		for (int i = 0; i < n; i++) c[i] = 0.f;
		for (int p = 0; p < m; p++) {
			const float b = B[p + j * m];
			const float *a = A + (size_t)p * n;
			for (int i = 0; i < n; i++) c[i] = fmaf(a[i], b, c[i]);
		}
	}
}
static void sgemm_nt(float *C, const float *A, const float *B, int n, int l, int m)
{
	for (int j = 0; j < l; j++) {
		float *c = C + (size_t)j * n;
		for (int i = 0; i < n; i++) c[i] = 0.f;
		for (int p = 0; p < m; p++) {
			const float b = B[j + p * l];
			const float *a = A + (size_t)p * n;
			for (int i = 0; i < n; i++) c[i] = fmaf(a[i], b, c[i]);
		}
	}
}

/* ------------------------------------------------------------------------ */
/* Block matching: estimateSimilarPatchesStep1 nldct.cpp:1143-1348 and      */
/* estimateSimilarPatchesStep2 nldct.cpp:1607-1789 (non-VBM3D_SEARCH,        */
/* non-CENTRED_SEARCH variants).                                            */
/* ------------------------------------------------------------------------ */

static inline float fmin_ref(float a, float b) { return (b < a) ? b : a; } /* std::min */
static inline float fmax_ref(float a, float b) { return (a < b) ? b : a; } /* std::max */
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* step: 1 -> distance on channel 0 of `im`; 2 -> all channels of `im`
 * (the basic image).  Writes index[0..nSimP) (tile-local reference indices,
 * t*whc + y*w + x) and returns nSimP. */
/* The reference's predictive search, compiled in with VBM3D_SEARCH
 * (src/core/nldct.h:20): estimateSimilarPatchesStep1 src/core/nldct.cpp:904-1130,
 * estimateSimilarPatchesStep2 src/core/nldct.cpp:1368-1589.  The reference
 * frame is searched in the full window; every other frame, in the union of
 * nSimilarPatchesPred windows of sizeSearchWindowPred around the
 * flow-propagated positions of the previous frame's best matches (insertion
 * order: window by window, each in raster order, positions already inserted
 * skipped, as the unordered_set test of :1044 does).  Distances on channel 0
 * (step 2: the channel-0 sum repeated once per channel, since the step-2
 * loop reads imBasic(x, y, t) without its channel, :1532-1535), minus
 * dsub^2*255 at the reference's (x, y) in every frame (:1049, :1539); the
 * nSimilarPatchesPred best of each frame by std::partial_sort are kept, and
 * the kept ones of all frames sorted again for the group (:1076, :1102).
 * Enabled by or_set_predictive(1). */
static int g_predictive = 0;
OR_EXPORT void or_set_predictive(int on) { g_predictive = on; }

/* The reference keeps nSimilarPatchesPred candidates of every frame's
 * search region (std::partial_sort to that middle, :1076); a region smaller
 * than that is undefined behaviour there and an error here. */
static int pred_check(const bm4dof_params *p, unsigned w, unsigned h, int nparts)
{
	if (!g_predictive || p->sizePatch == 0 || p->nSimilarPatches <= 1) return 0;
	/* the smallest tile bounds the search regions */
	int *tl = (int *)malloc((size_t)8 * (nparts > 0 ? nparts : 1) * sizeof(int));
	if (or_tiles(w, h, (int)p->boundary, nparts, tl)) { free(tl); return -1; }
	for (int n = 0; n < nparts; n++) {
		const unsigned tw = (unsigned)(tl[8 * n + 2] - tl[8 * n]), th = (unsigned)(tl[8 * n + 3] - tl[8 * n + 1]);
		if (tw < w) w = tw;
		if (th < h) h = th;
	}
	free(tl);
	const unsigned cx = w >= p->sizePatch ? w - p->sizePatch + 1 : 0;
	const unsigned cy = h >= p->sizePatch ? h - p->sizePatch + 1 : 0;
	const unsigned ws[2] = {p->sizeSearchWindow, p->sizeSearchWindowPred};
	for (int k = 0; k < 2; k++) {
		const unsigned a = ws[k] < cx ? ws[k] : cx, b = ws[k] < cy ? ws[k] : cy;
		if (p->nSimilarPatchesPred < 1 || (size_t)a * b < p->nSimilarPatchesPred)
			return fail("predictive search: a search region holds fewer than nSimilarPatchesPred patches");
	}
	return 0;
}

static unsigned bm_search_pred(const vid_t *im, const vid_t *fflow, const vid_t *bflow,
                               unsigned pidx, const bm4dof_params *prm, int step, unsigned *index)
{
	const int sPx = (int)prm->sizePatch, sPt = (int)prm->sizePatchTime;
	if (prm->nSimilarPatches <= 1) { index[0] = pidx; return 1; }
	const int npf = (int)prm->nSimilarPatchesPred;
	const float dsub = prm->dsub * prm->dsub * 255;
	const float tau = prm->tau * prm->tau * sPx * sPx * sPt;
	const int use_flow = fflow && fflow->w > 0;
	const int sWt_f = (int)prm->sizeSearchTimeRangeFwd, sWt_b = (int)prm->sizeSearchTimeRangeBwd;
	const int W = (int)im->w, H = (int)im->h, F = (int)im->f;
	const int px = (int)(pidx % im->w), py = (int)((pidx % im->wh) / im->w), pt = (int)(pidx / im->whc);
	const int shift_t = imin(0, pt - sWt_b) + imax(0, pt + sWt_f - F + sPt);
	const int rt0 = imax(0, pt - sWt_b - shift_t), rt1 = imin(F - sPt, pt + sWt_f - shift_t);
	const int sWt = rt1 - rt0 + 1;
	int sched[64];
	int ns = 0;
	sched[ns++] = pt;
	for (int qt = pt + 1; qt <= rt1; ++qt) sched[ns++] = qt;
	for (int qt = pt - 1; qt >= rt0; --qt) sched[ns++] = qt;
	const int wfull = (int)prm->sizeSearchWindow, wpred = (int)prm->sizeSearchWindowPred;
	const int cap = imax(wfull * wfull, npf * wpred * wpred) + 1;
	dpair *cand = (dpair *)malloc((size_t)cap * sizeof(dpair));
	dpair *kept = (dpair *)malloc((size_t)(npf * sWt + 1) * sizeof(dpair));
	int *tx = (int *)calloc((size_t)npf * sWt, sizeof(int)), *ty = (int *)calloc((size_t)npf * sWt, sizeof(int));
	int *rx0 = (int *)malloc(npf * sizeof(int)), *rx1 = (int *)malloc(npf * sizeof(int));
	int *ry0 = (int *)malloc(npf * sizeof(int)), *ry1 = (int *)malloc(npf * sizeof(int));
	const int reps = step == 1 ? 1 : (int)im->c;
	int nk = 0;
	for (int ii = 0; ii < sWt; ++ii) {
		const int qt = sched[ii], dt = qt - rt0;
		const int dir = imax(-1, imin(1, qt - pt));
		int nwin, sw;
		int cx0[64], cy0[64];
		if (dir != 0) {
			sw = wpred;
			nwin = npf;
			for (int i = 0; i < npf; i++) {
				const int cx = tx[i * sWt + dt - dir], cy = ty[i * sWt + dt - dir], ct = qt - dir;
				float fx = 0.f, fy = 0.f;
				if (use_flow) {
					const vid_t *fl = dir > 0 ? fflow : bflow;
					fx = fl->d[vidx(fl, cx, cy, ct, 0)];
					fy = fl->d[vidx(fl, cx, cy, ct, 1)];
				}
				cx0[i] = (int)fmax_ref(0.f, fmin_ref((float)W - 1.f, roundf((float)cx + fx)));
				cy0[i] = (int)fmax_ref(0.f, fmin_ref((float)H - 1.f, roundf((float)cy + fy)));
			}
		} else {
			sw = wfull;
			nwin = 1;
			cx0[0] = px;
			cy0[0] = py;
		}
		int nc = 0;
		for (int i = 0; i < nwin; i++) {
			int shx = imin(0, cx0[i] - (sw - 1) / 2), shy = imin(0, cy0[i] - (sw - 1) / 2);
			shx += imax(0, cx0[i] + (sw - 1) / 2 - W + sPx);
			shy += imax(0, cy0[i] + (sw - 1) / 2 - H + sPx);
			rx0[i] = imax(0, cx0[i] - (sw - 1) / 2 - shx);
			ry0[i] = imax(0, cy0[i] - (sw - 1) / 2 - shy);
			rx1[i] = imin(W - sPx, cx0[i] + (sw - 1) / 2 - shx);
			ry1[i] = imin(H - sPx, cy0[i] + (sw - 1) / 2 - shy);
			for (int qy = ry0[i]; qy <= ry1[i]; qy++)
				for (int qx = rx0[i]; qx <= rx1[i]; qx++) {
					int seen = 0;
					for (int j = 0; j < i && !seen; j++)
						seen = qx >= rx0[j] && qx <= rx1[j] && qy >= ry0[j] && qy <= ry1[j];
					if (seen) continue;
					float d = 0.f, dif;
					for (int r = 0; r < reps; r++)
						for (int ht = 0; ht < sPt; ht++)
							for (int hy = 0; hy < sPx; hy++)
								for (int hx = 0; hx < sPx; hx++) {
									dif = im->d[vidx(im, px + hx, py + hy, pt + ht, 0)] -
									      im->d[vidx(im, qx + hx, qy + hy, qt + ht, 0)];
									d += dif * dif;
								}
					d -= (qy == py && qx == px) ? dsub : 0;
					cand[nc].d = d;
					cand[nc].i = (uint32_t)vidx(im, qx, qy, qt, 0);
					cand[nc].rank = nc;
					nc++;
				}
		}
		partial_sort_libstdcxx(cand, npf, nc);
		for (int i = 0; i < npf; i++) {
			kept[nk] = cand[i];
			kept[nk].rank = nk;
			nk++;
			tx[i * sWt + dt] = (int)(cand[i].i % im->w);
			ty[i * sWt + dt] = (int)((cand[i].i % im->wh) / im->w);
		}
	}
	unsigned nSimP = umin(prm->nSimilarPatches, (unsigned)nk);
	partial_sort_libstdcxx(kept, nSimP, nk);
	unsigned n = 1;
	while (2 * n <= nSimP && kept[n].d < tau) n *= 2;
	for (unsigned k = 0; k < n; k++) index[k] = kept[k].i;
	free(cand); free(kept); free(tx); free(ty); free(rx0); free(rx1); free(ry0); free(ry1);
	return n;
}

static unsigned bm_search(const vid_t *im, const vid_t *fflow, const vid_t *bflow,
                          unsigned pidx, const bm4dof_params *prm, int step,
                          int tie_mode, dpair *dist, unsigned *index)
{
	const int sPx = (int)prm->sizePatch, sPt = (int)prm->sizePatchTime;
	unsigned nSimP = 1;
	if (prm->nSimilarPatches > 1) {
		const float tau = prm->tau * prm->tau * sPx * sPx * sPt;
		const int use_flow = fflow && fflow->w > 0;
		const int sWx = (int)prm->sizeSearchWindow, sWy = sWx;
		const int sWt_f = (int)prm->sizeSearchTimeRangeFwd, sWt_b = (int)prm->sizeSearchTimeRangeBwd;
		const int W = (int)im->w, H = (int)im->h, F = (int)im->f;
		const unsigned px = (unsigned)(pidx % im->w);
		const unsigned py = (unsigned)((pidx % im->wh) / im->w);
		const unsigned pt = (unsigned)(pidx / im->whc);
		if (g_predictive) return bm_search_pred(im, fflow, bflow, pidx, prm, step, index);

		int shift_t = imin(0, (int)pt - sWt_b) + imax(0, (int)pt + sWt_f - F + sPt);
		int rt0 = imax(0, (int)pt - sWt_b - shift_t);
		int rt1 = imin(F - sPt, (int)pt + sWt_f - shift_t);
		int sWt = rt1 - rt0 + 1;

		int sched[64], cx[64], cy[64], ct[64];
		int ns = 0;
		sched[ns++] = (int)pt;
		for (int qt = (int)pt + 1; qt <= rt1; ++qt) sched[ns++] = qt;
		for (int qt = (int)pt - 1; qt >= rt0; --qt) sched[ns++] = qt;

		float refpatch[4096];
		if (step == 1)
			for (int ht = 0, k = 0; ht < sPt; ht++)
				for (int hy = 0; hy < sPx; hy++)
					for (int hx = 0; hx < sPx; hx++, k++)
						refpatch[k] = im->d[vidx(im, px + hx, py + hy, pt + ht, 0)];

		int nsrch = 0;
		for (int ii = 0; ii < sWt; ++ii) {
			const int qt = sched[ii];
			const int dt = qt - rt0;
			const int dir = imax(-1, imin(1, qt - (int)pt));
			if (dir != 0) {
				const int cx0 = cx[dt - dir], cy0 = cy[dt - dir], ct0 = ct[dt - dir];
				float fx = 0.f, fy = 0.f;
				if (use_flow) {
					const vid_t *fl = dir > 0 ? fflow : bflow;
					fx = fl->d[vidx(fl, cx0, cy0, ct0, 0)];
					fy = fl->d[vidx(fl, cx0, cy0, ct0, 1)];
				}
				const float cx_f = (float)cx0 + fx;
				const float cy_f = (float)cy0 + fy;
				cx[dt] = (int)fmax_ref(0.f, fmin_ref((float)W - 1.f, roundf(cx_f)));
				cy[dt] = (int)fmax_ref(0.f, fmin_ref((float)H - 1.f, roundf(cy_f)));
				ct[dt] = qt;
			} else {
				cx[dt] = (int)px; cy[dt] = (int)py; ct[dt] = (int)pt;
			}
			int shift_x = imin(0, cx[dt] - (sWx - 1) / 2);
			int shift_y = imin(0, cy[dt] - (sWy - 1) / 2);
			shift_x += imax(0, cx[dt] + (sWx - 1) / 2 - W + sPx);
			shift_y += imax(0, cy[dt] + (sWy - 1) / 2 - H + sPx);
			const int rx0 = imax(0, cx[dt] - (sWx - 1) / 2 - shift_x);
			const int ry0 = imax(0, cy[dt] - (sWy - 1) / 2 - shift_y);
			const int rx1 = imin(W - sPx, cx[dt] + (sWx - 1) / 2 - shift_x);
			const int ry1 = imin(H - sPx, cy[dt] + (sWy - 1) / 2 - shift_y);

			for (int qy = ry0; qy <= ry1; qy++)
				for (int qx = rx0; qx <= rx1; qx++) {
					float d = 0.f, dif;
					if (step == 1) {
						for (int ht = 0, k = 0; ht < sPt; ht++)
							for (int hy = 0; hy < sPx; hy++)
								for (int hx = 0; hx < sPx; hx++, k++) {
									dif = refpatch[k] - im->d[vidx(im, qx + hx, qy + hy, qt + ht, 0)];
									d += dif * dif;
								}
					} else {
						for (unsigned c = 0; c < im->c; c++)
							for (int ht = 0; ht < sPt; ht++)
								for (int hy = 0; hy < sPx; hy++)
									for (int hx = 0; hx < sPx; hx++) {
										dif = im->d[vidx(im, px + hx, py + hy, pt + ht, c)]
										    - im->d[vidx(im, qx + hx, qy + hy, qt + ht, c)];
										d += dif * dif;
									}
					}
					dist[nsrch].d = d;
					dist[nsrch].i = (uint32_t)vidx(im, qx, qy, qt, 0);
					dist[nsrch].rank = (qx == (int)px && qy == (int)py && qt == (int)pt) ? -1 : nsrch;
					nsrch++;
				}
		}

		nSimP = umin(prm->nSimilarPatches, (unsigned)nsrch);
		if (tie_mode == 0) partial_sort_libstdcxx(dist, nSimP, nsrch);
		else qsort(dist, nsrch, sizeof *dist, cmp_total);

		{
			unsigned n = 1;
			while (2 * n <= nSimP && dist[n].d < tau) n *= 2;
			nSimP = n;
		}
		for (unsigned n = 0; n < nSimP; n++) index[n] = dist[n].i;
	} else
		index[0] = pidx;
	return nSimP;
}

/* Group gather, nldct.cpp:1330-1339 / 1776-1786: g[c][(pix)*N + n] */
static void gather_group(const vid_t *im, const unsigned *index, unsigned N,
                         int sPx, int sPt, float *g /* C*P*N */)
{
	const int P = sPx * sPx * sPt;
	for (unsigned c = 0; c < im->c; c++)
		for (int ht = 0, k = 0; ht < sPt; ht++)
			for (int hy = 0; hy < sPx; hy++)
				for (int hx = 0; hx < sPx; hx++)
					for (unsigned n = 0; n < N; n++, k++)
						g[(size_t)c * P * N + k] = im->d[c * im->wh + index[n] + ht * im->whc + hy * im->w + hx];
}

/* ------------------------------------------------------------------------ */
/* Group filtering.                                                          */
/* ------------------------------------------------------------------------ */

static void haar_fwd(float *z, int P, int N)
{
	const float isqrt2 = (float)(1. / sqrt(2.));
	for (int s = 2, h = 1; s <= N; s *= 2, h *= 2)
		for (int i = 0; i < P; i++)
			for (int n = 0; n < N; n += s) {
				const float z0 = z[n + i * N] * isqrt2;
				const float z1 = z[n + h + i * N] * isqrt2;
				z[n + i * N] = z0 + z1;
				z[n + h + i * N] = z0 - z1;
			}
}
static void haar_inv(float *z, int P, int N)
{
	const float isqrt2 = (float)(1. / sqrt(2.));
	for (int s = N, h = s / 2; s >= 2; s /= 2, h /= 2)
		for (int i = 0; i < P; i++)
			for (int n = 0; n < N; n += s) {
				const float z0 = z[n + i * N] * isqrt2;
				const float z1 = z[n + h + i * N] * isqrt2;
				z[n + i * N] = z0 + z1;
				z[n + h + i * N] = z0 - z1;
			}
}

/* computeBayesEstimateStep1 nldct.cpp:1962-2006 -> _vbm3d nldct.cpp:1799-1943.
 * g: C groups of P*N (in place).  Returns 1 if the equality early-out was
 * taken (group and weights untouched), else 0 and *weight. */
static int filter_step1(float *g, int C, int P, int N, const bases_t *b,
                        const bm4dof_params *prm, float *weight, float *tmp)
{
	if (N > 1) {
		for (int c = 0; c < C; c++)
			for (int j = 0; j < P; j++) {
				const float v = g[(size_t)c * P * N + j * N];
				for (int i = 1; i < N; i++)
					if (v != g[(size_t)c * P * N + j * N + i]) goto not_equal;
			}
		return 1;
	}
not_equal:;
	const float beta_sigma = prm->beta * prm->sigma;
	const float beta_sigma2 = beta_sigma * beta_sigma;
	const float betaM_sigma = prm->betaMean * prm->sigma;
	const float betaM_sigma2 = betaM_sigma * betaM_sigma;
	const float sigma2 = prm->sigma * prm->sigma;
	float nnz = 0;
	for (int c = 0; c < C; c++) {
		float *x = g + (size_t)c * P * N;
		sgemm_nn(tmp, x, b->basis, N, P, P);
		haar_fwd(tmp, P, N);
		if (prm->orderInvariance) {
			for (int i = 0; i < P; i++) {
				float *z = tmp + i * N;
				float var_i = 0.;
				for (int n = 1; n < N; ++n) var_i += z[n] * z[n];
				var_i /= (float)(N - 1);
				float var_0 = fmax_ref(z[0] * z[0] - sigma2, 0.f);
				float w_0 = var_0 / (var_0 + betaM_sigma2);
				z[0] *= w_0;
				var_i = fmax_ref(var_i - sigma2, 0.f);
				float w_i = var_i / (var_i + beta_sigma2);
				for (int n = 1; n < N; ++n) z[n] *= w_i;
				nnz += w_0 * w_0 + (float)(N - 1) * w_i * w_i;
			}
		} else {
			float *z = tmp;
			for (int k = 0; k < P; ++k)
				for (int i = 0; i < N; ++i, ++z)
					if (*z * *z > beta_sigma2 || (k == 0 && i == 0)) nnz += 1;
					else *z = 0.f;
		}
		haar_inv(tmp, P, N);
		sgemm_nt(x, tmp, b->binv, N, P, P);
	}
	*weight = 1 / sigma2 / (float)nnz;
	return 0;
}

/* computeBayesEstimateStep2 nldct.cpp:2196-2242 -> _vbm3d nldct.cpp:2015-2174.
 * gn, gb: C*P*N noisy / basic groups; gn filtered in place. */
static int filter_step2(float *gn, const float *gb, int C, int P, int N, const bases_t *b,
                        const bm4dof_params *prm, float *weight, float *tz, float *ty)
{
	if (N > 1) {
		for (int j = 0; j < P * C; j++) {
			const float v = gn[(size_t)j * N];
			for (int i = 1; i < N; i++)
				if (v != gn[(size_t)j * N + i]) goto not_equal;
		}
		return 1;
	}
not_equal:;
	const float beta_sigma = prm->beta * prm->sigma;
	const float beta_sigma2 = beta_sigma * beta_sigma;
	const float betaM_sigma = prm->betaMean * prm->sigma;
	const float betaM_sigma2 = betaM_sigma * betaM_sigma;
	const float sigma2 = prm->sigma * prm->sigma;
	float variance = 0.;
	for (int c = 0; c < C; c++) {
		float *xn = gn + (size_t)c * P * N;
		const float *xb = gb + (size_t)c * P * N;
		sgemm_nn(tz, xn, b->basis, N, P, P);
		sgemm_nn(ty, xb, b->basis, N, P, P);
		haar_fwd(tz, P, N);
		haar_fwd(ty, P, N);
		if (prm->orderInvariance) {
			for (int i = 0; i < P; i++) {
				float *y = ty + i * N, *z = tz + i * N;
				float var_0 = y[0] * y[0];
				float w_0 = var_0 / (var_0 + betaM_sigma2);
				z[0] *= w_0;
				float var_i = 0.;
				for (int n = 1; n < N; ++n) var_i += y[n] * y[n];
				var_i /= (float)(N - 1);
				float w_i = var_i / (var_i + beta_sigma2);
				for (int n = 1; n < N; ++n) z[n] *= w_i;
				variance += w_0 * w_0 + (float)(N - 1) * w_i * w_i;
			}
		} else {
			const float *y = ty;
			float *z = tz;
			for (int i = 0; i < P; ++i)
				for (int n = 0; n < N; ++n, ++y, ++z) {
					const float w = *y * *y / (*y * *y + beta_sigma2);
					*z *= w;
					variance += w * w;
				}
		}
		haar_inv(tz, P, N);
		sgemm_nt(xn, tz, b->binv, N, P, P);
	}
	*weight = (float)(1. / sigma2 / variance);
	return 0;
}

/* Exported single-group filters for the pin tests.  g is C*P*N; returns the
 * early-out flag (1) or 0; *weight receives the aggregation weight. */
OR_EXPORT int or_filter_group(const bm4dof_params *prm, int C, int N, float *gn,
                              const float *gb, float *weight)
{
	bases_t b;
	if (build_bases(prm, &b)) return -1;
	const int P = b.P;
	float *t1 = (float *)malloc((size_t)P * N * sizeof(float));
	float *t2 = (float *)malloc((size_t)P * N * sizeof(float));
	int r = prm->isFirstStep ? filter_step1(gn, C, P, N, &b, prm, weight, t1)
	                         : filter_step2(gn, gb, C, P, N, &b, prm, weight, t1, t2);
	free(t1); free(t2); free_bases(&b);
	return r;
}

/* ------------------------------------------------------------------------ */
/* processNlBayes on one tile, nldct.cpp:490-830                            */
/* ------------------------------------------------------------------------ */

typedef struct {
	int bx0, by0, bt0, bx1, by1, bt1;
} borders_t;

/* Reference-patch grid, nldct.cpp:513-553.  mask must be w*h*f bytes. */
static int build_mask(unsigned w, unsigned h, unsigned f, const bm4dof_params *prm,
                      const borders_t *bd, char *mask)
{
	const unsigned sWx = prm->sizeSearchWindow;
	const unsigned sWt = prm->sizeSearchTimeRangeFwd + prm->sizeSearchTimeRangeBwd + 1;
	const unsigned sPx = prm->sizePatch, sPt = prm->sizePatchTime;
	const unsigned stepx = prm->offSet, stepy = prm->offSet, stepf = prm->offSetTime;
	int n = 0;
	const int ori_x = bd->bx0 ? (int)(sPx - 1 + sWx / 2) : 0;
	const int ori_y = bd->by0 ? (int)(sPx - 1 + sWx / 2) : 0;
	const int ori_f = bd->bt0 ? (int)(sPt - 1 + sWt / 2) : 0;
	const int end_x = (int)w - (int)(bd->bx1 ? sPx - 1 + sWx / 2 : sPx - 1);
	const int end_y = (int)h - (int)(bd->by1 ? sPx - 1 + sWx / 2 : sPx - 1);
	const int end_f = (int)f - (int)(bd->bt1 ? sPt - 1 + sWt / 2 : sPt - 1);
	memset(mask, 0, (size_t)w * h * f);
	for (int t = ori_f, df = 0; t < end_f; t++, df++)
		for (int y = ori_y, dy = 0; y < end_y; y++, dy++)
			for (int x = ori_x, dx = 0; x < end_x; x++, dx++) {
				if ((df % (int)stepf == 0) || (!bd->bt1 && t == end_f - 1))
				if ((dy % (int)stepy == 0) || (!bd->by1 && y == end_y - 1) || (!bd->by0 && y == ori_y))
				if ((dx % (int)stepx == 0) || (!bd->bx1 && x == end_x - 1) || (!bd->bx0 && x == ori_x)) {
					mask[(size_t)t * w * h + (size_t)y * w + x] = 1;
					n++;
				}
			}
	return n;
}

/* One tile of one step.  noisy: tile noisy (working colour space);
 * basic: tile basic (step 2 only; the step-1 output otherwise); out: tile
 * output (resized).  Returns the number of groups processed. */
static unsigned process_tile(const vid_t *noisy, const vid_t *fflow, const vid_t *bflow,
                             vid_t *basic, vid_t *out, const bm4dof_params *prm,
                             const borders_t *bd, const bases_t *b, int tie_mode)
{
	const int step1 = prm->isFirstStep;
	const unsigned sWx = prm->sizeSearchWindow;
	const unsigned sWt = prm->sizeSearchTimeRangeFwd + prm->sizeSearchTimeRangeBwd + 1;
	const int sPx = (int)prm->sizePatch, sPt = (int)prm->sizePatchTime;
	const unsigned w = noisy->w, h = noisy->h, f = noisy->f, C = noisy->c;
	const int P = b->P;
	const unsigned patch_num = sWx * sWx * sWt;
	const size_t wh = (size_t)w * h, whf = wh * f;

	float *weight = (float *)calloc(whf ? whf : 1, sizeof(float));
	char *mask = (char *)malloc(whf ? whf : 1);
	build_mask(w, h, f, prm, bd, mask);

	dpair *dist = (dpair *)malloc((size_t)patch_num * sizeof(dpair));
	unsigned *index = (unsigned *)malloc((size_t)patch_num * sizeof(unsigned));
	const unsigned NMAX = prm->nSimilarPatches > 1 ? prm->nSimilarPatches : 1;
	float *gn = (float *)malloc((size_t)C * P * NMAX * sizeof(float));
	float *gb = (float *)malloc((size_t)C * P * NMAX * sizeof(float));
	float *t1 = (float *)malloc((size_t)P * NMAX * sizeof(float));
	float *t2 = (float *)malloc((size_t)P * NMAX * sizeof(float));
	/* aggreWeights persist across groups of the tile, initialised to 1
	 * (nldct.cpp:749, 818): the equality early-out reuses stale values. */
	float *aggw = (float *)malloc((size_t)patch_num * sizeof(float));
	for (unsigned k = 0; k < patch_num; k++) aggw[k] = 1.f;

	vid_t *dst = step1 ? basic : out;
	vid_init(dst, w, h, f, C);

	unsigned groups = 0;
	for (unsigned pt = 0; pt < f; pt++)
		for (unsigned py = 0; py < h; py++)
			for (unsigned px = 0; px < w; px++) {
				if (!mask[pt * wh + (size_t)py * w + px]) continue;
				groups++;
				const unsigned ij3 = (unsigned)vidx(noisy, px, py, pt, 0);
				unsigned N;
				float wgt;
				int early;
				if (step1) {
					N = bm_search(noisy, fflow, bflow, ij3, prm, 1, tie_mode, dist, index);
					gather_group(noisy, index, N, sPx, sPt, gn);
					early = filter_step1(gn, (int)C, P, (int)N, b, prm, &wgt, t1);
				} else {
					N = bm_search(basic, fflow, bflow, ij3, prm, 2, tie_mode, dist, index);
					gather_group(noisy, index, N, sPx, sPt, gn);
					gather_group(basic, index, N, sPx, sPt, gb);
					early = filter_step2(gn, gb, (int)C, P, (int)N, b, prm, &wgt, t1, t2);
				}
				if (!early)
					for (unsigned n = 0; n < N; n++) aggw[n] = wgt;
				/* computeAggregationStep{1,2}, nldct.cpp:2259-2425 */
				for (unsigned n = 0; n < N; n++) {
					const size_t ind = index[n];
					const size_t ind1 = (ind / dst->whc) * wh + ind % wh;
					for (int ht = 0, i = 0; ht < sPt; ht++)
						for (int hy = 0; hy < sPx; hy++)
							for (int hx = 0; hx < sPx; hx++, i++) {
								const float ww = aggw[n] * b->win[i];
								for (unsigned c = 0; c < C; c++)
									dst->d[ind + c * wh + ht * dst->whc + hy * w + hx] +=
										ww * gn[(size_t)c * P * N + (size_t)i * N + n];
								weight[ind1 + ht * wh + hy * w + hx] += ww;
							}
				}
				/* mask clear at index[0] with the multichannel index
				 * (nldct.cpp:2325-2329).  For C > 1 video this index is out of
				 * the 1-channel mask for t >= f/3; such writes are dropped
				 * here, the reference's behaviour there is undefined. */
				if (!prm->doPasteBoost && index[0] < whf) mask[index[0]] = 0;
			}

	/* computeWeightedAggregation nldct.cpp:2436-2448 */
	for (unsigned t = 0; t < f; t++)
		for (unsigned c = 0; c < C; c++)
			for (unsigned y = 0; y < h; y++)
				for (unsigned x = 0; x < w; x++) {
					const float wt = weight[t * wh + (size_t)y * w + x];
					float *o = dst->d + vidx(dst, x, y, t, c);
					if (wt > 0.f) *o /= wt;
					else *o /= noisy->d[vidx(noisy, x, y, t, c)];
				}

	free(weight); free(mask); free(dist); free(index); free(gn); free(gb);
	free(t1); free(t2); free(aggw);
	return groups;
}

/* One step over the whole video (runNlBayes body, nldct.cpp:352-475).
 * noisyw: noisy in the working colour space; basic: step-1 output (step 1)
 * or input basic in working space (step 2); out: this step's output in
 * working space. */
static int run_step(const vid_t *noisyw, const vid_t *fflow, const vid_t *bflow,
                    vid_t *basic, vid_t *out, const bm4dof_params *prm, int nparts,
                    int tie_mode, float *ratio)
{
	const int step1 = prm->isFirstStep;
	int *tl = (int *)malloc((size_t)8 * nparts * sizeof(int));
	if (or_tiles(noisyw->w, noisyw->h, (int)prm->boundary, nparts, tl)) { free(tl); return -1; }
	bases_t b;
	if (build_bases(prm, &b)) { free(tl); return -1; }

	vid_t *nsub = (vid_t *)calloc(nparts, sizeof(vid_t));
	vid_t *bsub = (vid_t *)calloc(nparts, sizeof(vid_t));
	vid_t *osub = (vid_t *)calloc(nparts, sizeof(vid_t));
	vid_t *fsub = (vid_t *)calloc(nparts, sizeof(vid_t));
	vid_t *ksub = (vid_t *)calloc(nparts, sizeof(vid_t));
	unsigned *groups = (unsigned *)calloc(nparts, sizeof(unsigned));
	const int use_flow = fflow && fflow->w > 0;

	for (int n = 0; n < nparts; n++) {
		crop_t cr = { tl[8*n], tl[8*n+1], tl[8*n+2], tl[8*n+3], tl[8*n+4], tl[8*n+5], tl[8*n+6], tl[8*n+7] };
		crop_video(noisyw, &nsub[n], &cr);
		if (!step1) crop_video(basic, &bsub[n], &cr);
		if (use_flow) { crop_video(fflow, &fsub[n], &cr); crop_video(bflow, &ksub[n], &cr); }
	}

#ifdef _OPENMP
	int nth = omp_get_max_threads();
	int chunk = nparts / nth > 0 ? nparts / nth : 1;
#pragma omp parallel for schedule(dynamic, chunk)
#endif
	for (int n = 0; n < nparts; n++) {
		const int *o = tl + 8 * n;
		borders_t bd;
		bd.bx0 = o[0] > 0; bd.by0 = o[1] > 0; bd.bt0 = 0;
		bd.bx1 = o[2] < (int)noisyw->w; bd.by1 = o[3] < (int)noisyw->h; bd.bt1 = 0;
		groups[n] = process_tile(&nsub[n], use_flow ? &fsub[n] : NULL, use_flow ? &ksub[n] : NULL,
		                         &bsub[n], &osub[n], prm, &bd, &b, tie_mode);
	}

	/* subBuildTight */
	vid_t *dst = step1 ? basic : out;
	for (int n = 0; n < nparts; n++) {
		const int *o = tl + 8 * n;
		const vid_t *src = step1 ? &bsub[n] : &osub[n];
		if (o[6] <= o[4] || o[7] <= o[5]) continue;
		for (unsigned t = 0; t < dst->f; t++)
			for (unsigned c = 0; c < dst->c; c++)
				for (int y = o[5]; y < o[7]; y++)
					memcpy(dst->d + vidx(dst, o[4], y, t, c),
					       src->d + vidx(src, o[4] - o[0], y - o[1], t, c),
					       (size_t)(o[6] - o[4]) * sizeof(float));
	}
	float r = 0.f;
	for (int n = 0; n < nparts; n++)
		r += 100.f * (float)groups[n] / (float)noisyw->whf;
	*ratio = r;

	for (int n = 0; n < nparts; n++) {
		vid_free(&nsub[n]);
		vid_free(&bsub[n]);
		vid_free(&osub[n]);
		if (use_flow) { vid_free(&fsub[n]); vid_free(&ksub[n]); }
	}
	free(nsub); free(bsub); free(osub); free(fsub); free(ksub); free(groups); free(tl);
	free_bases(&b);
	return 0;
}

/*
 * runNlBayes, nldct.cpp:267-478.  basic is in/out (input when step 1 is
 * skipped), final may be NULL when step 2 is skipped.  Returns 0 or -1
 * (message in or_last_error()).
 */
OR_EXPORT int or_run(const bm4dof_params *p1, const bm4dof_params *p2,
                     const float *noisy, unsigned w, unsigned h, unsigned f, unsigned c,
                     const float *fflow, const float *bflow, float *basic, float *final,
                     int nparts, int tie_mode, float *groups_ratio)
{
	if (!(c == 1 || c == 3 || c == 4))
		return fail("VideoNLB::runNlBayes: Wrong number of channels. Must be 1, 3 or 4.");
	const int step = p1->sizePatch ? (p2->sizePatch ? 0 : 1) : (p2->sizePatch ? 2 : -1);
	if (step == -1) return fail("VideoNLB::runNlBayes: Both patch sizes are zero.");
	if (nparts < 1) return fail("nparts must be >= 1");
	if ((step != 2 && pred_check(p1, w, h, nparts)) || (step != 1 && pred_check(p2, w, h, nparts))) return -1;
	if ((step != 2 && p1->doPasteBoost) || (step != 1 && p2->doPasteBoost))
		return fail("doPasteBoost is not supported (the reference never enables it)");
	if ((step != 2 && (p1->sizeSearchTimeRangeFwd + p1->sizeSearchTimeRangeBwd + 1 > 64 ||
	                   p1->sizePatch * p1->sizePatch * p1->sizePatchTime > 4096)) ||
	    (step != 1 && (p2->sizeSearchTimeRangeFwd + p2->sizeSearchTimeRangeBwd + 1 > 64 ||
	                   p2->sizePatch * p2->sizePatch * p2->sizePatchTime > 4096)))
		return fail("temporal search range above 64 frames or patch above 4096 samples");
	const size_t n = (size_t)w * h * f * c;

	vid_t vn, vf = {0}, vb = {0}, vbasic, vout;
	vid_init(&vn, w, h, f, c);
	int use_flow = fflow != NULL && bflow != NULL;
	if (use_flow) {
		vid_init(&vf, w, h, f, 2); memcpy(vf.d, fflow, vf.whcf * sizeof(float));
		vid_init(&vb, w, h, f, 2); memcpy(vb.d, bflow, vb.whcf * sizeof(float));
	}
	vid_init(&vbasic, w, h, f, c);
	vid_init(&vout, w, h, f, c);
	float ratio[2] = {0.f, 0.f};
	int rc = 0;

	if (step != 2) {
		memcpy(vn.d, noisy, n * sizeof(float));
		if (p1->colorSpace == BM4DOF_YUV) or_color_transform(vn.d, w, h, f, c, 1);
		rc = run_step(&vn, use_flow ? &vf : NULL, use_flow ? &vb : NULL, &vbasic, NULL, p1,
		              nparts, tie_mode, &ratio[0]);
		if (rc) goto done;
		if (p1->colorSpace == BM4DOF_YUV) or_color_transform(vbasic.d, w, h, f, c, 0);
	} else
		memcpy(vbasic.d, basic, n * sizeof(float));

	if (step != 1) {
		memcpy(vn.d, noisy, n * sizeof(float));
		if (p2->colorSpace == BM4DOF_YUV) {
			or_color_transform(vn.d, w, h, f, c, 1);
			or_color_transform(vbasic.d, w, h, f, c, 1);
		}
		rc = run_step(&vn, use_flow ? &vf : NULL, use_flow ? &vb : NULL, &vbasic, &vout, p2,
		              nparts, tie_mode, &ratio[1]);
		if (rc) goto done;
		if (p2->colorSpace == BM4DOF_YUV) {
			or_color_transform(vbasic.d, w, h, f, c, 0);
			or_color_transform(vout.d, w, h, f, c, 0);
		}
		if (final) memcpy(final, vout.d, n * sizeof(float));
	}
	memcpy(basic, vbasic.d, n * sizeof(float));
	if (groups_ratio) { groups_ratio[0] = ratio[0]; groups_ratio[1] = ratio[1]; }
done:
	vid_free(&vn); vid_free(&vbasic); vid_free(&vout);
	if (use_flow) { vid_free(&vf); vid_free(&vb); }
	return rc;
}

/*
 * Block matching only, at every grid position of every tile in processing
 * order (no mask clearing), for neighbour-set parity.  `image` is the image
 * the distance is taken on (step 1: noisy in working space; step 2: basic).
 * Outputs as bm4dof_bm() in include/bm4dof.h (global indices).  Returns the
 * group count, or -1.  Pass out arrays = NULL to only count.
 */
OR_EXPORT long or_bm(const bm4dof_params *prm, const float *image, unsigned w, unsigned h,
                     unsigned f, unsigned c, const float *fflow, const float *bflow,
                     int nparts, int tie_mode, uint32_t *out_index, uint32_t *out_nsim,
                     uint32_t *out_ref, uint32_t *out_tile)
{
	if (pred_check(prm, w, h, nparts)) return -1;
	int *tl = (int *)malloc((size_t)8 * nparts * sizeof(int));
	if (or_tiles(w, h, (int)prm->boundary, nparts, tl)) { free(tl); return -1; }
	vid_t full, ff = {0}, fb = {0};
	vid_init(&full, w, h, f, c);
	memcpy(full.d, image, full.whcf * sizeof(float));
	const int use_flow = fflow && bflow;
	if (use_flow) {
		vid_init(&ff, w, h, f, 2); memcpy(ff.d, fflow, ff.whcf * sizeof(float));
		vid_init(&fb, w, h, f, 2); memcpy(fb.d, bflow, fb.whcf * sizeof(float));
	}
	const unsigned sWx = prm->sizeSearchWindow;
	const unsigned sWt = prm->sizeSearchTimeRangeFwd + prm->sizeSearchTimeRangeBwd + 1;
	const unsigned npm = prm->nSimilarPatches;
	dpair *dist = (dpair *)malloc((size_t)sWx * sWx * sWt * sizeof(dpair));
	unsigned *index = (unsigned *)malloc((size_t)sWx * sWx * sWt * sizeof(unsigned) + 4);
	long g = 0;
	for (int n = 0; n < nparts; n++) {
		const int *o = tl + 8 * n;
		crop_t cr = { o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7] };
		vid_t sub, sf = {0}, sb = {0};
		crop_video(&full, &sub, &cr);
		if (use_flow) { crop_video(&ff, &sf, &cr); crop_video(&fb, &sb, &cr); }
		borders_t bd = { o[0] > 0, o[1] > 0, 0, o[2] < (int)w, o[3] < (int)h, 0 };
		char *mask = (char *)malloc(sub.whf);
		build_mask(sub.w, sub.h, sub.f, prm, &bd, mask);
		for (unsigned t = 0; t < sub.f; t++)
			for (unsigned y = 0; y < sub.h; y++)
				for (unsigned x = 0; x < sub.w; x++) {
					if (!mask[t * sub.wh + (size_t)y * sub.w + x]) continue;
					if (out_index) {
						const unsigned pidx = (unsigned)vidx(&sub, x, y, t, 0);
						unsigned N = bm_search(&sub, use_flow ? &sf : NULL, use_flow ? &sb : NULL,
						                       pidx, prm, prm->isFirstStep ? 1 : 2, tie_mode, dist, index);
						for (unsigned k = 0; k < npm; k++) {
							uint32_t v = 0xffffffffu;
							if (k < N) {
								const unsigned li = index[k];
								const unsigned lx = li % sub.w, ly = (li % sub.wh) / sub.w, lt = li / sub.whc;
								v = (uint32_t)vidx(&full, lx + o[0], ly + o[1], lt, 0);
							}
							out_index[(size_t)g * npm + k] = v;
						}
						out_nsim[g] = N;
						out_ref[g] = (uint32_t)vidx(&full, x + o[0], y + o[1], t, 0);
						out_tile[g] = (uint32_t)n;
					}
					g++;
				}
		free(mask);
		vid_free(&sub);
		if (use_flow) { vid_free(&sf); vid_free(&sb); }
	}
	free(dist); free(index); free(tl);
	vid_free(&full);
	if (use_flow) { vid_free(&ff); vid_free(&fb); }
	return g;
}
