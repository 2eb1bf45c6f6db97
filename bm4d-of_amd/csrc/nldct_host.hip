// Host side of the MI355X BM4D-OF denoiser: parameters, tiling, reference
// grid, bases, the per-step pipeline and the C-ABI of include/bm4dof.h.
#include "ctx.hpp"
#include "mlp_dev.hpp"


#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <memory>
#include <algorithm>
#include <stdexcept>

namespace bm4dof {

// Kernels (nldct_kernels.hip).
__global__ void color_kernel(float *, int, int, int, int, float, float, float);
__global__ void bm_kernel(const float *, const float *, const float *, const uint32_t *, const uint32_t *,
                          const TileBox *, StepConst, int, int, uint32_t *, uint32_t *, int, uint32_t *,
                          uint32_t *, int);
__global__ void bm_tie_kernel(const float *, const float *, const float *, const uint32_t *, const uint32_t *,
                              const TileBox *, StepConst, int, int, uint32_t *, uint32_t *, int, const uint32_t *,
                              const uint32_t *, int);
__global__ void bm_window_kernel(const float *, const float *, const uint32_t *, const uint32_t *,
                                 const TileBox *, StepConst, int, int4 *, int2 *);
#define BMF_ARGS const float *, const int4 *, const int2 *, const uint32_t *, const uint32_t *, \
	const TileBox *, StepConst, int, uint32_t *, uint32_t *, int, uint32_t *, uint32_t *, int
template <int PX, int PT, bool MC> __global__ void bm_reg_kernel(const float *, const int4 *, const int2 *,
                                                        const uint32_t *, const uint32_t *, const TileBox *,
                                                        StepConst, int, int, uint32_t *, uint32_t *, int,
                                                        uint32_t *, uint32_t *, int);
template <int PX, int PT, int NF> __global__ void bm_fast_kernel(BMF_ARGS);
__global__ void filter1_kernel(const float *, const uint32_t *, const uint32_t *, const uint32_t *,
                               const uint32_t *, StepConst, int, int, const float *, const float *,
                               const float *, float *, float *, uint32_t *, uint8_t *, float *,
                               uint32_t *);
__global__ void filter2_kernel(const float *, const float *, const uint32_t *, const uint32_t *,
                               const uint32_t *, const uint32_t *, StepConst, int, int,
                               const float *, const float *, const float *, float *, float *,
                               uint32_t *, uint8_t *, float *, uint32_t *);
#define FILTER1_ARGS const float *, const uint32_t *, const uint32_t *, const uint32_t *, \
	const uint32_t *, StepConst, int, int, const float *, const float *, const float *, float *, \
	float *, uint32_t *, uint8_t *, float *, uint32_t *, int, int, int64_t, float *, const float *
#define FILTER2_ARGS const float *, const float *, const uint32_t *, const uint32_t *, \
	const uint32_t *, const uint32_t *, StepConst, int, int, const float *, const float *, \
	const float *, float *, float *, uint32_t *, uint8_t *, float *, uint32_t *, int, int, int64_t, float *, \
	const float *, const float *
template <int PX> __global__ void filter1_fast_kernel(FILTER1_ARGS);
template <int PX> __global__ void coef_kernel(const float *, int, int, int, const float *, float *);
template <int PX> __global__ void filter2_fast_kernel(FILTER2_ARGS);
#define LEARNED_ARGS const float *, const uint32_t *, const uint32_t *, const uint32_t *, const uint32_t *, \
	StepConst, int, int, const uint16_t *, const float *, MlpInfer, float, float, float *, \
	float *, uint32_t *, uint8_t *, float *, uint8_t *, uint32_t *, int, int
#define PAIRS_ARGS const float *, const float *, const uint32_t *, const uint32_t *, StepConst, int, int, \
	const uint16_t *, float, float, float *, float *, uint8_t *
__global__ void split_basis_kernel(const float *, const float *, uint16_t *);
template <int PX> __global__ void filter_learned_kernel(LEARNED_ARGS);
template <int PX> __global__ void learned_pairs_kernel(PAIRS_ARGS);
__global__ void fallback_mark_kernel(const uint32_t *, const uint32_t *, const uint8_t *, StepConst, int,
                                     uint8_t *);
__global__ void fallback_blend_kernel(const uint8_t *, const float *, int64_t, int, int, float *);
__global__ void planes_kernel(const float *, int, int, int64_t, float *);
__global__ void emit_entries_kernel(const uint32_t *, const uint32_t *, const uint32_t *, int, int,
                                    int, uint32_t *);
// sortscan.hip
__global__ void fill_u32_kernel(uint32_t *, uint32_t, int64_t);
__global__ void fill_f32_kernel(float *, float, int64_t);
__global__ void fill_u8_kernel(uint8_t *, uint8_t, int64_t);
template <int OP> __global__ void scan_block_kernel(const uint32_t *, int64_t, int64_t, uint32_t *, uint32_t *);
template <int OP> __global__ void scan_top_kernel(uint32_t *, int64_t);
template <int OP> __global__ void scan_add_kernel(uint32_t *, int64_t, const uint32_t *);
template <int OP> __global__ void scan_block_w_kernel(const uint32_t *, int64_t, int64_t, uint32_t *, uint32_t *);
template <int OP> __global__ void scan_top_w_kernel(uint32_t *, int64_t);
__global__ void count_rows_kernel(const uint32_t *, const uint32_t *, int64_t, uint32_t *);
__global__ void scatter_rows_kernel(const uint32_t *, const uint32_t *, int64_t, const uint32_t *, uint32_t *,
                                    uint32_t *);
__global__ void sort_runs_kernel(const uint32_t *, int64_t, uint32_t *);
__global__ void add_rows_kernel(const uint32_t *, unsigned long long *);
__global__ void early_fix_kernel(const uint8_t *, const uint32_t *, const float *, const uint32_t *,
                                 const uint32_t *, const uint32_t *, const uint32_t *, int, int, const float *,
                                 int, int, float *, float *);
__global__ void filtered_index_kernel(const uint8_t *, int, uint32_t *);
__global__ void first_moved_kernel(const uint32_t *, const uint32_t *, int, int, uint32_t *);
__global__ void bm_pred_kernel(const float *, const float *, const float *, const uint32_t *, const uint32_t *,
                               const TileBox *, StepConst, int, int, int, float, int, int, uint32_t *,
                               uint32_t *, int);
__global__ void drop_groups_kernel(const uint32_t *, int, uint32_t *);
__global__ void first_index_kernel(const uint32_t *, int, int, uint32_t *);
#define AGG_ARGS const uint32_t *, const uint32_t *, const uint32_t *, const uint32_t *, \
	const float *, const float *, const float *, StepConst, const uint32_t *, int, int, float *, \
	float *, uint32_t *
template <int CC> __global__ void agg_exact_kernel(AGG_ARGS);
#define AGG2_ARGS const uint32_t *, const uint32_t *, const uint4 *, const float *, const float *, \
	StepConst, int, int, float *, float *, uint32_t *
__global__ void pack_rows_kernel(const uint32_t *, const uint32_t *, const float *, int64_t, uint32_t, uint32_t,
                                 uint4 *);
template <int CC, int CAP, int WP> __global__ void agg_chunk_kernel(AGG2_ARGS);
__global__ void bin_hist1_kernel(const uint32_t *, const uint32_t *, int64_t, StepConst, int, int, int, uint32_t *);
__global__ void bin_scatter1_kernel(const uint32_t *, const uint32_t *, int64_t, StepConst, int, int, int,
                                    const uint32_t *, uint32_t *, uint32_t *);
__global__ void bin_bounds_kernel(const uint32_t *, const uint32_t *, int64_t, int, uint32_t *);
__global__ void rs_hist_kernel(const uint32_t *, const uint32_t *, int, int, uint32_t *);
__global__ void rs_scatter_kernel(const uint32_t *, const uint32_t *, const uint32_t *, int, int, const uint32_t *,
                                  uint32_t *, uint32_t *);
#define AGGB_ARGS const uint32_t *, const uint32_t *, const uint4 *, const float *, const float *, StepConst, \
	int, int, int, float *, float *
template <int CC, int NBR, int PQ, int SQ> __global__ void agg_bin_fold_kernel(AGGB_ARGS);
__global__ void band_bounds_kernel(const uint32_t *, const uint32_t *, int, int, uint32_t *);
__global__ void normalize_kernel(const float *, const float *, const float *, StepConst,
                                 float *);

} // namespace bm4dof

using namespace bm4dof;

namespace {
thread_local std::string g_global_err;
}

void bm4dof::set_error(bm4dof_ctx *ctx, const std::string &s)
{
	if (ctx) ctx->err = s;
	g_global_err = s;
}

namespace {

void set_err(bm4dof_ctx *ctx, const std::string &s) { set_error(ctx, s); }

// ---------------------------------------------------------------- params

// initializeNlbParameters, src/core/nldct.cpp:81-148.  betaMean, left
// uninitialised by the reference, is set to beta.
void init_params(bm4dof_params *p, unsigned step, float sigma, unsigned frames, int verbose)
{
	const bool s1 = (step == 1);
	std::memset(p, 0, sizeof *p);
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

// ---------------------------------------------------------------- tiling

void determine_factor(unsigned n, unsigned &a, unsigned &b) // utilities.cpp:137-158
{
	if (n == 1) { a = 1; b = 1; return; }
	b = 2;
	while (n % b > 0) b++;
	a = n / b;
	if (b > a) { a = b; b = n / a; }
}

struct Tiling {
	int nW, nH, wTmp, hTmp;
	std::vector<TileBox> crop;
	std::vector<int> ix0, iy0, ix1, iy1; // interiors
};

// subDivideTight geometry, lib_videot.hpp:854-916.
Tiling make_tiling(int W, int H, int bdry, int nparts)
{
	Tiling t;
	unsigned a, b;
	determine_factor((unsigned)nparts, a, b);
	t.nW = (int)a;
	t.nH = (int)b;
	t.wTmp = (int)std::ceil((float)W / (float)t.nW);
	t.hTmp = (int)std::ceil((float)H / (float)t.nH);
	for (int p = 0; p < t.nH; p++)
		for (int q = 0; q < t.nW; q++) {
			TileBox c;
			c.ox = std::max(0, q * t.wTmp - bdry);
			c.oy = std::max(0, p * t.hTmp - bdry);
			c.ex = std::min(W, (q + 1) * t.wTmp + bdry);
			c.ey = std::min(H, (p + 1) * t.hTmp + bdry);
			if (c.ex <= c.ox || c.ey <= c.oy)
				throw ArgError("degenerate tiling: " + std::to_string(nparts) +
				               " parts leave a tile with no pixels (the reference reads out of "
				               "bounds here)");
			t.crop.push_back(c);
			t.ix0.push_back(q * t.wTmp);
			t.iy0.push_back(p * t.hTmp);
			t.ix1.push_back(std::min((q + 1) * t.wTmp, W));
			t.iy1.push_back(std::min((p + 1) * t.hTmp, H));
		}
	return t;
}

// Reference grid of one tile (processNlBayes, nldct.cpp:513-553).  The
// selection rule is a product of per-axis rules, so the raster-ordered grid
// is the cartesian product of three axis lists.
void axis_list(int len, int ori, int end, int step, bool b0, bool b1, std::vector<int> &out)
{
	out.clear();
	for (int v = ori, d = 0; v < end; v++, d++)
		if ((d % step == 0) || (!b1 && v == end - 1) || (b0 ? false : v == ori))
			out.push_back(v);
	(void)len;
}

void tile_grid(const bm4dof_params &p, int Wt, int Ht, int F, bool bx0, bool by0, bool bx1,
               bool by1, std::vector<int> &xs, std::vector<int> &ys, std::vector<int> &ts)
{
	const int sWx = (int)p.sizeSearchWindow;
	const int sWt = (int)(p.sizeSearchTimeRangeFwd + p.sizeSearchTimeRangeBwd + 1);
	const int sPx = (int)p.sizePatch, sPt = (int)p.sizePatchTime;
	const int ori_x = bx0 ? sPx - 1 + sWx / 2 : 0;
	const int ori_y = by0 ? sPx - 1 + sWx / 2 : 0;
	const int ori_f = 0; // tiles never split time (border_t0 = false)
	const int end_x = Wt - (bx1 ? sPx - 1 + sWx / 2 : sPx - 1);
	const int end_y = Ht - (by1 ? sPx - 1 + sWx / 2 : sPx - 1);
	const int end_f = F - (sPt - 1);
	(void)sWt;
	axis_list(Wt, ori_x, end_x, (int)p.offSet, bx0, bx1, xs);
	axis_list(Ht, ori_y, end_y, (int)p.offSet, by0, by1, ys);
	// time: (df % stepf == 0) || (!border_t1 && f == end_f - 1); border_t1 is false.
	ts.clear();
	for (int f = ori_f, d = 0; f < end_f; f++, d++)
		if ((d % (int)p.offSetTime == 0) || f == end_f - 1) ts.push_back(f);
}

// ---------------------------------------------------------------- bases

const double kBiorFwd[64] = { // nldct.cpp:640-647
	 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274,
	 0.219417649252501, 0.449283757993216, 0.449283757993216, 0.219417649252501,-0.219417649252501,-0.449283757993216,-0.449283757993216,-0.219417649252501,
	 0.569359398342846, 0.402347308162278,-0.402347308162278,-0.569359398342846,-0.083506045090284, 0.083506045090284,-0.083506045090284, 0.083506045090284,
	-0.083506045090284, 0.083506045090284,-0.083506045090284, 0.083506045090284, 0.569359398342846, 0.402347308162278,-0.402347308162278,-0.569359398342846,
	 0.707106781186547,-0.707106781186547, 0, 0, 0, 0, 0, 0,
	 0, 0, 0.707106781186547,-0.707106781186547, 0, 0, 0, 0,
	 0, 0, 0, 0, 0.707106781186547,-0.707106781186547, 0, 0,
	 0, 0, 0, 0, 0, 0, 0.707106781186547,-0.707106781186547 };
const double kBiorInv[64] = { // nldct.cpp:670-677
	0.353553390593274, 0.353553390593274, 0.5, 0.0, 0.707106781186547,-0.121533978016438,-0.000000000000000, 0.121533978016438,
	0.353553390593274, 0.353553390593274, 0.5,-0.0,-0.707106781186547,-0.121533978016438, 0.000000000000000, 0.121533978016438,
	0.353553390593274, 0.353553390593274,-0.5,-0.0, 0.121533978016438, 0.707106781186548,-0.121533978016438, 0.000000000000000,
	0.353553390593274, 0.353553390593274,-0.5, 0.0, 0.121533978016438,-0.707106781186548,-0.121533978016438, 0.000000000000000,
	0.353553390593274,-0.353553390593274, 0.0, 0.5, 0.000000000000000, 0.121533978016438, 0.707106781186547,-0.121533978016438,
	0.353553390593274,-0.353553390593274, 0.0, 0.5, 0.000000000000000, 0.121533978016438,-0.707106781186547,-0.121533978016438,
	0.353553390593274,-0.353553390593274, 0.0,-0.5,-0.121533978016438,-0.000000000000000, 0.121533978016438, 0.707106781186547,
	0.353553390593274,-0.353553390593274, 0.0,-0.5,-0.121533978016438,-0.000000000000000, 0.121533978016438,-0.707106781186547 };
const float kKaiser8[8] = {0.1924f, 0.2989f, 0.3846f, 0.4325f, 0.4325f, 0.3846f, 0.2989f, 0.1924f};
const float kKaiser8Rows[64] = { // nldct.cpp:706-713
	0.1924f, 0.2989f, 0.3846f, 0.4325f, 0.4325f, 0.3846f, 0.2989f, 0.1924f,
	0.2989f, 0.4642f, 0.5974f, 0.6717f, 0.6717f, 0.5974f, 0.4642f, 0.2989f,
	0.3846f, 0.5974f, 0.7688f, 0.8644f, 0.8644f, 0.7688f, 0.5974f, 0.3846f,
	0.4325f, 0.6717f, 0.8644f, 0.9718f, 0.9718f, 0.8644f, 0.6717f, 0.4325f,
	0.4325f, 0.6717f, 0.8644f, 0.9718f, 0.9718f, 0.8644f, 0.6717f, 0.4325f,
	0.3846f, 0.5974f, 0.7688f, 0.8644f, 0.8644f, 0.7688f, 0.5974f, 0.3846f,
	0.2989f, 0.4642f, 0.5974f, 0.6717f, 0.6717f, 0.5974f, 0.4642f, 0.2989f,
	0.1924f, 0.2989f, 0.3846f, 0.4325f, 0.4325f, 0.3846f, 0.2989f, 0.1924f };
const float kKaiser10Rows[100] = { // nldct.cpp:720-729
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

// Patch bases, processNlBayes set-up (nldct.cpp:571-737): basis[k*P+pix],
// binv[k*P+pix] (the matrix the inverse productMatrix reads with transB),
// window[pix].
void build_bases(const bm4dof_params &p, std::vector<float> &basis, std::vector<float> &binv,
                 std::vector<float> &win)
{
	const unsigned sPx = p.sizePatch, sPt = p.sizePatchTime;
	const size_t P = (size_t)sPx * sPx * sPt;
	basis.assign(P * P, 0.f);
	binv.assign(P * P, 0.f);
	win.assign(P, 1.f);
	if (p.transform == BM4DOF_DCT) {
		std::vector<float> cosx(sPx * sPx), cost(sPt * sPt);
		for (unsigned kx = 0, i = 0; kx < sPx; kx++)
			for (unsigned nx = 0; nx < sPx; nx++, i++)
				cosx[i] = (float)(std::cos((0.5 + (float)nx) * (float)kx * M_PI / (float)sPx) *
				                  std::sqrt((kx == 0 ? 1. : 2.) / (float)sPx));
		for (unsigned kt = 0, i = 0; kt < sPt; kt++)
			for (unsigned nt = 0; nt < sPt; nt++, i++)
				cost[i] = (float)(std::cos((0.5 + (float)nt) * (float)kt * M_PI / (float)sPt) *
				                  std::sqrt((kt == 0 ? 1. : 2.) / (float)sPt));
		size_t i = 0;
		for (unsigned kt = 0; kt < sPt; kt++)
			for (unsigned ky = 0; ky < sPx; ky++)
				for (unsigned kx = 0; kx < sPx; kx++)
					for (unsigned nt = 0; nt < sPt; nt++)
						for (unsigned ny = 0; ny < sPx; ny++)
							for (unsigned nx = 0; nx < sPx; nx++, i++) {
								const float v = cosx[kx * sPx + nx] * cosx[ky * sPx + ny];
								basis[i] = v * cost[kt * sPt + nt];
							}
		binv = basis;
	} else {
		if (sPx != 8)
			throw ArgError("VideoNLB::processNlBayes: bior1_5 transform is only implemented for "
			               "patches of size 8 x 8 x pt.");
		float bx[64], bxi[64];
		for (int k = 0; k < 64; k++) { bx[k] = (float)kBiorFwd[k]; bxi[k] = (float)kBiorInv[k]; }
		size_t i = 0;
		for (unsigned kt = 0; kt < sPt; kt++)
			for (unsigned ky = 0; ky < 8; ky++)
				for (unsigned kx = 0; kx < 8; kx++)
					for (unsigned nt = 0; nt < sPt; nt++)
						for (unsigned ny = 0; ny < 8; ny++)
							for (unsigned nx = 0; nx < 8; nx++, i++) {
								const float v = bx[kx * 8 + nx] * bx[ky * 8 + ny];
								basis[i] = v * (nt == kt ? 1.f : 0.f);
							}
		i = 0;
		for (unsigned nt = 0; nt < sPt; nt++)
			for (unsigned ny = 0; ny < 8; ny++)
				for (unsigned nx = 0; nx < 8; nx++)
					for (unsigned kt = 0; kt < sPt; kt++)
						for (unsigned ky = 0; ky < 8; ky++)
							for (unsigned kx = 0; kx < 8; kx++, i++) {
								const float v = bxi[kx * 8 + nx] * bxi[ky * 8 + ny];
								binv[i] = v * (nt == kt ? 1.f : 0.f);
							}
	}
	if (p.agg_window && sPx == 8)
		for (unsigned t = 0; t < sPt; t++) std::copy(kKaiser8Rows, kKaiser8Rows + 64, win.begin() + 64 * t);
	else if (p.agg_window && sPx == 10)
		for (unsigned t = 0; t < sPt; t++) std::copy(kKaiser10Rows, kKaiser10Rows + 100, win.begin() + 100 * t);
	(void)kKaiser8;
}

// ---------------------------------------------------------------- pipeline

struct Timer {
	bm4dof_ctx *ctx;
	explicit Timer(bm4dof_ctx *c) : ctx(c) {}
	void mark(const char *name) {
		hipEvent_t e;
		HIPCHK(hipEventCreate(&e));
		HIPCHK(hipEventRecord(e, ctx->stream));
		ctx->ev.push_back(e);
		ctx->timings.push_back({name, 0.f});
	}
};

void check_step_params(const bm4dof_params &p, unsigned C)
{
	if (p.doPasteBoost)
		throw ArgError("doPasteBoost is not supported (the reference never enables it)", BM4DOF_EUNSUPP);
	if (p.nSimilarPatches > 64)
		throw ArgError("nSimilarPatches above 64 is not supported", BM4DOF_EUNSUPP);
	const unsigned P = p.sizePatch * p.sizePatch * p.sizePatchTime;
	if (P == 0 || P > 256)
		throw ArgError("patch of more than 256 samples is not supported", BM4DOF_EUNSUPP);
	if (p.sizeSearchTimeRangeFwd + p.sizeSearchTimeRangeBwd + 1 > 64)
		throw ArgError("temporal search range above 64 frames is not supported", BM4DOF_EUNSUPP);
	if (p.offSet == 0 || p.offSetTime == 0) throw ArgError("offSet and offSetTime must be positive");
	if (p.sizeSearchWindow == 0) throw ArgError("sizeSearchWindow must be positive");
	// patch positions overlapping an 8 x 8 pixel block of one frame
	if ((p.sizePatch + 7) * (p.sizePatch + 7) * p.sizePatchTime > 512)
		throw ArgError("patch too large for the ordered aggregation ((px+7)^2*pt > 512)", BM4DOF_EUNSUPP);
	(void)C;
}

struct StepGeom {
	Tiling til;
	std::vector<uint32_t> gpos, gtile;
	std::vector<uint32_t> tile_first;       // first group of each tile
	std::vector<uint32_t> frame_first;      // first group of every (tile, reference frame)
};

// Reference positions of a step, tile by tile in the reference's processing
// order (nldct.cpp:756-760).  For C > 1 the mask clear after each group,
// io_mask(index[0]) = false (nldct.cpp:2325-2329), is applied: index[0] is
// the reference patch itself (it comes first in the neighbour order) but in
// the C-channel index t*whc + y*w + x, so it clears grid position
// (x, y, C*t) of the one-channel mask, which is then skipped; indices past
// the mask (t >= F/C) are dropped (undefined in the reference).  For C = 1
// the clear only touches the group's own position.  `mask_clear` false gives
// every grid position (the neighbour-list entry points).
StepGeom make_geometry(const bm4dof_params &p, int W, int H, int F, int nparts, int C = 1,
                       bool mask_clear = true)
{
	StepGeom g;
	g.til = make_tiling(W, H, (int)p.boundary, nparts);
	std::vector<int> xs, ys, ts;
	std::vector<uint8_t> mask;
	const uint32_t WH = (uint32_t)W * (uint32_t)H;
	for (int n = 0; n < nparts; n++) {
		const TileBox &c = g.til.crop[n];
		g.tile_first.push_back((uint32_t)g.gpos.size());
		const int wc = c.ex - c.ox, hc = c.ey - c.oy;
		tile_grid(p, wc, hc, F, c.ox > 0, c.oy > 0, c.ex < W, c.ey < H, xs, ys, ts);
		const bool clear = mask_clear && C > 1 && !p.doPasteBoost;
		const size_t whc = (size_t)wc * hc, whf = whc * F;
		if (clear) {
			mask.assign(whf, 0);
			for (int t : ts)
				for (int y : ys)
					for (int x : xs) mask[(size_t)t * whc + (size_t)y * wc + x] = 1;
		}
		for (int t : ts) {
			const uint32_t before = (uint32_t)g.gpos.size();
			for (int y : ys)
				for (int x : xs) {
					if (clear) {
						if (!mask[(size_t)t * whc + (size_t)y * wc + x]) continue;
						const size_t i0 = (size_t)t * whc * C + (size_t)y * wc + x;
						if (i0 < whf) mask[i0] = 0;
					}
					g.gpos.push_back((uint32_t)t * WH + (uint32_t)(y + c.oy) * (uint32_t)W + (uint32_t)(x + c.ox));
					g.gtile.push_back((uint32_t)n);
				}
			if (g.gpos.size() > before) g.frame_first.push_back(before);
		}
	}
	return g;
}

StepConst make_const(const bm4dof_params &p, int W, int H, int F, int C, bool use_flow,
                     const Tiling &t)
{
	StepConst sc;
	sc.W = W; sc.H = H; sc.F = F; sc.C = C;
	sc.WH = W * H;
	sc.WHC = W * H * C;
	sc.sPx = (int)p.sizePatch;
	sc.sPt = (int)p.sizePatchTime;
	sc.P = sc.sPx * sc.sPx * sc.sPt;
	sc.sWx = (int)p.sizeSearchWindow;
	sc.wtf = (int)p.sizeSearchTimeRangeFwd;
	sc.wtb = (int)p.sizeSearchTimeRangeBwd;
	sc.npmax = (int)std::max(1u, p.nSimilarPatches);
	sc.tau2 = p.tau * p.tau * sc.sPx * sc.sPx * sc.sPt;
	const float beta_sigma = p.beta * p.sigma;
	sc.beta_sigma2 = beta_sigma * beta_sigma;
	const float betaM_sigma = p.betaMean * p.sigma;
	sc.betaM_sigma2 = betaM_sigma * betaM_sigma;
	sc.sigma2 = p.sigma * p.sigma;
	sc.order_inv = p.orderInvariance ? 1 : 0;
	sc.use_flow = use_flow ? 1 : 0;
	sc.nW = t.nW;
	sc.wTmp = t.wTmp;
	sc.hTmp = t.hTmp;
	sc.pred_w = (int)p.sizeSearchWindowPred;
	sc.pred_np = (int)p.nSimilarPatchesPred;
	sc.pred_dsub = p.dsub * p.dsub * 255;   // nldct.cpp:920, 1387
	return sc;
}

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Colour transform on device, constants as transformColorSpace computes them.
void color(bm4dof_ctx *ctx, float *v, int W, int H, int F, int C, bool forward)
{
	if (C == 1) return;
	float a, b, c;
	if (C == 3) {
		a = 1.f / sqrtf(3.f);
		b = 1.f / sqrtf(2.f);
		c = forward ? 2.f * a * sqrtf(2.f) : a / b;
	} else {
		a = 0.5f;
		b = 1.f / sqrtf(2.f);
		c = 0.f;
	}
	const int64_t n = (int64_t)W * H * F;
	hipLaunchKernelGGL(color_kernel, dim3(std::min<int64_t>(nblk(n, 256), 65535 * 4)), dim3(256), 0,
	                   ctx->stream, v, W * H, C, F, forward ? 1 : 0, a, b, c);
	HIPCHK(hipGetLastError());
}

void fill_u32(bm4dof_ctx *ctx, uint32_t *p, uint32_t v, int64_t n, hipStream_t st = nullptr);

// The listed groups matched again with the reference's selection; a fixed
// grid that loops over the device-side count (nothing to read back).
static void bm_ties(bm4dof_ctx *ctx, const StepConst &sc, int step, const float *img, const float *ff,
                    const float *bf, const uint32_t *d_gpos, const uint32_t *d_gtile, const TileBox *d_tiles, int G,
                    uint32_t *d_idx, uint32_t *d_nsim, const uint32_t *d_ties, int maxCand, size_t lds)
{
	hipLaunchKernelGGL(bm_tie_kernel, dim3((unsigned)std::min(G, 2048)), dim3(256), lds, ctx->stream, img, ff, bf,
	                   d_gpos, d_gtile, d_tiles, sc, step, maxCand, d_idx, d_nsim, G, d_ties, d_ties + 1, G);
	HIPCHK(hipGetLastError());
}

// Block matching of all groups of a step.
void run_bm(bm4dof_ctx *ctx, const StepConst &sc, int step, const float *img, const float *ff,
            const float *bf, const uint32_t *d_gpos, const uint32_t *d_gtile, const TileBox *d_tiles,
            int G, uint32_t *d_idx, uint32_t *d_nsim)
{
	const int sWt = sc.wtf + sc.wtb + 1;
	const int maxCand = sc.sWx * sc.sWx * sWt;
	const size_t lds = (size_t)maxCand * 12 + (size_t)sc.C * sc.P * 4;
	if (lds > 60 * 1024) throw ArgError("search region too large for the block-matching kernel", BM4DOF_EUNSUPP);
	if (G == 0) return;
	if (ctx->vbm3d_search && sc.npmax > 1) {
		// predictive search: libstdc++'s selection throughout, no tie pass
		const int cap = std::max(sc.sWx * sc.sWx, sc.pred_np * sc.pred_w * sc.pred_w);
		const int kcap = sc.pred_np * sWt;
		const size_t lds_p = ((size_t)2 * cap + (size_t)4 * kcap + (size_t)sc.P) * 4;
		if (lds_p > 60 * 1024) throw ArgError("predictive search regions too large", BM4DOF_EUNSUPP);
		hipLaunchKernelGGL(bm_pred_kernel, dim3(G), dim3(64), lds_p, ctx->stream, img, ff, bf, d_gpos, d_gtile,
		                   d_tiles, sc, step == 1 ? 1 : sc.C, sc.pred_w, sc.pred_np, sc.pred_dsub, cap, kcap,
		                   d_idx, d_nsim, G);
		HIPCHK(hipGetLastError());
		return;
	}
	// groups whose list depends on the order of equal distances: counted in
	// ties[0], listed in ties[1..G], then matched again in the reference's
	// own order (bm_tie_kernel)
	uint32_t *d_ties = ctx->ties.as<uint32_t>((size_t)G + 1);
	fill_u32(ctx, d_ties, 0u, 1);
	// one-wave kernels, patches 4x4x4 or 8x8x1: step 1 matches on channel 0,
	// step 2 sums all channels (the register kernel takes any channel count,
	// the LDS-staged one a single channel)
	const int PX = sc.sPx, PT = sc.sPt;
	const bool shape_ok = (PX == 4 && PT == 4) || (PX == 8 && PT == 1);
	const int NF = ctx->bm_frames;   // frames staged per round: 1, 2 or 4
	const size_t stage_floats = std::max<size_t>((size_t)NF * (sc.sWx + PX - 1) * (sc.sWx + PX - 1) * PT,
	                                             (size_t)maxCand);
	const size_t lds_fast = ((size_t)maxCand + stage_floats) * 4;
	const size_t lds_reg = (size_t)maxCand * 8;
	if (ctx->fast_bm == 2 && shape_ok && lds_reg <= 56 * 1024) {
		int4 *d_wins = ctx->bm_wins.as<int4>((size_t)G * sWt);
		int2 *d_wmeta = ctx->bm_wmeta.as<int2>((size_t)G);
		hipLaunchKernelGGL(bm_window_kernel, dim3(nblk(G, 256)), dim3(256), 0, ctx->stream, ff, bf, d_gpos,
		                   d_gtile, d_tiles, sc, G, d_wins, d_wmeta);
		HIPCHK(hipGetLastError());
		// row chunks per candidate column: enough work items for a wave
		const int rchunks = std::max(1, std::min(sc.sWx, 64 / std::max(1, sWt * sc.sWx)));
		const int nch = step == 1 ? 1 : sc.C;
		auto kern = nch == 1 ? (PX == 4 ? bm_reg_kernel<4, 4, false> : bm_reg_kernel<8, 1, false>)
		                     : (PX == 4 ? bm_reg_kernel<4, 4, true> : bm_reg_kernel<8, 1, true>);
		hipLaunchKernelGGL(kern, dim3(G), dim3(64), lds_reg, ctx->stream, img, d_wins, d_wmeta, d_gpos, d_gtile,
		                   d_tiles, sc, rchunks, nch, d_idx, d_nsim, G, d_ties, d_ties + 1, G);
		HIPCHK(hipGetLastError());
		bm_ties(ctx, sc, step, img, ff, bf, d_gpos, d_gtile, d_tiles, G, d_idx, d_nsim, d_ties, maxCand, lds);
		return;
	}
	if (ctx->fast_bm && shape_ok && (step == 1 || sc.C == 1) && lds_fast <= 56 * 1024) {
		int4 *d_wins = ctx->bm_wins.as<int4>((size_t)G * sWt);
		int2 *d_wmeta = ctx->bm_wmeta.as<int2>((size_t)G);
		hipLaunchKernelGGL(bm_window_kernel, dim3(nblk(G, 256)), dim3(256), 0, ctx->stream, ff, bf, d_gpos,
		                   d_gtile, d_tiles, sc, G, d_wins, d_wmeta);
		HIPCHK(hipGetLastError());
		auto kern = PX == 4 ? bm_fast_kernel<4, 4, 2> : bm_fast_kernel<8, 1, 2>;
		if (NF == 1) kern = PX == 4 ? bm_fast_kernel<4, 4, 1> : bm_fast_kernel<8, 1, 1>;
		else if (NF == 4) kern = PX == 4 ? bm_fast_kernel<4, 4, 4> : bm_fast_kernel<8, 1, 4>;
		hipLaunchKernelGGL(kern, dim3(G), dim3(64), lds_fast,
		                   ctx->stream, img, d_wins, d_wmeta, d_gpos, d_gtile, d_tiles, sc, (int)stage_floats,
		                   d_idx, d_nsim, G, d_ties, d_ties + 1, G);
		HIPCHK(hipGetLastError());
		bm_ties(ctx, sc, step, img, ff, bf, d_gpos, d_gtile, d_tiles, G, d_idx, d_nsim, d_ties, maxCand, lds);
		return;
	}
	hipLaunchKernelGGL(bm_kernel, dim3(G), dim3(256), lds, ctx->stream, img, ff, bf, d_gpos, d_gtile,
	                   d_tiles, sc, step, maxCand, d_idx, d_nsim, G, d_ties, d_ties + 1, G);
	HIPCHK(hipGetLastError());
	bm_ties(ctx, sc, step, img, ff, bf, d_gpos, d_gtile, d_tiles, G, d_idx, d_nsim, d_ties, maxCand, lds);
}

void fill_u32(bm4dof_ctx *ctx, uint32_t *p, uint32_t v, int64_t n, hipStream_t st)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(fill_u32_kernel, dim3(nblk(n, 256)), dim3(256), 0, st ? st : ctx->stream, p, v, n);
	HIPCHK(hipGetLastError());
}
void fill_f32(bm4dof_ctx *ctx, float *p, float v, int64_t n)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(fill_f32_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, p, v, n);
	HIPCHK(hipGetLastError());
}

// Grid of a second-stream kernel: at most side_blocks one-wave blocks (the
// kernels loop over their work), so that they take few of the wave slots
// the filter beside them uses; 0 = one block per unit of work.
static unsigned side_grid(const bm4dof_ctx *ctx, int64_t blocks)
{
	if (ctx->side_blocks > 0) blocks = std::min<int64_t>(blocks, ctx->side_blocks);
	return (unsigned)std::max<int64_t>(1, blocks);
}

// out[0..n) = scan of in[0..n_in) (identity past n_in): op 0 exclusive sum,
// op 1 inclusive maximum.  Needs ctx->scan_tmp reserved (reserve_scan).
void reserve_scan(bm4dof_ctx *ctx, int64_t n)
{
	(void)ctx->scan_tmp.as<uint32_t>((size_t)(n / 1024 + 2));
	(void)ctx->scan_tmp2.as<uint32_t>((size_t)(n / 1024 + 2));
}
// On the side stream (side = true) it uses its own block-sum buffer.
void dscan(bm4dof_ctx *ctx, int op, const uint32_t *in, int64_t n_in, int64_t n, uint32_t *out,
           bool side = false)
{
	if (n <= 0) return;
	const int64_t nb = (n + 1023) / 1024;
	uint32_t *bsum = (side ? ctx->scan_tmp2 : ctx->scan_tmp).as<uint32_t>((size_t)nb + 2);
	hipStream_t st = side ? ctx->side : ctx->stream;
	if (side) {
		// one-wave blocks on the second stream, so that they get dispatched
		// beside the one-wave blocks of the filter
		if (op == 0) {
			hipLaunchKernelGGL(scan_block_w_kernel<0>, dim3(side_grid(ctx, nb)), dim3(64), 0, st, in, n_in, n, out, bsum);
			hipLaunchKernelGGL(scan_top_w_kernel<0>, dim3(1), dim3(64), 0, st, bsum, nb);
			hipLaunchKernelGGL(scan_add_kernel<0>, dim3(side_grid(ctx, nblk(n, 64))), dim3(64), 0, st, out, n, bsum);
		} else {
			hipLaunchKernelGGL(scan_block_w_kernel<1>, dim3(side_grid(ctx, nb)), dim3(64), 0, st, in, n_in, n, out, bsum);
			hipLaunchKernelGGL(scan_top_w_kernel<1>, dim3(1), dim3(64), 0, st, bsum, nb);
			hipLaunchKernelGGL(scan_add_kernel<1>, dim3(side_grid(ctx, nblk(n, 64))), dim3(64), 0, st, out, n, bsum);
		}
	} else if (op == 0) {
		hipLaunchKernelGGL(scan_block_kernel<0>, dim3((unsigned)nb), dim3(256), 0, st, in, n_in, n, out, bsum);
		hipLaunchKernelGGL(scan_top_kernel<0>, dim3(1), dim3(256), 0, st, bsum, nb);
		hipLaunchKernelGGL(scan_add_kernel<0>, dim3(nblk(n, 256)), dim3(256), 0, st, out, n, bsum);
	} else {
		hipLaunchKernelGGL(scan_block_kernel<1>, dim3((unsigned)nb), dim3(256), 0, st, in, n_in, n, out, bsum);
		hipLaunchKernelGGL(scan_top_kernel<1>, dim3(1), dim3(256), 0, st, bsum, nb);
		hipLaunchKernelGGL(scan_add_kernel<1>, dim3(nblk(n, 256)), dim3(256), 0, st, out, n, bsum);
	}
	HIPCHK(hipGetLastError());
}

// Binned aggregation: shapes it takes (64-sample patches 4x4x4 or 8x8x1,
// one block overlapping at most 2 x 2 tile interiors) and the list sizes
// (at most 8 bins per row, pair counts in 32 bits; rows and frames below
// 2^16, as the row records hold them in 16 bits each).
bool agg_binned_ok(const bm4dof_ctx *ctx, const StepConst &sc, int64_t Mub)
{
	return ctx->agg_binned && ctx->agg_chunked && sc.wTmp >= 8 && sc.hTmp >= 8 &&
	       ((sc.sPx == 4 && sc.sPt == 4) || (sc.sPx == 8 && sc.sPt == 1)) &&
	       (sc.C == 1 || sc.C == 3 || sc.C == 4) && sc.H < 65536 && sc.F < 65536 &&
	       8 * Mub < ((int64_t)1 << 32) - RS_TILE;
}

struct BinGrid { int nbx, nby, nbins; };
BinGrid bin_grid(const StepConst &sc)
{
	BinGrid g;
	g.nbx = (sc.W + 7) / 8;
	g.nby = (sc.H + 7) / 8;
	g.nbins = g.nbx * g.nby * ((sc.F + AGG_TB - 1) / AGG_TB);
	return g;
}

// Device buffers of the bin lists for Mub rows: the first pass's digit
// offsets per row tile (and the pair count after them), pairs (two key /
// value buffers, 8 Mub each), digit counts per radix tile, list starts per
// bin.  The sorted values end in bin_list (v0).
struct BinBufs {
	uint32_t *off1, *k0, *v0, *k1, *v1, *hist, *hoff, *boff;
	int nt1, nblk;   // tiles of the first pass (rows) and of the later ones (pairs)
};
BinBufs bin_bufs(bm4dof_ctx *ctx, const StepConst &sc, int64_t Mub)
{
	const BinGrid bg = bin_grid(sc);
	const size_t npair = (size_t)8 * Mub;
	BinBufs q;
	q.nblk = (int)((npair + RS_TILE - 1) / RS_TILE);
	q.nt1 = (int)((Mub + BIN_RT - 1) / BIN_RT);
	q.off1 = ctx->bin_cnt.as<uint32_t>((size_t)256 * q.nt1 + 1);
	q.k0 = ctx->bin_k0.as<uint32_t>(npair);
	q.v0 = ctx->bin_list.as<uint32_t>(npair);
	q.k1 = ctx->bin_k1.as<uint32_t>(npair);
	q.v1 = ctx->bin_v1.as<uint32_t>(npair);
	q.hist = ctx->bin_hist.as<uint32_t>((size_t)256 * q.nblk);
	q.hoff = ctx->bin_hoff.as<uint32_t>((size_t)256 * q.nblk + 1);
	q.boff = ctx->bin_off.as<uint32_t>((size_t)bg.nbins + 1);
	return q;
}

// Each bin's rows in increasing e (on the side stream, beside the filter):
// (bin, e) pairs in increasing e, a stable radix sort by bin, list starts.
void agg_bin_lists(bm4dof_ctx *ctx, const StepConst &sc, int64_t Mub, const uint32_t *epos, const uint32_t *Mp)
{
	const BinGrid bg = bin_grid(sc);
	const BinBufs q = bin_bufs(ctx, sc, Mub);
	hipStream_t st = ctx->side;
	int bits = 1;
	while ((1ll << bits) < (int64_t)bg.nbins) bits++;
	const int passes = (bits + 7) / 8;
	// the passes alternate between the buffers and end in bin_k0 / bin_list
	uint32_t *kout = passes % 2 ? q.k0 : q.k1, *vout = passes % 2 ? q.v0 : q.v1;
	// pass 0 from the rows; off1[256 nt1] = the pair count
	hipLaunchKernelGGL(bin_hist1_kernel, dim3(side_grid(ctx, q.nt1)), dim3(64), 0, st, epos, Mp, Mub, sc, bg.nbx,
	                   bg.nby, q.nt1, q.hist);
	dscan(ctx, 0, q.hist, (int64_t)256 * q.nt1, (int64_t)256 * q.nt1 + 1, q.off1, true);
	const uint32_t *np = q.off1 + (size_t)256 * q.nt1;
	hipLaunchKernelGGL(bin_scatter1_kernel, dim3(side_grid(ctx, q.nt1)), dim3(64), 0, st, epos, Mp, Mub, sc,
	                   bg.nbx, bg.nby, q.nt1, q.off1, kout, vout);
	HIPCHK(hipGetLastError());
	uint32_t *kin = kout, *vin = vout;
	kout = kin == q.k0 ? q.k1 : q.k0;
	vout = vin == q.v0 ? q.v1 : q.v0;
	for (int ps = 1; ps < passes; ps++) {
		hipLaunchKernelGGL(rs_hist_kernel, dim3(side_grid(ctx, q.nblk)), dim3(RS_THREADS), 0, st, kin, np, 8 * ps,
		                   q.nblk, q.hist);
		dscan(ctx, 0, q.hist, (int64_t)256 * q.nblk, (int64_t)256 * q.nblk, q.hoff, true);
		hipLaunchKernelGGL(rs_scatter_kernel, dim3(side_grid(ctx, q.nblk)), dim3(RS_THREADS), 0, st, kin, vin, np, 8 * ps, q.nblk,
		                   q.hoff, kout, vout);
		HIPCHK(hipGetLastError());
		std::swap(kin, kout);
		std::swap(vin, vout);
	}
	const int64_t nub = 8 * Mub;
	hipLaunchKernelGGL(bin_bounds_kernel, dim3(side_grid(ctx, (nub + 1 + 1023) / 1024)), dim3(64), 0, st,
	                   q.k0, np, nub, bg.nbins, q.boff);
	HIPCHK(hipGetLastError());
}

// Adds a chunk's rows, in row order, into the running float sums.  The row
// count is on the device (*Mp); Mub bounds it.  Bands (row ranges processed
// one after the other, banded kernel only) start at the first group of every
// reference frame of every tile inside the chunk.
void aggregate_chunk(bm4dof_ctx *ctx, const StepConst &sc, const std::vector<uint32_t> &band_groups,
                     int g0, int64_t Mub, const uint32_t *Mp, int t0c, int t1c, const uint32_t *d_base,
                     const uint32_t *kstart, const uint32_t *svals, const uint32_t *epos,
                     const uint32_t *etile, const float *ew, const float *est, const float *win,
                     float *num, float *wsum, uint32_t *d_nsplit)
{
	if (Mub <= 0 || t1c <= t0c) return;
	const dim3 grid((sc.W + 7) / 8, (sc.H + 7) / 8, t1c - t0c);
	// chunked kernel: a block then spans at most 2 x 2 tile interiors
	const bool chunked = ctx->agg_chunked && sc.wTmp >= 8 && sc.hTmp >= 8;
	if (chunked) {
		uint4 *d_erec = ctx->erec.as<uint4>((size_t)Mub);
		hipLaunchKernelGGL(pack_rows_kernel, dim3(nblk(Mub, 256)), dim3(256), 0, ctx->stream, epos, etile, ew, Mub,
		                   (uint32_t)sc.W, (uint32_t)sc.WH, d_erec);
		HIPCHK(hipGetLastError());
		if (t0c == 0 && t1c == sc.F && agg_binned_ok(ctx, sc, Mub)) {
			// the bin lists were built on the side stream (agg_bin_lists)
			const BinGrid bg = bin_grid(sc);
			const BinBufs q = bin_bufs(ctx, sc, Mub);
			const uint32_t *boff = q.boff, *blist = q.v0;
			auto fold = agg_bin_fold_kernel<1, 32, 16, 4>;
			if (sc.sPx == 4) {
				if (sc.C == 3) fold = agg_bin_fold_kernel<3, 16, 16, 4>;
				else if (sc.C == 4) fold = agg_bin_fold_kernel<4, 16, 16, 4>;
			} else {
				fold = sc.C == 1 ? agg_bin_fold_kernel<1, 32, 16, 16>
				                 : sc.C == 3 ? agg_bin_fold_kernel<3, 16, 16, 16> : agg_bin_fold_kernel<4, 16, 16, 16>;
			}
			const unsigned nb8 = (unsigned)(((size_t)bg.nbins + 7) / 8 * 8);
			hipLaunchKernelGGL(fold, dim3(nb8), dim3(64), 0, ctx->stream, boff, blist, d_erec, est, win, sc, bg.nbx,
			                   bg.nby, bg.nbins, num, wsum);
			HIPCHK(hipGetLastError());
			ctx->agg_blocks += (uint64_t)bg.nbins;
			ctx->agg_region[0] = 8;
			ctx->agg_region[1] = 8;
			ctx->agg_region[2] = AGG_TB;
			return;
		}
		const dim3 cgrid((unsigned)(((size_t)(t1c - t0c) * ((sc.W + 7) / 8) * ((sc.H + 7) / 8) + 7) / 8 * 8));
		// rows held in LDS per e-range: 512 / 768 / 1024 / 1536 (one channel,
		// patches of up to 64 samples); window table of 64 or 256 samples
		auto kern = agg_chunk_kernel<1, 1024, 64>;
		if (sc.P > 64)
			kern = sc.C == 1 ? agg_chunk_kernel<1, 1024, 256>
			                 : sc.C == 3 ? agg_chunk_kernel<3, 1024, 256> : agg_chunk_kernel<4, 1024, 256>;
		else if (sc.C == 1 && ctx->agg_cap == 512) kern = agg_chunk_kernel<1, 512, 64>;
		else if (sc.C == 1 && ctx->agg_cap == 768) kern = agg_chunk_kernel<1, 768, 64>;
		else if (sc.C == 1 && ctx->agg_cap == 1536) kern = agg_chunk_kernel<1, 1536, 64>;
		else if (sc.C == 3) kern = agg_chunk_kernel<3, 1024, 64>;
		else if (sc.C == 4) kern = agg_chunk_kernel<4, 1024, 64>;
		hipLaunchKernelGGL(kern, cgrid, dim3(64), 0, ctx->stream, kstart, svals, d_erec, est, win, sc, t0c,
		                   t1c - t0c, num, wsum, d_nsplit);
		HIPCHK(hipGetLastError());
		ctx->agg_blocks += (uint64_t)grid.x * grid.y * grid.z;
		ctx->agg_region[0] = 8;
		ctx->agg_region[1] = 8;
		ctx->agg_region[2] = 1;
		return;
	}
	const int nb = (int)band_groups.size() + 1;
	uint32_t *d_bands = ctx->elist.as<uint32_t>(2 * (size_t)nb + 2);
	uint32_t *d_gfirst = d_bands + nb + 1;
	fill_u32(ctx, d_bands, 0u, 1);
	HIPCHK(hipMemcpyAsync(d_bands + nb, Mp, 4, hipMemcpyDeviceToDevice, ctx->stream));
	if (nb > 1) {
		HIPCHK(hipMemcpyAsync(d_gfirst, band_groups.data(), band_groups.size() * 4, hipMemcpyHostToDevice,
		                      ctx->stream));
		hipLaunchKernelGGL(band_bounds_kernel, dim3(nblk(nb - 1, 256)), dim3(256), 0, ctx->stream, d_base,
		                   d_gfirst, nb - 1, g0, d_bands + 1);
		HIPCHK(hipGetLastError());
	}
	if (sc.C == 1)
		hipLaunchKernelGGL(agg_exact_kernel<1>, grid, dim3(64), 0, ctx->stream, kstart, svals, epos, etile, ew,
		                   est, win, sc, d_bands, nb, t0c, num, wsum, d_nsplit);
	else if (sc.C == 3)
		hipLaunchKernelGGL(agg_exact_kernel<3>, grid, dim3(64), 0, ctx->stream, kstart, svals, epos, etile, ew,
		                   est, win, sc, d_bands, nb, t0c, num, wsum, d_nsplit);
	else
		hipLaunchKernelGGL(agg_exact_kernel<4>, grid, dim3(64), 0, ctx->stream, kstart, svals, epos, etile, ew,
		                   est, win, sc, d_bands, nb, t0c, num, wsum, d_nsplit);
	HIPCHK(hipGetLastError());
	// the banded kernel needs the band list on the host side: keep it alive
	HIPCHK(hipStreamSynchronize(ctx->stream));
	ctx->agg_blocks += (uint64_t)grid.x * grid.y * grid.z;
	ctx->agg_region[0] = 8;
	ctx->agg_region[1] = 8;
	ctx->agg_region[2] = nb;
}

// The learned single pass runs step 1's grid and matching with the network's
// gains in place of the threshold.
struct LearnedArgs {
	MlpInfer net;
	float sigma, gvar_thr;
	uint8_t *gflag;      // per group: gain variance above gvar_thr
	uint32_t *nflag;     // count of flagged groups
};

// Host-side plan of one step: parameters, geometry and derived sizes.
// The reference keeps nSimilarPatchesPred candidates of every frame's search
// region in its predictive search (std::partial_sort to that middle,
// nldct.cpp:1076); a smaller region is undefined behaviour there and an error
// here.  Regions are clamped to the tiles: the smallest tile bounds them.
void check_predictive(const bm4dof_ctx *ctx, const bm4dof_params &p, int W, int H, const StepGeom &geo)
{
	if (!ctx->vbm3d_search || p.nSimilarPatches <= 1) return;
	if (p.nSimilarPatchesPred < 1 || p.nSimilarPatchesPred > 64 || p.sizeSearchWindowPred < 1)
		throw ArgError("predictive search: nSimilarPatchesPred must be in 1..64 and sizeSearchWindowPred positive",
		               BM4DOF_EUNSUPP);
	int mw = W, mh = H;
	for (const TileBox &c : geo.til.crop) { mw = std::min(mw, c.ex - c.ox); mh = std::min(mh, c.ey - c.oy); }
	const long cx = std::max(0, mw - (int)p.sizePatch + 1), cy = std::max(0, mh - (int)p.sizePatch + 1);
	for (unsigned ws : {p.sizeSearchWindow, p.sizeSearchWindowPred})
		if (std::min<long>(ws, cx) * std::min<long>(ws, cy) < (long)p.nSimilarPatchesPred)
			throw ArgError("predictive search: a search region holds fewer than nSimilarPatchesPred patches");
}

struct StepPlan {
	bm4dof_params p;
	int W = 0, H = 0, F = 0, C = 0, nparts = 0;
	bool step1 = true, use_flow = false, learned = false, fast = false, fast_shape = false;
	std::vector<float> hb, hbi, hw;
	StepGeom geo;
	StepConst sc;
	int G = 0, np = 0, chunk_groups = 1;
	size_t WHF = 0, N_all = 0, filt_lds = 0;
};

StepPlan plan_step(bm4dof_ctx *ctx, const bm4dof_params &p, int W, int H, int F, int C, bool use_flow, int nparts,
                   bool learned, bool full_grid = false)
{
	StepPlan pl;
	pl.p = p;
	pl.W = W; pl.H = H; pl.F = F; pl.C = C; pl.nparts = nparts;
	pl.step1 = p.isFirstStep != 0;
	pl.use_flow = use_flow;
	pl.learned = learned;
	check_step_params(p, (unsigned)C);
	build_bases(p, pl.hb, pl.hbi, pl.hw);
	pl.geo = make_geometry(p, W, H, F, nparts, C, !full_grid);
	check_predictive(ctx, p, W, H, pl.geo);
	pl.sc = make_const(p, W, H, F, C, use_flow, pl.geo.til);
	pl.G = (int)pl.geo.gpos.size();
	pl.np = pl.sc.npmax;
	pl.WHF = (size_t)W * H * F;
	pl.N_all = pl.WHF * C;
	// chunks of consecutive groups bounded by the estimate buffer budget
	const size_t row_bytes = (size_t)C * pl.sc.P * 4;
	const size_t max_rows = std::max<size_t>(pl.np, ctx->est_budget / row_bytes);
	pl.chunk_groups = (int)std::min<size_t>((size_t)std::max(pl.G, 1), max_rows / pl.np);
	pl.filt_lds = pl.step1 ? ((size_t)C * pl.sc.P * pl.np + (size_t)pl.sc.P * pl.np) * 4
	                       : ((size_t)C * pl.sc.P * pl.np + 2 * (size_t)pl.sc.P * pl.np) * 4;
	if (pl.filt_lds > 64 * 1024) throw ArgError("group too large for the filter kernel", BM4DOF_EUNSUPP);
	// register-resident filters: 64-sample patches, at most 32 neighbours
	// (np < 64: the power-of-two cut keeps at most 32), one launch per channel
	pl.fast_shape = pl.sc.P == 64 && pl.np < 64 &&
	                ((pl.sc.sPx == 4 && pl.sc.sPt == 4) || (pl.sc.sPx == 8 && pl.sc.sPt == 1));
	pl.fast = ctx->fast_filter && pl.fast_shape && !pl.sc.order_inv;
	if (learned && !pl.fast_shape)
		throw ArgError("learned pass: 4x4x4 or 8x8x1 patches and fewer than 64 similar patches", BM4DOF_EUNSUPP);
	return pl;
}

// Device geometry of a plan (bases, tiles, group positions and tiles, first
// group per tile) in buffers that outlive the step's launches.
struct StepDev {
	float *basis = nullptr, *binv = nullptr, *win = nullptr;
	uint16_t *bsp = nullptr;   // fp16 split tables of 64-sample bases (matrix-core transforms)
	TileBox *tiles = nullptr;
	uint32_t *gpos = nullptr, *gtile = nullptr, *tile_first = nullptr;
};

// Uploads a plan's geometry into `geo_buf` (grown as needed).  Synchronous.
StepDev upload_plan(bm4dof_ctx *ctx, const StepPlan &pl, DevBuf &geo_buf)
{
	const size_t nb = pl.hb.size(), nbi = pl.hbi.size(), nw = pl.hw.size(), nt = pl.geo.til.crop.size();
	const size_t G = (size_t)std::max(pl.G, 1), ntf = pl.geo.tile_first.size();
	auto al = [](size_t b) { return (b + 255) / 256 * 256; };
	const size_t o_basis = 0, o_binv = o_basis + al(nb * 4), o_win = o_binv + al(nbi * 4),
	             o_tiles = o_win + al(nw * 4), o_gpos = o_tiles + al(nt * sizeof(TileBox)),
	             o_gtile = o_gpos + al(G * 4), o_tf = o_gtile + al(G * 4), o_bsp = o_tf + al(std::max<size_t>(ntf, 1) * 4),
	             total = o_bsp + al(nb == 4096 ? 4 * nb * 2 : 0);
	char *b = (char *)geo_buf.get(total);
	StepDev d;
	d.basis = (float *)(b + o_basis);
	d.binv = (float *)(b + o_binv);
	d.win = (float *)(b + o_win);
	d.tiles = (TileBox *)(b + o_tiles);
	d.gpos = (uint32_t *)(b + o_gpos);
	d.gtile = (uint32_t *)(b + o_gtile);
	d.tile_first = (uint32_t *)(b + o_tf);
	d.bsp = nb == 4096 ? (uint16_t *)(b + o_bsp) : nullptr;
	HIPCHK(hipMemcpy(d.basis, pl.hb.data(), nb * 4, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d.binv, pl.hbi.data(), nbi * 4, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d.win, pl.hw.data(), nw * 4, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d.tiles, pl.geo.til.crop.data(), nt * sizeof(TileBox), hipMemcpyHostToDevice));
	if (pl.G > 0) {
		HIPCHK(hipMemcpy(d.gpos, pl.geo.gpos.data(), (size_t)pl.G * 4, hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(d.gtile, pl.geo.gtile.data(), (size_t)pl.G * 4, hipMemcpyHostToDevice));
	}
	if (ntf) HIPCHK(hipMemcpy(d.tile_first, pl.geo.tile_first.data(), ntf * 4, hipMemcpyHostToDevice));
	if (d.bsp) {
		hipLaunchKernelGGL(split_basis_kernel, dim3(16), dim3(256), 0, ctx->stream, d.basis, d.binv, d.bsp);
		HIPCHK(hipGetLastError());
		HIPCHK(hipStreamSynchronize(ctx->stream));
	}
	return d;
}

// Grows every shared work buffer a plan's launches use, so that enqueueing
// the step allocates nothing (required while a stream is being captured).
// Threads of coef_kernel over one channel plane (one wave per 16 anchors
// along x), 0 when the patch does not fit the video.
int64_t coef_waves(const StepConst &sc)
{
	const int nx = sc.W - sc.sPx + 1, ny = sc.H - sc.sPx + 1, nt = sc.F - sc.sPt + 1;
	return nx > 0 && ny > 0 && nt > 0 ? (int64_t)((nx + 15) / 16) * ny * nt : 0;
}

// Floats of one image's coefficient table for a step that uses them (the
// register filters of a plain step, within coef_budget), else 0.
size_t coef_floats(const StepPlan &pl, size_t budget)
{
	const int64_t waves = coef_waves(pl.sc);
	if (!pl.fast || pl.learned || waves <= 0 || waves * 64 >= ((int64_t)1 << 32)) return 0;
	const size_t per = (size_t)pl.C * pl.WHF * pl.sc.P;
	return per * sizeof(float) * (pl.step1 ? 1 : 2) <= budget ? per : 0;
}

void reserve_step(bm4dof_ctx *ctx, const StepPlan &pl)
{
	const size_t G = (size_t)std::max(pl.G, 1), np = (size_t)pl.np;
	const size_t Mub = (size_t)std::min(pl.chunk_groups, std::max(pl.G, 1)) * np;
	const int sWt = pl.sc.wtf + pl.sc.wtb + 1;
	(void)ctx->idx.as<uint32_t>(G * np);
	(void)ctx->nsim.as<uint32_t>(G);
	(void)ctx->early.as<uint8_t>(G);
	(void)ctx->gweight.as<float>(G);
	(void)ctx->counters.as<uint32_t>(16);
	(void)ctx->stat64.as<unsigned long long>(4);
	(void)ctx->num.as<float>(pl.N_all);
	(void)ctx->wsum.as<float>(pl.WHF);
	(void)ctx->lstart.as<uint32_t>(pl.WHF + 1);
	(void)ctx->lend.as<uint32_t>(pl.WHF + 1);
	(void)ctx->base.as<uint32_t>((size_t)pl.chunk_groups + 1);
	(void)ctx->keys.as<uint32_t>(Mub);
	(void)ctx->svals.as<uint32_t>(Mub);
	(void)ctx->etile.as<uint32_t>(Mub);
	(void)ctx->entry_w.as<float>(Mub);
	(void)ctx->est.as<float>(Mub * pl.C * pl.sc.P);
	(void)ctx->erec.as<uint4>(Mub);
	if (agg_binned_ok(ctx, pl.sc, (int64_t)Mub)) (void)bin_bufs(ctx, pl.sc, (int64_t)Mub);
	(void)ctx->lastf.as<uint32_t>(G);
	(void)ctx->fidx.as<uint32_t>(G);
	(void)ctx->bm_wins.as<int4>(G * sWt);
	(void)ctx->ties.as<uint32_t>(G + 1);
	(void)ctx->bm_wmeta.as<int2>(G);
	if (pl.learned) {
		(void)ctx->gflag.as<uint8_t>(G);
		(void)ctx->fmask.as<uint8_t>(pl.WHF);
	}
	if (pl.C > 1 && (pl.learned || pl.fast)) {
		(void)ctx->planes.as<float>(pl.N_all);
		(void)ctx->gacc.as<float>(G);
		if (!pl.step1) (void)ctx->planes_b.as<float>(pl.N_all);
	}
	// per-anchor coefficient tables within coef_budget; left out when the
	// device cannot hold them (the filters then transform group by group)
	const size_t per = coef_floats(pl, ctx->coef_budget);
	if (per > 0) {
		try {
			(void)ctx->coef_n.as<float>(per);
			if (!pl.step1) (void)ctx->coef_b.as<float>(per);
		} catch (const ArgError &e) {
			if (e.code != BM4DOF_ENOMEM) throw;
			ctx->coef_n.release();
			ctx->coef_b.release();
		}
	}
	int64_t nscan = std::max(pl.WHF + 1, G + 1);
	if (agg_binned_ok(ctx, pl.sc, (int64_t)Mub)) {
		const BinBufs q = bin_bufs(ctx, pl.sc, (int64_t)Mub);   // the bin sort's scans
		nscan = std::max<int64_t>(nscan, (int64_t)256 * std::max(q.nt1, q.nblk) + 1);
	}
	reserve_scan(ctx, nscan);
}

// The per-anchor coefficient tables of a step on stream st (the ones
// reserve_step allocated; nothing allocates here): the noisy planes' table,
// and in step 2 the basic planes' one, unless the tables do not fit
// coef_budget (then both pointers stay null).
static void coef_tables(bm4dof_ctx *ctx, const StepPlan &pl, const StepDev &dv, const float *d_nplanes,
                        const float *d_bplanes, hipStream_t st, const float *&d_coef_n, const float *&d_coef_b)
{
	const StepConst &sc = pl.sc;
	const bool step1 = pl.step1;
	const int C = pl.C;
	const size_t WHF = pl.WHF;
	const int nimg = step1 ? 1 : 2;
	const size_t per = coef_floats(pl, ctx->coef_budget);
	const int64_t waves = coef_waves(sc);
	d_coef_n = d_coef_b = nullptr;
	if (per > 0 && ctx->coef_n.n >= per * sizeof(float) && (step1 || ctx->coef_b.n >= per * sizeof(float))) {
		const bool have_n = !step1 && ctx->coef_n_keep && ctx->coef_n_ok && ctx->coef_n_whf == WHF &&
		                    ctx->coef_n_shape[0] == sc.sPx && ctx->coef_n_shape[1] == sc.sPt &&
		                    ctx->coef_n_shape[2] == C && ctx->coef_n_basis == pl.hb;
		ctx->coef_n_ok = false;
		for (int i = 0; i < nimg; i++) {
			float *tab = (float *)(i ? ctx->coef_b : ctx->coef_n).p;
			const float *src = i ? d_bplanes : d_nplanes;
			(i ? d_coef_b : d_coef_n) = tab;
			if (i == 0 && have_n) continue;   // the first step's table of the same planes
			for (int c = 0; c < C; c++)
				hipLaunchKernelGGL(sc.sPx == 4 ? coef_kernel<4> : coef_kernel<8>, dim3((unsigned)waves), dim3(64), 0,
				                   st, src + (size_t)c * WHF, sc.W, sc.H, sc.F, dv.basis, tab + (size_t)c * WHF * sc.P);
		}
		HIPCHK(hipGetLastError());
		ctx->coef_n_ok = true;
		ctx->coef_n_whf = WHF;
		ctx->coef_n_shape[0] = sc.sPx;
		ctx->coef_n_shape[1] = sc.sPt;
		ctx->coef_n_shape[2] = C;
		ctx->coef_n_basis = pl.hb;
	}
}

// Launches one step on ctx->stream.  With `tm` it records stage events; a
// captured enqueue passes none.  Needs reserve_step(plan) first.
// parts: 1 set-up and block matching, 2 the rest (filter, aggregation,
// normalisation), 3 both.
void enqueue_step(bm4dof_ctx *ctx, const StepPlan &pl, const StepDev &dv, const float *noisy_w,
                  const float *basic_w, const float *ff, const float *bf, float *out_w, Timer *tm,
                  const LearnedArgs *la, int parts = 3)
{
	const bool step1 = pl.step1;
	const StepConst &sc = pl.sc;
	const int G = pl.G, np = pl.np, C = pl.C;
	const size_t WHF = pl.WHF, N_all = pl.N_all;
	auto mark = [&](const char *a, const char *b1, const char *b2) {
		if (tm) tm->mark(la ? a : step1 ? b1 : b2);
	};
	uint32_t *d_idx = ctx->idx.as<uint32_t>((size_t)std::max(G, 1) * np);
	uint32_t *d_nsim = ctx->nsim.as<uint32_t>(std::max(G, 1));
	uint8_t *d_early = ctx->early.as<uint8_t>(std::max(G, 1));
	float *d_gw = ctx->gweight.as<float>(std::max(G, 1));
	uint32_t *d_cnt = ctx->counters.as<uint32_t>(16);
	unsigned long long *d_rows = ctx->stat64.as<unsigned long long>(4) + (step1 ? 0 : 1);
	float *d_num = ctx->num.as<float>(N_all);
	float *d_wsum = ctx->wsum.as<float>(WHF);
	uint32_t *d_kstart = ctx->lstart.as<uint32_t>(WHF + 1);
	uint32_t *d_pcnt = ctx->lend.as<uint32_t>(WHF + 1);
	if (parts & 1) {
		fill_u32(ctx, d_cnt, 0u, 4);   // early-out count (slots 4.. are run totals)
		fill_f32(ctx, d_num, 0.f, (int64_t)N_all);
		fill_f32(ctx, d_wsum, 0.f, (int64_t)WHF);
		mark("l_setup", "s1_setup", "s2_setup");
		// a gray step's coefficient tables depend on the images only: on the
		// second stream beside the block matching
		ctx->coef_on_side = false;
		if (pl.fast && !la && C == 1 && ctx->coef_beside_bm) {
			HIPCHK(hipEventRecord(ctx->ev_fork, ctx->stream));
			HIPCHK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
			coef_tables(ctx, pl, dv, noisy_w, basic_w, ctx->side, ctx->coef_side_n, ctx->coef_side_b);
			HIPCHK(hipEventRecord(ctx->ev_coef, ctx->side));
			ctx->coef_on_side = true;
		}
		run_bm(ctx, sc, step1 ? 1 : 2, step1 ? noisy_w : basic_w, ff, bf, dv.gpos, dv.gtile, dv.tiles, G, d_idx,
		       d_nsim);
		mark("l_bm", "s1_bm", "s2_bm");
	}
	if (!(parts & 2)) return;
	uint8_t *d_gflag = la ? ctx->gflag.as<uint8_t>(std::max(G, 1)) : nullptr;
	// the register-resident filters read one-channel planes
	const float *d_nplanes = noisy_w, *d_bplanes = basic_w;
	float *d_gacc = nullptr;
	if (C > 1 && (la || pl.fast)) {
		float *pn = ctx->planes.as<float>(N_all);
		hipLaunchKernelGGL(planes_kernel, dim3(nblk((int64_t)N_all, 256)), dim3(256), 0, ctx->stream, noisy_w, C,
		                   sc.WH, (int64_t)N_all, pn);
		d_nplanes = pn;
		if (!step1 && !la) {
			float *pb = ctx->planes_b.as<float>(N_all);
			hipLaunchKernelGGL(planes_kernel, dim3(nblk((int64_t)N_all, 256)), dim3(256), 0, ctx->stream, basic_w,
			                   C, sc.WH, (int64_t)N_all, pb);
			d_bplanes = pb;
		}
		d_gacc = ctx->gacc.as<float>(std::max(G, 1));
		HIPCHK(hipGetLastError());
	}
	// forward transforms of every patch anchor, once per image, when the
	// tables fit coef_budget: the register filters then read a group row's
	// coefficients instead of transforming the patch again in every group
	// that holds it (bit for bit the same sums); a gray step launched them
	// beside its block matching
	const float *d_coef_n = nullptr, *d_coef_b = nullptr;
	if (pl.fast && !la) {
		if (ctx->coef_on_side) {
			HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->ev_coef, 0));
			d_coef_n = ctx->coef_side_n;
			d_coef_b = ctx->coef_side_b;
			ctx->coef_on_side = false;
		} else
			coef_tables(ctx, pl, dv, d_nplanes, d_bplanes, ctx->stream, d_coef_n, d_coef_b);
		mark("l_coef", "s1_coef", "s2_coef");
	}

	for (int g0 = 0; g0 < G; g0 += pl.chunk_groups) {
		const int Gc = std::min(pl.chunk_groups, G - g0);
		const int64_t Mub = (int64_t)Gc * np;
		// row offsets of the chunk's groups; base[Gc] = rows of the chunk
		uint32_t *d_base = ctx->base.as<uint32_t>((size_t)Gc + 1);
		dscan(ctx, 0, d_nsim + g0, Gc, (int64_t)Gc + 1, d_base);
		const uint32_t *d_M = d_base + Gc;
		uint32_t *d_epos = ctx->keys.as<uint32_t>(Mub);
		uint32_t *d_svals = ctx->svals.as<uint32_t>(Mub);
		uint32_t *d_etile = ctx->etile.as<uint32_t>(Mub);
		float *d_ew = ctx->entry_w.as<float>(Mub);
		float *d_est = ctx->est.as<float>((size_t)Mub * C * sc.P);
		// rows listed by patch position, in row order within a position: on
		// the side stream, beside the filter (only the aggregation reads them)
		HIPCHK(hipEventRecord(ctx->ev_fork, ctx->stream));
		HIPCHK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
		hipLaunchKernelGGL(emit_entries_kernel, dim3(side_grid(ctx, nblk(Mub, 64))), dim3(64), 0, ctx->side, d_idx, d_nsim,
		                   d_base, np, g0, Gc, d_epos);
		HIPCHK(hipGetLastError());
		hipLaunchKernelGGL(add_rows_kernel, dim3(1), dim3(64), 0, ctx->side, d_M, d_rows);
		if (agg_binned_ok(ctx, sc, Mub)) {
			agg_bin_lists(ctx, sc, Mub, d_epos, d_M);
		} else {
			fill_u32(ctx, d_pcnt, 0u, (int64_t)WHF + 1, ctx->side);
			hipLaunchKernelGGL(count_rows_kernel, dim3(nblk(Mub, 256)), dim3(256), 0, ctx->side, d_epos, d_M, Mub,
			                   d_pcnt);
			dscan(ctx, 0, d_pcnt, (int64_t)WHF, (int64_t)WHF + 1, d_kstart, true);
			fill_u32(ctx, d_pcnt, 0u, (int64_t)WHF + 1, ctx->side);
			hipLaunchKernelGGL(scatter_rows_kernel, dim3(nblk(Mub, 256)), dim3(256), 0, ctx->side, d_epos, d_M, Mub,
			                   d_kstart, d_pcnt, d_svals);
			hipLaunchKernelGGL(sort_runs_kernel, dim3(nblk((int64_t)WHF, 256)), dim3(256), 0, ctx->side, d_kstart,
			                   (int64_t)WHF, d_svals);
		}
		HIPCHK(hipGetLastError());
		HIPCHK(hipEventRecord(ctx->ev_join, ctx->side));

		if (la)
			for (int c = 0; c < C; c++)   // channel planes, the first one sets the weights
				hipLaunchKernelGGL(sc.sPx == 4 ? filter_learned_kernel<4> : filter_learned_kernel<8>, dim3(Gc),
				                   dim3(64), 0, ctx->stream, d_nplanes + (size_t)c * WHF, d_idx, d_nsim, d_base,
				                   dv.gtile, sc, g0, Gc, dv.bsp, dv.win, la->net, la->sigma,
				                   la->gvar_thr, d_est, d_ew, d_etile, d_early, d_gw, d_gflag, la->nflag, C, c);
		else if (pl.fast && step1)
			for (int c = 0; c < C; c++)   // channel planes; the last launch weights all
				hipLaunchKernelGGL(sc.sPx == 4 ? filter1_fast_kernel<4> : filter1_fast_kernel<8>, dim3(Gc),
				                   dim3(64), 0, ctx->stream, d_nplanes, d_idx, d_nsim, d_base, dv.gtile, sc, g0,
				                   Gc, dv.basis, dv.binv, dv.win, d_est, d_ew, d_etile, d_early, d_gw, d_cnt, C, c,
				                   (int64_t)WHF, d_gacc, d_coef_n);
		else if (pl.fast)
			for (int c = 0; c < C; c++)
				hipLaunchKernelGGL(sc.sPx == 4 ? filter2_fast_kernel<4> : filter2_fast_kernel<8>, dim3(Gc),
				                   dim3(64), 0, ctx->stream, d_nplanes, d_bplanes, d_idx, d_nsim, d_base,
				                   dv.gtile, sc, g0, Gc, dv.basis, dv.binv, dv.win, d_est, d_ew, d_etile, d_early,
				                   d_gw, d_cnt, C, c, (int64_t)WHF, d_gacc, d_coef_n, d_coef_b);
		else if (step1)
			hipLaunchKernelGGL(filter1_kernel, dim3(Gc), dim3(256), pl.filt_lds, ctx->stream, noisy_w,
			                   d_idx, d_nsim, d_base, dv.gtile, sc, g0, Gc, dv.basis, dv.binv, dv.win, d_est,
			                   d_ew, d_etile, d_early, d_gw, d_cnt);
		else
			hipLaunchKernelGGL(filter2_kernel, dim3(Gc), dim3(256), pl.filt_lds, ctx->stream, noisy_w,
			                   basic_w, d_idx, d_nsim, d_base, dv.gtile, sc, g0, Gc, dv.basis, dv.binv,
			                   dv.win, d_est, d_ew, d_etile, d_early, d_gw, d_cnt);
		HIPCHK(hipGetLastError());
		if (!la) {
			// equality early-outs take the stale weights of their tile
			const int gend = g0 + Gc;
			uint32_t *d_fidx = ctx->fidx.as<uint32_t>((size_t)std::max(G, 1));
			uint32_t *d_lastf = ctx->lastf.as<uint32_t>((size_t)std::max(G, 1));
			hipLaunchKernelGGL(filtered_index_kernel, dim3(nblk(gend, 256)), dim3(256), 0, ctx->stream, d_early,
			                   gend, d_fidx);
			dscan(ctx, 1, d_fidx, gend, gend, d_lastf);
			hipLaunchKernelGGL(early_fix_kernel, dim3(nblk(Gc, 64)), dim3(64), 0, ctx->stream, d_early, d_nsim, d_gw,
			                   d_lastf, dv.gtile, dv.tile_first, d_base, g0, Gc, dv.win, C, sc.P, d_est, d_ew);
			HIPCHK(hipGetLastError());
		}
		mark("l_filter", "s1_filter", "s2_filter");
		// the stage "sort": what the row sort on the second stream still
		// takes once the filter is done
		HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
		mark("l_sort", "s1_sort", "s2_sort");

		std::vector<uint32_t> band_groups;
		for (uint32_t gf : pl.geo.frame_first)
			if ((int)gf > g0 && (int)gf < g0 + Gc) band_groups.push_back(gf);
		aggregate_chunk(ctx, sc, band_groups, g0, Mub, d_M, 0, pl.F, d_base, d_kstart, d_svals, d_epos, d_etile,
		                d_ew, d_est, dv.win, d_num, d_wsum, d_cnt + 4);
		mark("l_aggregate", "s1_aggregate", "s2_aggregate");
	}
	hipLaunchKernelGGL(normalize_kernel, dim3(nblk((int64_t)N_all, 256)), dim3(256), 0, ctx->stream,
	                   d_num, d_wsum, noisy_w, sc, out_w);
	HIPCHK(hipGetLastError());
	mark("l_normalize", "s1_normalize", "s2_normalize");
	if (la && G > 0) {
		// pixels under flagged groups, for the plain two-step fallback
		uint8_t *d_mask = ctx->fmask.as<uint8_t>(WHF);
		hipLaunchKernelGGL(fill_u8_kernel, dim3(nblk((int64_t)WHF, 256)), dim3(256), 0, ctx->stream, d_mask,
		                   (uint8_t)0, (int64_t)WHF);
		hipLaunchKernelGGL(fallback_mark_kernel, dim3(G), dim3(64), 0, ctx->stream, d_idx, d_nsim, d_gflag, sc,
		                   G, d_mask);
		HIPCHK(hipGetLastError());
	}
}

// Plans of recent step shapes, each with its own device copy, so that a
// repeated run of the same shape plans and uploads nothing.  The key holds
// everything plan_step reads.
struct CachedPlan {
	std::string key;
	StepPlan pl;
	DevBuf geo;
	StepDev dv;
	~CachedPlan() { geo.release(); }
};
struct PlanCache {
	std::vector<std::unique_ptr<CachedPlan>> v;   // most recent first
};
constexpr size_t kPlanCacheSize = 4;

std::string plan_key(const bm4dof_ctx *ctx, const bm4dof_params &p, int W, int H, int F, int C, bool use_flow,
                     int nparts, bool learned, bool full_grid)
{
	const long long f[] = {W, H, F, C, use_flow, nparts, learned, ctx->fast_filter, (long long)ctx->est_budget,
	                       full_grid, ctx->vbm3d_search};
	std::string k(reinterpret_cast<const char *>(&p), sizeof p);
	k.append(reinterpret_cast<const char *>(f), sizeof f);
	return k;
}

const CachedPlan &cached_plan(bm4dof_ctx *ctx, const bm4dof_params &p, int W, int H, int F, int C,
                              bool use_flow, int nparts, bool learned, bool full_grid = false)
{
	if (!ctx->plan_cache) ctx->plan_cache = std::make_shared<PlanCache>();
	auto &v = static_cast<PlanCache *>(ctx->plan_cache.get())->v;
	const std::string key = plan_key(ctx, p, W, H, F, C, use_flow, nparts, learned, full_grid);
	for (size_t i = 0; i < v.size(); i++)
		if (v[i]->key == key) {
			std::rotate(v.begin(), v.begin() + i, v.begin() + i + 1);
			return *v[0];
		}
	auto e = std::make_unique<CachedPlan>();
	e->key = key;
	e->pl = plan_step(ctx, p, W, H, F, C, use_flow, nparts, learned, full_grid);
	e->dv = upload_plan(ctx, e->pl, e->geo);
	if (v.size() >= kPlanCacheSize) {
		HIPCHK(hipStreamSynchronize(ctx->stream));   // the evicted plan may still be in use
		v.pop_back();
	}
	v.insert(v.begin(), std::move(e));
	return *v[0];
}

// The reference's scan of a step (processNlBayes, nldct.cpp:753-790): per
// tile, grid positions in raster order are processed while their mask is
// set, and each processed group clears the mask at its first neighbour's
// index taken in the C-channel layout (nldct.cpp:2325-2329; for C > 1 that
// index points to frame C*t of the one-channel mask, and past the mask it is
// dropped, as in make_geometry).  Groups the reference skips get no rows;
// per_tile receives the processed count of every tile.  Synchronous.
void replay_grid(bm4dof_ctx *ctx, const StepPlan &pl, const StepDev &dv, std::vector<uint32_t> &per_tile)
{
	const int G = pl.G, np = pl.np;
	uint32_t *d_first = ctx->fidx.as<uint32_t>((size_t)G);
	hipLaunchKernelGGL(first_index_kernel, dim3(nblk(G, 256)), dim3(256), 0, ctx->stream, ctx->idx.as<uint32_t>(1),
	                   np, G, d_first);
	HIPCHK(hipGetLastError());
	std::vector<uint32_t> first((size_t)G);
	HIPCHK(hipMemcpyAsync(first.data(), d_first, (size_t)G * 4, hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(hipStreamSynchronize(ctx->stream));
	const uint32_t W = (uint32_t)pl.W, WH = (uint32_t)pl.W * (uint32_t)pl.H;
	const int C = pl.C, nparts = (int)pl.geo.til.crop.size();
	std::vector<uint32_t> drop;
	std::vector<uint8_t> mask;
	per_tile.assign(nparts, 0);
	for (int n = 0; n < nparts; n++) {
		const TileBox &c = pl.geo.til.crop[n];
		const size_t wc = (size_t)(c.ex - c.ox), hc = (size_t)(c.ey - c.oy), whc = wc * hc, whf = whc * pl.F;
		const uint32_t g0 = pl.geo.tile_first[n];
		const uint32_t g1 = n + 1 < nparts ? pl.geo.tile_first[n + 1] : (uint32_t)G;
		auto local = [&](uint32_t q, size_t cs) {   // tile-local index of a position, channel stride cs
			const uint32_t t = q / WH, r = q - t * WH, y = r / W, x = r - y * W;
			return (size_t)t * cs + (size_t)(y - (uint32_t)c.oy) * wc + (size_t)(x - (uint32_t)c.ox);
		};
		mask.assign(whf, 0);
		for (uint32_t g = g0; g < g1; g++) mask[local(pl.geo.gpos[g], whc)] = 1;
		for (uint32_t g = g0; g < g1; g++) {
			if (!mask[local(pl.geo.gpos[g], whc)]) {
				drop.push_back(g);
				continue;
			}
			per_tile[n]++;
			const size_t i0 = local(first[g], whc * (size_t)C);
			if (i0 < whf) mask[i0] = 0;
		}
	}
	if (!drop.empty()) {
		uint32_t *d_drop = ctx->lastf.as<uint32_t>(drop.size());
		HIPCHK(hipMemcpyAsync(d_drop, drop.data(), drop.size() * 4, hipMemcpyHostToDevice, ctx->stream));
		hipLaunchKernelGGL(drop_groups_kernel, dim3(nblk((int64_t)drop.size(), 256)), dim3(256), 0, ctx->stream,
		                   d_drop, (int)drop.size(), ctx->nsim.as<uint32_t>(1));
		HIPCHK(hipGetLastError());
		HIPCHK(hipStreamSynchronize(ctx->stream));
	}
}

// One step, eager: (cached) plan, reserve and enqueue.
// `ratio` (may be NULL) receives the reference's groupsRatio term of the step:
// per tile 100 * groups / (W*H*F), summed over the tiles in order
// (nldct.cpp:403-404, 471-472).
uint64_t run_step(bm4dof_ctx *ctx, const bm4dof_params &p, int W, int H, int F, int C,
                  const float *noisy_w, const float *basic_w, const float *ff, const float *bf,
                  float *out_w, int nparts, Timer &tm, const LearnedArgs *la = nullptr, float *ratio = nullptr)
{
	const bool use_flow = ff != nullptr && bf != nullptr;
	const float *ffu = use_flow ? ff : nullptr, *bfu = use_flow ? bf : nullptr;
	const CachedPlan *cp = &cached_plan(ctx, p, W, H, F, C, use_flow, nparts, la != nullptr);
	reserve_step(ctx, cp->pl);
	enqueue_step(ctx, cp->pl, cp->dv, noisy_w, basic_w, ffu, bfu, out_w, &tm, la, 1);
	// The grid was laid out with every group's first neighbour being its own
	// patch.  Equal distances in heap order can put another patch first; the
	// reference then clears that position of its processing mask.  Replay
	// its scan on the full grid when that happens.
	std::vector<uint32_t> per_tile;
	if (cp->pl.np > 1 && cp->pl.G > 0) {
		uint32_t *d_moved = ctx->counters.as<uint32_t>(16) + 15;
		fill_u32(ctx, d_moved, 0u, 1);
		hipLaunchKernelGGL(first_moved_kernel, dim3(nblk(cp->pl.G, 256)), dim3(256), 0, ctx->stream,
		                   ctx->idx.as<uint32_t>(1), cp->dv.gpos, cp->pl.np, cp->pl.G, d_moved);
		HIPCHK(hipGetLastError());
		uint32_t moved = 0;
		HIPCHK(hipMemcpyAsync(&moved, d_moved, 4, hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
		if (moved > 0) {
			if (C > 1) {
				// the laid-out grid already left out the positions the colour
				// clears remove: match the full grid
				cp = &cached_plan(ctx, p, W, H, F, C, use_flow, nparts, la != nullptr, true);
				reserve_step(ctx, cp->pl);
				enqueue_step(ctx, cp->pl, cp->dv, noisy_w, basic_w, ffu, bfu, out_w, &tm, la, 1);
			}
			replay_grid(ctx, cp->pl, cp->dv, per_tile);
		}
	}
	const StepPlan &pl = cp->pl;
	uint64_t groups = 0;
	if (ratio) *ratio = 0.f;
	for (int t = 0; t < nparts; t++) {
		const uint32_t g0 = pl.geo.tile_first[t];
		const uint32_t g1 = t + 1 < nparts ? pl.geo.tile_first[t + 1] : (uint32_t)pl.geo.gpos.size();
		const uint32_t n = per_tile.empty() ? g1 - g0 : per_tile[t];
		groups += n;
		if (ratio) *ratio += 100.f * (float)n / (float)((size_t)W * H * F);
	}
	ctx->step_groups[pl.step1 ? 0 : 1] += groups;
	enqueue_step(ctx, pl, cp->dv, noisy_w, basic_w, ffu, bfu, out_w, &tm, la, 2);
	return groups;
}

// Run statistics: zeroed on the device at a run's start, read at its end.
void stats_begin(bm4dof_ctx *ctx)
{
	ctx->agg_blocks = ctx->agg_overflow_blocks = 0;
	ctx->step_groups[0] = ctx->step_groups[1] = ctx->step_rows[0] = ctx->step_rows[1] = 0;
	fill_u32(ctx, ctx->counters.as<uint32_t>(16), 0u, 16);
	fill_u32(ctx, (uint32_t *)ctx->stat64.as<unsigned long long>(4), 0u, 8);
}
void stats_end(bm4dof_ctx *ctx)
{
	unsigned long long rows[2] = {0, 0};
	uint32_t cnt[16];
	HIPCHK(hipMemcpyAsync(rows, ctx->stat64.as<unsigned long long>(4), 16, hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(hipMemcpyAsync(cnt, ctx->counters.as<uint32_t>(16), 64, hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(hipStreamSynchronize(ctx->stream));
	ctx->step_rows[0] = rows[0];
	ctx->step_rows[1] = rows[1];
	ctx->agg_overflow_blocks = cnt[4];
}

// The two steps of runNlBayes on device buffers (nldct.cpp:352-470), with
// the YUV transform of colour video around each step (nldct.cpp:363, 401,
// 422-423, 467-468): step 0 both, 1 the first only, 2 the second only.
static void two_step(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2, int step,
                     const float *noisy, int W, int H, int F, int C, const float *ff, const float *bf,
                     float *basic, float *final, int nparts, Timer &tm, float *gr)
{
	const size_t n = (size_t)W * H * F * C;
	// the second step reads the first one's noisy coefficient table when both
	// steps filter the same noisy planes (same colour space)
	ctx->coef_n_ok = ctx->coef_n_keep = false;
	if (step != 2) {
		const float *nw = noisy;
		if (C > 1 && p1->colorSpace == BM4DOF_YUV) {
			float *t = ctx->noisy_w.as<float>(n);
			HIPCHK(hipMemcpyAsync(t, noisy, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
			color(ctx, t, W, H, F, C, true);
			nw = t;
		}
		run_step(ctx, *p1, W, H, F, C, nw, nullptr, ff, bf, basic, nparts, tm, nullptr, &gr[0]);
		if (C > 1 && p1->colorSpace == BM4DOF_YUV) color(ctx, basic, W, H, F, C, false);
	}
	if (step != 1) {
		const float *nw = noisy;
		if (C > 1 && p2->colorSpace == BM4DOF_YUV) {
			float *t = ctx->noisy_w.as<float>(n);
			HIPCHK(hipMemcpyAsync(t, noisy, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
			color(ctx, t, W, H, F, C, true);
			color(ctx, basic, W, H, F, C, true);
			nw = t;
		}
		ctx->coef_n_keep = step == 0 && (C == 1 || p1->colorSpace == p2->colorSpace);
		run_step(ctx, *p2, W, H, F, C, nw, basic, ff, bf, final, nparts, tm, nullptr, &gr[1]);
		ctx->coef_n_ok = ctx->coef_n_keep = false;
		if (C > 1 && p2->colorSpace == BM4DOF_YUV) {
			color(ctx, basic, W, H, F, C, false);
			color(ctx, final, W, H, F, C, false);
		}
	}
}

// Whole runNlBayes on device buffers (nldct.cpp:267-478).
void run_dev(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2, const float *noisy,
             int W, int H, int F, int C, const float *ff, const float *bf, float *basic, float *final,
             int nparts, float *ratio)
{
	if (!(C == 1 || C == 3 || C == 4))
		throw ArgError("VideoNLB::runNlBayes: Wrong number of channels. Must be 1, 3 or 4.");
	const int step = p1->sizePatch ? (p2->sizePatch ? 0 : 1) : (p2->sizePatch ? 2 : -1);
	if (step == -1) throw ArgError("VideoNLB::runNlBayes: Both patch sizes are zero.");
	if (nparts < 1) throw ArgError("nparts must be >= 1");
	if (nparts > 255) throw ArgError("nparts above 255 is not supported", BM4DOF_EUNSUPP);
	if (W <= 0 || H <= 0 || F <= 0) throw ArgError("empty video");
	if ((uint64_t)W * H * F * C >= (1ull << 32))
		throw ArgError("video above 2^32 samples is not supported", BM4DOF_EUNSUPP);
	if (step != 1 && !final) throw ArgError("final output buffer is NULL");
	HIPCHK(hipSetDevice(ctx->device));
	for (auto e : ctx->ev) (void)hipEventDestroy(e);
	ctx->ev.clear();
	ctx->timings.clear();
	stats_begin(ctx);
	Timer tm(ctx);
	tm.mark("start");
	float gr[2] = {0.f, 0.f};
	two_step(ctx, p1, p2, step, noisy, W, H, F, C, ff, bf, basic, final, nparts, tm, gr);
	HIPCHK(hipStreamSynchronize(ctx->stream));
	stats_end(ctx);
	if (ratio) { ratio[0] = gr[0]; ratio[1] = gr[1]; }
	// timings
	for (size_t i = 1; i < ctx->ev.size(); i++) {
		float ms = 0.f;
		HIPCHK(hipEventElapsedTime(&ms, ctx->ev[i - 1], ctx->ev[i]));
		ctx->timings[i].second = ms;
	}
}

// Learned single pass on device buffers: step-1 grid and matching on the
// noisy video, network gains, step-2 style weights.  When gvar_thr >= 0,
// pixels under groups whose gain variance exceeds it take the plain
// two-step result (p1, p2) instead; *flag_frac receives the fraction of
// flagged groups.
void run_learned_dev(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2, const float *noisy,
                     int W, int H, int F, int C, const float *ff, const float *bf, float *out, int nparts,
                     float gvar_thr, float *flag_frac)
{
	if (!(C == 1 || C == 3 || C == 4))
		throw ArgError("VideoNLB::runNlBayes: Wrong number of channels. Must be 1, 3 or 4.");
	if (nparts < 1 || nparts > 255) throw ArgError("nparts must be in 1..255");
	if (W <= 0 || H <= 0 || F <= 0) throw ArgError("empty video");
	if ((uint64_t)W * H * F * C >= (1ull << 32))
		throw ArgError("video above 2^32 samples is not supported", BM4DOF_EUNSUPP);
	if (!p1->sizePatch) throw ArgError("learned pass: sizePatch of the first-step parameters is zero");
	if (gvar_thr >= 0.f && (!p2 || !p2->sizePatch)) throw ArgError("fallback needs second-step parameters");
	HIPCHK(hipSetDevice(ctx->device));
	for (auto e : ctx->ev) (void)hipEventDestroy(e);
	ctx->ev.clear();
	ctx->timings.clear();
	stats_begin(ctx);
	Timer tm(ctx);
	tm.mark("start");
	LearnedArgs la;
	la.net = mlp_net(ctx);
	la.sigma = p1->sigma;
	la.gvar_thr = gvar_thr >= 0.f ? gvar_thr : 3.4e38f;
	la.nflag = ctx->counters.as<uint32_t>(16) + 8;
	la.gflag = nullptr;
	bm4dof_params pl = *p1;
	pl.isFirstStep = 1;
	const size_t n = (size_t)W * H * F * C;
	// colour video is filtered in the first step's colour space
	const bool yuv = C > 1 && p1->colorSpace == BM4DOF_YUV;
	const float *nw = noisy;
	if (yuv) {
		float *t = ctx->noisy_w.as<float>(n);
		HIPCHK(hipMemcpyAsync(t, noisy, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
		color(ctx, t, W, H, F, C, true);
		nw = t;
	}
	const uint64_t G = run_step(ctx, pl, W, H, F, C, nw, nullptr, ff, bf, out, nparts, tm, &la);
	if (yuv) color(ctx, out, W, H, F, C, false);
	uint32_t nflag = 0;
	HIPCHK(hipMemcpyAsync(&nflag, la.nflag, 4, hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(hipStreamSynchronize(ctx->stream));
	if (flag_frac) *flag_frac = G ? (float)nflag / (float)G : 0.f;
	if (gvar_thr >= 0.f && nflag > 0) {
		// the plain two-step result for the pixels under flagged groups
		float *pb = ctx->basic_w.as<float>(n);
		float *pf = ctx->out_w.as<float>(n);
		float gr[2];
		two_step(ctx, p1, p2, 0, noisy, W, H, F, C, ff, bf, pb, pf, nparts, tm, gr);
		hipLaunchKernelGGL(fallback_blend_kernel, dim3(nblk((int64_t)n, 256)), dim3(256), 0, ctx->stream,
		                   ctx->fmask.as<uint8_t>((size_t)W * H * F), pf, (int64_t)n, W * H, C, out);
		HIPCHK(hipGetLastError());
		tm.mark("fallback_blend");
	}
	HIPCHK(hipStreamSynchronize(ctx->stream));
	stats_end(ctx);
	for (size_t i = 1; i < ctx->ev.size(); i++) {
		float ms = 0.f;
		HIPCHK(hipEventElapsedTime(&ms, ctx->ev[i - 1], ctx->ev[i]));
		ctx->timings[i].second = ms;
	}
}

// Training pairs of the learned pass: for every `stride`-th group of the
// step-1 grid (matching on the noisy video), 32 rows of features and of the
// step-2 Wiener gains of the basic video at the same neighbours.  Device
// outputs of max_rows rows; rows beyond a group's N are flagged 0.
int64_t learned_pairs_dev(bm4dof_ctx *ctx, const bm4dof_params *p1, const float *noisy, const float *basic,
                          int W, int H, int F, const float *ff, const float *bf, int nparts, int stride,
                          float beta2, float *xout, float *tout, uint8_t *rflag, int64_t max_rows)
{
	if (stride < 1) throw ArgError("stride must be >= 1");
	if (nparts < 1 || nparts > 255) throw ArgError("nparts must be in 1..255");
	HIPCHK(hipSetDevice(ctx->device));
	bm4dof_params p = *p1;
	p.isFirstStep = 1;
	check_step_params(p, 1);
	const bool use_flow = ff && bf;
	std::vector<float> hb, hbi, hw;
	build_bases(p, hb, hbi, hw);
	StepGeom geo = make_geometry(p, W, H, F, nparts);
	StepConst sc = make_const(p, W, H, F, 1, use_flow, geo.til);
	if (!(sc.P == 64 && sc.npmax < 64 && ((sc.sPx == 4 && sc.sPt == 4) || (sc.sPx == 8 && sc.sPt == 1))))
		throw ArgError("learned pass: 4x4x4 or 8x8x1 patches and fewer than 64 similar patches", BM4DOF_EUNSUPP);
	const int G = (int)geo.gpos.size();
	const int nslot = (G + stride - 1) / stride;
	const int64_t rows = (int64_t)nslot * 32;
	if (rows > max_rows) throw ArgError("output holds fewer rows than the sampled groups need");
	if (G == 0) return 0;
	float *d_basis = ctx->basis.as<float>(hb.size());
	float *d_binv = ctx->binv.as<float>(hbi.size());
	uint16_t *d_bsp = ctx->bsplit.as<uint16_t>(4 * hb.size());
	HIPCHK(hipMemcpyAsync(d_basis, hb.data(), hb.size() * 4, hipMemcpyHostToDevice, ctx->stream));
	HIPCHK(hipMemcpyAsync(d_binv, hbi.data(), hbi.size() * 4, hipMemcpyHostToDevice, ctx->stream));
	hipLaunchKernelGGL(split_basis_kernel, dim3(16), dim3(256), 0, ctx->stream, d_basis, d_binv, d_bsp);
	TileBox *d_tiles = ctx->tiles.as<TileBox>(geo.til.crop.size());
	HIPCHK(hipMemcpyAsync(d_tiles, geo.til.crop.data(), geo.til.crop.size() * sizeof(TileBox),
	                      hipMemcpyHostToDevice, ctx->stream));
	uint32_t *d_gpos = ctx->gpos.as<uint32_t>(G);
	uint32_t *d_gtile = ctx->gtile.as<uint32_t>(G);
	HIPCHK(hipMemcpyAsync(d_gpos, geo.gpos.data(), (size_t)G * 4, hipMemcpyHostToDevice, ctx->stream));
	HIPCHK(hipMemcpyAsync(d_gtile, geo.gtile.data(), (size_t)G * 4, hipMemcpyHostToDevice, ctx->stream));
	uint32_t *d_idx = ctx->idx.as<uint32_t>((size_t)G * sc.npmax);
	uint32_t *d_nsim = ctx->nsim.as<uint32_t>(G);
	run_bm(ctx, sc, 1, noisy, use_flow ? ff : nullptr, use_flow ? bf : nullptr, d_gpos, d_gtile, d_tiles, G,
	       d_idx, d_nsim);
	const float bs = beta2 * p.sigma;
	hipLaunchKernelGGL(sc.sPx == 4 ? learned_pairs_kernel<4> : learned_pairs_kernel<8>, dim3(nslot), dim3(64), 0,
	                   ctx->stream, noisy, basic, d_idx, d_nsim, sc, G, stride, d_bsp, p.sigma, bs * bs, xout,
	                   tout, rflag);
	HIPCHK(hipGetLastError());
	HIPCHK(hipStreamSynchronize(ctx->stream));
	return rows;
}

// --------------------------------------------------------------- autotuner
// Execution options that do not change the output (block-matching kernel and
// its staging depth, aggregation e-range capacity) are timed on a probe
// sequence and the fastest set is kept per shape class (w, h, patch, search
// window).
struct TuneChoice { int fast_bm, bm_frames, agg_cap; };

void apply_choice(bm4dof_ctx *ctx, const TuneChoice &c)
{
	ctx->fast_bm = c.fast_bm;
	ctx->bm_frames = c.bm_frames;
	ctx->agg_cap = c.agg_cap;
}

TuneChoice current_choice(const bm4dof_ctx *ctx) { return {ctx->fast_bm, ctx->bm_frames, ctx->agg_cap}; }

// Times both steps of one gray sequence (device buffers) under every
// candidate; returns the fastest choice and fills `report`.
TuneChoice autotune(bm4dof_ctx *ctx, const bm4dof_params &p1, const bm4dof_params &p2, const float *noisy,
                    int W, int H, int F, const float *ff, const float *bf, float *basic, float *final,
                    int nparts, std::string &report)
{
	static const TuneChoice cands[] = {{2, 2, 768}, {2, 2, 1024}, {2, 2, 1536},
	                                   {1, 1, 1024}, {1, 2, 1024}, {1, 4, 1024}};
	const TuneChoice saved = current_choice(ctx);
	TuneChoice best = saved;
	float best_ms = 3.4e38f;
	char line[160];
	for (const TuneChoice &c : cands) {
		apply_choice(ctx, c);
		run_dev(ctx, &p1, &p2, noisy, W, H, F, 1, ff, bf, basic, final, nparts, nullptr);   // warm
		run_dev(ctx, &p1, &p2, noisy, W, H, F, 1, ff, bf, basic, final, nparts, nullptr);
		float ms = 0.f;
		for (auto &t : ctx->timings) ms += t.second;
		snprintf(line, sizeof line, "fast_bm=%d bm_frames=%d agg_cap=%d: %.3f ms\n", c.fast_bm, c.bm_frames,
		         c.agg_cap, ms);
		report += line;
		if (ms < best_ms) { best_ms = ms; best = c; }
	}
	apply_choice(ctx, saved);
	snprintf(line, sizeof line, "chosen fast_bm=%d bm_frames=%d agg_cap=%d (%.3f ms)\n", best.fast_bm,
	         best.bm_frames, best.agg_cap, best_ms);
	report += line;
	return best;
}

} // namespace

// --------------------------------------------------------------- batches
// Many gray sequences of uneven size and noise level, both steps each, as
// one launch set: planned and uploaded once, then either enqueued eagerly or
// captured into one hipGraph that every bm4dof_batch_run replays.
struct bm4dof_batch {
	bm4dof_ctx *ctx = nullptr;
	int flags = 0, nparts = 2;
	std::vector<bm4dof_seq> seqs;
	std::vector<StepPlan> plans;          // 2 per sequence
	std::vector<DevBuf> geo;              // 2 per sequence
	std::vector<StepDev> dev;             // 2 per sequence
	std::vector<TuneChoice> choice;       // per sequence
	// per sequence and step: the count of groups whose first neighbour is not
	// their own patch (first_moved_kernel, written inside the launch set)
	DevBuf moved;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	std::vector<void *> captured_bufs;    // the context's buffer addresses the graph holds
	int last_path = 0, last_reruns = 0, captures = 0;
	float last_ms = 0.f;
	std::string report;
};

namespace {

std::vector<void *> buffer_addresses(bm4dof_ctx *ctx)
{
	std::vector<void *> a;
	for (DevBuf *d : ctx->buffers()) a.push_back(d->p);
	return a;
}

// Both steps of every sequence, in order.  Between a step's block matching
// and its filter, first_moved_kernel counts the groups whose first neighbour
// is another patch than their own (equal distances in heap order): the
// reference then clears another position of its processing mask
// (nldct.cpp:2325-2329) and skips groups, which only the eager path replays
// (run_step), so bm4dof_batch_run runs such sequences again eagerly.
void batch_enqueue(bm4dof_batch *b)
{
	bm4dof_ctx *ctx = b->ctx;
	uint32_t *d_moved = b->moved.as<uint32_t>(2 * b->seqs.size());
	uint32_t *d_idx = ctx->idx.as<uint32_t>(1);
	for (size_t i = 0; i < b->seqs.size(); i++) {
		const bm4dof_seq &q = b->seqs[i];
		apply_choice(ctx, b->choice[i]);
		const float *ff = (q.fflow && q.bflow) ? q.fflow : nullptr;
		const float *bf = (q.fflow && q.bflow) ? q.bflow : nullptr;
		for (int s = 0; s < 2; s++) {
			const StepPlan &pl = b->plans[2 * i + s];
			const StepDev &dv = b->dev[2 * i + s];
			const float *basic_in = s ? q.basic : nullptr;
			float *out = s ? q.final : q.basic;
			enqueue_step(ctx, pl, dv, q.noisy, basic_in, ff, bf, out, nullptr, nullptr, 1);
			fill_u32(ctx, d_moved + 2 * i + s, 0u, 1);
			if (pl.np > 1 && pl.G > 0) {
				hipLaunchKernelGGL(first_moved_kernel, dim3(nblk(pl.G, 256)), dim3(256), 0, ctx->stream, d_idx,
				                   dv.gpos, pl.np, pl.G, d_moved + 2 * i + s);
				HIPCHK(hipGetLastError());
			}
			enqueue_step(ctx, pl, dv, q.noisy, basic_in, ff, bf, out, nullptr, nullptr, 2);
		}
	}
}

// (Re)captures the batch's launch set into a graph: every buffer the
// launches use is reserved first, so nothing allocates inside the capture.
void batch_capture(bm4dof_batch *b)
{
	bm4dof_ctx *ctx = b->ctx;
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	b->exec = nullptr;
	b->graph = nullptr;
	for (const StepPlan &pl : b->plans) reserve_step(ctx, pl);
	(void)b->moved.as<uint32_t>(2 * b->seqs.size());
	HIPCHK(hipStreamSynchronize(ctx->stream));
	const TuneChoice saved = current_choice(ctx);
	HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
	try {
		batch_enqueue(b);
	} catch (...) {
		hipGraph_t g = nullptr;
		(void)hipStreamEndCapture(ctx->stream, &g);
		if (g) (void)hipGraphDestroy(g);
		apply_choice(ctx, saved);
		throw;
	}
	apply_choice(ctx, saved);
	HIPCHK(hipStreamEndCapture(ctx->stream, &b->graph));
	HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
	b->captured_bufs = buffer_addresses(ctx);
	b->captures++;
}

} // namespace

namespace {

// End of a denoiser call: the coefficient tables (several GB) go back to the
// device unless the caller asked to keep them for the next call.
void end_run(bm4dof_ctx *ctx)
{
	if (!ctx->keep_tables) {
		ctx->coef_n.release();
		ctx->coef_b.release();
	}
	ctx->coef_n_ok = false;
}

} // namespace

// ====================================================================== C-ABI

extern "C" {

__attribute__((visibility("default")))
void bm4dof_init_params(bm4dof_params *p, unsigned step, float sigma, unsigned width,
                        unsigned height, unsigned frames, unsigned channels, int verbose)
{
	(void)width; (void)height; (void)channels;
	init_params(p, step, sigma, frames, verbose);
}

__attribute__((visibility("default")))
void bm4dof_set_search_window(bm4dof_params *p, unsigned sw)
{
	p->sizeSearchWindow = sw;
	p->boundary = 2 * (sw / 2) + (p->sizePatch - 1);
	p->nSimilarPatches = std::min(p->nSimilarPatches,
	                              sw * sw * (p->sizeSearchTimeRangeFwd + p->sizeSearchTimeRangeBwd + 1));
}

__attribute__((visibility("default")))
void bm4dof_set_patch_size(bm4dof_params *p, unsigned frames, unsigned px)
{
	p->sizePatch = px;
	p->boundary = 2 * (p->sizeSearchWindow / 2) + (p->sizePatch - 1);
	if (p->sizePatch != 8) p->transform = BM4DOF_DCT;
	if (frames > 1)
		p->offSet = p->isFirstStep ? (unsigned)((double)(float)px / 4. * 3.) : px / 2;
}

__attribute__((visibility("default")))
void bm4dof_set_n_similar(bm4dof_params *p, unsigned np)
{
	p->nSimilarPatches = std::min(np, p->sizeSearchWindow * p->sizeSearchWindow *
	                                      (p->sizeSearchTimeRangeFwd + p->sizeSearchTimeRangeBwd + 1));
	if (p->sizeSearchTimeRangeFwd + p->sizeSearchTimeRangeFwd == 0)
		p->nSimilarPatchesPred = p->nSimilarPatches;
}

__attribute__((visibility("default")))
const char *bm4dof_global_error(void) { return g_global_err.c_str(); }

__attribute__((visibility("default")))
int bm4dof_ctx_create(bm4dof_ctx **out, int device)
{
	if (!out) return BM4DOF_EINVAL;
	*out = nullptr;
	bm4dof_ctx *ctx = new bm4dof_ctx();
	ctx->device = device;
	int rc = guarded(ctx, [&] {
		int n = 0;
		HIPCHK(hipGetDeviceCount(&n));
		if (device < 0 || device >= n) throw ArgError("no HIP device " + std::to_string(device));
		HIPCHK(hipSetDevice(device));
		HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
		HIPCHK(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
		HIPCHK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
		HIPCHK(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
		HIPCHK(hipEventCreateWithFlags(&ctx->ev_coef, hipEventDisableTiming));
		HIPCHK(hipEventCreateWithFlags(&ctx->ev_tvpoll, hipEventDisableTiming));
		// TV-L1 polls (no allocation in the solve itself, which the
		// allocation-free flow entry point shares); without it the solve
		// waits for the stream at each poll
		if (hipHostMalloc((void **)&ctx->tv_pin, (1 + 2 * (size_t)8192) * sizeof(int), hipHostMallocDefault) ==
		    hipSuccess)
			ctx->tv_pin_pairs = 8192;
		else
			ctx->tv_pin = nullptr;
		tv_scratch_init(device);
		// coefficient tables: at most 2/5 of the device memory free now
		size_t free_b = 0, total_b = 0;
		HIPCHK(hipMemGetInfo(&free_b, &total_b));
		ctx->coef_budget = std::min(ctx->coef_budget, free_b / 5 * 2);
	});
	if (rc) {
		g_global_err = ctx->err;
		delete ctx;
		return rc;
	}
	*out = ctx;
	return BM4DOF_OK;
}

__attribute__((visibility("default")))
void bm4dof_ctx_destroy(bm4dof_ctx *ctx)
{
	if (!ctx) return;
	(void)hipSetDevice(ctx->device);
	if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
	if (ctx->side) (void)hipStreamSynchronize(ctx->side);
	for (auto e : ctx->ev) (void)hipEventDestroy(e);
	ctx->plan_cache.reset();
	for (DevBuf *b : ctx->buffers()) b->release();
	if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
	if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
	if (ctx->ev_coef) (void)hipEventDestroy(ctx->ev_coef);
	if (ctx->ev_tvpoll) (void)hipEventDestroy(ctx->ev_tvpoll);
	if (ctx->tv_pin) (void)hipHostFree(ctx->tv_pin);
	if (ctx->side) (void)hipStreamDestroy(ctx->side);
	if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

__attribute__((visibility("default")))
const char *bm4dof_last_error(const bm4dof_ctx *ctx) { return ctx ? ctx->err.c_str() : g_global_err.c_str(); }

__attribute__((visibility("default")))
int bm4dof_ctx_set_option(bm4dof_ctx *ctx, const char *key, long long value)
{
	if (!ctx) return BM4DOF_EINVAL;
	return guarded(ctx, [&] {
		const std::string k = key ? key : "";
		if (k == "fast_filter") ctx->fast_filter = value != 0;
		else if (k == "agg_chunked") ctx->agg_chunked = value != 0;
		else if (k == "agg_binned") ctx->agg_binned = value != 0;
		else if (k == "coef_beside_bm") ctx->coef_beside_bm = value != 0;
		else if (k == "side_blocks") ctx->side_blocks = (int)value;
		else if (k == "fast_bm") {
			if (value < 0 || value > 2) throw ArgError("fast_bm must be 0, 1 or 2");
			ctx->fast_bm = (int)value;
		}
		else if (k == "bm_frames") {
			if (value != 1 && value != 2 && value != 4) throw ArgError("bm_frames must be 1, 2 or 4");
			ctx->bm_frames = (int)value;
		} else if (k == "coef_budget") {
			if (value < 0) throw ArgError("coef_budget must be >= 0");
			ctx->coef_budget = (size_t)value;
		} else if (k == "keep_tables") {
			ctx->keep_tables = value != 0;
		} else if (k == "vbm3d_search") {
			ctx->vbm3d_search = value != 0;
		} else if (k == "agg_cap") {
			if (value != 512 && value != 768 && value != 1024 && value != 1536)
				throw ArgError("agg_cap must be 512, 768, 1024 or 1536");
			ctx->agg_cap = (int)value;
		}
		else throw ArgError("unknown option '" + k + "'");
	});
}

__attribute__((visibility("default")))
int bm4dof_run_dev(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2,
                   const float *noisy, unsigned w, unsigned h, unsigned f, unsigned c,
                   const float *ff, const float *bf, float *basic, float *final, int nparts,
                   float groups_ratio[2])
{
	if (!ctx || !p1 || !p2 || !noisy || !basic) { set_err(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		run_dev(ctx, p1, p2, noisy, (int)w, (int)h, (int)f, (int)c, (ff && bf) ? ff : nullptr,
		        (ff && bf) ? bf : nullptr, basic, final, nparts, groups_ratio);
		end_run(ctx);
	});
}

__attribute__((visibility("default")))
int bm4dof_autotune(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2, const float *noisy,
                    unsigned w, unsigned h, unsigned f, const float *fflow, const float *bflow, float *basic,
                    float *final, int nparts, char *report, int report_len)
{
	if (!ctx || !p1 || !p2 || !noisy || !basic || !final) { set_err(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		std::string rep;
		const TuneChoice c = autotune(ctx, *p1, *p2, noisy, (int)w, (int)h, (int)f, (fflow && bflow) ? fflow : nullptr,
		                              (fflow && bflow) ? bflow : nullptr, basic, final, nparts, rep);
		apply_choice(ctx, c);
		if (report && report_len > 0) {
			strncpy(report, rep.c_str(), (size_t)report_len - 1);
			report[report_len - 1] = 0;
		}
	});
}

__attribute__((visibility("default")))
int bm4dof_batch_create(bm4dof_ctx *ctx, const bm4dof_seq *seqs, const bm4dof_params *p1s,
                        const bm4dof_params *p2s, int n, int nparts, int flags, bm4dof_batch **out)
{
	if (!ctx || !seqs || !p1s || !p2s || !out || n <= 0) { set_err(ctx, "bad batch arguments"); return BM4DOF_EINVAL; }
	*out = nullptr;
	bm4dof_batch *b = new bm4dof_batch;
	b->ctx = ctx;
	b->flags = flags;
	b->nparts = nparts;
	const int rc = guarded(ctx, [&] {
		HIPCHK(hipSetDevice(ctx->device));
		if (nparts < 1 || nparts > 255) throw ArgError("nparts must be in 1..255");
		b->seqs.assign(seqs, seqs + n);
		b->plans.reserve(2 * n);
		b->geo.resize(2 * n);
		b->dev.resize(2 * n);
		const TuneChoice saved = current_choice(ctx);
		b->choice.assign(n, saved);
		std::vector<std::pair<std::string, TuneChoice>> tuned;
		for (int i = 0; i < n; i++) {
			const bm4dof_seq &q = seqs[i];
			if (!q.noisy || !q.basic || !q.final) throw ArgError("sequence " + std::to_string(i) + ": NULL buffer");
			if (!p1s[i].sizePatch || !p2s[i].sizePatch) throw ArgError("batches run both steps");
			const int W = (int)q.w, H = (int)q.h, F = (int)q.f;
			if (W <= 0 || H <= 0 || F <= 0) throw ArgError("sequence " + std::to_string(i) + ": empty");
			if ((uint64_t)W * H * F >= (1ull << 32)) throw ArgError("sequence above 2^32 samples", BM4DOF_EUNSUPP);
			const bool flow = q.fflow && q.bflow;
			bm4dof_params a = p1s[i], c = p2s[i];
			a.isFirstStep = 1;
			c.isFirstStep = 0;
			b->plans.push_back(plan_step(ctx, a, W, H, F, 1, flow, nparts, false));
			b->plans.push_back(plan_step(ctx, c, W, H, F, 1, flow, nparts, false));
			for (int s = 0; s < 2; s++) {
				const StepPlan &pl = b->plans[2 * i + s];
				if ((flags & BM4DOF_BATCH_GRAPH) && (pl.sc.wTmp < 8 || pl.sc.hTmp < 8 || !ctx->agg_chunked))
					throw ArgError("a captured batch needs tiles of at least 8 x 8 pixels", BM4DOF_EUNSUPP);
				b->dev[2 * i + s] = upload_plan(ctx, pl, b->geo[2 * i + s]);
				reserve_step(ctx, pl);
			}
			if (flags & BM4DOF_BATCH_AUTOTUNE) {
				char key[96];
				// shape class: frame size and patch / search parameters (the
				// frame count scales the work but not the kernels' balance)
				snprintf(key, sizeof key, "%dx%d px%u pt%u wx%u np%u", W, H, a.sizePatch, a.sizePatchTime,
				         a.sizeSearchWindow, a.nSimilarPatches);
				auto it = std::find_if(tuned.begin(), tuned.end(), [&](const auto &t) { return t.first == key; });
				if (it == tuned.end()) {
					std::string rep;
					const TuneChoice ch = autotune(ctx, a, c, q.noisy, W, H, F, flow ? q.fflow : nullptr,
					                               flow ? q.bflow : nullptr, q.basic, q.final, nparts, rep);
					b->report += std::string(key) + ":\n" + rep;
					tuned.push_back({key, ch});
					b->choice[i] = ch;
				} else
					b->choice[i] = it->second;
			}
		}
		if (flags & BM4DOF_BATCH_GRAPH) batch_capture(b);
		apply_choice(ctx, saved);
	});
	if (rc) {
		bm4dof_batch_destroy(b);
		return rc;
	}
	*out = b;
	return BM4DOF_OK;
}

__attribute__((visibility("default")))
int bm4dof_batch_run(bm4dof_batch *b)
{
	if (!b || !b->ctx) return BM4DOF_EINVAL;
	bm4dof_ctx *ctx = b->ctx;
	return guarded(ctx, [&] {
		HIPCHK(hipSetDevice(ctx->device));
		const TuneChoice saved = current_choice(ctx);
		// the graph holds the addresses of the context's buffers at its
		// capture: when other work on the context has moved one since, capture
		// again
		int path = 0;
		if (b->flags & BM4DOF_BATCH_GRAPH) {
			path = 1;
			if (!b->exec || buffer_addresses(ctx) != b->captured_bufs) {
				batch_capture(b);
				path = 2;
			}
		}
		hipEvent_t e0, e1;
		HIPCHK(hipEventCreate(&e0));
		HIPCHK(hipEventCreate(&e1));
		HIPCHK(hipEventRecord(e0, ctx->stream));
		if (path) HIPCHK(hipGraphLaunch(b->exec, ctx->stream));
		else batch_enqueue(b);
		// sequences where equal distances put another patch first in some
		// group: the reference's processing-mask scan skips groups there,
		// which the eager path replays; run them again that way
		const size_t n = b->seqs.size();
		std::vector<uint32_t> mv(2 * n);
		HIPCHK(hipMemcpyAsync(mv.data(), b->moved.as<uint32_t>(2 * n), 8 * n, hipMemcpyDeviceToHost,
		                      ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
		int reruns = 0;
		for (size_t i = 0; i < n; i++) {
			if (!mv[2 * i] && !mv[2 * i + 1]) continue;
			const bm4dof_seq &q = b->seqs[i];
			const bool flow = q.fflow && q.bflow;
			apply_choice(ctx, b->choice[i]);
			run_dev(ctx, &b->plans[2 * i].p, &b->plans[2 * i + 1].p, q.noisy, (int)q.w, (int)q.h, (int)q.f, 1,
			        flow ? q.fflow : nullptr, flow ? q.bflow : nullptr, q.basic, q.final, b->nparts, nullptr);
			reruns++;
		}
		HIPCHK(hipEventRecord(e1, ctx->stream));
		HIPCHK(hipEventSynchronize(e1));
		HIPCHK(hipEventElapsedTime(&b->last_ms, e0, e1));
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		b->last_path = path;
		b->last_reruns = reruns;
		apply_choice(ctx, saved);
	});
}

__attribute__((visibility("default")))
int bm4dof_batch_stats(const bm4dof_batch *b, int64_t *out, int max_slots)
{
	if (!b || !out) return 0;
	const int64_t v[3] = {b->last_path, b->last_reruns, b->captures};
	int k = 0;
	for (; k < 3 && k < max_slots; k++) out[k] = v[k];
	return k;
}

__attribute__((visibility("default")))
float bm4dof_batch_last_ms(const bm4dof_batch *b) { return b ? b->last_ms : 0.f; }

__attribute__((visibility("default")))
const char *bm4dof_batch_report(const bm4dof_batch *b) { return b ? b->report.c_str() : ""; }

__attribute__((visibility("default")))
void bm4dof_batch_destroy(bm4dof_batch *b)
{
	if (!b) return;
	if (b->ctx) (void)hipSetDevice(b->ctx->device);
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	for (auto &g : b->geo) g.release();
	b->moved.release();
	delete b;
}

__attribute__((visibility("default")))
int bm4dof_run_learned_dev(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2,
                           const float *noisy, unsigned w, unsigned h, unsigned f, unsigned c,
                           const float *ff, const float *bf, float *out, int nparts, float gvar_thr,
                           float *flag_frac)
{
	if (!ctx || !p1 || !noisy || !out) { set_err(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		run_learned_dev(ctx, p1, p2, noisy, (int)w, (int)h, (int)f, (int)c, (ff && bf) ? ff : nullptr,
		                (ff && bf) ? bf : nullptr, out, nparts, gvar_thr, flag_frac);
		end_run(ctx);
	});
}

__attribute__((visibility("default")))
int bm4dof_run_learned(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2,
                       const float *noisy, unsigned w, unsigned h, unsigned f, unsigned c,
                       const float *ff, const float *bf, float *out, int nparts, float gvar_thr,
                       float *flag_frac)
{
	if (!ctx || !p1 || !noisy || !out) { set_err(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		HIPCHK(hipSetDevice(ctx->device));
		const size_t n = (size_t)w * h * f * c;
		const size_t nf = (size_t)w * h * f * 2;
		const bool flow = ff && bf;
		float *dn = ctx->io_noisy.as<float>(n);
		float *dout = ctx->io_final.as<float>(n);
		float *dff = flow ? ctx->io_fflow.as<float>(nf) : nullptr;
		float *dbf = flow ? ctx->io_bflow.as<float>(nf) : nullptr;
		HIPCHK(hipMemcpyAsync(dn, noisy, n * 4, hipMemcpyHostToDevice, ctx->stream));
		if (flow) {
			HIPCHK(hipMemcpyAsync(dff, ff, nf * 4, hipMemcpyHostToDevice, ctx->stream));
			HIPCHK(hipMemcpyAsync(dbf, bf, nf * 4, hipMemcpyHostToDevice, ctx->stream));
		}
		run_learned_dev(ctx, p1, p2, dn, (int)w, (int)h, (int)f, (int)c, dff, dbf, dout, nparts, gvar_thr,
		                flag_frac);
		HIPCHK(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
		end_run(ctx);
	});
}

__attribute__((visibility("default")))
int bm4dof_learned_pairs(bm4dof_ctx *ctx, const bm4dof_params *p1, const float *noisy, const float *basic,
                         unsigned w, unsigned h, unsigned f, const float *ff, const float *bf, int nparts,
                         int stride, float beta2, float *x_out, float *t_out, unsigned char *row_flag,
                         int64_t max_rows, int64_t *rows)
{
	if (!ctx || !p1 || !noisy || !basic || !x_out || !t_out || !row_flag) {
		set_err(ctx, "NULL argument");
		return BM4DOF_EINVAL;
	}
	return guarded(ctx, [&] {
		const int64_t r = learned_pairs_dev(ctx, p1, noisy, basic, (int)w, (int)h, (int)f, (ff && bf) ? ff : nullptr,
		                                    (ff && bf) ? bf : nullptr, nparts, stride, beta2, x_out, t_out,
		                                    (uint8_t *)row_flag, max_rows);
		if (rows) *rows = r;
	});
}

__attribute__((visibility("default")))
int bm4dof_run(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2,
               const float *noisy, unsigned w, unsigned h, unsigned f, unsigned c,
               const float *ff, const float *bf, float *basic, float *final, int nparts,
               float groups_ratio[2])
{
	if (!ctx || !p1 || !p2 || !noisy || !basic) { set_err(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		HIPCHK(hipSetDevice(ctx->device));
		const size_t n = (size_t)w * h * f * c;
		const size_t nf = (size_t)w * h * f * 2;
		const bool flow = ff && bf;
		const int step = p1->sizePatch ? (p2->sizePatch ? 0 : 1) : (p2->sizePatch ? 2 : -1);
		float *dn = ctx->io_noisy.as<float>(n);
		float *db = ctx->io_basic.as<float>(n);
		float *dfin = step != 1 ? ctx->io_final.as<float>(n) : nullptr;
		float *dff = flow ? ctx->io_fflow.as<float>(nf) : nullptr;
		float *dbf = flow ? ctx->io_bflow.as<float>(nf) : nullptr;
		HIPCHK(hipMemcpyAsync(dn, noisy, n * 4, hipMemcpyHostToDevice, ctx->stream));
		if (step == 2) HIPCHK(hipMemcpyAsync(db, basic, n * 4, hipMemcpyHostToDevice, ctx->stream));
		if (flow) {
			HIPCHK(hipMemcpyAsync(dff, ff, nf * 4, hipMemcpyHostToDevice, ctx->stream));
			HIPCHK(hipMemcpyAsync(dbf, bf, nf * 4, hipMemcpyHostToDevice, ctx->stream));
		}
		run_dev(ctx, p1, p2, dn, (int)w, (int)h, (int)f, (int)c, dff, dbf, db, dfin, nparts, groups_ratio);
		HIPCHK(hipMemcpyAsync(basic, db, n * 4, hipMemcpyDeviceToHost, ctx->stream));
		if (step != 1 && final) HIPCHK(hipMemcpyAsync(final, dfin, n * 4, hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
		end_run(ctx);
	});
}

__attribute__((visibility("default")))
int bm4dof_bm_count(bm4dof_ctx *ctx, const bm4dof_params *p, unsigned w, unsigned h, unsigned f,
                    unsigned c, int nparts, int64_t *n_groups)
{
	(void)c;
	return guarded(ctx, [&] {
		StepGeom geo = make_geometry(*p, (int)w, (int)h, (int)f, nparts, 1, false);
		*n_groups = (int64_t)geo.gpos.size();
	});
}

__attribute__((visibility("default")))
int bm4dof_bm(bm4dof_ctx *ctx, const bm4dof_params *p, const float *image, unsigned w, unsigned h,
              unsigned f, unsigned c, const float *ff, const float *bf, int nparts,
              uint32_t *out_index, uint32_t *out_nsim, uint32_t *out_ref, uint32_t *out_tile)
{
	if (!ctx || !p || !image) { set_err(ctx, "NULL argument"); return BM4DOF_EINVAL; }
	return guarded(ctx, [&] {
		HIPCHK(hipSetDevice(ctx->device));
		check_step_params(*p, c);
		const int W = (int)w, H = (int)h, F = (int)f, C = (int)c;
		const bool flow = ff && bf;
		StepGeom geo = make_geometry(*p, W, H, F, nparts, C, false);
		check_predictive(ctx, *p, W, H, geo);
		StepConst sc = make_const(*p, W, H, F, C, flow, geo.til);
		const int G = (int)geo.gpos.size();
		const size_t n = (size_t)W * H * F * C, nf = (size_t)W * H * F * 2;
		float *dimg = ctx->bm_image.as<float>(n);
		HIPCHK(hipMemcpyAsync(dimg, image, n * 4, hipMemcpyHostToDevice, ctx->stream));
		float *dff = nullptr, *dbf = nullptr;
		if (flow) {
			dff = ctx->io_fflow.as<float>(nf);
			dbf = ctx->io_bflow.as<float>(nf);
			HIPCHK(hipMemcpyAsync(dff, ff, nf * 4, hipMemcpyHostToDevice, ctx->stream));
			HIPCHK(hipMemcpyAsync(dbf, bf, nf * 4, hipMemcpyHostToDevice, ctx->stream));
		}
		TileBox *d_tiles = ctx->tiles.as<TileBox>(geo.til.crop.size());
		HIPCHK(hipMemcpyAsync(d_tiles, geo.til.crop.data(), geo.til.crop.size() * sizeof(TileBox),
		                      hipMemcpyHostToDevice, ctx->stream));
		uint32_t *d_gpos = ctx->gpos.as<uint32_t>(std::max(G, 1));
		uint32_t *d_gtile = ctx->gtile.as<uint32_t>(std::max(G, 1));
		HIPCHK(hipMemcpyAsync(d_gpos, geo.gpos.data(), (size_t)G * 4, hipMemcpyHostToDevice, ctx->stream));
		HIPCHK(hipMemcpyAsync(d_gtile, geo.gtile.data(), (size_t)G * 4, hipMemcpyHostToDevice, ctx->stream));
		uint32_t *d_idx = ctx->idx.as<uint32_t>((size_t)std::max(G, 1) * sc.npmax);
		uint32_t *d_nsim = ctx->nsim.as<uint32_t>(std::max(G, 1));
		run_bm(ctx, sc, p->isFirstStep ? 1 : 2, dimg, dff, dbf, d_gpos, d_gtile, d_tiles, G, d_idx, d_nsim);
		std::vector<uint32_t> hidx((size_t)G * sc.npmax), hn(G);
		HIPCHK(hipMemcpyAsync(hidx.data(), d_idx, hidx.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(hipMemcpyAsync(hn.data(), d_nsim, hn.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
		const uint32_t WH = (uint32_t)(W * H);
		auto enc = [&](uint32_t pos) { // 3-index position -> reference patch index
			const uint32_t t = pos / WH;
			return t * WH * (uint32_t)C + (pos - t * WH);
		};
		for (int g = 0; g < G; g++) {
			for (int k = 0; k < sc.npmax; k++) {
				const size_t o = (size_t)g * sc.npmax + k;
				out_index[o] = k < (int)hn[g] ? enc(hidx[o]) : 0xffffffffu;
			}
			out_nsim[g] = hn[g];
			out_ref[g] = enc(geo.gpos[g]);
			out_tile[g] = geo.gtile[g];
		}
	});
}

__attribute__((visibility("default")))
int bm4dof_last_stats(const bm4dof_ctx *ctx, int64_t *stats, int max_slots)
{
	if (!ctx || !stats) return 0;
	const int64_t v[10] = {(int64_t)ctx->agg_blocks, (int64_t)ctx->agg_overflow_blocks,
	                       ctx->agg_region[0], ctx->agg_region[1], ctx->agg_region[2],
	                       (int64_t)ctx->step_groups[0], (int64_t)ctx->step_rows[0],
	                       (int64_t)ctx->step_groups[1], (int64_t)ctx->step_rows[1],
	                       (int64_t)ctx->tv_pixel_iters};
	int n = 0;
	for (; n < 10 && n < max_slots; n++) stats[n] = v[n];
	return n;
}

__attribute__((visibility("default")))
int bm4dof_last_timings(const bm4dof_ctx *ctx, float *ms, const char **names, int max_slots)
{
	if (!ctx) return 0;
	int n = 0;
	for (size_t i = 1; i < ctx->timings.size() && n < max_slots; i++, n++) {
		if (ms) ms[n] = ctx->timings[i].second;
		if (names) names[n] = ctx->timings[i].first;
	}
	return n;
}

} // extern "C"
