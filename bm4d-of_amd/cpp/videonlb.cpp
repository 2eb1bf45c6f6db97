// VideoNLB host interface over libbm4dof.so (see videonlb.hpp).
#include "videonlb.hpp"

#include "../../include/bm4dof.h"
#include "../../include/bm4dof_io.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace {

bm4dof_params to_c(const VideoNLB::nlbParams &p)
{
	bm4dof_params c;
	std::memset(&c, 0, sizeof c);
	c.sigma = p.sigma;
	c.sizePatch = p.sizePatch;
	c.sizePatchTime = p.sizePatchTime;
	c.nSimilarPatches = p.nSimilarPatches;
	c.sizeSearchWindow = p.sizeSearchWindow;
	c.sizeSearchTimeRangeFwd = p.sizeSearchTimeRangeFwd;
	c.sizeSearchTimeRangeBwd = p.sizeSearchTimeRangeBwd;
	c.boundary = p.boundary;
	c.offSet = p.offSet;
	c.offSetTime = p.offSetTime;
	c.beta = p.beta;
	c.betaMean = p.betaMean;
	c.tau = p.tau;
	c.orderInvariance = p.orderInvariance;
	c.isFirstStep = p.isFirstStep;
	c.doPasteBoost = p.doPasteBoost;
	c.verbose = p.verbose;
	c.colorSpace = p.colorSpace == VideoNLB::YUV ? BM4DOF_YUV : BM4DOF_RGB;
	c.transform = p.transform == VideoNLB::bior1_5 ? BM4DOF_BIOR1_5 : BM4DOF_DCT;
	c.sizeSearchWindowPred = p.sizeSearchWindowPred;
	c.nSimilarPatchesPred = p.nSimilarPatchesPred;
	c.dsub = p.dsub;
	c.agg_window = p.agg_window;
	return c;
}

VideoNLB::nlbParams from_c(const bm4dof_params &c)
{
	VideoNLB::nlbParams p;
	p.sigma = c.sigma;
	p.sizePatch = c.sizePatch;
	p.sizePatchTime = c.sizePatchTime;
	p.nSimilarPatches = c.nSimilarPatches;
	p.sizeSearchWindow = c.sizeSearchWindow;
	p.sizeSearchTimeRangeFwd = c.sizeSearchTimeRangeFwd;
	p.sizeSearchTimeRangeBwd = c.sizeSearchTimeRangeBwd;
	p.boundary = c.boundary;
	p.offSet = c.offSet;
	p.offSetTime = c.offSetTime;
	p.beta = c.beta;
	p.betaMean = c.betaMean;
	p.tau = c.tau;
	p.orderInvariance = c.orderInvariance != 0;
	p.isFirstStep = c.isFirstStep != 0;
	p.doPasteBoost = c.doPasteBoost != 0;
	p.verbose = c.verbose != 0;
	p.colorSpace = c.colorSpace == BM4DOF_YUV ? VideoNLB::YUV : VideoNLB::RGB;
	p.transform = c.transform == BM4DOF_BIOR1_5 ? VideoNLB::bior1_5 : VideoNLB::dct;
	p.sizeSearchWindowPred = c.sizeSearchWindowPred;
	p.nSimilarPatchesPred = c.nSimilarPatchesPred;
	p.dsub = c.dsub;
	p.agg_window = c.agg_window != 0;
	return p;
}

int g_predictive = -1;   // -1: not set, read BM4DOF_VBM3D_SEARCH

bool predictive_on()
{
	if (g_predictive < 0) {
		const char *e = std::getenv("BM4DOF_VBM3D_SEARCH");
		g_predictive = e && std::atoi(e) ? 1 : 0;
	}
	return g_predictive != 0;
}

// One context for the process, created on first use on BM4DOF_DEVICE (0).
bm4dof_ctx *context()
{
	static std::mutex m;
	static bm4dof_ctx *ctx = nullptr;
	std::lock_guard<std::mutex> lk(m);
	if (!ctx) {
		const char *d = std::getenv("BM4DOF_DEVICE");
		if (bm4dof_ctx_create(&ctx, d ? std::atoi(d) : 0) != BM4DOF_OK)
			throw std::runtime_error(bm4dof_global_error());
	}
	bm4dof_ctx_set_option(ctx, "vbm3d_search", predictive_on() ? 1 : 0);
	return ctx;
}

int nparts()
{
	if (const char *s = std::getenv("BM4DOF_NPARTS")) return std::max(1, std::atoi(s));
	if (const char *s = std::getenv("OMP_NUM_THREADS")) return 2 * std::max(1, std::atoi(s));
	return 2;
}

} // namespace

template <> void Video<float>::loadVideo(const std::string &path, unsigned first, unsigned last, unsigned step)
{
	clear();
	if (step == 0) throw std::runtime_error("Video<T>::loadVideo: frame step is zero");
	const unsigned frames = (last - first + 1) / step;
	char name[1024];
	unsigned t = 0;
	for (unsigned f = first; f <= last && t < frames; f += step, t++) {
		std::snprintf(name, sizeof name, path.c_str(), f);
		float *img = nullptr;
		int w, h, c;
		if (bm4dof_io_read(name, &img, &w, &h, &c) != 0)
			throw std::runtime_error("Video<T>::loadVideo: loading of " + std::string(name) + " failed");
		if (c > 2) {
			// a colour file whose first three channels agree is gray
			size_t k = 0;
			const size_t wh = (size_t)w * h;
			while (k < wh && img[k] == img[wh + k] && img[k] == img[2 * wh + k]) k++;
			c = k == wh ? 1 : 3;
		}
		if (t == 0) {
			sz = VideoSize((unsigned)w, (unsigned)h, frames, (unsigned)c);
			data.assign(sz.whcf, 0.f);
		} else if ((unsigned)w != sz.width || (unsigned)h != sz.height || (unsigned)c != sz.channels) {
			bm4dof_io_free(img);
			throw std::runtime_error("Video<T>::loadVideo: size of " + std::string(name) +
			                         " does not match video size");
		}
		std::memcpy(&data[(size_t)t * sz.whc], img, (size_t)sz.whc * sizeof(float));
		bm4dof_io_free(img);
	}
}

template <>
void Video<float>::saveVideo(const std::string &path, unsigned first, unsigned step, float pmin, float pmax) const
{
	(void)pmin;   // the reference does not clip either (lib_image.cpp:108-132)
	(void)pmax;
	char name[1024];
	for (unsigned t = 0; t < sz.frames; t++) {
		std::snprintf(name, sizeof name, path.c_str(), first + t * step);
		if (bm4dof_io_write(name, &data[(size_t)t * sz.whc], (int)sz.width, (int)sz.height, (int)sz.channels) != 0)
			throw std::runtime_error("Video<T>::saveVideo: writing of " + std::string(name) + " failed");
	}
}

namespace VideoNLB {

void initializeNlbParameters(nlbParams &params, const unsigned step, const float sigma, const VideoSize &size,
                             const bool verbose)
{
	bm4dof_params c;
	bm4dof_init_params(&c, step, sigma, size.width, size.height, size.frames, size.channels, verbose);
	params = from_c(c);
}

void setSizeSearchWindow(nlbParams &prms, unsigned sizeSearchWindow)
{
	bm4dof_params c = to_c(prms);
	bm4dof_set_search_window(&c, sizeSearchWindow);
	prms = from_c(c);
}

void setSizePatch(nlbParams &prms, const VideoSize &size, unsigned sizePatch)
{
	bm4dof_params c = to_c(prms);
	bm4dof_set_patch_size(&c, size.frames, sizePatch);
	prms = from_c(c);
}

void setNSimilarPatches(nlbParams &prms, unsigned nSimilarPatches)
{
	bm4dof_params c = to_c(prms);
	bm4dof_set_n_similar(&c, nSimilarPatches);
	prms = from_c(c);
}

void setPredictiveSearch(bool on) { g_predictive = on ? 1 : 0; }
bool predictiveSearch() { return predictive_on(); }

void printNlbParameters(const nlbParams &p)
{
	// the reference prints the backward range on both sides (nldct.cpp:223)
	const int px = (int)p.sizePatch, pt = (int)p.sizePatchTime, wx = (int)p.sizeSearchWindow;
	const int wtb = (int)p.sizeSearchTimeRangeBwd;
	std::printf("\x1b[37;01mParameters for step %d:\x1b[0m\n", p.isFirstStep ? 1 : 2);
	std::printf("\tPatch search:\n");
	std::printf("\t\tPatch size                  = %dx%dx%d\n", px, px, pt);
	std::printf("\t\tNumber of patches           = %d\n", (int)p.nSimilarPatches);
	std::printf("\t\tSpatial search window       = %dx%d\n", wx, wx);
	std::printf("\t\tTemporal search window      = [-%d,%d]\n", wtb, wtb);
	std::printf("\t\tDistance threshold (tau)    = %g\n", p.tau);
	if (predictive_on()) {   // lines of a VBM3D_SEARCH build (nldct.cpp:233-239)
		const int wxpred = (int)p.sizeSearchWindowPred;
		std::printf("\t\tPred. spatial search window = %dx%d\n", wxpred, wxpred);
		std::printf("\t\tPred. distance bias (dsub)  = %g\n", p.dsub);
		std::printf("\t\tSimilar patches per frame   = %d\n", (int)p.nSimilarPatchesPred);
	}
	std::printf("\tGroup filtering:\n");
	std::printf("\t\tTransform                   = %s\n", p.transform == dct ? "dct" : "bior1.5");
	std::printf("\t\tBeta                        = %g\n", p.beta);
	std::printf("\t\tOrder invariance            = %s\n", p.orderInvariance ? "yes" : "no");
	std::printf("\tAggregation:\n");
	std::printf("\t\tOffset                      = %d\n", (int)p.offSet);
	std::printf("\t\tPasteBoost                  = %s\n", p.doPasteBoost ? "active" : "inactive");
	std::printf("\t\tAggregation window          = %s\n", p.agg_window ? "yes" : "no");
	std::printf("\n");
}

std::vector<float> runNlBayes(const Video<float> &i_imNoisy, const Video<float> &i_fflow,
                              const Video<float> &i_bflow, Video<float> &o_imBasic, Video<float> &o_imFinal,
                              const nlbParams p_params1, const nlbParams p_params2)
{
	const VideoSize &s = i_imNoisy.sz;
	if (!(s.channels == 1 || s.channels == 3 || s.channels == 4))
		throw std::runtime_error("VideoNLB::runNlBayes: Wrong number of channels. Must be 1, 3 or 4.");
	const bool use_oflow = i_fflow.sz.width > 0;
	if (use_oflow && (i_fflow.sz.channels != 2 || i_fflow.sz.width != s.width || i_fflow.sz.height != s.height ||
	                  i_fflow.sz.frames != s.frames))
		throw std::runtime_error("VideoNLB::runNlBayes: Wrong size of fwd optical flow.");
	if (use_oflow && (i_bflow.sz.channels != 2 || i_bflow.sz.width != s.width || i_bflow.sz.height != s.height ||
	                  i_bflow.sz.frames != s.frames))
		throw std::runtime_error("VideoNLB::runNlBayes: Wrong size of bwd optical flow.");
	const int step = p_params1.sizePatch ? (p_params2.sizePatch ? 0 : 1) : (p_params2.sizePatch ? 2 : -1);
	if (step == -1) throw std::runtime_error("VideoNLB::runNlBayes: Both patch sizes are zero.");
	if (step != 2) o_imBasic.resize(s);
	if (step != 1) o_imFinal.resize(s);
	if (step == 2 && s != o_imBasic.sz)
		throw std::runtime_error("VideoNLB::runNlBayes: sizes of noisy and basic videos don't match");
	if (p_params1.verbose && step != 2) printNlbParameters(p_params1);
	if (p_params2.verbose && step != 1) printNlbParameters(p_params2);
	const bm4dof_params c1 = to_c(p_params1), c2 = to_c(p_params2);
	bm4dof_ctx *ctx = context();
	const int np = nparts();
	// step headers of nldct.cpp:352-357, 411-416 (the per-tile progress
	// lines the reference draws between them have no counterpart here)
	if (p_params1.verbose && step != 2) {
		std::printf("1st Step\n");
		for (int p = 0; p < np; ++p) std::printf("\n");
	}
	if (p_params2.verbose && step != 1) {
		if (step == 2)
			for (int p = 0; p <= np; ++p) std::printf("\n");
		std::printf("\x1b[%dF2nd Step\n", np + 1);
		for (int p = 0; p < np; ++p) std::printf("\x1b[2K\n");
	}
	std::fflush(stdout);
	float ratio[2] = {0.f, 0.f};
	const int rc = bm4dof_run(ctx, &c1, &c2, i_imNoisy.data.data(), s.width, s.height, s.frames, s.channels,
	                          use_oflow ? i_fflow.data.data() : nullptr, use_oflow ? i_bflow.data.data() : nullptr,
	                          o_imBasic.data.data(), step != 1 ? o_imFinal.data.data() : nullptr, np, ratio);
	if (rc != BM4DOF_OK) throw std::runtime_error(bm4dof_last_error(ctx));
	if (p_params2.verbose && step != 1) std::printf("\n");
	return {ratio[0], ratio[1]};
}

} // namespace VideoNLB
