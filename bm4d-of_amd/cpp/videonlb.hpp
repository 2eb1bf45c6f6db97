// C++ host interface of the denoiser, with the names, argument meaning and
// error behaviour of the reference's VideoNLB / Video API, so a caller
// written against src/core/nldct.h and src/utils/lib_videot.hpp compiles
// against this header.  The work runs on the GPU through libbm4dof.so's
// C-ABI (include/bm4dof.h); this layer only converts types.
//
//   VideoSize, Video<float>            src/utils/lib_videot.hpp:40-181
//   VideoNLB::nlbParams and enums      src/core/nldct.h:33-82
//   initializeNlbParameters, set*      src/core/nldct.cpp:81-210
//   printNlbParameters                 src/core/nldct.cpp:219-251
//   runNlBayes                         src/core/nldct.cpp:267-478
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

struct VideoSize {
	unsigned width = 0, height = 0, frames = 0, channels = 0;
	unsigned wh = 0, whc = 0, whcf = 0, whf = 0;

	VideoSize() { update_fields(); }
	VideoSize(unsigned w, unsigned h, unsigned f, unsigned c) : width(w), height(h), frames(f), channels(c)
	{
		update_fields();
	}
	void update_fields()
	{
		wh = width * height;
		whc = wh * channels;
		whcf = whc * frames;
		whf = wh * frames;
	}
	bool operator==(const VideoSize &o) const
	{
		return width == o.width && height == o.height && channels == o.channels && frames == o.frames;
	}
	bool operator!=(const VideoSize &o) const { return !(*this == o); }
};

// Video with frames of planar channels: data[t*whc + c*wh + y*w + x].
template <class T> class Video {
public:
	VideoSize sz;
	std::vector<T> data;

	Video() = default;
	Video(unsigned w, unsigned h, unsigned f, unsigned c = 1, T v = T(0)) : sz(w, h, f, c), data(sz.whcf, v) {}
	explicit Video(const VideoSize &s, T v = T(0)) : sz(s), data(s.whcf, v) {}
	Video(const std::string &path, unsigned first, unsigned last, unsigned step = 1) { loadVideo(path, first, last, step); }

	void clear()
	{
		sz = VideoSize();
		data.clear();
	}
	void resize(unsigned w, unsigned h, unsigned f, unsigned c = 1) { resize(VideoSize(w, h, f, c)); }
	void resize(const VideoSize &s)
	{
		if (s != sz) {
			sz = s;
			data.assign(sz.whcf, T(0));
		}
	}
	unsigned index(unsigned x, unsigned y, unsigned t, unsigned c = 0) const
	{
		return t * sz.whc + c * sz.wh + y * sz.width + x;
	}
	T &operator()(unsigned i) { return data[i]; }
	T operator()(unsigned i) const { return data[i]; }
	T &operator()(unsigned x, unsigned y, unsigned t, unsigned c = 0) { return data[index(x, y, t, c)]; }
	T operator()(unsigned x, unsigned y, unsigned t, unsigned c = 0) const { return data[index(x, y, t, c)]; }

	// Frames first, first+step, ..., last of a printf path pattern; PNG,
	// TIFF or .flo frames; a 3- or 4-channel frame whose first three
	// channels are equal reads as gray (src/utils/lib_image.cpp:66-73).
	void loadVideo(const std::string &path, unsigned first, unsigned last, unsigned step = 1);
	// One file per frame, numbered from `first` by `step`.  .png quantises.
	void saveVideo(const std::string &path, unsigned first, unsigned step = 1, T pmin = 0, T pmax = 255) const;
};

namespace VideoNLB {

enum ColorSpace { RGB, YUV };
enum Transform { dct, bior1_5 };

struct nlbParams {
	float sigma;
	unsigned sizePatch;
	unsigned sizePatchTime;
	unsigned nSimilarPatches;
	unsigned sizeSearchWindow;
	unsigned sizeSearchTimeRangeFwd;
	unsigned sizeSearchTimeRangeBwd;
	unsigned boundary;
	unsigned offSet;
	unsigned offSetTime;
	float beta;
	float betaMean;
	float tau;
	bool orderInvariance;
	bool isFirstStep;
	bool doPasteBoost;
	bool verbose;
	ColorSpace colorSpace;
	Transform transform;
	unsigned sizeSearchWindowPred;
	unsigned nSimilarPatchesPred;
	float dsub;
	bool agg_window;
};

void initializeNlbParameters(nlbParams &params, const unsigned step, const float sigma, const VideoSize &size,
                             const bool verbose);
void setSizeSearchWindow(nlbParams &prms, unsigned sizeSearchWindow);
void setSizePatch(nlbParams &prms, const VideoSize &size, unsigned sizePatch);
void setNSimilarPatches(nlbParams &prms, unsigned nSimilarPatches);
void printNlbParameters(const nlbParams &params);

// Both steps (or one, selected by a zero sizePatch) on the GPU.  Returns the
// percentage of processed groups per step.  The tiling emulation uses
// nParts = BM4DOF_NPARTS if set, else 2 x OMP_NUM_THREADS if set, else 2
// (the reference uses 2 x its OpenMP threads, nldct.cpp:318-323).
std::vector<float> runNlBayes(const Video<float> &i_imNoisy, const Video<float> &i_fflow,
                              const Video<float> &i_bflow, Video<float> &o_imBasic, Video<float> &o_imFinal,
                              const nlbParams p_params1, const nlbParams p_params2);

// The reference selects its predictive patch search at compile time
// (VBM3D_SEARCH, src/core/nldct.h:20; off by default): here a run-time
// switch, for the block matching of runNlBayes and for printNlbParameters'
// extra lines.  Also set by BM4DOF_VBM3D_SEARCH=1 in the environment.
void setPredictiveSearch(bool on);
bool predictiveSearch();

} // namespace VideoNLB
