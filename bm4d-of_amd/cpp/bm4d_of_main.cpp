// bm4d-of: the denoiser's command-line tool, with the flags, defaults, checks,
// messages and override order of src/main_vnlb.cpp:30-264.  The denoising
// runs on the GPU (libbm4dof.so); reading and writing frames uses
// libbm4dof_io.so.
//
// Behaviour kept from the reference tool on purpose, so that the same command
// line gives the same output:
//   * -t1/-t2 change tau only with -print-prms: on the denoising path the
//     reference applies them inside a block compiled only with VBM3D_SEARCH,
//     which it does not define (main_vnlb.cpp:233-240, nldct.h:20);
//   * -vbm3d-search (no flag of the reference: its compile-time switch made a
//     run-time one) selects the predictive patch search of a VBM3D_SEARCH
//     build; then -wxf1/-wxf2/-npf1/-npf2/-dsub1/-dsub2 apply and -t1/-t2
//     apply on the denoising path too, as in such a build
//     (main_vnlb.cpp:71-78, 165-172, 227-236);
//   * optical flow is used only when last - first > step (unsigned
//     arithmetic, main_vnlb.cpp:136);
//   * -b1 sets both beta and betaMean; -bm1 then overrides betaMean.
#include "cli_option.hpp"
#include "videonlb.hpp"

#include <cstdio>
#include <cstdlib>
#include <exception>

namespace {

enum Mode { BSIC_DENO, BSIC_ONLY, DENO_ONLY };

struct Overrides {
	int wx1, wx2, px1, px2, pt1, pt2, np1, np2;
	float tau1, tau2, beta1, beta2, bm1, bm2;
	bool no_paste1, no_paste2, no_step1, no_step2, agg_win1, agg_win2, order_inv;
	int wxf1, wxf2, npf1, npf2;
	float dsub1, dsub2;
	bool predictive;
};

// Command-line values over the defaults, in the reference's order
// (main_vnlb.cpp:147-186 and 206-246).
void apply(VideoNLB::nlbParams &p1, VideoNLB::nlbParams &p2, const VideoSize &sz, const Overrides &o, bool with_tau)
{
	if (o.wx1 >= 0) VideoNLB::setSizeSearchWindow(p1, (unsigned)o.wx1);
	if (o.wx2 >= 0) VideoNLB::setSizeSearchWindow(p2, (unsigned)o.wx2);
	if (o.px1 >= 0) VideoNLB::setSizePatch(p1, sz, (unsigned)o.px1);
	if (o.px2 >= 0) VideoNLB::setSizePatch(p2, sz, (unsigned)o.px2);
	if (o.np1 >= 0) VideoNLB::setNSimilarPatches(p1, (unsigned)o.np1);
	if (o.np2 >= 0) VideoNLB::setNSimilarPatches(p2, (unsigned)o.np2);
	if (o.beta1 >= 0) p1.beta = p1.betaMean = o.beta1;
	if (o.beta2 >= 0) p2.beta = p2.betaMean = o.beta2;
	if (o.bm1 >= 0) p1.betaMean = o.bm1;
	if (o.bm2 >= 0) p2.betaMean = o.bm2;
	if (o.pt1 >= 0) p1.sizePatchTime = (unsigned)o.pt1;
	if (o.pt2 >= 0) p2.sizePatchTime = (unsigned)o.pt2;
	if (o.predictive) {
		if (o.wxf1 >= 0) p1.sizeSearchWindowPred = (unsigned)o.wxf1;
		if (o.wxf2 >= 0) p2.sizeSearchWindowPred = (unsigned)o.wxf2;
		if (o.npf1 >= 0) p1.nSimilarPatchesPred = (unsigned)o.npf1;
		if (o.npf2 >= 0) p2.nSimilarPatchesPred = (unsigned)o.npf2;
		if (o.dsub1 >= 0) p1.dsub = o.dsub1;
		if (o.dsub2 >= 0) p2.dsub = o.dsub2;
	}
	if ((with_tau || o.predictive) && o.tau1 >= 0) p1.tau = o.tau1;
	if ((with_tau || o.predictive) && o.tau2 >= 0) p2.tau = o.tau2;
	if (o.agg_win1) p1.agg_window = true;
	if (o.agg_win2) p2.agg_window = true;
	if (o.no_paste1) p1.doPasteBoost = false;
	if (o.no_paste2) p2.doPasteBoost = false;
	if (o.no_step1) p1.offSet = 1;
	if (o.no_step2) p2.offSet = 1;
	if (o.order_inv) p1.orderInvariance = p2.orderInvariance = true;
}

int usage_error(const char *prog, const char *msg)
{
	std::fprintf(stderr, "%s: %s\nTry `%s --help' for more information.\n", prog, msg, prog);
	return EXIT_FAILURE;
}

int run(int argc, char **argv)
{
	CliOptions opt(argc, argv, "BM4D-OF video denoising",
	               " NOTE: Input (<) and output (>) sequences are specified by their paths in printf format.\n");

	const std::string input_path = opt.get("-i", "", "< input sequence");
	const std::string inbsc_path = opt.get("-b", "", "< input basic sequence");
	const std::string final_path = opt.get("-deno", "deno_%03d.png", "> denoised sequence");
	const std::string basic_path = opt.get("-bsic", "bsic_%03d.png", "> basic denoised sequence");
	const unsigned first_frame = (unsigned)opt.get("-f", 0, "first frame");
	const unsigned last_frame = (unsigned)opt.get("-l", 0, "last frame");
	const unsigned frame_step = (unsigned)opt.get("-s", 1, "frame step");
	const std::string fflow_path = opt.get("-fof", "", "< input forward  optical flow");
	const std::string bflow_path = opt.get("-bof", "", "< input backward optical flow");
	const float sigma = opt.get("-sigma", 0.f, "< standard deviation of the noise");
	const bool verbose = opt.get("-verbose", true, "< verbose output");
	Overrides o;
	o.order_inv = opt.get("-order-inv", false, "< invariance to patch order in stack");
	const unsigned print_prms = (unsigned)opt.get("-print-prms", 0, "< prints parameters for given channels");
	const int wt1 = opt.get("-wt1", 0, "< search window temporal radius, step 1");
	const int wt2 = opt.get("-wt2", 0, "< search window temporal radius, step 2");
	o.wx1 = opt.get("-wx1", -1, "< search window spatial radius, step 1");
	o.wx2 = opt.get("-wx2", -1, "< search window spatial radius, step 2");
	o.px1 = opt.get("-px1", -1, "< spatial patch size, step 1");
	o.px2 = opt.get("-px2", -1, "< spatial patch size, step 2");
	o.pt1 = opt.get("-pt1", 1, "< temporal patch size, step 1");
	o.pt2 = opt.get("-pt2", 1, "< temporal patch size, step 2");
	o.np1 = opt.get("-np1", -1, "< number of similar patches, step 1");
	o.np2 = opt.get("-np2", -1, "< number of similar patches, step 2");
	o.tau1 = opt.get("-t1", -1.f, "< step 1 distance threshold");
	o.tau2 = opt.get("-t2", -1.f, "< step 2 distance threshold");
	o.predictive = opt.get("-vbm3d-search", false, "< predictive patch search (as built with VBM3D_SEARCH)");
	o.wxf1 = opt.get("-wxf1", -1, "< search window for predictive search, step 1");
	o.wxf2 = opt.get("-wxf2", -1, "< search window for predictive search, step 2");
	o.npf1 = opt.get("-npf1", -1, "< number of similar patches per frame, step 1");
	o.npf2 = opt.get("-npf2", -1, "< number of similar patches per frame, step 2");
	o.dsub1 = opt.get("-dsub1", -1.f, "< distance bias, step 1");
	o.dsub2 = opt.get("-dsub2", -1.f, "< distance bias, step 2");
	VideoNLB::setPredictiveSearch(o.predictive || VideoNLB::predictiveSearch());
	o.predictive = VideoNLB::predictiveSearch();
	o.beta1 = opt.get("-b1", -1.f, "< noise correction factor beta, step 1");
	o.beta2 = opt.get("-b2", -1.f, "< noise correction factor beta, step 2");
	o.bm1 = opt.get("-bm1", -1.f, "< noise correction factor beta for mean, step 1");
	o.bm2 = opt.get("-bm2", -1.f, "< noise correction factor beta for mean, step 2");
	o.no_paste1 = opt.get("-no-paste1", false, "< disable paste trick, step 1");
	o.no_paste2 = opt.get("-no-paste2", false, "< disable paste trick, step 2");
	o.no_step1 = opt.get("-no-step1", false, "< disable patch skipping, step 1");
	o.no_step2 = opt.get("-no-step2", false, "< disable patch skipping, step 2");
	o.agg_win1 = opt.get("-agg-win1", false, "< aggregation window, step 1");
	o.agg_win2 = opt.get("-agg-win2", false, "< aggregation window, step 2");

	// checks of main_vnlb.cpp:89-130
	const char *prog = argv[0];
	if (input_path.empty()) return usage_error(prog, "no input sequence.");
	if (o.px1 == 0 && o.px2 > 0 && inbsc_path.empty())
		return usage_error(prog, "if px1 = 0 and px2 > 0, a basic sequence path must be given.");
	if ((o.px1 < 0 && o.px1 != -1) || (o.px2 < 0 && o.px2 != -1) || o.pt1 < 0 || o.pt2 < 0)
		return usage_error(prog, "px1, px2, pt1 and pt2 cannot be negative.");
	if ((o.np1 < 0 && o.np1 != -1) || (o.np2 < 0 && o.np2 != -1))
		return usage_error(prog, "np1 and np2 cannot be negative.");
	if ((o.wx1 < 0 && o.wx1 != -1) || (o.wx2 < 0 && o.wx2 != -1) || wt1 < 0 || wt2 < 0)
		return usage_error(prog, "wx1, wx2, wt1 and wt2 cannot be negative.");
	if (o.px1 > 0 && !inbsc_path.empty())
		std::fprintf(stderr, "\x1b[33;1mWarning:\x1b[0m a basic sequence path ignored since px1 > 0.\n");
	if (fflow_path.empty() != bflow_path.empty()) {
		std::fprintf(stderr, "Only one oflow path provided.\nTry `%s --help' for more information.\n", prog);
		return EXIT_FAILURE;
	}
	if (o.px1 == 0 && o.px2 == 0) {
		std::fprintf(stderr, "Given patch sizes for both steps are 0. Nothing to do.\n");
		return EXIT_FAILURE;
	}

	const Mode mode = o.px1 != 0 ? (o.px2 != 0 ? BSIC_DENO : BSIC_ONLY) : DENO_ONLY;
	const bool use_oflow = !fflow_path.empty() && (last_frame - first_frame > frame_step);

	if (print_prms) {
		VideoSize tmp;
		tmp.channels = print_prms;
		VideoNLB::nlbParams p1, p2;
		VideoNLB::initializeNlbParameters(p1, 1, sigma, tmp, verbose);
		VideoNLB::initializeNlbParameters(p2, 2, sigma, tmp, verbose);
		apply(p1, p2, tmp, o, true);
		VideoNLB::printNlbParameters(p1);
		VideoNLB::printNlbParameters(p2);
		return EXIT_SUCCESS;
	}

	Video<float> noisy, basic, final_, fflow, bflow;
	noisy.loadVideo(input_path, first_frame, last_frame, frame_step);
	if (mode == DENO_ONLY) basic.loadVideo(inbsc_path, first_frame, last_frame, frame_step);
	if (use_oflow) fflow.loadVideo(fflow_path, first_frame, last_frame, frame_step);
	if (use_oflow) bflow.loadVideo(bflow_path, first_frame, last_frame, frame_step);

	if (verbose) std::printf("Running BM4D-OF on the noisy video\n");
	VideoNLB::nlbParams p1, p2;
	VideoNLB::initializeNlbParameters(p1, 1, sigma, noisy.sz, verbose);
	VideoNLB::initializeNlbParameters(p2, 2, sigma, noisy.sz, verbose);
	apply(p1, p2, noisy.sz, o, false);

	const std::vector<float> ratio = VideoNLB::runNlBayes(noisy, fflow, bflow, basic, final_, p1, p2);
	if (verbose)
		std::printf("Done. Processed %5.2f%% of possible patch groups in 1st step, and\n"
		            "%5.2f%% in 2nd step.\n",
		            ratio[0], ratio[1]);

	if (verbose) std::printf("\nSaving output sequences\n");
	if (p2.sizePatch) final_.saveVideo(final_path, first_frame, frame_step);
	basic.saveVideo(basic_path, first_frame, frame_step);
	return EXIT_SUCCESS;
}

} // namespace

int main(int argc, char **argv)
{
	// The reference lets its std::runtime_error escape (terminate, abort);
	// here the message is printed and the exit status is 1.
	try {
		return run(argc, argv);
	} catch (const std::exception &e) {
		std::fprintf(stderr, "%s\n", e.what());
		return EXIT_FAILURE;
	}
}
