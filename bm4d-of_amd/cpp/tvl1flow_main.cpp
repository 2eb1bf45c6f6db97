// tvl1flow: TV-L1 optical flow between two images, with the positional
// arguments, defaults, parameter checks and messages of
// lib/tvl1flow/main.c:73-196.  The flow is computed on the GPU
// (Dual_TVL1_optic_flow_multiscale of libbm4dof.so); the images are read as
// gray with the reference's rule (bm4dof_io_read_gray).
//
//   tvl1flow I0 I1 [out nproc tau lambda theta nscales fscale zfactor
//                    nwarps epsilon verbose]
//
// nproc is accepted and ignored (it sets the reference's OpenMP threads).
#include "../../include/bm4dof.h"
#include "../../include/bm4dof_io.h"
#include "tvl1_cli.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

// defaults of lib/tvl1flow/main.c:25-35
const char *const DEF_OUT = "flow.flo";
const int DEF_NPROC = 0;
const float DEF_TAU = 0.25f, DEF_LAMBDA = 0.15f, DEF_THETA = 0.3f, DEF_ZFACTOR = 0.5f, DEF_EPSILON = 0.01f;
const int DEF_NSCALES = 100, DEF_FSCALE = 0, DEF_NWARPS = 5, DEF_VERBOSE = 0;

float *read_gray(const char *name, int *w, int *h)
{
	float *f = nullptr;
	if (bm4dof_io_read_gray(name, &f, w, h) != 0) {
		std::fprintf(stderr, "ERROR: could not read image from file \"%s\"\n", name);
		return nullptr;
	}
	return f;
}

} // namespace

int main(int argc, char **argv)
{
	if (argc < 3) {
		std::fprintf(stderr,
		             "Usage: %s I0 I1 [out nproc tau lambda theta nscales fscale zfactor nwarps epsilon verbose]\n",
		             argv[0]);
		return EXIT_FAILURE;
	}
	int i = 3;
	auto next = [&](void) -> const char * { return argc > i ? argv[i++] : (i++, nullptr); };
	const char *in0 = argv[1], *in1 = argv[2];
	const char *a;
	const char *out = (a = next()) ? a : DEF_OUT;
	int nproc = (a = next()) ? std::atoi(a) : DEF_NPROC;
	float tau = (a = next()) ? (float)std::atof(a) : DEF_TAU;
	float lambda = (a = next()) ? (float)std::atof(a) : DEF_LAMBDA;
	float theta = (a = next()) ? (float)std::atof(a) : DEF_THETA;
	int nscales = (a = next()) ? std::atoi(a) : DEF_NSCALES;
	int fscale = (a = next()) ? std::atoi(a) : DEF_FSCALE;
	float zfactor = (a = next()) ? (float)std::atof(a) : DEF_ZFACTOR;
	int nwarps = (a = next()) ? std::atoi(a) : DEF_NWARPS;
	float epsilon = (a = next()) ? (float)std::atof(a) : DEF_EPSILON;
	const int verbose = (a = next()) ? std::atoi(a) : DEF_VERBOSE;

	// parameter checks of main.c:102-149: out-of-range values take the default
	if (nproc < 0) {
		nproc = DEF_NPROC;
		if (verbose) std::fprintf(stderr, "warning: nproc changed to %d\n", nproc);
	}
	if (tau <= 0 || tau > 0.25) {
		tau = DEF_TAU;
		if (verbose) std::fprintf(stderr, "warning: tau changed to %g\n", tau);
	}
	if (lambda <= 0) {
		lambda = DEF_LAMBDA;
		if (verbose) std::fprintf(stderr, "warning: lambda changed to %g\n", lambda);
	}
	if (theta <= 0) {
		theta = DEF_THETA;
		if (verbose) std::fprintf(stderr, "warning: theta changed to %g\n", theta);
	}
	if (nscales <= 0) {
		nscales = DEF_NSCALES;
		if (verbose) std::fprintf(stderr, "warning: nscales changed to %d\n", nscales);
	}
	if (zfactor <= 0 || zfactor >= 1) {
		zfactor = DEF_ZFACTOR;
		if (verbose) std::fprintf(stderr, "warning: zfactor changed to %g\n", zfactor);
	}
	if (nwarps <= 0) {
		nwarps = DEF_NWARPS;
		if (verbose) std::fprintf(stderr, "warning: nwarps changed to %d\n", nwarps);
	}
	if (epsilon <= 0) {
		epsilon = DEF_EPSILON;
		if (verbose) std::fprintf(stderr, "warning: epsilon changed to %f\n", epsilon);
	}

	int nx = 0, ny = 0, nx2 = 0, ny2 = 0;
	float *I0 = read_gray(in0, &nx, &ny);
	float *I1 = read_gray(in1, &nx2, &ny2);
	if (!I0 || !I1 || nx != nx2 || ny != ny2) {
		if (I0 && I1)
			std::fprintf(stderr, "ERROR: input images size mismatch %dx%d != %dx%d\n", nx, ny, nx2, ny2);
		bm4dof_io_free(I0);
		bm4dof_io_free(I1);
		return EXIT_FAILURE;
	}
	// the coarsest scale stays at least 16 pixels across (main.c:159-165)
	nscales = tvl1_cap_nscales(nx, ny, zfactor, nscales);
	if (nscales < fscale) fscale = nscales;
	if (verbose)
		std::fprintf(stderr, "nproc=%d tau=%f lambda=%f theta=%f nscales=%d zfactor=%f nwarps=%d epsilon=%g\n", nproc,
		             tau, lambda, theta, nscales, zfactor, nwarps, epsilon);

	std::vector<float> uv((size_t)2 * nx * ny);
	Dual_TVL1_optic_flow_multiscale(I0, I1, uv.data(), uv.data() + (size_t)nx * ny, nx, ny, tau, lambda, theta,
	                                nscales, fscale, zfactor, nwarps, epsilon, verbose != 0);
	bm4dof_io_free(I0);
	bm4dof_io_free(I1);
	if (bm4dof_io_write(out, uv.data(), nx, ny, 2) != 0) {
		std::fprintf(stderr, "ERROR: %s\n", bm4dof_io_error());
		return EXIT_FAILURE;
	}
	return EXIT_SUCCESS;
}
