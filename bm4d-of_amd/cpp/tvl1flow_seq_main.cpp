// tvl1flow-seq: forward and backward TV-L1 flow of a whole sequence, the
// work of scripts/tvl1flow-seq.sh:1-61 done in one process on the GPU.
//
//   tvl1flow-seq SEQ FIRST LAST OUT [all|fwd|bwd]
//
// SEQ is a printf pattern of the frames.  Written files, as the script names
// them:  OUT/tvl1_%03d_b.flo  for i in FIRST+1..LAST: flow I_i -> I_{i-1},
// then a copy of FIRST+1 as FIRST;  OUT/tvl1_%03d_f.flo  for i in
// FIRST..LAST-1: flow I_i -> I_{i+1}, then a copy of LAST-1 as LAST.
// Existing flow files are kept, as the script keeps them.  Parameters are
// the script's: tau 0.25, lambda 0.2, theta 0.3, nscales 100 (capped by the
// image size as tvl1flow caps it), fscale 1, zfactor 0.5, nwarps 5,
// epsilon 0.01.  Frames are read as gray, as tvl1flow reads them.
//
// When every requested flow is missing and both directions are asked for,
// all pairs are solved together (bm4dof_tvl1_sequence); otherwise each
// missing pair is solved alone (bm4dof_tvl1).  Each pair's flow is the same
// either way.  A sequence of one frame is an error here (the script would
// fail on its copy step).
#include "../../include/bm4dof.h"
#include "../../include/bm4dof_io.h"
#include "tvl1_cli.hpp"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

std::string fmt(const std::string &pattern, int i)
{
	char buf[4096];
	std::snprintf(buf, sizeof buf, pattern.c_str(), i);
	return buf;
}

bool exists(const std::string &p)
{
	struct stat st;
	return ::stat(p.c_str(), &st) == 0;
}

bool mkdir_p(const std::string &dir)
{
	std::string cur;
	for (size_t i = 0; i <= dir.size(); i++) {
		if (i == dir.size() || dir[i] == '/') {
			if (!cur.empty() && ::mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return false;
		}
		if (i < dir.size()) cur += dir[i];
	}
	return true;
}

// interleaved write of a (u, v) plane pair
bool write_flow(const std::string &path, const float *u, const float *v, int nx, int ny)
{
	std::vector<float> uv((size_t)2 * nx * ny);
	std::memcpy(uv.data(), u, sizeof(float) * nx * ny);
	std::memcpy(uv.data() + (size_t)nx * ny, v, sizeof(float) * nx * ny);
	if (bm4dof_io_write(path.c_str(), uv.data(), nx, ny, 2) != 0) {
		std::fprintf(stderr, "ERROR: %s\n", bm4dof_io_error());
		return false;
	}
	return true;
}

bool copy_file(const std::string &from, const std::string &to)
{
	FILE *a = std::fopen(from.c_str(), "rb");
	if (!a) return false;
	FILE *b = std::fopen(to.c_str(), "wb");
	if (!b) {
		std::fclose(a);
		return false;
	}
	char buf[1 << 16];
	size_t n;
	bool ok = true;
	while ((n = std::fread(buf, 1, sizeof buf, a)) > 0) ok &= std::fwrite(buf, 1, n, b) == n;
	std::fclose(a);
	ok &= std::fclose(b) == 0;
	return ok;
}

int fail(bm4dof_ctx *ctx, const char *what)
{
	std::fprintf(stderr, "tvl1flow-seq: %s: %s\n", what, ctx ? bm4dof_last_error(ctx) : bm4dof_global_error());
	return EXIT_FAILURE;
}

} // namespace

int main(int argc, char **argv)
{
	if (argc < 5) {
		std::fprintf(stderr, "Usage: %s SEQ FIRST LAST OUT [all|fwd|bwd]\n", argv[0]);
		return EXIT_FAILURE;
	}
	const std::string seq = argv[1], out = argv[4], dir = argc > 5 ? argv[5] : "all";
	const int ffr = std::atoi(argv[2]), lfr = std::atoi(argv[3]);
	const bool do_b = dir == "bwd" || dir == "all", do_f = dir == "fwd" || dir == "all";
	if (!mkdir_p(out)) {
		std::fprintf(stderr, "tvl1flow-seq: cannot create %s\n", out.c_str());
		return EXIT_FAILURE;
	}
	for (int i = ffr; i <= lfr; i++)
		if (!exists(fmt(seq, i))) {
			std::printf("ERROR: %s not found\n", fmt(seq, i).c_str());
			return EXIT_FAILURE;
		}
	const int F = lfr - ffr + 1;
	if (F < 2) {
		std::fprintf(stderr, "tvl1flow-seq: at least two frames are needed\n");
		return EXIT_FAILURE;
	}

	// frames, gray
	int nx = 0, ny = 0;
	std::vector<float> frames;
	for (int i = ffr; i <= lfr; i++) {
		float *img = nullptr;
		int w, h;
		if (bm4dof_io_read_gray(fmt(seq, i).c_str(), &img, &w, &h) != 0) {
			std::fprintf(stderr, "ERROR: could not read image from file \"%s\"\n", fmt(seq, i).c_str());
			return EXIT_FAILURE;
		}
		if (i == ffr) {
			nx = w;
			ny = h;
			frames.resize((size_t)nx * ny * F);
		} else if (w != nx || h != ny) {
			std::fprintf(stderr, "ERROR: input images size mismatch %dx%d != %dx%d\n", nx, ny, w, h);
			bm4dof_io_free(img);
			return EXIT_FAILURE;
		}
		std::memcpy(&frames[(size_t)(i - ffr) * nx * ny], img, sizeof(float) * nx * ny);
		bm4dof_io_free(img);
	}

	bm4dof_tvl1_params prm;
	prm.tau = 0.25f;
	prm.lambda = 0.2f;
	prm.theta = 0.3f;
	prm.zfactor = 0.5f;
	prm.nscales = tvl1_cap_nscales(nx, ny, prm.zfactor, 100);
	prm.fscale = prm.nscales < 1 ? prm.nscales : 1;
	prm.warps = 5;
	prm.epsilon = 0.01f;
	prm.verbose = 0;

	const std::string bpat = out + "/tvl1_%03d_b.flo", fpat = out + "/tvl1_%03d_f.flo";
	bool all_missing = do_b && do_f;
	for (int i = ffr + 1; i <= lfr && all_missing; i++) all_missing = !exists(fmt(bpat, i));
	for (int i = ffr; i < lfr && all_missing; i++) all_missing = !exists(fmt(fpat, i));

	bm4dof_ctx *ctx = nullptr;
	const char *dev = std::getenv("BM4DOF_DEVICE");
	if (bm4dof_ctx_create(&ctx, dev ? std::atoi(dev) : 0) != BM4DOF_OK) return fail(nullptr, "no device");
	const size_t n = (size_t)nx * ny;
	int rc = EXIT_SUCCESS;
	if (all_missing) {
		std::vector<float> ff(2 * n * F), bf(2 * n * F);
		if (bm4dof_tvl1_sequence(ctx, &prm, frames.data(), nx, ny, F, ff.data(), bf.data(), 0) != BM4DOF_OK)
			rc = fail(ctx, "flow");
		for (int i = ffr + 1; i <= lfr && rc == EXIT_SUCCESS; i++) {
			const float *b = &bf[2 * n * (i - ffr)];
			if (!write_flow(fmt(bpat, i), b, b + n, nx, ny)) rc = EXIT_FAILURE;
		}
		for (int i = ffr; i < lfr && rc == EXIT_SUCCESS; i++) {
			const float *f = &ff[2 * n * (i - ffr)];
			if (!write_flow(fmt(fpat, i), f, f + n, nx, ny)) rc = EXIT_FAILURE;
		}
	} else {
		std::vector<float> u(n), v(n);
		auto pair = [&](int i0, int i1, const std::string &path) {
			if (rc != EXIT_SUCCESS || exists(path)) return;
			if (bm4dof_tvl1(ctx, &prm, &frames[n * (i0 - ffr)], &frames[n * (i1 - ffr)], u.data(), v.data(), nx, ny, 0,
			                nullptr) != BM4DOF_OK)
				rc = fail(ctx, "flow");
			else if (!write_flow(path, u.data(), v.data(), nx, ny))
				rc = EXIT_FAILURE;
		};
		if (do_b)
			for (int i = ffr + 1; i <= lfr; i++) pair(i, i - 1, fmt(bpat, i));
		if (do_f)
			for (int i = ffr; i < lfr; i++) pair(i, i + 1, fmt(fpat, i));
	}
	// end copies of the script
	if (rc == EXIT_SUCCESS && do_b && !copy_file(fmt(bpat, ffr + 1), fmt(bpat, ffr))) rc = EXIT_FAILURE;
	if (rc == EXIT_SUCCESS && do_f && !copy_file(fmt(fpat, lfr - 1), fmt(fpat, lfr))) rc = EXIT_FAILURE;
	bm4dof_ctx_destroy(ctx);
	return rc;
}
