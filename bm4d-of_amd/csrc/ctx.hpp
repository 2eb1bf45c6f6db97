// Host-side context, device buffers and error handling shared by the
// denoiser and the TV-L1 host code.
#pragma once

#include "common.hpp"
#include "../../include/bm4dof.h"

#include <atomic>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace bm4dof {

struct HipError : std::runtime_error {
	explicit HipError(const std::string &s) : std::runtime_error(s) {}
};
struct ArgError : std::runtime_error {
	int code;
	ArgError(const std::string &s, int c = BM4DOF_EINVAL) : std::runtime_error(s), code(c) {}
};

#define HIPCHK(x)                                                                         \
	do {                                                                                  \
		hipError_t e_ = (x);                                                              \
		if (e_ != hipSuccess)                                                             \
			throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));               \
	} while (0)

// Grow-only device buffer.
struct DevBuf {
	void *p = nullptr;
	size_t n = 0;
	void *get(size_t bytes) {
		if (bytes <= n && p) return p;
		if (p) HIPCHK(hipFree(p));
		p = nullptr;
		n = 0;
		if (bytes == 0) bytes = 256;
		if (hipMalloc(&p, bytes) != hipSuccess) {
			(void)hipGetLastError();
			throw ArgError("device allocation of " + std::to_string(bytes) + " bytes failed",
			               BM4DOF_ENOMEM);
		}
		n = bytes;
		return p;
	}
	template <class T> T *as(size_t count) { return (T *)get(count * sizeof(T)); }
	void release() {
		if (p) (void)hipFree(p);
		p = nullptr;
		n = 0;
	}
};

} // namespace bm4dof

struct bm4dof_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	// second stream (default priority) for the position sort of a step's
	// rows, which only the aggregation needs: it runs beside the filter
	hipStream_t side = nullptr;
	hipEvent_t ev_fork = nullptr, ev_join = nullptr;
	// coefficient tables of the current gray step launched on the second
	// stream beside its block matching (ev_coef marks their end)
	hipEvent_t ev_coef = nullptr;
	// TV-L1: pinned host words the solve reads its active counts and
	// iteration counts into while the GPU runs on (1 + 2 tv_pin_pairs ints,
	// allocated at context creation), and the event of a poll
	int *tv_pin = nullptr;
	int tv_pin_pairs = 0;
	hipEvent_t ev_tvpoll = nullptr;
	bool coef_on_side = false;
	const float *coef_side_n = nullptr, *coef_side_b = nullptr;
	std::string err;
	// step buffers
	bm4dof::DevBuf gpos, gtile, tiles, idx, nsim, base, keys, vals, skeys, svals, sort_tmp, lstart, lend,
	    rank_of, etile, entry_w, est, early, gweight, counters, stale, elist, num, wsum,
	    basis, binv, win, noisy_w, basic_w, out_w, io_noisy, io_fflow, io_bflow, io_basic,
	    io_final, planes, planes_b, coef_n, coef_b, gacc, bsplit, ties, bm_image, bm_wins, bm_wmeta, erec, gflag, fmask, lp_idx, stat64, lastf, fidx, geo, scan_tmp,
	    scan_tmp2, bin_cnt, bin_k0, bin_k1, bin_v1, bin_hist, bin_hoff, bin_off, bin_list;
	// timings of the last run
	std::vector<std::pair<const char *, float>> timings;
	std::vector<hipEvent_t> ev;
	size_t est_budget = (size_t)24 << 30;
	// bytes of per-anchor coefficient tables a step may hold (0: none); at
	// context creation the smaller of 96 GiB and 2/5 of the free device memory
	size_t coef_budget = (size_t)96 << 30;
	// keep the tables allocated between calls (else released at each call's end)
	bool keep_tables = false;
	// the noisy coefficient table of the last step that filled it (patch
	// shape, channels, size, basis), and whether the next step of the same
	// run may read it: set around the second step when both steps see the
	// same noisy planes
	bool coef_n_ok = false, coef_n_keep = false;
	int coef_n_shape[4] = {0, 0, 0, 0};
	size_t coef_n_whf = 0;
	std::vector<float> coef_n_basis;
	// ordered-aggregation statistics of the last run
	uint64_t agg_blocks = 0, agg_overflow_blocks = 0;
	int agg_region[3] = {0, 0, 0};
	uint64_t step_groups[2] = {0, 0}, step_rows[2] = {0, 0};   // groups / group rows per step
	bool fast_filter = true;   // register-resident filter where the group shape allows it
	int fast_bm = 2;           // block matching: 2 register windows, 1 LDS-staged windows, 0 general
	int bm_frames = 2;         // frames of search window staged per round of the staged matching
	bool agg_chunked = true;   // chunked (mask transpose) ordered aggregation where tiles allow it
	int agg_cap = 1024;        // rows per e-range of the chunked aggregation (one channel)
	bool agg_binned = true;
	bool coef_beside_bm = true;   // gray steps: coefficient tables on the second stream beside block matching
	int side_blocks = 2048;    // most blocks of a second-stream kernel (0: no limit)    // binned aggregation (row lists per 8 x 8 x 4 bin) where the shape allows it
	bool vbm3d_search = false; // block matching of a VBM3D_SEARCH build (predictive search)
	// learned gain network (mlp.hip)
	bm4dof::DevBuf mlp_theta, mlp_m, mlp_v, mlp_wq, mlp_inf, mlp_ws, mlp_grad, mlp_xb, mlp_tb, mlp_sse;
	int mlp_step = 0;
	bool mlp_ready = false;
	// step plans (grids, bases, tiles) kept on the device for repeated runs of
	// the same shape; owned by the host code that defines them
	std::shared_ptr<void> plan_cache;
	// TV-L1: sum over pairs, scales and warps of executed iterations x pixels
	// of the last flow computation
	uint64_t tv_pixel_iters = 0;
	// TV-L1 buffers
	bm4dof::DevBuf tv_frames, tv_pairs, tv_part, tv_mm, tv_state, tv_work, tv_pyr, tv_taps;

	// Every device buffer of the context: released by bm4dof_ctx_destroy;
	// their addresses are what a captured batch graph depends on.
	std::vector<bm4dof::DevBuf *> buffers()
	{
		return {&gpos, &gtile, &tiles, &idx, &nsim, &base, &keys, &vals, &skeys, &svals, &sort_tmp, &lstart,
		        &lend, &rank_of, &etile, &entry_w, &est, &early, &gweight, &counters, &stale, &elist, &num,
		        &wsum, &basis, &binv, &win, &noisy_w, &basic_w, &out_w, &io_noisy, &io_fflow, &io_bflow,
		        &io_basic, &io_final, &planes, &planes_b, &coef_n, &coef_b, &gacc, &bsplit, &ties, &bm_image,
		        &bm_wins, &bm_wmeta, &erec, &gflag, &fmask, &lp_idx, &stat64, &lastf, &fidx, &geo, &scan_tmp,
		        &scan_tmp2, &bin_cnt, &bin_k0, &bin_k1, &bin_v1, &bin_hist, &bin_hoff, &bin_off,
		        &bin_list, &mlp_theta, &mlp_m, &mlp_v, &mlp_wq, &mlp_inf,
		        &mlp_ws, &mlp_grad, &mlp_xb, &mlp_tb, &mlp_sse, &tv_frames, &tv_pairs, &tv_part, &tv_mm,
		        &tv_state, &tv_work, &tv_pyr, &tv_taps};
	}
};

namespace bm4dof {

// Records an error message on ctx (may be NULL) and as the thread's global
// error (bm4dof_global_error).
void set_error(bm4dof_ctx *ctx, const std::string &s);

// Records, for `device` (current), the address of the fixed scratch of
// bm4dof_tvl1_noalloc (tvl1.hip).  Called once per context creation.
void tv_scratch_init(int device);

// Runs fn, mapping exceptions to the C-ABI's error codes.
template <class Fn> int guarded(bm4dof_ctx *ctx, Fn fn)
{
	try {
		fn();
		return BM4DOF_OK;
	} catch (const ArgError &e) {
		set_error(ctx, e.what());
		return e.code;
	} catch (const HipError &e) {
		set_error(ctx, e.what());
		return BM4DOF_EHIP;
	} catch (const std::exception &e) {
		set_error(ctx, e.what());
		return BM4DOF_EINVAL;
	}
}

} // namespace bm4dof

