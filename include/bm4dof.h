/*
 * bm4dof.h -- C-ABI of the MI355X BM4D-OF / NL-DCT video denoiser and its
 * TV-L1 optical-flow input.
 *
 * Every entry point takes plain pointers and sizes.  Videos use the
 * reference's planar layout  data[t*w*h*c + ch*w*h + y*w + x]
 * (src/utils/lib_videot.hpp:88-92); flow videos have c = 2 (u, v).
 *
 * Entry points that take `const float*` video arguments without a `_dev`
 * suffix take HOST memory; the `_dev` variants take device memory already
 * resident in HBM on the context's device.
 *
 * Errors: functions return 0 on success and a negative BM4DOF_E* code on
 * failure; bm4dof_last_error() returns the message, which for the checks the
 * reference performs is the reference's own std::runtime_error text.
 */
#ifndef BM4DOF_H_INCLUDED
#define BM4DOF_H_INCLUDED

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
#define BM4DOF_BOOL bool
extern "C" {
#else
#define BM4DOF_BOOL _Bool
#endif

/* enum ColorSpace {RGB, YUV}       src/core/nldct.h:33 */
enum { BM4DOF_RGB = 0, BM4DOF_YUV = 1 };
/* enum Transform {dct, bior1_5}    src/core/nldct.h:34 */
enum { BM4DOF_DCT = 0, BM4DOF_BIOR1_5 = 1 };

/* Error codes. */
enum {
	BM4DOF_OK         =  0,
	BM4DOF_EINVAL     = -1, /* argument check failed (reference runtime_error) */
	BM4DOF_EHIP       = -2, /* HIP runtime error                               */
	BM4DOF_ENOMEM     = -3, /* device allocation failed                        */
	BM4DOF_EUNSUPP    = -4  /* configuration outside what the kernels handle   */
};

/*
 * Fixed-width POD mirror of `struct nlbParams` (src/core/nldct.h:55-82),
 * field for field and in the same order.  bool fields are int32 (0/1), the
 * two enums are int32 with the reference's values.
 */
typedef struct bm4dof_params {
	float    sigma;
	uint32_t sizePatch;
	uint32_t sizePatchTime;
	uint32_t nSimilarPatches;
	uint32_t sizeSearchWindow;
	uint32_t sizeSearchTimeRangeFwd;
	uint32_t sizeSearchTimeRangeBwd;
	uint32_t boundary;
	uint32_t offSet;
	uint32_t offSetTime;
	float    beta;
	float    betaMean;
	float    tau;
	int32_t  orderInvariance;
	int32_t  isFirstStep;
	int32_t  doPasteBoost;
	int32_t  verbose;
	int32_t  colorSpace;            /* BM4DOF_RGB / BM4DOF_YUV          */
	int32_t  transform;             /* BM4DOF_DCT / BM4DOF_BIOR1_5      */
	uint32_t sizeSearchWindowPred;  /* predictive search ("vbm3d_search")  */
	uint32_t nSimilarPatchesPred;   /* predictive search ("vbm3d_search")  */
	float    dsub;                  /* predictive search ("vbm3d_search")  */
	int32_t  agg_window;
} bm4dof_params;

/* ---- parameter set-up: same semantics as the reference, bit for bit ---- */

/* initializeNlbParameters  src/core/nldct.cpp:81-148 (step = 1 or 2) */
void bm4dof_init_params(bm4dof_params *p, unsigned step, float sigma,
                        unsigned width, unsigned height, unsigned frames,
                        unsigned channels, int verbose);
/* setSizeSearchWindow      src/core/nldct.cpp:160-168 */
void bm4dof_set_search_window(bm4dof_params *p, unsigned sizeSearchWindow);
/* setSizePatch             src/core/nldct.cpp:179-188 */
void bm4dof_set_patch_size(bm4dof_params *p, unsigned frames, unsigned sizePatch);
/* setNSimilarPatches       src/core/nldct.cpp:199-210 */
void bm4dof_set_n_similar(bm4dof_params *p, unsigned nSimilarPatches);

/* ---- denoiser context ---- */

typedef struct bm4dof_ctx bm4dof_ctx;

/* Create a context bound to HIP device `device`.  One context per host
 * thread; calls on one context are not re-entrant (the reference's
 * runNlBayes is single-caller, src/core/nldct.cpp:267). */
int  bm4dof_ctx_create(bm4dof_ctx **out, int device);
void bm4dof_ctx_destroy(bm4dof_ctx *ctx);
/* Message of the last failed call on ctx (never NULL). */
const char *bm4dof_last_error(const bm4dof_ctx *ctx);
/* Message of the last failed call that had no context (ctx_create). */
const char *bm4dof_global_error(void);

/*
 * Execution options of a context (no counterpart in the reference, whose
 * execution is fixed).  All but "vbm3d_search" select between kernels that
 * compute the same result; outputs do not depend on them.  Keys:
 *   "fast_filter"  1 (default): register-resident group filter where the
 *                  group shape allows it (64-sample patches, fewer than 64
 *                  similar patches, no orderInvariance; colour one launch
 *                  per channel); 0: the general LDS filter for every group.
 *   "fast_bm"      block matching for patches 4x4x4 or 8x8x1: 2 (default):
 *                  one wave per group, candidate windows in registers (any
 *                  channel count); 1: one wave per group on LDS-staged
 *                  search windows (one channel in the distance); 0: the
 *                  general kernel for every shape.
 *   "bm_frames"    frames of search window staged in LDS per round of the
 *                  LDS-staged block matching: 1, 2 (default) or 4.
 *   "agg_chunked"  1 (default): ordered aggregation by sorted row chunks
 *                  with per-pixel masks (tiles at least 8 pixels wide and
 *                  high); 0: the banded aggregation for every block.
 *   "agg_cap"      rows sorted per step of the chunked aggregation, for one
 *                  channel: 512, 768, 1024 (default) or 1536.
 *   "agg_binned"   1 (default): where the chunked aggregation applies and
 *                  patches have at most 64 samples (at most 8x8 pixels and
 *                  4 frames, 1, 3 or 4 channels), each bin of 8 x 8 pixels
 *                  x 4 frames gets its rows as one list in row order, built
 *                  beside the filter, and one wave folds the list into the
 *                  bin's four frames; 0: the chunked aggregation sorts the
 *                  rows of every 8 x 8 x 1 block itself.
 *   "coef_beside_bm" 1 (default): a gray step's coefficient tables are
 *                  computed on the second stream beside its block matching;
 *                  0: after it, on the main stream.
 *   "side_blocks"  most blocks (one wave each) of a kernel that builds the
 *                  bin lists on the second stream, beside the filter; the
 *                  kernels loop over their work.  2048 (default); 0: one
 *                  block per unit of work.
 *   "coef_budget"  bytes of device memory a step of the register filters may
 *                  hold in per-anchor coefficient tables (the forward
 *                  transform of every patch position, computed once per
 *                  image; step 1 one table, step 2 two, each C*W*H*F*64
 *                  floats; default: the smaller of 96 GiB and 2/5 of the
 *                  device memory free when the context was created).  Steps
 *                  whose tables do not fit the budget or the free device
 *                  memory transform the patches group by group; the output
 *                  is the same either way.  0 disables the tables.
 *   "vbm3d_search" 0 (default): block matching as the reference builds it;
 *                  1: the predictive search of a build with VBM3D_SEARCH
 *                  defined (src/core/nldct.h:20; nldct.cpp:904-1130,
 *                  1368-1589), driven by the parameters' sizeSearchWindowPred,
 *                  nSimilarPatchesPred and dsub.  This one changes results:
 *                  it selects the search a differently built reference runs.
 *   "keep_tables"  0 (default): the tables are freed at the end of every
 *                  bm4dof_run / _run_dev / _run_learned(_dev) call; 1: they
 *                  stay allocated for the next call on the context (saves
 *                  their allocation when videos of one shape are denoised
 *                  repeatedly).
 * Returns BM4DOF_EINVAL for an unknown key.
 */
int bm4dof_ctx_set_option(bm4dof_ctx *ctx, const char *key, long long value);

/*
 * Whole two-step denoiser.  Replaces
 *   std::vector<float> VideoNLB::runNlBayes(noisy, fflow, bflow, basic,
 *                                           final, prms1, prms2)
 *   src/core/nldct.h:239-247, src/core/nldct.cpp:267-478.
 *
 * Step selection is the reference's: p1->sizePatch == 0 skips step 1 and
 * then `basic` is an INPUT (DENO_ONLY), p2->sizePatch == 0 skips step 2.
 * fflow/bflow may be NULL (no flow: zero motion), else w*h*f*2 floats each.
 * `nparts` is the number of spatial tiles the reference would use
 * (2 * omp_get_max_threads(), src/core/nldct.cpp:318-323); the output
 * depends on it and parity with a CPU run needs the same value.
 * groups_ratio (may be NULL) receives the two percentages the reference
 * returns (src/core/nldct.cpp:403-404, 471-472).
 * Host buffers.  `final` may be NULL when step 2 is skipped.
 */
int bm4dof_run(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2,
               const float *noisy, unsigned width, unsigned height,
               unsigned frames, unsigned channels,
               const float *fflow, const float *bflow,
               float *basic, float *final, int nparts, float groups_ratio[2]);

/* Same as bm4dof_run on device buffers already resident in HBM on the
 * context's device.  Runs on the context's stream and returns when the
 * result is complete. */
int bm4dof_run_dev(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2,
                   const float *noisy_dev, unsigned width, unsigned height,
                   unsigned frames, unsigned channels,
                   const float *fflow_dev, const float *bflow_dev,
                   float *basic_dev, float *final_dev, int nparts,
                   float groups_ratio[2]);

/*
 * Block matching of one step only, for parity checks of the neighbour sets.
 * Mirrors estimateSimilarPatchesStep1 (src/core/nldct.cpp:1143-1348) /
 * estimateSimilarPatchesStep2 (src/core/nldct.cpp:1607-1789) applied at every
 * reference position of every tile, in the reference's processing order.
 *   step        1 or 2 (selects which image the distance is taken on)
 *   image       step 1: noisy video (already in the working colour space);
 *               step 2: basic video
 * Output arrays are sized by bm4dof_bm_count(): out_index[g*np + n] holds the
 * reference's patch index  t*w*h*c + y*w + x  (global coordinates) of the
 * n-th neighbour of group g, out_nsim[g] its group size, out_ref[g] the
 * reference position (same encoding) and out_tile[g] the tile number.
 * Host buffers.
 */
int bm4dof_bm_count(bm4dof_ctx *ctx, const bm4dof_params *p, unsigned width,
                    unsigned height, unsigned frames, unsigned channels,
                    int nparts, int64_t *n_groups);
int bm4dof_bm(bm4dof_ctx *ctx, const bm4dof_params *p, const float *image,
              unsigned width, unsigned height, unsigned frames, unsigned channels,
              const float *fflow, const float *bflow, int nparts,
              uint32_t *out_index, uint32_t *out_nsim, uint32_t *out_ref,
              uint32_t *out_tile);

/* Per-stage timings of the last bm4dof_run / bm4dof_run_dev call, in ms,
 * measured with HIP events on the context's stream.  `names` receives a
 * static string naming each slot (per step: setup, bm, coef, filter, sort,
 * aggregate, normalize; "sort" is the time the stream still waits, after the
 * filter, for the row sort that runs on the second stream).  Returns the
 * number of slots filled. */
int bm4dof_last_timings(const bm4dof_ctx *ctx, float *ms, const char **names, int max_slots);

/* Statistics of the last bm4dof_run / bm4dof_run_dev call: stats[0] =
 * pixel blocks of the ordered aggregation, stats[1] = blocks whose rows
 * did not fit the block capacity and were processed in several row bands,
 * stats[2..4] = block shape (x, y, frames) of the last aggregation launch,
 * stats[5..6] = groups and group rows (patches over all groups) of step 1,
 * stats[7..8] = the same for step 2; stats[9] = executed TV-L1 iterations
 * times pixels, summed over pairs, scales and warps, of the last flow
 * computation on the context.  Returns the number of slots filled. */
int bm4dof_last_stats(const bm4dof_ctx *ctx, int64_t *stats, int max_slots);

/* ---- batches and the autotuner ----
 *
 * No counterpart in the reference (it denoises one video per process).  A
 * batch runs both steps on many one-channel sequences of uneven size and
 * noise level as one launch set: planned and uploaded at creation, then
 * either enqueued on every run or captured once into a hipGraph
 * (BM4DOF_BATCH_GRAPH) that every run replays.  Each sequence's output is
 * bitwise the output of bm4dof_run_dev on it alone (same nparts).  Where
 * equal distances put a patch other than the reference one first in a
 * neighbour list, the reference's processing-mask scan skips groups
 * (src/core/nldct.cpp:2325-2329); the launch set counts such groups on the
 * device for every sequence, and bm4dof_batch_run then runs the sequences
 * that have any once more through the path of bm4dof_run_dev, which replays
 * that scan.  With BM4DOF_BATCH_AUTOTUNE the execution options that cannot
 * change results ("fast_bm", "bm_frames", "agg_cap") are timed on the first
 * sequence of every shape class and the fastest set is used for that class.
 * Buffers are device pointers and must stay valid for the batch's lifetime;
 * a captured batch holds its scratch memory in the context, so runs on that
 * context of other work in between are allowed but not concurrently.  The
 * graph holds the context buffers' addresses of its capture: when such work
 * has moved one of them, the next bm4dof_batch_run captures the launch set
 * again before replaying it.
 */
typedef struct bm4dof_seq {
	const float *noisy;          /* w*h*f, one channel */
	const float *fflow, *bflow;  /* w*h*f*2 each, or both NULL */
	float *basic, *final;        /* outputs, w*h*f each */
	unsigned w, h, f;
} bm4dof_seq;
typedef struct bm4dof_batch bm4dof_batch;
#define BM4DOF_BATCH_GRAPH 1
#define BM4DOF_BATCH_AUTOTUNE 2
int   bm4dof_batch_create(bm4dof_ctx *ctx, const bm4dof_seq *seqs, const bm4dof_params *p1s,
                          const bm4dof_params *p2s, int n, int nparts, int flags, bm4dof_batch **out);
int   bm4dof_batch_run(bm4dof_batch *b);
float bm4dof_batch_last_ms(const bm4dof_batch *b);
const char *bm4dof_batch_report(const bm4dof_batch *b);
/* Statistics of the last bm4dof_batch_run: out[0] = how the launch set ran
 * (0 enqueued, 1 graph replayed, 2 graph captured again and replayed),
 * out[1] = sequences run again for the processing-mask scan, out[2] =
 * captures so far.  Returns the number of slots filled. */
int   bm4dof_batch_stats(const bm4dof_batch *b, int64_t *out, int max_slots);
void  bm4dof_batch_destroy(bm4dof_batch *b);
/* Times both steps on one sequence (device buffers, one channel) under every
 * candidate option set, keeps the fastest in ctx, and writes a report. */
int   bm4dof_autotune(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2, const float *noisy,
                      unsigned w, unsigned h, unsigned f, const float *fflow, const float *bflow, float *basic,
                      float *final, int nparts, char *report, int report_len);

/* ---- learned single pass ----
 *
 * No counterpart in the reference.  A network of 66 -> 128 -> 128 -> 64
 * units (ReLU, sigmoid; bfloat16 weights, float32 accumulation on the
 * matrix cores) maps, for every row of a group of the first step's grid,
 * the Haar-domain coefficients of the noisy patches (features:
 * coefficient / sigma, log(sigma / 20), log1p of the row's energy over
 * 64 sigma^2) to 64 gains that stand in for the second step's
 * Wiener gains y^2 / (y^2 + (beta sigma)^2) (src/core/nldct.cpp:2131), so
 * one pass replaces the two.
 *
 * theta: BM4DOF_MLP_NPARAMS floats, W1[66][128], b1[128], W2[128][128],
 * b2[128], W3[128][64], b3[64], each W input-major (W[k][j] at k*cols+j).
 */
#define BM4DOF_MLP_NPARAMS 33344

/* Load a network (host theta); resets the Adam state. */
int bm4dof_mlp_set(bm4dof_ctx *ctx, const float *theta);
/* Copy the current master parameters to host theta. */
int bm4dof_mlp_get(const bm4dof_ctx *ctx, float *theta);
/* Gains of `rows` feature rows (x: rows*66, gains: rows*64), on the matrix
 * cores; buffers on the host or (on_device != 0) on the device. */
int bm4dof_mlp_forward(bm4dof_ctx *ctx, const float *x, int64_t rows, float *gains, int on_device);
/* Mean squared error loss and its gradient (host buffers; grad may be
 * NULL) of one batch, by the training kernels. */
int bm4dof_mlp_grad(bm4dof_ctx *ctx, const float *x, const float *t, int rows, float *loss, float *grad);
/* `steps` Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) on batches of
 * `batch` rows drawn from the device data set x (rows*66), t (rows*64);
 * batch == rows takes the whole set in order.  losses (host, may be NULL)
 * receives each step's batch loss before its update. */
int bm4dof_mlp_train(bm4dof_ctx *ctx, const float *x, const float *t, int64_t rows, int batch, int steps,
                     float lr, unsigned seed, float *losses);
/* Training pairs: for every stride-th group of p1's grid (matching on the
 * noisy video), 32 rows of features (x_out, rows*66) and of the Wiener gains
 * of `basic` at the same neighbours with factor beta2 (t_out, rows*64);
 * row_flag marks the rows of real neighbours.  Device buffers of max_rows
 * rows; *rows receives 32 * (number of sampled groups).  One channel. */
int bm4dof_learned_pairs(bm4dof_ctx *ctx, const bm4dof_params *p1, const float *noisy, const float *basic,
                         unsigned w, unsigned h, unsigned f, const float *fflow, const float *bflow,
                         int nparts, int stride, float beta2, float *x_out, float *t_out,
                         unsigned char *row_flag, int64_t max_rows, int64_t *rows);
/* The learned single pass (one channel): p1's grid, matching and patch
 * size with the network's gains.  With gvar_thr >= 0, pixels under groups
 * whose predicted gains have a variance above gvar_thr take the plain
 * two-step result (p1, p2) instead; *flag_frac (may be NULL) receives the
 * fraction of such groups.  Host buffers / device buffers (_dev). */
int bm4dof_run_learned(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2,
                       const float *noisy, unsigned w, unsigned h, unsigned f, unsigned c,
                       const float *fflow, const float *bflow, float *out, int nparts, float gvar_thr,
                       float *flag_frac);
int bm4dof_run_learned_dev(bm4dof_ctx *ctx, const bm4dof_params *p1, const bm4dof_params *p2,
                           const float *noisy, unsigned w, unsigned h, unsigned f, unsigned c,
                           const float *fflow, const float *bflow, float *out, int nparts, float gvar_thr,
                           float *flag_frac);

/* ---- TV-L1 optical flow ---- */

/*
 * Drop-in for the reference's flow library entry point, same signature and
 * argument meaning:
 *   void Dual_TVL1_optic_flow_multiscale(float *I0, float *I1, float *u1,
 *        float *u2, int nx, int ny, float tau, float lambda, float theta,
 *        int nscales, int fscale, float zfactor, int warps, float epsilon,
 *        bool verbose)   lib/tvl1flow/tvl1flow_lib.c:345-361
 * Host buffers; the caller allocates u1, u2 (nx*ny floats each), as
 * lib/tvl1flow/main.c:173-174 does.  Runs on device 0 through a process-wide
 * context.  Like the reference it has no error return; failures abort with a
 * message on stderr.
 */
void Dual_TVL1_optic_flow_multiscale(float *I0, float *I1, float *u1, float *u2,
                                     int nx, int ny, float tau, float lambda,
                                     float theta, int nscales, int fscale,
                                     float zfactor, int warps, float epsilon,
                                     BM4DOF_BOOL verbose);

/* TV-L1 parameters, in the argument order of the tvl1flow command line
 * (lib/tvl1flow/main.c:86-99). */
typedef struct bm4dof_tvl1_params {
	float tau;
	float lambda;
	float theta;
	int   nscales;
	int   fscale;
	float zfactor;
	int   warps;
	float epsilon;
	int   verbose;
} bm4dof_tvl1_params;

/*
 * Flow of one frame pair with an explicit context (device memory or host
 * memory, see `on_device`).  Same computation as
 * Dual_TVL1_optic_flow_multiscale.  iters (may be NULL) receives the
 * iteration count of every (scale, warp) solve, nscales*warps ints,
 * scale-major from the finest scale; unused slots are -1.
 */
int bm4dof_tvl1(bm4dof_ctx *ctx, const bm4dof_tvl1_params *prm,
                const float *I0, const float *I1, float *u1, float *u2,
                int nx, int ny, int on_device, int *iters);

/*
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops with a fixed memory
 * budget).  The flow of one frame pair on DEVICE buffers (I0, I1, u1, u2:
 * nx*ny floats each), the same computation as bm4dof_tvl1 with on_device = 1
 * and bitwise the same result, with every scratch buffer carved from one
 * fixed device region of BM4DOF_TVL1_NOALLOC_BYTES per GPU (part of the
 * library's code object, held until the process ends) instead of the
 * context's buffers; no device memory is allocated or freed.  The scratch
 * needed is about 86 bytes per pixel at zfactor 0.5 (a 960x540 pair fits;
 * a larger zfactor needs more): a pair that does not fit is rejected with
 * BM4DOF_EINVAL and a message giving both sizes.  Every call on a GPU shares
 * that GPU's region, which keeps the last call's contents; calls that
 * overlap in time on one GPU (two threads, two contexts) race and must be
 * serialised by the caller.
 */
#define BM4DOF_TVL1_NOALLOC_BYTES (48u << 20)
int bm4dof_tvl1_noalloc(bm4dof_ctx *ctx, const bm4dof_tvl1_params *prm,
                        const float *I0, const float *I1, float *u1, float *u2,
                        int nx, int ny, int *iters);

/*
 * Forward and backward flow of a whole sequence, replacing the per-pair
 * process loop of scripts/tvl1flow-seq.sh:27-61 in-process:
 *   fwd(i) = flow(I_i -> I_{i+1}),  bwd(i) = flow(I_i -> I_{i-1}),
 *   bwd(0) = bwd(1),  fwd(F-1) = fwd(F-2).
 * frames: F gray frames (w*h*F floats).  fflow/bflow: w*h*F*2 floats in the
 * denoiser's flow layout (u plane, then v plane, per frame).  F == 1 gives
 * zero flow.  All pairs of the sequence are solved together.
 */
int bm4dof_tvl1_sequence(bm4dof_ctx *ctx, const bm4dof_tvl1_params *prm,
                         const float *frames, int nx, int ny, int nframes,
                         float *fflow, float *bflow, int on_device);

#ifdef __cplusplus
}
#endif

#endif /* BM4DOF_H_INCLUDED */
