// HIP kernels of the BM4D-OF / NL-DCT denoiser for gfx950 (MI355X).
//
// Arithmetic that decides a discontinuous outcome is evaluated in the
// reference's exact operation order with IEEE float operations (this file
// is compiled with -ffp-contract=off, so only the explicit fused
// multiply-adds below are fused):
//   * patch distances, sequential over (c, ht, hy, hx), separately rounded
//     as the reference's own loops (nldct.cpp:1262-1274, 1723-1737);
//   * the forward patch transform, a sequential-K dot product per
//     coefficient with a fused multiply-add per term (nldct.cpp:1840-1844 via
//     productMatrix/cblas_sgemm: the reference leaves the order to the BLAS
//     it links; an FMA BLAS kernel accumulates K = 64 this way);
//   * the Haar butterflies and the hard threshold (nldct.cpp:1847-1907);
//   * the inverse transform, sequential in the coefficient index, fused
//     likewise (nldct.cpp:1927-1931, 2155-2159);
//   * the Wiener energy sum_w^2 in (coefficient, patch) order
//     (nldct.cpp:2123-2135) and the aggregation sums in group order
//     (nldct.cpp:2284-2300, 2376-2393), separately rounded.
// The GPU output is therefore bitwise equal to the sequential oracle.
#include "common.hpp"
#include "mlp_dev.hpp"

namespace bm4dof {

#define BM_THREADS 256
#define BM_MAX_FRAMES 64
#define FILTER_THREADS 256

// std::min / std::max as the reference calls them on floats.
__device__ __forceinline__ float fmin_ref(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ uint32_t video_addr(int x, int y, int t, int c, const StepConst &sc)
{
	return (uint32_t)t * (uint32_t)sc.WHC + (uint32_t)c * (uint32_t)sc.WH + (uint32_t)y * (uint32_t)sc.W + (uint32_t)x;
}

// ---------------------------------------------------------------------------
// Colour transform, VideoUtils::transformColorSpace (lib_videot.hpp:728-838).
// k: coefficients computed on the host exactly as the reference does.
// ---------------------------------------------------------------------------
__global__ void color_kernel(float *__restrict__ v, int WH, int C, int F, int forward,
                             float a, float b, float c)
{
	const int64_t n = (int64_t)WH * F;
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const int64_t t = i / WH, k = i - t * WH;
		float *p = v + t * (int64_t)WH * C + k;
		if (C == 3) {
			const float x0 = p[0], x1 = p[WH], x2 = p[2 * (int64_t)WH];
			if (forward) {
				p[0] = a * (x0 + x1 + x2);
				p[WH] = b * (x0 - x2);
				p[2 * (int64_t)WH] = c * (0.25f * x0 - 0.5f * x1 + 0.25f * x2);
			} else {
				p[0] = a * x0 + b * x1 + c * 0.5f * x2;
				p[WH] = a * x0 - c * x2;
				p[2 * (int64_t)WH] = a * x0 - b * x1 + c * 0.5f * x2;
			}
		} else if (C == 4) {
			const float Gr = p[0], R = p[WH], B = p[2 * (int64_t)WH], Gb = p[3 * (int64_t)WH];
			if (forward) {
				p[0] = a * (Gr + R + B + Gb);
				p[WH] = b * (R - B);
				p[2 * (int64_t)WH] = a * (-Gr + R + B - Gb);
				p[3 * (int64_t)WH] = b * (-Gr + Gb);
			} else {
				p[0] = a * Gr - a * B - b * Gb;
				p[WH] = a * Gr + b * R + a * B;
				p[2 * (int64_t)WH] = a * Gr - b * R + a * B;
				p[3 * (int64_t)WH] = a * Gr - a * B + b * Gb;
			}
		}
	}
}

// ---------------------------------------------------------------------------
// Block matching: one workgroup per reference patch.
// estimateSimilarPatchesStep1 (nldct.cpp:1143-1348) / Step2 (:1607-1789).
//
// Candidate order s follows the reference's scan (frame schedule pt, pt+1..,
// pt-1.., then raster qy, qx).  Selection keeps the kk = largest power of
// two <= min(np, #candidates) smallest distances in the total order
// (distance, reference patch first, s), which equals the reference's
// std::partial_sort order whenever the distances the list depends on are
// distinct, then applies the power-of-two tau cut (nldct.cpp:1285-1289).
// Groups where they are not (bm_has_tie) are listed and matched again with
// the selection of the reference itself: libstdc++'s heap select and
// sort_heap on the scan order (bm_tie_kernel).
// ---------------------------------------------------------------------------
struct BmFrame { int qt, rx0, ry0, nx, ny, base; };

// True when the kept list could come out otherwise under std::partial_sort,
// which leaves equal distances in heap order: two kept distances are equal,
// or a candidate left out has the distance of the last kept one.  n is the
// count kept after the tau cut; out_key holds the kk smallest, sorted; skey
// the nsrch distances (all of them, or every one up to a bound at least the
// kk-th smallest).  One wave, every lane.
__device__ __forceinline__ bool bm_has_tie(const uint32_t *__restrict__ skey, int nsrch,
                                           const uint32_t *__restrict__ out_key, int nSimP, float tau2)
{
	const int lane = threadIdx.x & 63;
	int n = 1;
	while (2 * n <= nSimP && __uint_as_float(out_key[n]) < tau2) n *= 2;
	const uint32_t last = out_key[n - 1];
	const bool eq = lane + 1 < n && out_key[lane] == out_key[lane + 1];
	int le = 0;
	for (int s = lane; s < nsrch; s += 64) le += skey[s] <= last ? 1 : 0;
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) le += __shfl_xor(le, o, 64);
	return __any(eq) || le > n;
}

// The same test over candidates held as (distance << 32 | rank + 1).
__device__ __forceinline__ bool bm_has_tie(const uint64_t *__restrict__ cand, int ncand,
                                           const uint32_t *__restrict__ out_key, int nSimP, float tau2)
{
	const int lane = threadIdx.x & 63;
	int n = 1;
	while (2 * n <= nSimP && __uint_as_float(out_key[n]) < tau2) n *= 2;
	const uint32_t last = out_key[n - 1];
	const bool eq = lane + 1 < n && out_key[lane] == out_key[lane + 1];
	int le = 0;
	for (int s = lane; s < ncand; s += 64) le += (uint32_t)(cand[s] >> 32) <= last ? 1 : 0;
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) le += __shfl_xor(le, o, 64);
	return __any(eq) || le > n;
}

__device__ __forceinline__ void bm_tie_append(int g, uint32_t *__restrict__ tie_count,
                                              uint32_t *__restrict__ tie_list, int tie_cap)
{
	const uint32_t i = atomicAdd(tie_count, 1u);
	if (i < (uint32_t)tie_cap) tie_list[i] = (uint32_t)g;
}

// libstdc++'s std::partial_sort(first, first + middle, first + len) under
// `<` on the key (__heap_select, then __sort_heap), on parallel arrays; one
// thread.  Same steps as the library's __adjust_heap / __push_heap /
// __make_heap / __pop_heap.
__device__ void ls_push_heap(uint32_t *key, int *val, int hole, int top, uint32_t vk, int vv)
{
	int parent = (hole - 1) / 2;
	while (hole > top && __uint_as_float(key[parent]) < __uint_as_float(vk)) {
		key[hole] = key[parent];
		val[hole] = val[parent];
		hole = parent;
		parent = (hole - 1) / 2;
	}
	key[hole] = vk;
	val[hole] = vv;
}
__device__ void ls_adjust_heap(uint32_t *key, int *val, int hole, int len, uint32_t vk, int vv)
{
	const int top = hole;
	int second = hole;
	while (second < (len - 1) / 2) {
		second = 2 * (second + 1);
		if (__uint_as_float(key[second]) < __uint_as_float(key[second - 1])) second--;
		key[hole] = key[second];
		val[hole] = val[second];
		hole = second;
	}
	if ((len & 1) == 0 && second == (len - 2) / 2) {
		second = 2 * (second + 1);
		key[hole] = key[second - 1];
		val[hole] = val[second - 1];
		hole = second - 1;
	}
	ls_push_heap(key, val, hole, top, vk, vv);
}
__device__ void ls_pop_heap(uint32_t *key, int *val, int len, int result)
{
	const uint32_t vk = key[result];
	const int vv = val[result];
	key[result] = key[0];
	val[result] = val[0];
	ls_adjust_heap(key, val, 0, len, vk, vv);
}
__device__ void ls_partial_sort(uint32_t *key, int *val, int middle, int len)
{
	if (middle >= 2)
		for (int parent = (middle - 2) / 2;; parent--) {
			ls_adjust_heap(key, val, parent, middle, key[parent], val[parent]);
			if (parent == 0) break;
		}
	for (int i = middle; i < len; ++i)
		if (__uint_as_float(key[i]) < __uint_as_float(key[0])) ls_pop_heap(key, val, middle, i);
	for (int l = middle; l > 1;) {
		--l;
		ls_pop_heap(key, val, l, l);
	}
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
	const int lane = threadIdx.x & 63;
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		const int o = __shfl_up(v, d, 64);
		if (lane >= d) v += o;
	}
	return v;
}

// One group, one workgroup.  LIBSTD: the selection is libstdc++'s
// partial_sort on the scan order (for the listed groups of bm_tie_kernel);
// else the total order, listing the groups bm_has_tie finds (tie_list).
template <bool LIBSTD>
__device__ void bm_group(int g, const float *__restrict__ img, const float *__restrict__ fflow,
                         const float *__restrict__ bflow, const uint32_t *__restrict__ gpos,
                         const uint32_t *__restrict__ gtile, const TileBox *__restrict__ tiles,
                         StepConst sc, int step, int maxCand, uint32_t *__restrict__ out_idx,
                         uint32_t *__restrict__ out_nsim, int G, uint32_t *__restrict__ tie_count,
                         uint32_t *__restrict__ tie_list, int tie_cap)
{
	extern __shared__ unsigned char smem[];
	uint32_t *skey = (uint32_t *)smem;               // maxCand
	int *srank = (int *)(skey + maxCand);            // maxCand
	int *stie = srank + maxCand;                     // maxCand
	float *sref = (float *)(stie + maxCand);         // C*P

	__shared__ BmFrame fr[BM_MAX_FRAMES];
	__shared__ int s_cx[BM_MAX_FRAMES], s_cy[BM_MAX_FRAMES];
	__shared__ int s_nframes, s_nsrch, s_rt0;
	__shared__ uint32_t hist[256];
	__shared__ uint32_t s_bucket, s_below;
	__shared__ uint32_t sel_key[64];
	__shared__ int sel_rank[64];
	__shared__ uint32_t out_key[64];
	__shared__ int out_rank[64];
	__shared__ int s_na, s_nb, s_nfin;

	if (g >= G) return;
	const int tid = threadIdx.x;
	const uint32_t pos = gpos[g];
	int gx, gy, gt;
	pos_decode(pos, sc.W, sc.WH, gx, gy, gt);
	const TileBox tb = tiles[gtile[g]];
	const int Wt = tb.ex - tb.ox, Ht = tb.ey - tb.oy;
	const int px = gx - tb.ox, py = gy - tb.oy, pt = gt;
	const int sPx = sc.sPx, sPt = sc.sPt, P = sc.P;
	uint32_t *oidx = out_idx + (size_t)g * sc.npmax;

	if (sc.npmax <= 1) {
		if (tid == 0) { oidx[0] = pos; out_nsim[g] = 1; }
		return;
	}

	// Temporal range (nldct.cpp:1173-1186) and flow trajectory
	// (nldct.cpp:1212-1236): thread 0 walks forward, thread 1 backward.
	const int F = sc.F;
	const int shift_t = min(0, pt - sc.wtb) + max(0, pt + sc.wtf - F + sPt);
	const int rt0 = max(0, pt - sc.wtb - shift_t);
	const int rt1 = min(F - sPt, pt + sc.wtf - shift_t);
	if (tid < 2) {
		const int dir = tid == 0 ? 1 : -1;
		const float *fl = tid == 0 ? fflow : bflow;
		int cx = px, cy = py, ct = pt;
		if (tid == 0) { s_cx[pt - rt0] = px; s_cy[pt - rt0] = py; s_rt0 = rt0; }
		for (int qt = pt + dir; dir > 0 ? qt <= rt1 : qt >= rt0; qt += dir) {
			float fx = 0.f, fy = 0.f;
			if (sc.use_flow) {
				const size_t base = (size_t)ct * 2 * sc.WH + (size_t)(cy + tb.oy) * sc.W + (cx + tb.ox);
				fx = fl[base];
				fy = fl[base + sc.WH];
			}
			const float cx_f = (float)cx + fx;
			const float cy_f = (float)cy + fy;
			cx = (int)fmax_ref(0.f, fmin_ref((float)Wt - 1.f, roundf(cx_f)));
			cy = (int)fmax_ref(0.f, fmin_ref((float)Ht - 1.f, roundf(cy_f)));
			ct = qt;
			s_cx[qt - rt0] = cx;
			s_cy[qt - rt0] = cy;
		}
	}
	// Reference patch (step 1: channel 0 of the noisy video; step 2: all
	// channels of the basic video).
	const int nref = step == 1 ? P : sc.C * P;
	for (int k = tid; k < nref; k += BM_THREADS) {
		const int c = k / P, i = k - c * P;
		const int ht = i / (sPx * sPx), r = i - ht * sPx * sPx;
		const int hy = r / sPx, hx = r - hy * sPx;
		sref[k] = img[video_addr(gx + hx, gy + hy, gt + ht, c, sc)];
	}
	__syncthreads();
	if (tid == 0) {
		const int sWx = sc.sWx, half = (sWx - 1) / 2;
		int nf = 0, base = 0;
		for (int ii = 0; ii < rt1 - rt0 + 1; ii++) {
			const int qt = ii == 0 ? pt : (ii <= rt1 - pt ? pt + ii : pt - (ii - (rt1 - pt)));
			const int dt = qt - rt0;
			const int cx = s_cx[dt], cy = s_cy[dt];
			int shift_x = min(0, cx - half);
			int shift_y = min(0, cy - half);
			shift_x += max(0, cx + half - Wt + sPx);
			shift_y += max(0, cy + half - Ht + sPx);
			const int rx0 = max(0, cx - half - shift_x);
			const int ry0 = max(0, cy - half - shift_y);
			const int rx1 = min(Wt - sPx, cx + half - shift_x);
			const int ry1 = min(Ht - sPx, cy + half - shift_y);
			BmFrame f;
			f.qt = qt; f.rx0 = rx0; f.ry0 = ry0;
			f.nx = max(0, rx1 - rx0 + 1); f.ny = max(0, ry1 - ry0 + 1);
			f.base = base;
			base += f.nx * f.ny;
			fr[nf++] = f;
		}
		s_nframes = nf;
		s_nsrch = base;
		s_na = 0; s_nb = 0;
	}
	__syncthreads();
	const int nsrch = s_nsrch, nframes = s_nframes;

	// Distances.
	for (int s = tid; s < nsrch; s += BM_THREADS) {
		int ii = 0;
		while (ii + 1 < nframes && fr[ii + 1].base <= s) ii++;
		const BmFrame f = fr[ii];
		const int o = s - f.base;
		const int qy = f.ry0 + o / f.nx, qx = f.rx0 + o % f.nx, qt = f.qt;
		const int ax = qx + tb.ox, ay = qy + tb.oy;
		float d = 0.f;
		if (step == 1) {
			int k = 0;
			for (int ht = 0; ht < sPt; ht++)
				for (int hy = 0; hy < sPx; hy++) {
					const float *row = img + video_addr(ax, ay + hy, qt + ht, 0, sc);
					for (int hx = 0; hx < sPx; hx++, k++) {
						const float dif = sref[k] - row[hx];
						d = d + dif * dif;
					}
				}
		} else {
			int k = 0;
			for (int c = 0; c < sc.C; c++)
				for (int ht = 0; ht < sPt; ht++)
					for (int hy = 0; hy < sPx; hy++) {
						const float *row = img + video_addr(ax, ay + hy, qt + ht, c, sc);
						for (int hx = 0; hx < sPx; hx++, k++) {
							const float dif = sref[k] - row[hx];
							d = d + dif * dif;
						}
					}
		}
		skey[s] = __float_as_uint(d);
		srank[s] = (!LIBSTD && qx == px && qy == py && qt == pt) ? -1 : s;
	}
	__syncthreads();

	const int nSimP = min(sc.npmax, nsrch);
	int kk = 1;
	while (kk * 2 <= nSimP) kk *= 2;

	if (LIBSTD) {
		if (tid == 0) ls_partial_sort(skey, srank, nSimP, nsrch);
		__syncthreads();
		if (tid < kk) {
			out_key[tid] = skey[tid];
			out_rank[tid] = srank[tid];
		}
		__syncthreads();
	} else {
	// Radix select of the kk-th smallest key, 8 bits per pass.
	uint32_t prefix = 0, pmask = 0;
	uint32_t remaining = (uint32_t)kk;
	for (int pass = 0; pass < 4; pass++) {
		const int shift = 24 - 8 * pass;
		hist[tid] = 0;
		__syncthreads();
		for (int s = tid; s < nsrch; s += BM_THREADS) {
			const uint32_t k = skey[s];
			if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
		}
		__syncthreads();
		if (tid < 64) {
			const uint32_t c0 = hist[4 * tid], c1 = hist[4 * tid + 1], c2 = hist[4 * tid + 2], c3 = hist[4 * tid + 3];
			const int local = (int)(c0 + c1 + c2 + c3);
			const int incl = wave_incl_scan(local);
			const int excl = incl - local;
			if ((uint32_t)excl < remaining && remaining <= (uint32_t)incl) {
				uint32_t cum = (uint32_t)excl;
				const uint32_t cs[4] = {c0, c1, c2, c3};
				for (int b = 0; b < 4; b++) {
					if (cum + cs[b] >= remaining) { s_bucket = 4 * tid + b; s_below = cum; break; }
					cum += cs[b];
				}
			}
		}
		__syncthreads();
		prefix |= s_bucket << shift;
		pmask |= 255u << shift;
		remaining -= s_below;
		__syncthreads();
	}
	const uint32_t T = prefix;
	const int na_expected = kk - (int)remaining;

	for (int s = tid; s < nsrch; s += BM_THREADS) {
		const uint32_t k = skey[s];
		if (k < T) {
			const int slot = atomicAdd(&s_na, 1);
			sel_key[slot] = k;
			sel_rank[slot] = srank[s];
		} else if (k == T) {
			const int slot = atomicAdd(&s_nb, 1);
			stie[slot] = srank[s];
		}
	}
	__syncthreads();
	const int nb = s_nb;
	for (int j = tid; j < nb; j += BM_THREADS) {
		const int r = stie[j];
		int c = 0;
		for (int i = 0; i < nb; i++) c += stie[i] < r;
		if (c < (int)remaining) {
			sel_key[na_expected + c] = T;
			sel_rank[na_expected + c] = r;
		}
	}
	__syncthreads();
	if (tid < kk) {
		const uint32_t k = sel_key[tid];
		const int r = sel_rank[tid];
		int p = 0;
		for (int j = 0; j < kk; j++) {
			const uint32_t kj = sel_key[j];
			p += (kj < k) || (kj == k && sel_rank[j] < r);
		}
		out_key[p] = k;
		out_rank[p] = r;
	}
	__syncthreads();
	if (tie_list && tid < 64) {
		const bool tie = bm_has_tie(skey, nsrch, out_key, nSimP, sc.tau2);
		if (tie && tid == 0) bm_tie_append(g, tie_count, tie_list, tie_cap);
	}
	}
	if (tid == 0) {
		int n = 1;
		while (2 * n <= nSimP && __uint_as_float(out_key[n]) < sc.tau2) n *= 2;
		s_nfin = n;
		out_nsim[g] = n;
	}
	__syncthreads();
	const int nfin = s_nfin;
	if (tid < nfin) {
		const int r = out_rank[tid];
		uint32_t q;
		if (r < 0) q = pos;
		else {
			int ii = 0;
			while (ii + 1 < nframes && fr[ii + 1].base <= r) ii++;
			const BmFrame f = fr[ii];
			const int o = r - f.base;
			const int qy = f.ry0 + o / f.nx, qx = f.rx0 + o % f.nx;
			q = (uint32_t)f.qt * (uint32_t)sc.WH + (uint32_t)(qy + tb.oy) * (uint32_t)sc.W + (uint32_t)(qx + tb.ox);
		}
		oidx[tid] = q;
	}
}

__global__ __launch_bounds__(BM_THREADS)
void bm_kernel(const float *__restrict__ img, const float *__restrict__ fflow,
               const float *__restrict__ bflow, const uint32_t *__restrict__ gpos,
               const uint32_t *__restrict__ gtile, const TileBox *__restrict__ tiles,
               StepConst sc, int step, int maxCand, uint32_t *__restrict__ out_idx,
               uint32_t *__restrict__ out_nsim, int G, uint32_t *__restrict__ tie_count,
               uint32_t *__restrict__ tie_list, int tie_cap)
{
	bm_group<false>((int)blockIdx.x, img, fflow, bflow, gpos, gtile, tiles, sc, step, maxCand, out_idx, out_nsim,
	                G, tie_count, tie_list, tie_cap);
}

// The listed groups again, with the reference's own selection (their
// earlier lists are overwritten).
__global__ __launch_bounds__(BM_THREADS)
void bm_tie_kernel(const float *__restrict__ img, const float *__restrict__ fflow,
                   const float *__restrict__ bflow, const uint32_t *__restrict__ gpos,
                   const uint32_t *__restrict__ gtile, const TileBox *__restrict__ tiles,
                   StepConst sc, int step, int maxCand, uint32_t *__restrict__ out_idx,
                   uint32_t *__restrict__ out_nsim, int G, const uint32_t *__restrict__ tie_count,
                   const uint32_t *__restrict__ tie_list, int tie_cap)
{
	const uint32_t n = min(*tie_count, (uint32_t)tie_cap);
	for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
		bm_group<true>((int)tie_list[i], img, fflow, bflow, gpos, gtile, tiles, sc, step, maxCand, out_idx,
		               out_nsim, G, nullptr, nullptr, 0);
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------
// Predictive block matching of a VBM3D_SEARCH build of the reference
// (estimateSimilarPatchesStep1 nldct.cpp:904-1130, Step2 :1368-1589), one
// wave per reference patch.  Frames in the schedule pt, pt+1.., pt-1..: the
// reference frame is searched in the full window; every other frame in the
// union of npf windows of wpred around the positions its previous frame's
// npf best matches reach along the flow (read at the match, rounded half
// away from zero, clamped to the tile).  Union order: window by window, each
// in raster order, a position already inside an earlier window skipped (what
// the reference's unordered_set insertion test lets through, :1044).
// Distances run on channel 0 (step 2 repeats the channel-0 sum once per
// channel: its loop reads imBasic(x, y, t) inside the channel loop, :1532),
// minus dsub at the reference's (x, y) in every frame (:1049).  The npf best
// of each frame (std::partial_sort, :1076) are kept, the kept ones of all
// frames sorted again (:1102), and the power-of-two tau cut applied
// (:1117-1121).  Both selections are libstdc++'s heap select + sort_heap
// (ls_partial_sort), run by lane 0.
//
// Dynamic LDS: cap keys + cap values (candidates of one frame), kcap keys +
// kcap values (kept), kcap x and y (the matches of every frame, the starts of
// the next frame's windows), P floats (reference patch); kcap = npf x frames.
// ---------------------------------------------------------------------------
#define PRED_MAX_WIN 64
__global__ __launch_bounds__(64)
void bm_pred_kernel(const float *__restrict__ img, const float *__restrict__ fflow,
                    const float *__restrict__ bflow, const uint32_t *__restrict__ gpos,
                    const uint32_t *__restrict__ gtile, const TileBox *__restrict__ tiles, StepConst sc,
                    int reps, int wpred, int npf, float dsub, int cap, int kcap,
                    uint32_t *__restrict__ out_idx, uint32_t *__restrict__ out_nsim, int G)
{
	extern __shared__ unsigned char smem[];
	uint32_t *ckey = (uint32_t *)smem;
	int *cval = (int *)(ckey + cap);
	uint32_t *kkey = (uint32_t *)(cval + cap);
	int *kval = (int *)(kkey + kcap);
	int *t_cx = kval + kcap, *t_cy = t_cx + kcap;   // [frame - rt0][i]
	float *sref = (float *)(t_cy + kcap);
	__shared__ int w_rx0[PRED_MAX_WIN], w_rx1[PRED_MAX_WIN], w_ry0[PRED_MAX_WIN], w_ry1[PRED_MAX_WIN];
	__shared__ int s_n;

	const int g = blockIdx.x;
	if (g >= G) return;
	const int lane = threadIdx.x;
	const uint32_t pos = gpos[g];
	int gx, gy, gt;
	pos_decode(pos, sc.W, sc.WH, gx, gy, gt);
	const TileBox tb = tiles[gtile[g]];
	const int Wt = tb.ex - tb.ox, Ht = tb.ey - tb.oy;
	const int px = gx - tb.ox, py = gy - tb.oy, pt = gt;
	const int sPx = sc.sPx, sPt = sc.sPt, P = sc.P;
	uint32_t *oidx = out_idx + (size_t)g * sc.npmax;
	if (sc.npmax <= 1) {
		if (lane == 0) { oidx[0] = pos; out_nsim[g] = 1; }
		return;
	}
	const int F = sc.F;
	const int shift_t = min(0, pt - sc.wtb) + max(0, pt + sc.wtf - F + sPt);
	const int rt0 = max(0, pt - sc.wtb - shift_t), rt1 = min(F - sPt, pt + sc.wtf - shift_t);
	const int sWt = rt1 - rt0 + 1;
	for (int k = lane; k < P; k += 64) {
		const int ht = k / (sPx * sPx), r = k - ht * sPx * sPx;
		const int hy = r / sPx, hx = r - hy * sPx;
		sref[k] = img[video_addr(gx + hx, gy + hy, gt + ht, 0, sc)];
	}
	int nk = 0;
	for (int ii = 0; ii < sWt; ii++) {
		const int qt = ii == 0 ? pt : (ii <= rt1 - pt ? pt + ii : pt - (ii - (rt1 - pt)));
		const int dir = qt > pt ? 1 : (qt < pt ? -1 : 0);
		const int dt = qt - rt0;
		const int sw = dir ? wpred : sc.sWx, half = (sw - 1) / 2;
		const int nwin = dir ? npf : 1;
		// window centres, then the windows shifted into the tile
		// (nldct.cpp:1019-1036, 1238-1259)
		for (int i = lane; i < nwin; i += 64) {
			int cx = px, cy = py;
			if (dir) {
				const int ct = qt - dir;
				cx = t_cx[(dt - dir) * npf + i];
				cy = t_cy[(dt - dir) * npf + i];
				float fx = 0.f, fy = 0.f;
				if (sc.use_flow) {
					const float *fl = dir > 0 ? fflow : bflow;
					const size_t base = (size_t)ct * 2 * sc.WH + (size_t)(cy + tb.oy) * sc.W + (cx + tb.ox);
					fx = fl[base];
					fy = fl[base + sc.WH];
				}
				cx = (int)fmax_ref(0.f, fmin_ref((float)Wt - 1.f, roundf((float)cx + fx)));
				cy = (int)fmax_ref(0.f, fmin_ref((float)Ht - 1.f, roundf((float)cy + fy)));
			}
			int shx = min(0, cx - half), shy = min(0, cy - half);
			shx += max(0, cx + half - Wt + sPx);
			shy += max(0, cy + half - Ht + sPx);
			w_rx0[i] = max(0, cx - half - shx);
			w_ry0[i] = max(0, cy - half - shy);
			w_rx1[i] = min(Wt - sPx, cx + half - shx);
			w_ry1[i] = min(Ht - sPx, cy + half - shy);
		}
		__syncthreads();
		// the union, in insertion order
		int nc = 0;
		for (int i = 0; i < nwin; i++) {
			const int nx = w_rx1[i] - w_rx0[i] + 1, ny = w_ry1[i] - w_ry0[i] + 1;
			const int cells = max(0, nx) * max(0, ny);
			for (int c0 = 0; c0 < cells; c0 += 64) {
				const int c = c0 + lane;
				bool keep = false;
				int qx = 0, qy = 0;
				if (c < cells) {
					qy = w_ry0[i] + c / nx;
					qx = w_rx0[i] + c % nx;
					keep = true;
					for (int j = 0; j < i && keep; j++)
						keep = !(qx >= w_rx0[j] && qx <= w_rx1[j] && qy >= w_ry0[j] && qy <= w_ry1[j]);
				}
				const unsigned long long b = __ballot(keep);
				if (keep) {
					const int slot = nc + (int)__popcll(b & ((1ull << lane) - 1ull));
					if (slot < cap) cval[slot] = qx | (qy << 16);
				}
				nc += (int)__popcll(b);
			}
		}
		nc = min(nc, cap);
		__syncthreads();
		// distances in the reference's order, then the bias
		for (int s = lane; s < nc; s += 64) {
			const int qx = cval[s] & 0xffff, qy = cval[s] >> 16;
			const int ax = qx + tb.ox, ay = qy + tb.oy;
			float d = 0.f;
			for (int r = 0; r < reps; r++) {
				int k = 0;
				for (int ht = 0; ht < sPt; ht++)
					for (int hy = 0; hy < sPx; hy++) {
						const float *row = img + video_addr(ax, ay + hy, qt + ht, 0, sc);
						for (int hx = 0; hx < sPx; hx++, k++) {
							const float dif = sref[k] - row[hx];
							d = d + dif * dif;
						}
					}
			}
			if (qy == py && qx == px) d = d - dsub;
			ckey[s] = __float_as_uint(d);
		}
		__syncthreads();
		if (lane == 0) {
			ls_partial_sort(ckey, cval, npf, nc);
			for (int i = 0; i < npf; i++) {
				const int qx = cval[i] & 0xffff, qy = cval[i] >> 16;
				t_cx[dt * npf + i] = qx;
				t_cy[dt * npf + i] = qy;
				kkey[nk + i] = ckey[i];
				kval[nk + i] = (int)((uint32_t)qt * (uint32_t)sc.WH + (uint32_t)(qy + tb.oy) * (uint32_t)sc.W +
				                     (uint32_t)(qx + tb.ox));
			}
		}
		nk += npf;
		__syncthreads();
	}
	if (lane == 0) {
		const int nSimP = min(sc.npmax, nk);
		ls_partial_sort(kkey, kval, nSimP, nk);
		int n = 1;
		while (2 * n <= nSimP && __uint_as_float(kkey[n]) < sc.tau2) n *= 2;
		s_n = n;
		out_nsim[g] = (uint32_t)n;
	}
	__syncthreads();
	for (int k = lane; k < s_n; k += 64) oidx[k] = (uint32_t)kval[k];
}

// ---------------------------------------------------------------------------
// Block matching, staged form (one channel, patches of PX x PX x PT).
//
// bm_window_kernel: one thread per group computes the flow trajectory and
// the shifted search window of every frame of the temporal range
// (nldct.cpp:1173-1259), in the reference's frame schedule, as
// (qt, rx0 | ry0 << 16, nx | ny << 16, first candidate index).
// bm_fast_kernel: one wave per group; the search regions of NF frames at a
// time are staged in LDS, lane = candidate, and each distance is the
// sequential (ht, hy, hx) sum of the reference (nldct.cpp:1262-1274,
// 1723-1737).  Selection as bm_kernel.
// ---------------------------------------------------------------------------
__global__ void bm_window_kernel(const float *__restrict__ fflow, const float *__restrict__ bflow,
                                 const uint32_t *__restrict__ gpos, const uint32_t *__restrict__ gtile,
                                 const TileBox *__restrict__ tiles, StepConst sc, int G,
                                 int4 *__restrict__ wins, int2 *__restrict__ wmeta)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= G) return;
	const int sWt = sc.wtf + sc.wtb + 1;
	int gx, gy, gt;
	pos_decode(gpos[g], sc.W, sc.WH, gx, gy, gt);
	const TileBox tb = tiles[gtile[g]];
	const int Wt = tb.ex - tb.ox, Ht = tb.ey - tb.oy;
	const int px = gx - tb.ox, py = gy - tb.oy, pt = gt;
	const int F = sc.F, sPx = sc.sPx, sPt = sc.sPt;
	const int shift_t = min(0, pt - sc.wtb) + max(0, pt + sc.wtf - F + sPt);
	const int rt0 = max(0, pt - sc.wtb - shift_t);
	const int rt1 = min(F - sPt, pt + sc.wtf - shift_t);
	// The frame schedule is pt, pt+1..rt1, pt-1..rt0: the trajectory is
	// walked in that same order and each frame's window emitted on the way.
	const int half = (sc.sWx - 1) / 2;
	int base = 0, ii = 0;
	auto emit = [&](int qt, int cx, int cy) {
		int shift_x = min(0, cx - half);
		int shift_y = min(0, cy - half);
		shift_x += max(0, cx + half - Wt + sPx);
		shift_y += max(0, cy + half - Ht + sPx);
		const int rx0 = max(0, cx - half - shift_x);
		const int ry0 = max(0, cy - half - shift_y);
		const int rx1 = min(Wt - sPx, cx + half - shift_x);
		const int ry1 = min(Ht - sPx, cy + half - shift_y);
		const int nx = max(0, rx1 - rx0 + 1), ny = max(0, ry1 - ry0 + 1);
		wins[(size_t)g * sWt + ii] = make_int4(qt, rx0 | (ry0 << 16), nx | (ny << 16), base);
		base += nx * ny;
		ii++;
	};
	emit(pt, px, py);
	for (int dir = 1; dir >= -1; dir -= 2) {
		const float *fl = dir > 0 ? fflow : bflow;
		int cx = px, cy = py, ct = pt;
		for (int qt = pt + dir; dir > 0 ? qt <= rt1 : qt >= rt0; qt += dir) {
			float fx = 0.f, fy = 0.f;
			if (sc.use_flow) {
				const size_t b = (size_t)ct * 2 * sc.WH + (size_t)(cy + tb.oy) * sc.W + (cx + tb.ox);
				fx = fl[b];
				fy = fl[b + sc.WH];
			}
			const float cx_f = (float)cx + fx;
			const float cy_f = (float)cy + fy;
			cx = (int)fmax_ref(0.f, fmin_ref((float)Wt - 1.f, roundf(cx_f)));
			cy = (int)fmax_ref(0.f, fmin_ref((float)Ht - 1.f, roundf(cy_f)));
			ct = qt;
			emit(qt, cx, cy);
		}
	}
	const int nf = ii;
	wmeta[g] = make_int2(nf, base);
}

// Selection of the kk smallest of the keys skey[0..nsrch) in the total order
// (key, reference patch first, scan index) into out_key / out_rank[0..kk),
// in that order (rank -1 = the reference patch, s_self).  One wave: a radix
// select of the kk-th key over four 8-bit digits, then the keys below it and
// the lowest-ranked ties at it, sorted.  stie: scratch of nsrch ints.
__device__ void bm_radix_select(const uint32_t *__restrict__ skey, int *__restrict__ stie, int nsrch, int kk,
                                int s_self, uint32_t *__restrict__ out_key, int *__restrict__ out_rank)
{
	__shared__ uint32_t hist[256];
	__shared__ uint32_t s_bucket, s_below;
	__shared__ uint32_t sel_key[64];
	__shared__ int sel_rank[64];
	__shared__ int s_na, s_nb;
	const int lane = threadIdx.x;
	if (lane == 0) { s_na = 0; s_nb = 0; }
	__syncthreads();
	uint32_t prefix = 0, pmask = 0;
	uint32_t remaining = (uint32_t)kk;
	for (int pass = 0; pass < 4; pass++) {
		const int shift = 24 - 8 * pass;
#pragma unroll
		for (int b = 0; b < 4; b++) hist[4 * lane + b] = 0;
		__syncthreads();
		for (int s2 = lane; s2 < nsrch; s2 += 64) {
			const uint32_t k = skey[s2];
			if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
		}
		__syncthreads();
		{
			const uint32_t c0 = hist[4 * lane], c1 = hist[4 * lane + 1], c2 = hist[4 * lane + 2], c3 = hist[4 * lane + 3];
			const int local = (int)(c0 + c1 + c2 + c3);
			const int incl = wave_incl_scan(local);
			const int excl = incl - local;
			if ((uint32_t)excl < remaining && remaining <= (uint32_t)incl) {
				uint32_t cum = (uint32_t)excl;
				const uint32_t cs[4] = {c0, c1, c2, c3};
				for (int b = 0; b < 4; b++) {
					if (cum + cs[b] >= remaining) { s_bucket = 4 * lane + b; s_below = cum; break; }
					cum += cs[b];
				}
			}
		}
		__syncthreads();
		prefix |= s_bucket << shift;
		pmask |= 255u << shift;
		remaining -= s_below;
		__syncthreads();
	}
	const uint32_t T = prefix;
	const int na_expected = kk - (int)remaining;
	for (int s2 = lane; s2 < nsrch; s2 += 64) {
		const uint32_t k = skey[s2];
		const int rank = s2 == s_self ? -1 : s2;
		if (k < T) {
			const int slot = atomicAdd(&s_na, 1);
			sel_key[slot] = k;
			sel_rank[slot] = rank;
		} else if (k == T) {
			const int slot = atomicAdd(&s_nb, 1);
			stie[slot] = rank;
		}
	}
	__syncthreads();
	const int nb = s_nb;
	for (int j = lane; j < nb; j += 64) {
		const int r = stie[j];
		int c = 0;
		for (int i = 0; i < nb; i++) c += stie[i] < r;
		if (c < (int)remaining) {
			sel_key[na_expected + c] = T;
			sel_rank[na_expected + c] = r;
		}
	}
	__syncthreads();
	if (lane < kk) {
		const uint32_t k = sel_key[lane];
		const int r = sel_rank[lane];
		int p = 0;
		for (int j = 0; j < kk; j++) {
			const uint32_t kj = sel_key[j];
			p += (kj < k) || (kj == k && sel_rank[j] < r);
		}
		out_key[p] = k;
		out_rank[p] = r;
	}
	__syncthreads();
}

// Power-of-two tau cut (nldct.cpp:1285-1289) on the sorted selection and the
// neighbour positions t*W*H + y*W + x of the kept candidates.
__device__ void bm_emit(const uint32_t *__restrict__ out_key, const int *__restrict__ out_rank, int nSimP,
                        const StepConst &sc, const int4 *__restrict__ gw, int nframes, uint32_t pos,
                        const TileBox &tb, uint32_t *__restrict__ oidx, uint32_t *__restrict__ onsim)
{
	const int lane = threadIdx.x;
	int n = 1;
	while (2 * n <= nSimP && __uint_as_float(out_key[n]) < sc.tau2) n *= 2;
	if (lane == 0) *onsim = n;
	if (lane < n) {
		const int r = out_rank[lane];
		uint32_t q;
		if (r < 0) q = pos;
		else {
			int ii = 0;
			while (ii + 1 < nframes && gw[ii + 1].w <= r) ii++;
			const int4 f = gw[ii];
			const int o = r - f.w;
			const int nx = f.z & 0xffff;
			const int qy = (f.y >> 16) + o / nx, qx = (f.y & 0xffff) + o % nx;
			q = (uint32_t)f.x * (uint32_t)sc.WH + (uint32_t)(qy + tb.oy) * (uint32_t)sc.W + (uint32_t)(qx + tb.ox);
		}
		oidx[lane] = q;
	}
}

template <int PX, int PT, int NF>
__global__ __launch_bounds__(64)
void bm_fast_kernel(const float *__restrict__ img, const int4 *__restrict__ wins,
                    const int2 *__restrict__ wmeta, const uint32_t *__restrict__ gpos,
                    const uint32_t *__restrict__ gtile, const TileBox *__restrict__ tiles,
                    StepConst sc, int stage_floats, uint32_t *__restrict__ out_idx,
                    uint32_t *__restrict__ out_nsim, int G, uint32_t *__restrict__ tie_count,
                    uint32_t *__restrict__ tie_list, int tie_cap)
{
	constexpr int PP = PX * PX * PT;
	extern __shared__ unsigned char smem[];
	const int sWt = sc.wtf + sc.wtb + 1;
	const int maxCand = sc.sWx * sc.sWx * sWt;
	uint32_t *skey = (uint32_t *)smem;                    // maxCand
	float *stage = (float *)(skey + maxCand);             // stage_floats (two frames' regions)
	int *stie = (int *)stage;                             // ties, after the distances
	__shared__ uint32_t out_key[64];
	__shared__ int out_rank[64];

	const int g = blockIdx.x;
	if (g >= G) return;
	const int lane = threadIdx.x;
	const uint32_t pos = gpos[g];
	uint32_t *oidx = out_idx + (size_t)g * sc.npmax;
	if (sc.npmax <= 1) {
		if (lane == 0) { oidx[0] = pos; out_nsim[g] = 1; }
		return;
	}
	int gx, gy, gt;
	pos_decode(pos, sc.W, sc.WH, gx, gy, gt);
	const TileBox tb = tiles[gtile[g]];
	const int px = gx - tb.ox, py = gy - tb.oy, pt = gt;
	const int2 meta = wmeta[g];
	const int nframes = meta.x, nsrch = meta.y;
	const int4 *gw = wins + (size_t)g * sWt;

	// reference patch (channel 0)
	float ref[PP];
#pragma unroll
	for (int ht = 0; ht < PT; ht++)
#pragma unroll
		for (int hy = 0; hy < PX; hy++)
#pragma unroll
			for (int hx = 0; hx < PX; hx++)
				ref[(ht * PX + hy) * PX + hx] = img[(size_t)(gt + ht) * sc.WHC + (size_t)(gy + hy) * sc.W + gx + hx];

	for (int ii = 0; ii < nframes; ii += NF) {
		// stage frames ii .. ii+NF-1
		int4 f[NF];
		int rw[NF], rh[NF], off[NF], cst[NF], ncand[NF];
		int tot = 0, n01 = 0;
#pragma unroll
		for (int j = 0; j < NF; j++) {
			f[j] = ii + j < nframes ? gw[ii + j] : make_int4(0, 0, 0, 0);
			const int nx = f[j].z & 0xffff, ny = f[j].z >> 16;
			rw[j] = nx + PX - 1;
			rh[j] = ny + PX - 1;
			ncand[j] = (ii + j < nframes) ? nx * ny : 0;
			off[j] = tot;
			cst[j] = n01;
			tot += ncand[j] ? rw[j] * rh[j] * PT : 0;
			n01 += ncand[j];
		}
		__syncthreads();
		for (int i = lane; i < tot; i += 64) {
			int j = 0;
#pragma unroll
			for (int jj = 1; jj < NF; jj++) if (ncand[jj] && i >= off[jj]) j = jj;
			const int r = i - off[j];
			const int ht = r / (rw[j] * rh[j]);
			const int q = r - ht * rw[j] * rh[j];
			const int yy = q / rw[j], xx = q - yy * rw[j];
			const int ax = tb.ox + (f[j].y & 0xffff) + xx, ay = tb.oy + (f[j].y >> 16) + yy;
			stage[i] = img[(size_t)(f[j].x + ht) * sc.WHC + (size_t)ay * sc.W + ax];
		}
		__syncthreads();
		for (int o = lane; o < n01; o += 64) {
			int j = 0;
#pragma unroll
			for (int jj = 1; jj < NF; jj++) if (ncand[jj] && o >= cst[jj]) j = jj;
			const int lo = o - cst[j];
			const int nx = f[j].z & 0xffff;
			const int cy = lo / nx, cx = lo - cy * nx;
			const float *b = stage + off[j] + cy * rw[j] + cx;
			float d = 0.f;
#pragma unroll
			for (int ht = 0; ht < PT; ht++)
#pragma unroll
				for (int hy = 0; hy < PX; hy++) {
					const float *r = b + (ht * rh[j] + hy) * rw[j];
#pragma unroll
					for (int hx = 0; hx < PX; hx++) {
						const float dif = ref[(ht * PX + hy) * PX + hx] - r[hx];
						d = d + dif * dif;
					}
				}
			skey[f[j].w + lo] = __float_as_uint(d);
		}
	}
	// index of the reference patch itself (frame pt is the first of the list)
	int s_self;
	{
		const int4 f0 = gw[0];
		s_self = f0.w + (py - (f0.y >> 16)) * (f0.z & 0xffff) + (px - (f0.y & 0xffff));
	}
	const int nSimP = min(sc.npmax, nsrch);
	int kk = 1;
	while (kk * 2 <= nSimP) kk *= 2;
	__syncthreads();
	bm_radix_select(skey, stie, nsrch, kk, s_self, out_key, out_rank);
	__syncthreads();
	const bool tie = tie_list ? bm_has_tie(skey, nsrch, out_key, nSimP, sc.tau2) : false;
	if (tie && threadIdx.x == 0) bm_tie_append(g, tie_count, tie_list, tie_cap);
	bm_emit(out_key, out_rank, nSimP, sc, gw, nframes, pos, tb, oidx, out_nsim + g);
}

// ---------------------------------------------------------------------------
// Block matching, register-window form (one channel, patches PX x PX x PT,
// PX = 4 or 8).  One wave per group; a work item is one candidate column
// (frame j, column cx) and a range of candidate rows.  The item walks down
// its rows with the PX rows x PT frames of PX samples that the current
// candidate covers held in registers: each step loads one new row per frame
// (vector loads from global memory, cached) and computes one distance, the
// sequential (ht, hy, hx) sum of the reference (nldct.cpp:1262-1274,
// 1723-1737), with the reference patch in scalar registers.  Distances go to
// LDS by scan index.  Selection: every lane keeps its two smallest
// distances; the kk-th smallest of these 128 bounds the kk-th smallest
// distance from above (they are kk distances of distinct candidates at or
// below it), so only the candidates at or below that bound are ranked, in
// the total order of bm_radix_select (distance, then scan index with the
// reference patch first) as one 64-bit key; more than BMR_CAP of them (or
// fewer than kk finite values) falls back to the radix select over all
// distances.  Same output as bm_fast_kernel / bm_kernel.
// ---------------------------------------------------------------------------

#define BMR_CAP 128

template <int PX, int PT>
__device__ __forceinline__ void bmr_load_row(float (&dst)[PT][PX], const float *__restrict__ p, int frame_stride)
{
#pragma unroll
	for (int ht = 0; ht < PT; ht++)
#pragma unroll
		for (int q = 0; q < PX / 4; q++) {
			float4 v;
			__builtin_memcpy(&v, p + (size_t)ht * frame_stride + 4 * q, sizeof v);
			dst[ht][4 * q + 0] = v.x;
			dst[ht][4 * q + 1] = v.y;
			dst[ht][4 * q + 2] = v.z;
			dst[ht][4 * q + 3] = v.w;
		}
}

template <int PX, int PT, bool MC>
__global__ __launch_bounds__(64)
void bm_reg_kernel(const float *__restrict__ img, const int4 *__restrict__ wins,
                   const int2 *__restrict__ wmeta, const uint32_t *__restrict__ gpos,
                   const uint32_t *__restrict__ gtile, const TileBox *__restrict__ tiles,
                   StepConst sc, int rchunks, int nch, uint32_t *__restrict__ out_idx,
                   uint32_t *__restrict__ out_nsim, int G, uint32_t *__restrict__ tie_count,
                   uint32_t *__restrict__ tie_list, int tie_cap)
{
	constexpr int PP = PX * PX * PT;
	extern __shared__ unsigned char smem[];
	const int sWt = sc.wtf + sc.wtb + 1;
	const int maxCand = sc.sWx * sc.sWx * sWt;
	uint32_t *skey = (uint32_t *)smem;          // maxCand distances
	int *stie = (int *)(skey + maxCand);         // maxCand (radix fallback)
	__shared__ uint64_t cand[BMR_CAP];   // (distance << 32) | (scan index + 1; 0: the reference patch)
	__shared__ uint32_t out_key[64];
	__shared__ int out_rank[64];

	const int g = blockIdx.x;
	if (g >= G) return;
	const int lane = threadIdx.x;
	const uint32_t pos = gpos[g];
	uint32_t *oidx = out_idx + (size_t)g * sc.npmax;
	if (sc.npmax <= 1) {
		if (lane == 0) { oidx[0] = pos; out_nsim[g] = 1; }
		return;
	}
	int gx, gy, gt;
	pos_decode(pos, sc.W, sc.WH, gx, gy, gt);
	const TileBox tb = tiles[gtile[g]];
	const int px = gx - tb.ox, py = gy - tb.oy;
	const int2 meta = wmeta[g];
	const int nframes = meta.x, nsrch = meta.y;
	const int4 *gw = wins + (size_t)g * sWt;

	// items: (frame j, column cx, row chunk r), frame-major; lane j <
	// nframes holds frame j's item count and the inclusive sum up to it, so
	// that an item finds its frame without a chain of dependent loads
	const int my_nit = lane < nframes ? (gw[lane].z & 0xffff) * rchunks : 0;
	const int my_cum = wave_incl_scan(my_nit);
	int total = 0;
	if (nframes <= 64) total = __shfl(my_cum, nframes - 1, 64);
	else
		for (int j = 0; j < nframes; j++) total += (gw[j].z & 0xffff) * rchunks;
	uint32_t lmin = 0xffffffffu, lmin2 = 0xffffffffu;   // the lane's two smallest distances
	// channels in order (step 1: channel 0; step 2: all): a channel's sums
	// continue the previous channel's, which the same lane stored in skey
	// (MC: more than one channel; the one-channel form keeps no such state)
	const int nc = MC ? nch : 1;
	for (int c = 0; c < nc; c++) {
		const bool last = c == nc - 1;
		const float *imc = img + (size_t)c * sc.WH;
		// reference patch of this channel, wave-uniform
		float ref[PP];
#pragma unroll
		for (int ht = 0; ht < PT; ht++)
#pragma unroll
			for (int hy = 0; hy < PX; hy++)
#pragma unroll
				for (int hx = 0; hx < PX; hx++)
					ref[(ht * PX + hy) * PX + hx] =
					    imc[(size_t)(gt + ht) * sc.WHC + (size_t)(gy + hy) * sc.W + gx + hx];
		for (int it = lane; it < total; it += 64) {
			int j = 0, loc = it;
			int4 f;
			if (nframes <= 64) {
				// (readlane reads lane jj whatever the branch's active lanes)
				int base = 0;
				for (int jj = 0; jj + 1 < nframes; jj++) {
					const int c = __builtin_amdgcn_readlane(my_cum, jj);   // first item of frame jj + 1
					if (it >= c) { j = jj + 1; base = c; }
				}
				loc = it - base;
				f = gw[j];
			} else {
				f = gw[0];
				for (;;) {
					const int nit = (f.z & 0xffff) * rchunks;
					if (loc < nit) break;
					loc -= nit;
					f = gw[++j];
				}
			}
			const int nx = f.z & 0xffff, ny = f.z >> 16;
			const int r = loc / nx, cx = loc - r * nx;
			const int y0 = r * ny / rchunks, y1 = (r + 1) * ny / rchunks;
			if (y0 >= y1) continue;
			const float *p = imc + (size_t)f.x * sc.WHC + (size_t)(tb.oy + (f.y >> 16)) * sc.W +
			                 (tb.ox + (f.y & 0xffff) + cx);
			uint32_t *kout = skey + f.w + cx;
			float w[PX][PT][PX];
#pragma unroll
			for (int k = 0; k < PX - 1; k++) bmr_load_row<PX, PT>(w[k], p + (size_t)(y0 + k) * sc.W, sc.WHC);
			for (int y = y0; y < y1; y += PX) {
#pragma unroll
				for (int u = 0; u < PX; u++) {
					if (y + u < y1) {
						bmr_load_row<PX, PT>(w[(u + PX - 1) % PX], p + (size_t)(y + u + PX - 1) * sc.W, sc.WHC);
						float d = MC && c ? __uint_as_float(kout[(y + u) * nx]) : 0.f;
#pragma unroll
						for (int ht = 0; ht < PT; ht++)
#pragma unroll
							for (int hy = 0; hy < PX; hy++)
#pragma unroll
								for (int hx = 0; hx < PX; hx++) {
									const float dif = ref[(ht * PX + hy) * PX + hx] - w[(u + hy) % PX][ht][hx];
									d = d + dif * dif;
								}
						const uint32_t kb = __float_as_uint(d);
						kout[(y + u) * nx] = kb;
						if (!MC || last) {
							lmin2 = min(lmin2, max(lmin, kb));
							lmin = min(lmin, kb);
						}
					}
				}
			}
		}
	}
	// index of the reference patch itself (frame pt is the first of the list)
	int s_self;
	{
		const int4 f0 = gw[0];
		s_self = f0.w + (py - (f0.y >> 16)) * (f0.z & 0xffff) + (px - (f0.y & 0xffff));
	}
	const int nSimP = min(sc.npmax, nsrch);
	int kk = 1;
	while (kk * 2 <= nSimP) kk *= 2;

	// bound: the kk-th smallest of the lanes' two smallest distances (a lane
	// with fewer candidates holds ~0, which then means no bound), by a
	// bitonic sort of the 128 values (value i = lane + 64 k in v[k])
	uint32_t T0;
	{
		uint32_t v[2] = {lmin, lmin2};
#pragma unroll
		for (int size = 2; size <= 128; size <<= 1)
#pragma unroll
			for (int j = size >> 1; j > 0; j >>= 1) {
				if (j == 64) {
					const uint32_t lo = min(v[0], v[1]), hi = max(v[0], v[1]);
					v[0] = lo;
					v[1] = hi;
				} else {
#pragma unroll
					for (int k = 0; k < 2; k++) {
						const int i = lane + 64 * k;
						const uint32_t o = (uint32_t)__shfl_xor((int)v[k], j, 64);
						const bool keep_min = ((i & j) == 0) == ((i & size) == 0);
						v[k] = keep_min ? min(v[k], o) : max(v[k], o);
					}
				}
			}
		T0 = (uint32_t)__shfl((int)v[0], kk - 1, 64);
	}
	int cnt = 0;
	bool fits = T0 != 0xffffffffu;
	if (fits) {
		// candidates at or below the bound, in scan order
		for (int b = 0; b < nsrch; b += 64) {
			const int s2 = b + lane;
			const uint32_t k = s2 < nsrch ? skey[s2] : 0xffffffffu;
			const bool take = s2 < nsrch && k <= T0;
			const uint64_t m = __ballot(take);
			const int before = __popcll(m & ((1ull << lane) - 1ull));
			if (take && cnt + before < BMR_CAP)
				cand[cnt + before] = ((uint64_t)k << 32) | (uint32_t)(s2 == s_self ? 0 : s2 + 1);
			cnt += __popcll(m);
		}
		fits = cnt <= BMR_CAP;
	}
	__syncthreads();
	if (fits) {
		for (int e = lane; e < cnt; e += 64) {
			const uint64_t c = cand[e];
			int p = 0;
			for (int j = 0; j < cnt; j++) p += cand[j] < c;
			if (p < kk) {
				out_key[p] = (uint32_t)(c >> 32);
				out_rank[p] = (int)(uint32_t)c - 1;
			}
		}
		__syncthreads();
	} else {
		bm_radix_select(skey, stie, nsrch, kk, s_self, out_key, out_rank);
	}
	__syncthreads();
	// every distance up to the bound is among the ranked candidates
	const bool tie = tie_list ? (fits ? bm_has_tie(cand, cnt, out_key, nSimP, sc.tau2)
	                                  : bm_has_tie(skey, nsrch, out_key, nSimP, sc.tau2))
	                          : false;
	if (tie && lane == 0) bm_tie_append(g, tie_count, tie_list, tie_cap);
	bm_emit(out_key, out_rank, nSimP, sc, gw, nframes, pos, tb, oidx, out_nsim + g);
}

#define BMR_ARGS const float *, const int4 *, const int2 *, const uint32_t *, const uint32_t *, \
	const TileBox *, StepConst, int, int, uint32_t *, uint32_t *, int, uint32_t *, uint32_t *, int
template __global__ void bm_reg_kernel<4, 4, false>(BMR_ARGS);
template __global__ void bm_reg_kernel<8, 1, false>(BMR_ARGS);
template __global__ void bm_reg_kernel<4, 4, true>(BMR_ARGS);
template __global__ void bm_reg_kernel<8, 1, true>(BMR_ARGS);

#define BMF_ARGS const float *, const int4 *, const int2 *, const uint32_t *, const uint32_t *, \
	const TileBox *, StepConst, int, uint32_t *, uint32_t *, int, uint32_t *, uint32_t *, int
template __global__ void bm_fast_kernel<4, 4, 1>(BMF_ARGS);
template __global__ void bm_fast_kernel<4, 4, 2>(BMF_ARGS);
template __global__ void bm_fast_kernel<4, 4, 4>(BMF_ARGS);
template __global__ void bm_fast_kernel<8, 1, 1>(BMF_ARGS);
template __global__ void bm_fast_kernel<8, 1, 2>(BMF_ARGS);
template __global__ void bm_fast_kernel<8, 1, 4>(BMF_ARGS);

// ---------------------------------------------------------------------------
// Group filtering, one workgroup per group.
//   step 1: computeBayesEstimateStep1 (nldct.cpp:1962-2006) ->
//           computeBayesEstimateStep1_vbm3d (:1799-1943)
//   step 2: computeBayesEstimateStep2 (:2196-2242) -> _vbm3d (:2015-2174)
// Each neighbour's filtered patch, already multiplied by its aggregation
// weight and window (the factor computeAggregationStep{1,2} applies,
// nldct.cpp:2296-2299, 2385-2393), is written to row e = base[g-g0] + n of
// est: rows follow the reference's processing order (group, then neighbour).
// ---------------------------------------------------------------------------

// Z[k*N+n] = sum_pix X[pix*N+n] * B[k*P+pix], sequential in pix.
__device__ __forceinline__ void fwd_transform(const float *__restrict__ X, const float *__restrict__ B,
                                              float *__restrict__ Z, int P, int N)
{
	for (int o = threadIdx.x; o < P * N; o += FILTER_THREADS) {
		const int k = o / N, n = o - k * N;
		const float *b = B + (size_t)k * P;
		float acc = 0.f;
		for (int p = 0; p < P; p++) acc = __builtin_fmaf(X[p * N + n], b[p], acc);
		Z[o] = acc;
	}
}

// Xh[pix*N+n] = sum_k Z[k*N+n] * Bi[k*P+pix], sequential in k.
__device__ __forceinline__ void inv_transform(const float *__restrict__ Z, const float *__restrict__ Bi,
                                              float *__restrict__ Xh, int P, int N)
{
	for (int o = threadIdx.x; o < P * N; o += FILTER_THREADS) {
		const int pix = o / N, n = o - pix * N;
		float acc = 0.f;
		for (int k = 0; k < P; k++) acc = __builtin_fmaf(Z[k * N + n], Bi[(size_t)k * P + pix], acc);
		Xh[o] = acc;
	}
}

__device__ __forceinline__ void haar(float *Z, int P, int N, bool forward)
{
	const float isqrt2 = 0.70710678118654752440f; // (float)(1./sqrt(2.))
	for (int it = 0, s = forward ? 2 : N; forward ? s <= N : s >= 2; it++, s = forward ? s * 2 : s / 2) {
		const int h = s / 2;
		const int nb = N / s;
		for (int o = threadIdx.x; o < P * nb; o += FILTER_THREADS) {
			const int i = o / nb, n = (o - i * nb) * s;
			const float z0 = Z[n + i * N] * isqrt2;
			const float z1 = Z[n + h + i * N] * isqrt2;
			Z[n + i * N] = z0 + z1;
			Z[n + h + i * N] = z0 - z1;
		}
		__syncthreads();
	}
}

__device__ __forceinline__ void load_group(const float *__restrict__ img, const uint32_t *__restrict__ gidx,
                                           int N, int c, const StepConst &sc, float *__restrict__ X)
{
	const int sPx = sc.sPx;
	for (int o = threadIdx.x; o < sc.P * N; o += FILTER_THREADS) {
		const int pix = o / N, n = o - pix * N;
		const int ht = pix / (sPx * sPx), r = pix - ht * sPx * sPx;
		const int hy = r / sPx, hx = r - hy * sPx;
		int x, y, t;
		pos_decode(gidx[n], sc.W, sc.WH, x, y, t);
		X[o] = img[video_addr(x + hx, y + hy, t + ht, c, sc)];
	}
}

// 1 when all N patches of the group are equal in every channel
// (nldct.cpp:1986-1999, 2222-2235).
__device__ bool group_all_equal(const float *__restrict__ img, const uint32_t *__restrict__ gidx,
                                int N, const StepConst &sc)
{
	const int sPx = sc.sPx;
	int differs = 0;
	if (N > 1) {
		for (int o = threadIdx.x; o < sc.C * sc.P && !differs; o += FILTER_THREADS) {
			const int c = o / sc.P, pix = o - c * sc.P;
			const int ht = pix / (sPx * sPx), r = pix - ht * sPx * sPx;
			const int hy = r / sPx, hx = r - hy * sPx;
			int x, y, t;
			pos_decode(gidx[0], sc.W, sc.WH, x, y, t);
			const float v = img[video_addr(x + hx, y + hy, t + ht, c, sc)];
			for (int n = 1; n < N; n++) {
				pos_decode(gidx[n], sc.W, sc.WH, x, y, t);
				if (v != img[video_addr(x + hx, y + hy, t + ht, c, sc)]) { differs = 1; break; }
			}
		}
	} else
		differs = 1;
	return !__syncthreads_or(differs);
}

// Writes the raw (unfiltered) group of an early-out group; the caller
// replaces its weights later with the stale values the reference reuses.
__device__ void write_raw_group(const float *__restrict__ img, const uint32_t *__restrict__ gidx,
                                const uint32_t *__restrict__ rrow, int N, const StepConst &sc,
                                float *__restrict__ est)
{
	const int sPx = sc.sPx, P = sc.P;
	for (int o = threadIdx.x; o < sc.C * P * N; o += FILTER_THREADS) {
		const int n = o / (sc.C * P), r = o - n * sc.C * P;
		const int c = r / P, pix = r - c * P;
		const int ht = pix / (sPx * sPx), q = pix - ht * sPx * sPx;
		const int hy = q / sPx, hx = q - hy * sPx;
		int x, y, t;
		pos_decode(gidx[n], sc.W, sc.WH, x, y, t);
		est[((size_t)rrow[n] * sc.C + c) * P + pix] = img[video_addr(x + hx, y + hy, t + ht, c, sc)];
	}
}

__global__ __launch_bounds__(FILTER_THREADS)
void filter1_kernel(const float *__restrict__ img, const uint32_t *__restrict__ idx,
                    const uint32_t *__restrict__ nsim, const uint32_t *__restrict__ base,
                    const uint32_t *__restrict__ gtile, StepConst sc, int g0, int Gc,
                    const float *__restrict__ basis, const float *__restrict__ binv,
                    const float *__restrict__ win, float *__restrict__ est,
                    float *__restrict__ entry_w, uint32_t *__restrict__ etile,
                    uint8_t *__restrict__ early, float *__restrict__ gweight,
                    uint32_t *__restrict__ n_early)
{
	extern __shared__ float fsm[];
	__shared__ uint32_t s_rrow[64];
	__shared__ int s_nnz_i;
	__shared__ float s_terms[256];
	const int gl = blockIdx.x;
	if (gl >= Gc) return;
	const int g = g0 + gl;
	const int N = (int)nsim[g];
	const int P = sc.P, C = sc.C;
	const uint32_t *gidx = idx + (size_t)g * sc.npmax;
	float *X = fsm;                  // C*P*N (filtered estimates of all channels)
	float *Z = X + (size_t)C * P * N; // P*N
	if (threadIdx.x < N) {
		s_rrow[threadIdx.x] = base[gl] + threadIdx.x;
		etile[base[gl] + threadIdx.x] = gtile[g];
	}
	if (threadIdx.x == 0) s_nnz_i = 0;
	__syncthreads();

	if (group_all_equal(img, gidx, N, sc)) {
		write_raw_group(img, gidx, s_rrow, N, sc, est);
		if (threadIdx.x == 0) { early[g] = 1; gweight[g] = 0.f; atomicAdd(n_early, 1u); }
		return;
	}

	float nnz_f = 0.f; // orderInvariance: float sum as in the reference
	for (int c = 0; c < C; c++) {
		float *Xc = X + (size_t)c * P * N;
		load_group(img, gidx, N, c, sc, Xc);
		__syncthreads();
		fwd_transform(Xc, basis, Z, P, N);
		__syncthreads();
		haar(Z, P, N, true);
		if (!sc.order_inv) {
			int cnt = 0;
			for (int o = threadIdx.x; o < P * N; o += FILTER_THREADS) {
				const float z = Z[o];
				if (z * z > sc.beta_sigma2 || o == 0) cnt++;
				else Z[o] = 0.f;
			}
			if (cnt) atomicAdd(&s_nnz_i, cnt);
		} else {
			for (int i = threadIdx.x; i < P; i += FILTER_THREADS) {
				float *z = Z + i * N;
				float var_i = 0.f;
				for (int n = 1; n < N; ++n) var_i = var_i + z[n] * z[n];
				var_i = var_i / (float)(N - 1);
				const float var_0 = fmax_ref(z[0] * z[0] - sc.sigma2, 0.f);
				const float w_0 = var_0 / (var_0 + sc.betaM_sigma2);
				z[0] = z[0] * w_0;
				var_i = fmax_ref(var_i - sc.sigma2, 0.f);
				const float w_i = var_i / (var_i + sc.beta_sigma2);
				for (int n = 1; n < N; ++n) z[n] = z[n] * w_i;
				s_terms[i] = w_0 * w_0 + (float)(N - 1) * w_i * w_i;
			}
			__syncthreads();
			if (threadIdx.x == 0)
				for (int i = 0; i < P; i++) nnz_f = nnz_f + s_terms[i];
		}
		__syncthreads();
		haar(Z, P, N, false);
		inv_transform(Z, binv, Xc, P, N);
		__syncthreads();
	}
	__shared__ float s_w;
	if (threadIdx.x == 0) {
		const float nnz = sc.order_inv ? nnz_f : (float)s_nnz_i;
		s_w = 1 / sc.sigma2 / nnz;
		gweight[g] = s_w;
		early[g] = 0;
	}
	__syncthreads();
	const float w = s_w;
	for (int o = threadIdx.x; o < C * P * N; o += FILTER_THREADS) {
		const int n = o / (C * P), r = o - n * C * P;
		const int c = r / P, pix = r - c * P;
		const float ww = w * win[pix];
		est[((size_t)s_rrow[n] * C + c) * P + pix] = ww * X[((size_t)c * P + pix) * N + n];
	}
	for (int n = threadIdx.x; n < N; n += FILTER_THREADS) entry_w[s_rrow[n]] = w;
}

__global__ __launch_bounds__(FILTER_THREADS)
void filter2_kernel(const float *__restrict__ noisy, const float *__restrict__ basic,
                    const uint32_t *__restrict__ idx, const uint32_t *__restrict__ nsim,
                    const uint32_t *__restrict__ base, const uint32_t *__restrict__ gtile,
                    StepConst sc, int g0, int Gc,
                    const float *__restrict__ basis, const float *__restrict__ binv,
                    const float *__restrict__ win, float *__restrict__ est,
                    float *__restrict__ entry_w, uint32_t *__restrict__ etile,
                    uint8_t *__restrict__ early, float *__restrict__ gweight,
                    uint32_t *__restrict__ n_early)
{
	extern __shared__ float fsm[];
	__shared__ uint32_t s_rrow[64];
	__shared__ float s_terms[256];
	const int gl = blockIdx.x;
	if (gl >= Gc) return;
	const int g = g0 + gl;
	const int N = (int)nsim[g];
	const int P = sc.P, C = sc.C;
	const uint32_t *gidx = idx + (size_t)g * sc.npmax;
	float *X = fsm;                   // C*P*N noisy group, then estimates
	float *Zn = X + (size_t)C * P * N; // P*N
	float *Yb = Zn + (size_t)P * N;   // P*N (basic group, then its coefficients)
	if (threadIdx.x < N) {
		s_rrow[threadIdx.x] = base[gl] + threadIdx.x;
		etile[base[gl] + threadIdx.x] = gtile[g];
	}
	__syncthreads();

	if (group_all_equal(noisy, gidx, N, sc)) {
		write_raw_group(noisy, gidx, s_rrow, N, sc, est);
		if (threadIdx.x == 0) { early[g] = 1; gweight[g] = 0.f; atomicAdd(n_early, 1u); }
		return;
	}

	float part = 0.f;
	float var_oi = 0.f;
	for (int c = 0; c < C; c++) {
		float *Xc = X + (size_t)c * P * N;
		load_group(noisy, gidx, N, c, sc, Xc);
		load_group(basic, gidx, N, c, sc, Zn);
		__syncthreads();
		fwd_transform(Zn, basis, Yb, P, N);   // basic coefficients
		__syncthreads();
		fwd_transform(Xc, basis, Zn, P, N);   // noisy coefficients
		__syncthreads();
		haar(Zn, P, N, true);
		haar(Yb, P, N, true);
		if (!sc.order_inv) {
			// w^2 terms are kept (in Yb) and summed in the reference's order
			// below: the float sum sets the aggregation weight.
			for (int o = threadIdx.x; o < P * N; o += FILTER_THREADS) {
				const float y = Yb[o];
				const float w = y * y / (y * y + sc.beta_sigma2);
				Zn[o] = Zn[o] * w;
				Yb[o] = w * w;
			}
			__syncthreads();
			if (threadIdx.x == 0) {
				float v = part;
				for (int o = 0; o < P * N; o++) v = v + Yb[o];
				part = v;
			}
		} else {
			for (int i = threadIdx.x; i < P; i += FILTER_THREADS) {
				const float *y = Yb + i * N;
				float *z = Zn + i * N;
				const float var_0 = y[0] * y[0];
				const float w_0 = var_0 / (var_0 + sc.betaM_sigma2);
				z[0] = z[0] * w_0;
				float var_i = 0.f;
				for (int n = 1; n < N; ++n) var_i = var_i + y[n] * y[n];
				var_i = var_i / (float)(N - 1);
				const float w_i = var_i / (var_i + sc.beta_sigma2);
				for (int n = 1; n < N; ++n) z[n] = z[n] * w_i;
				s_terms[i] = w_0 * w_0 + (float)(N - 1) * w_i * w_i;
			}
			__syncthreads();
			if (threadIdx.x == 0)
				for (int i = 0; i < P; i++) var_oi = var_oi + s_terms[i];
		}
		__syncthreads();
		haar(Zn, P, N, false);
		inv_transform(Zn, binv, Xc, P, N);
		__syncthreads();
	}
	__shared__ float s_w;
	if (threadIdx.x == 0) {
		const float variance = sc.order_inv ? var_oi : part;
		s_w = (float)(1. / (double)sc.sigma2 / (double)variance);
		gweight[g] = s_w;
		early[g] = 0;
	}
	__syncthreads();
	const float w = s_w;
	for (int o = threadIdx.x; o < C * P * N; o += FILTER_THREADS) {
		const int n = o / (C * P), r = o - n * C * P;
		const int c = r / P, pix = r - c * P;
		const float ww = w * win[pix];
		est[((size_t)s_rrow[n] * C + c) * P + pix] = ww * X[((size_t)c * P + pix) * N + n];
	}
	for (int n = threadIdx.x; n < N; n += FILTER_THREADS) entry_w[s_rrow[n]] = w;
}

// ---------------------------------------------------------------------------
// Register-resident filter for one-channel groups of 64-sample patches
// (px=4/pt=4 or px=8/pt=1) with at most 32 neighbours: one wave per group.
//
// Forward transform: lane k holds row k of the basis in registers; the
// samples of patch n are wave-uniform and arrive through scalar loads, so
// each coefficient z[k][n] is a register dot product in pixel order.
// Haar along the group and the threshold / Wiener shrinkage then run inside
// each lane (lane k owns coefficient k of every patch).  Inverse transform:
// lane pix holds column pix of the inverse basis; the coefficients of
// patch n are read from LDS as wave-uniform (broadcast) values, sequential
// in k.  Same arithmetic, same order as filter1_kernel / filter2_kernel.
// ---------------------------------------------------------------------------

#define FAST_P 64
#define FAST_NMAX 32

// Haar transform across the rows of lane-held columns, levels up to N.
// Rows n >= N must be +0 on entry and stay so: the butterflies that start
// at such a row pair two of them (n + h > n >= N) and give +0 again, so
// they run unguarded, without a select per row.
__device__ __forceinline__ void haar_lane(float (&z)[FAST_NMAX], int N, bool forward)
{
	const float isqrt2 = 0.70710678118654752440f;
#pragma unroll
	for (int l = 0; l < 5; l++) {
		const int s = forward ? (2 << l) : (32 >> l);   // 2..32 / 32..2
		const int h = s >> 1;
		if (s <= N) {
#pragma unroll
			for (int n = 0; n < FAST_NMAX; n += s) {
				const float z0 = z[n] * isqrt2;
				const float z1 = z[n + h] * isqrt2;
				z[n] = z0 + z1;
				z[n + h] = z0 - z1;
			}
		}
	}
}

__device__ __forceinline__ int wave_sum_int(int v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
	return v;
}

// Forward transform of the N patches at anchors gidx[0..N) of `img`
// (one channel, anchor address = position) into zs[n*64 + lane].
// PAIRS: two patches per round (two independent sums in flight) or one.
template <int PX, bool PAIRS = true>
__device__ __forceinline__ void fast_forward(const float *__restrict__ img, const uint32_t *__restrict__ gidx,
                                             int N, int W, int WH, const float (&brow)[FAST_P],
                                             float *__restrict__ zs)
{
	constexpr int PT = FAST_P / (PX * PX);
	int n = 0;
	for (; PAIRS && n + 1 < N; n += 2) {
		const float *src0 = img + gidx[n], *src1 = img + gidx[n + 1];
		float acc0 = 0.f, acc1 = 0.f;
#pragma unroll
		for (int ht = 0; ht < PT; ht++)
#pragma unroll
			for (int hy = 0; hy < PX; hy++) {
				const float *r0 = src0 + ht * WH + hy * W, *r1 = src1 + ht * WH + hy * W;
#pragma unroll
				for (int hx = 0; hx < PX; hx++) {
					const float b = brow[(ht * PX + hy) * PX + hx];
					acc0 = __builtin_fmaf(r0[hx], b, acc0);
					acc1 = __builtin_fmaf(r1[hx], b, acc1);
				}
			}
		zs[n * FAST_P + threadIdx.x] = acc0;
		zs[(n + 1) * FAST_P + threadIdx.x] = acc1;
	}
	for (; n < N; n++) {
		const float *src = img + gidx[n];
		float acc = 0.f;
#pragma unroll
		for (int ht = 0; ht < PT; ht++)
#pragma unroll
			for (int hy = 0; hy < PX; hy++) {
				const float *r = src + ht * WH + hy * W;
#pragma unroll
				for (int hx = 0; hx < PX; hx++) acc = __builtin_fmaf(r[hx], brow[(ht * PX + hy) * PX + hx], acc);
			}
		zs[n * FAST_P + threadIdx.x] = acc;
	}
}

// Inverse transform of the coefficients zs[n*64 + k] and the weighted est
// rows: est[(row0+n)*rs + ro + pix] = (w*win[pix]) * sum_k z[k][n] Bi[k][pix]
// (rs = floats per est row, ro = offset of this channel in the row).
template <bool PAIRS = true>
__device__ __forceinline__ void fast_inverse(const float *__restrict__ zs, const float *__restrict__ binv,
                                             int N, float ww, uint32_t row0, float *__restrict__ est,
                                             int rs = FAST_P, int ro = 0)
{
	float bcol[FAST_P];
#pragma unroll
	for (int k = 0; k < FAST_P; k++) bcol[k] = binv[k * FAST_P + threadIdx.x];
	int n = 0;
	for (; PAIRS && n + 1 < N; n += 2) {
		const float4 *zr0 = reinterpret_cast<const float4 *>(zs + n * FAST_P);
		const float4 *zr1 = reinterpret_cast<const float4 *>(zs + (n + 1) * FAST_P);
		float acc0 = 0.f, acc1 = 0.f;
#pragma unroll
		for (int q = 0; q < FAST_P / 4; q++) {
			const float4 z0 = zr0[q], z1 = zr1[q];
			acc0 = __builtin_fmaf(z0.x, bcol[4 * q + 0], acc0);
			acc1 = __builtin_fmaf(z1.x, bcol[4 * q + 0], acc1);
			acc0 = __builtin_fmaf(z0.y, bcol[4 * q + 1], acc0);
			acc1 = __builtin_fmaf(z1.y, bcol[4 * q + 1], acc1);
			acc0 = __builtin_fmaf(z0.z, bcol[4 * q + 2], acc0);
			acc1 = __builtin_fmaf(z1.z, bcol[4 * q + 2], acc1);
			acc0 = __builtin_fmaf(z0.w, bcol[4 * q + 3], acc0);
			acc1 = __builtin_fmaf(z1.w, bcol[4 * q + 3], acc1);
		}
		est[(size_t)(row0 + n) * rs + ro + threadIdx.x] = ww * acc0;
		est[(size_t)(row0 + n + 1) * rs + ro + threadIdx.x] = ww * acc1;
	}
	for (; n < N; n++) {
		const float4 *zr = reinterpret_cast<const float4 *>(zs + n * FAST_P);
		float acc = 0.f;
#pragma unroll
		for (int q = 0; q < FAST_P / 4; q++) {
			const float4 z = zr[q];
			acc = __builtin_fmaf(z.x, bcol[4 * q + 0], acc);
			acc = __builtin_fmaf(z.y, bcol[4 * q + 1], acc);
			acc = __builtin_fmaf(z.z, bcol[4 * q + 2], acc);
			acc = __builtin_fmaf(z.w, bcol[4 * q + 3], acc);
		}
		est[(size_t)(row0 + n) * rs + ro + threadIdx.x] = ww * acc;
	}
}

// The inverse over the columns k set in kmask only, for groups whose other
// columns are zero in every row (step 1 after the hard threshold): a
// skipped term is fma(+0, Bi, acc) = acc bit for bit for every acc but -0,
// and a sum that starts at +0 reaches -0 only by an underflow below 2^-149,
// which sums of image data times basis entries do not come near, so the
// result equals fast_inverse's.  Lane k holds z[0..N) of column k on entry
// (registers); the compacted rows are written over zs.
__device__ __forceinline__ void fast_inverse_sparse(const float (&z)[FAST_NMAX], float *__restrict__ zs,
                                                    const float *__restrict__ binv, int N, float ww,
                                                    uint32_t row0, float *__restrict__ est, uint64_t kmask,
                                                    int rs = FAST_P, int ro = 0)
{
	const int lane = threadIdx.x;
	const int kc = __popcll(kmask);
	// compacted column index of this lane's column
	if ((kmask >> lane) & 1ull) {
		const int j = __popcll(kmask & ((1ull << lane) - 1ull));
#pragma unroll
		for (int n = 0; n < FAST_NMAX; n++) if (n < N) zs[n * FAST_P + j] = z[n];
	}
	// inverse basis rows of the kept columns, in column order
	float bsel[FAST_P];
	uint64_t m = kmask;
#pragma unroll
	for (int j = 0; j < FAST_P; j++) {
		bsel[j] = 0.f;
		if (j < kc) {
			const int k = __ffsll((unsigned long long)m) - 1;
			m &= m - 1ull;
			bsel[j] = binv[k * FAST_P + lane];
		}
	}
	__syncthreads();
	int n = 0;
	for (; n + 1 < N; n += 2) {
		const float *z0 = zs + n * FAST_P, *z1 = zs + (n + 1) * FAST_P;
		float acc0 = 0.f, acc1 = 0.f;
#pragma unroll
		for (int j = 0; j < FAST_P; j += 4) {
			if (j < kc) {
				const float4 a = *reinterpret_cast<const float4 *>(z0 + j);
				const float4 b = *reinterpret_cast<const float4 *>(z1 + j);
				acc0 = __builtin_fmaf(a.x, bsel[j], acc0);
				acc1 = __builtin_fmaf(b.x, bsel[j], acc1);
				if (j + 1 < kc) { acc0 = __builtin_fmaf(a.y, bsel[j + 1], acc0); acc1 = __builtin_fmaf(b.y, bsel[j + 1], acc1); }
				if (j + 2 < kc) { acc0 = __builtin_fmaf(a.z, bsel[j + 2], acc0); acc1 = __builtin_fmaf(b.z, bsel[j + 2], acc1); }
				if (j + 3 < kc) { acc0 = __builtin_fmaf(a.w, bsel[j + 3], acc0); acc1 = __builtin_fmaf(b.w, bsel[j + 3], acc1); }
			}
		}
		est[(size_t)(row0 + n) * rs + ro + lane] = ww * acc0;
		est[(size_t)(row0 + n + 1) * rs + ro + lane] = ww * acc1;
	}
	for (; n < N; n++) {
		const float *z0 = zs + n * FAST_P;
		float acc = 0.f;
#pragma unroll
		for (int j = 0; j < FAST_P; j += 4) {
			if (j < kc) {
				const float4 a = *reinterpret_cast<const float4 *>(z0 + j);
				acc = __builtin_fmaf(a.x, bsel[j], acc);
				if (j + 1 < kc) acc = __builtin_fmaf(a.y, bsel[j + 1], acc);
				if (j + 2 < kc) acc = __builtin_fmaf(a.z, bsel[j + 2], acc);
				if (j + 3 < kc) acc = __builtin_fmaf(a.w, bsel[j + 3], acc);
			}
		}
		est[(size_t)(row0 + n) * rs + ro + lane] = ww * acc;
	}
}

// Forward transform of every patch anchor of a one-channel plane, once:
// coef[pos*64 + k] = sum_p B[k][p] x[pos + off(p)] in the order of
// fast_forward (the same dot product whatever group the patch is in, so the
// filters' rows read it bit for bit).  One wave per run of anchors along x,
// lane = coefficient, basis row in registers, four anchors per round (their
// patches overlap: one load of a row segment serves all four).
template <int PX>
__global__ __launch_bounds__(64)
void coef_kernel(const float *__restrict__ img, int W, int H, int F, const float *__restrict__ basis,
                 float *__restrict__ coef)
{
	constexpr int PT = FAST_P / (PX * PX);
	constexpr int RUN = 16;   // anchors per wave
	const int nx = W - PX + 1, ny = H - PX + 1, nt = F - PT + 1;
	const int runs_x = (nx + RUN - 1) / RUN;
	const int64_t wid = blockIdx.x;
	if (wid >= (int64_t)runs_x * ny * nt) return;
	const int rx = (int)(wid % runs_x);
	const int64_t rest = wid / runs_x;
	const int y = (int)(rest % ny), t = (int)(rest / ny);
	const int x0 = rx * RUN, x1 = min(x0 + RUN, nx);
	const int lane = threadIdx.x;
	const int64_t WH = (int64_t)W * H;
	float brow[FAST_P];
#pragma unroll
	for (int p = 0; p < FAST_P; p++) brow[p] = basis[lane * FAST_P + p];
	const int64_t pos0 = (int64_t)t * WH + (int64_t)y * W;
	int x = x0;
	for (; x + 3 < x1; x += 4) {
		const float *s0 = img + pos0 + x;
		float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
		for (int ht = 0; ht < PT; ht++)
#pragma unroll
			for (int hy = 0; hy < PX; hy++) {
				const float *r = s0 + ht * WH + hy * W;
				float seg[PX + 3];
#pragma unroll
				for (int i = 0; i < PX + 3; i++) seg[i] = r[i];
#pragma unroll
				for (int hx = 0; hx < PX; hx++) {
					const float b = brow[(ht * PX + hy) * PX + hx];
#pragma unroll
					for (int a = 0; a < 4; a++) acc[a] = __builtin_fmaf(seg[hx + a], b, acc[a]);
				}
			}
#pragma unroll
		for (int a = 0; a < 4; a++) coef[(pos0 + x + a) * FAST_P + lane] = acc[a];
	}
	for (; x < x1; x++) {
		const float *s0 = img + pos0 + x;
		float acc = 0.f;
#pragma unroll
		for (int ht = 0; ht < PT; ht++)
#pragma unroll
			for (int hy = 0; hy < PX; hy++) {
				const float *r = s0 + ht * WH + hy * W;
#pragma unroll
				for (int hx = 0; hx < PX; hx++) acc = __builtin_fmaf(r[hx], brow[(ht * PX + hy) * PX + hx], acc);
			}
		coef[(pos0 + x) * FAST_P + lane] = acc;
	}
}
template __global__ void coef_kernel<4>(const float *, int, int, int, const float *, float *);
template __global__ void coef_kernel<8>(const float *, int, int, int, const float *, float *);

// Coefficient rows of the N patches at anchors gidx[0..N) from a
// coef_kernel table: lane k gets z[n] = coef[gidx[n]*64 + k] (rows n >= N zero).
// VIDX: the N row indices come from one vector load (lane n < N holds
// gidx[n]) and are read back lane by lane, so that the row loads do not
// wait for one scalar load each; rows n >= N load row gidx of lane n = 0
// (anchor 0) and are set to 0.  (The step-2 filter keeps the scalar loads:
// there the 32 indices held in scalar registers spill and the filter is
// slower.)
template <bool VIDX = false>
__device__ __forceinline__ void fast_rows(const float *__restrict__ coef, const uint32_t *__restrict__ gidx, int N,
                                          float (&z)[FAST_NMAX])
{
	static_assert(FAST_NMAX <= 64, "one index per lane");
	if (VIDX) {
		const uint32_t my_idx = (int)threadIdx.x < N ? gidx[threadIdx.x] : 0u;
#pragma unroll
		for (int n = 0; n < FAST_NMAX; n++) {
			const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)my_idx, n);
			const float v = coef[(size_t)r * FAST_P + threadIdx.x];
			z[n] = n < N ? v : 0.f;
		}
	} else {
#pragma unroll
		for (int n = 0; n < FAST_NMAX; n++) z[n] = n < N ? coef[(size_t)gidx[n] * FAST_P + threadIdx.x] : 0.f;
	}
}

// Whether all N coefficient rows are bitwise equal (every lane must call
// it).  Equal patches give equal rows: the same sums in the same order, and
// a +-0 term cannot change a sum that starts at +0 (it never becomes -0).
// So rows that differ rule out the equality early-out without reading the
// patches.
__device__ __forceinline__ bool fast_rows_equal(const float (&z)[FAST_NMAX], int N)
{
	const uint32_t z0 = __float_as_uint(z[0]);
	bool d = false;
#pragma unroll
	for (int n = 1; n < FAST_NMAX; n++) d |= n < N && __float_as_uint(z[n]) != z0;
	return !__any(d);
}

// Equality early-out (nldct.cpp:1986-1999, 2222-2235) with lane = sample:
// true when N > 1 and, in every one of the C channel planes (cstride
// apart), every patch equals patch 0; such a group is written raw (rows of
// C*64 samples) and flagged for the stale-weight fixup.
template <int PX>
__device__ __forceinline__ bool fast_early_out(const float *__restrict__ img, const uint32_t *__restrict__ gidx,
                                               int N, int W, int WH, int C, int64_t cstride, int g,
                                               uint32_t row0, float *__restrict__ est, uint8_t *__restrict__ early,
                                               float *__restrict__ gweight, uint32_t *__restrict__ n_early)
{
	if (N <= 1) return false;
	const int l = threadIdx.x;
	const int off = (l / (PX * PX)) * WH + ((l / PX) % PX) * W + (l % PX);
	int differs = 0;
	for (int c = 0; c < C; c++) {
		const float *im = img + c * cstride;
		const float v0 = im[gidx[0] + off];
		for (int n = 1; n < N; n++) differs |= (im[gidx[n] + off] != v0);
	}
	if (__any(differs)) return false;
	for (int c = 0; c < C; c++)
		for (int n = 0; n < N; n++)
			est[((size_t)(row0 + n) * C + c) * FAST_P + l] = img[c * cstride + gidx[n] + off];
	if (l == 0) { early[g] = 1; gweight[g] = 0.f; atomicAdd(n_early, 1u); }
	return true;
}

// Multichannel groups are filtered one channel per launch (c = 0..C-1, on
// channel planes cstride apart).  The group weight counts every channel
// (nldct.cpp:1820-1936, 2038-2168), so gacc[g] carries the running count or
// sum from launch to launch; launches before the last store the inverse
// transform unweighted and the last one scales them by its weight: the same
// products (w * window) * value as one pass over all channels.
__device__ __forceinline__ void fast_scale_channels(int N, int C, float ww, uint32_t row0, float *__restrict__ est)
{
	for (int n = 0; n < N; n++)
		for (int c = 0; c < C - 1; c++) {
			float *e = est + ((size_t)(row0 + n) * C + c) * FAST_P + threadIdx.x;
			*e = ww * *e;
		}
}

template <int PX>
__global__ __launch_bounds__(64)
void filter1_fast_kernel(const float *__restrict__ img, const uint32_t *__restrict__ idx,
                         const uint32_t *__restrict__ nsim, const uint32_t *__restrict__ base,
                         const uint32_t *__restrict__ gtile, StepConst sc, int g0, int Gc,
                         const float *__restrict__ basis, const float *__restrict__ binv,
                         const float *__restrict__ win, float *__restrict__ est,
                         float *__restrict__ entry_w, uint32_t *__restrict__ etile,
                         uint8_t *__restrict__ early, float *__restrict__ gweight,
                         uint32_t *__restrict__ n_early, int C, int c, int64_t cstride,
                         float *__restrict__ gacc, const float *__restrict__ coef)
{
	__shared__ __align__(16) float zs[FAST_NMAX * FAST_P];
	const int gl = blockIdx.x;
	if (gl >= Gc) return;
	const int g = g0 + gl;
	const int lane = threadIdx.x;
	const int N = (int)nsim[g];
	const uint32_t row0 = base[gl];
	const uint32_t *gidx = idx + (size_t)g * sc.npmax;
	float z[FAST_NMAX];
	if (coef) fast_rows<true>(coef + c * cstride * FAST_P, gidx, N, z);
	if (c == 0) {
		if (lane < N) etile[row0 + lane] = gtile[g];
		if ((!coef || fast_rows_equal(z, N)) &&
		    fast_early_out<PX>(img, gidx, N, sc.W, sc.WH, C, cstride, g, row0, est, early, gweight, n_early))
			return;
	} else if (early[g])
		return;

	if (!coef) {
		{
			float brow[FAST_P];
#pragma unroll
			for (int p = 0; p < FAST_P; p++) brow[p] = basis[lane * FAST_P + p];
			fast_forward<PX>(img + c * cstride, gidx, N, sc.W, sc.WH, brow, zs);
		}
		__syncthreads();
#pragma unroll
		for (int n = 0; n < FAST_NMAX; n++) z[n] = n < N ? zs[n * FAST_P + lane] : 0.f;
	}
	haar_lane(z, N, true);
	// rows n >= N are +0: 0 > (beta sigma)^2 fails, so they count nothing
	// and stay +0 without a guard
	int cnt = 0;
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) {
		if (z[n] * z[n] > sc.beta_sigma2 || (lane == 0 && n == 0)) cnt++;
		else z[n] = 0.f;
	}
	haar_lane(z, N, false);
	// columns with a coefficient left in any row (the others are +0 throughout)
	bool nzc = false;
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) nzc |= z[n] != 0.f;
	const uint64_t kmask = __ballot(nzc);
	const int nnz_i = wave_sum_int(cnt) + (c ? (int)gacc[g] : 0);
	const float nnz = (float)nnz_i;
	const float w = 1 / sc.sigma2 / nnz;
	const bool last = c == C - 1;
	if (lane == 0) {
		if (!last) gacc[g] = (float)nnz_i;
		else gweight[g] = w;
		early[g] = 0;
	}
	if (last && lane < N) entry_w[row0 + lane] = w;
	// launches before the last store unweighted rows
	const float ww = last ? w * win[lane] : 1.f;
	fast_inverse_sparse(z, zs, binv, N, ww, row0, est, kmask, C * FAST_P, c * FAST_P);
	if (last && C > 1) fast_scale_channels(N, C, w * win[lane], row0, est);
}

template <int PX>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
void filter2_fast_kernel(const float *__restrict__ noisy, const float *__restrict__ basic,
                         const uint32_t *__restrict__ idx, const uint32_t *__restrict__ nsim,
                         const uint32_t *__restrict__ base, const uint32_t *__restrict__ gtile,
                         StepConst sc, int g0, int Gc,
                         const float *__restrict__ basis, const float *__restrict__ binv,
                         const float *__restrict__ win, float *__restrict__ est,
                         float *__restrict__ entry_w, uint32_t *__restrict__ etile,
                         uint8_t *__restrict__ early, float *__restrict__ gweight,
                         uint32_t *__restrict__ n_early, int C, int c, int64_t cstride,
                         float *__restrict__ gacc, const float *__restrict__ coef_n,
                         const float *__restrict__ coef_b)
{
	__shared__ __align__(16) float zs[FAST_NMAX * FAST_P];
	const int gl = blockIdx.x;
	if (gl >= Gc) return;
	const int g = g0 + gl;
	const int lane = threadIdx.x;
	const int N = (int)nsim[g];
	const uint32_t row0 = base[gl];
	const uint32_t *gidx = idx + (size_t)g * sc.npmax;
	float z[FAST_NMAX], y[FAST_NMAX];
	if (coef_n) fast_rows(coef_n + c * cstride * FAST_P, gidx, N, z);
	if (c == 0) {
		if (lane < N) etile[row0 + lane] = gtile[g];
		if ((!coef_n || fast_rows_equal(z, N)) &&
		    fast_early_out<PX>(noisy, gidx, N, sc.W, sc.WH, C, cstride, g, row0, est, early, gweight, n_early))
			return;
	} else if (early[g])
		return;

	if (coef_n)
		fast_rows(coef_b + c * cstride * FAST_P, gidx, N, y);
	else {
		{
			float brow[FAST_P];
#pragma unroll
			for (int p = 0; p < FAST_P; p++) brow[p] = basis[lane * FAST_P + p];
			fast_forward<PX>(basic + c * cstride, gidx, N, sc.W, sc.WH, brow, zs);
			__syncthreads();
#pragma unroll
			for (int n = 0; n < FAST_NMAX; n++) y[n] = n < N ? zs[n * FAST_P + lane] : 0.f;
			__syncthreads();
			fast_forward<PX>(noisy + c * cstride, gidx, N, sc.W, sc.WH, brow, zs);
		}
		__syncthreads();
#pragma unroll
		for (int n = 0; n < FAST_NMAX; n++) z[n] = n < N ? zs[n * FAST_P + lane] : 0.f;
	}
	haar_lane(z, N, true);
	haar_lane(y, N, true);
	// Wiener shrinkage; y[n] becomes the energy term w^2
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) {
		if (n < N) {
			const float w = y[n] * y[n] / (y[n] * y[n] + sc.beta_sigma2);
			z[n] = z[n] * w;
			y[n] = w * w;
		}
	}
	// sum of w^2 in the reference's order, channel, coefficient, then patch:
	// lane k folds its N terms into the running sum handed on by lane k-1.
	// Rows n >= N hold +0 and the sum is >= +0 (or NaN), so adding them
	// leaves it unchanged bit for bit: no select per term.
	float v = c ? gacc[g] : 0.f;
	for (int k = 0; k < FAST_P; k++) {
		float t = v;
#pragma unroll
		for (int n = 0; n < FAST_NMAX; n++) t = t + y[n];
		v = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, t), k));
	}
	haar_lane(z, N, false);
	__syncthreads();
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) if (n < N) zs[n * FAST_P + lane] = z[n];
	if (c < C - 1) {
		if (lane == 0) { gacc[g] = v; early[g] = 0; }
		__syncthreads();
		fast_inverse(zs, binv, N, 1.f, row0, est, C * FAST_P, c * FAST_P);
		return;
	}
	const float w = (float)(1. / (double)sc.sigma2 / (double)v);
	if (lane == 0) { gweight[g] = w; early[g] = 0; }
	if (lane < N) entry_w[row0 + lane] = w;
	__syncthreads();
	fast_inverse(zs, binv, N, w * win[lane], row0, est, C * FAST_P, c * FAST_P);
	if (C > 1) fast_scale_channels(N, C, w * win[lane], row0, est);
}


// ---------------------------------------------------------------------------
// Learned single pass (no counterpart in the reference).  Groups and
// neighbours as step 1 (noisy image); per group row n the network maps the
// Haar-domain coefficients z[.][n] (features: z/sigma, log(sigma/20) and the
// row statistic log1p(sum_k z[k][n]^2 / (64 sigma^2))) to 64 gains that replace the
// step-2 Wiener gains; aggregation weight 1/sigma^2/sum g^2 as in step 2
// (nldct.cpp:2167).  Groups whose gain variance exceeds `gvar_thr` are
// flagged (gflag) for the plain two-step fallback.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Patch transforms of the learned pass on the matrix cores.  The learned
// pass has no reference order to keep, so its two dense 64 x 64 transforms
// run as 32 x 64 x 64 products on v_mfma_f32_32x32x8_f16 with each fp32
// operand split into fp16 hi + lo (x = hi + lo to 2^-22 relative) and three
// products per step (hi*hi + hi*lo + lo*hi), accumulated in fp32.
// Split tables (per plan, 64 x 64 each, fp16 bits): [0] basis hi [k][p],
// [1] basis lo, [2] inverse basis transposed hi [p][k], [3] its lo.
// Lane layout as the network's (mlp_dev.hpp): lane l, h = l / 32, j = l % 32.
// ---------------------------------------------------------------------------

typedef _Float16 tr_f16x4 __attribute__((ext_vector_type(4)));
constexpr int TR_ZS = 68;   // padded LDS row stride of the inverse's operand

__device__ __forceinline__ void tr_split(float4 x, tr_f16x4 &hi, tr_f16x4 &lo)
{
	hi[0] = (_Float16)x.x; lo[0] = (_Float16)(x.x - (float)hi[0]);
	hi[1] = (_Float16)x.y; lo[1] = (_Float16)(x.y - (float)hi[1]);
	hi[2] = (_Float16)x.z; lo[2] = (_Float16)(x.z - (float)hi[2]);
	hi[3] = (_Float16)x.w; lo[3] = (_Float16)(x.w - (float)hi[3]);
}

// Power of two that brings the wave's largest |value| to [2^13, 2^14): the
// scaled operands stay inside fp16 whatever the input range, and scaling by
// a power of two is exact both ways.
__device__ __forceinline__ int tr_scale_exp(const float4 (&x)[8])
{
	float m = 0.f;
#pragma unroll
	for (int s = 0; s < 8; s++)
		m = fmaxf(m, fmaxf(fmaxf(fabsf(x[s].x), fabsf(x[s].y)), fmaxf(fabsf(x[s].z), fabsf(x[s].w))));
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
	int e;
	(void)frexpf(m, &e);
	return 14 - e;
}

__device__ __forceinline__ void tr_split_scaled(float4 x, int sh, tr_f16x4 &hi, tr_f16x4 &lo)
{
	tr_split(make_float4(ldexpf(x.x, sh), ldexpf(x.y, sh), ldexpf(x.z, sh), ldexpf(x.w, sh)), hi, lo);
}

__device__ __forceinline__ tr_f16x4 tr_load4(const uint16_t *p)
{
	tr_f16x4 r;
	__builtin_memcpy(&r, p, sizeof r);
	return r;
}

// zs[n*64 + k] = sum_p X[n][p] B[k][p] for the N patches at gidx (rows
// n >= N zero).  One wave; every lane must call it.
template <int PX>
__device__ __forceinline__ void mfma_forward(const float *__restrict__ img, const uint32_t *__restrict__ gidx,
                                             int N, int W, int WH, const uint16_t *__restrict__ bsp,
                                             float *__restrict__ zs)
{
	const int l = threadIdx.x, h = l >> 5, j = l & 31;
	// A operand: patch j, samples p = 8s + 4h .. +3 (one row of a 4 x 4 x 4
	// patch, half a row of an 8 x 8 x 1 patch)
	float4 x[8];
#pragma unroll
	for (int s = 0; s < 8; s++) {
		x[s] = make_float4(0.f, 0.f, 0.f, 0.f);
		if (j < N) {
			const int off = PX == 4 ? (s >> 1) * WH + ((2 * s + h) & 3) * W : s * W + 4 * h;
			__builtin_memcpy(&x[s], img + gidx[j] + off, sizeof x[s]);
		}
	}
	const int sh = tr_scale_exp(x);
	tr_f16x4 ah[8], al[8];
#pragma unroll
	for (int s = 0; s < 8; s++) tr_split_scaled(x[s], sh, ah[s], al[s]);
#pragma unroll
	for (int t = 0; t < 2; t++) {
		mlp_f32x16 c = {};
		const uint16_t *bh = bsp + (32 * t + j) * FAST_P + 4 * h, *bl = bh + FAST_P * FAST_P;
#pragma unroll
		for (int s = 0; s < 8; s++) {
			const tr_f16x4 b_hi = tr_load4(bh + 8 * s), b_lo = tr_load4(bl + 8 * s);
			c = __builtin_amdgcn_mfma_f32_32x32x8f16(ah[s], b_hi, c, 0, 0, 0);
			c = __builtin_amdgcn_mfma_f32_32x32x8f16(ah[s], b_lo, c, 0, 0, 0);
			c = __builtin_amdgcn_mfma_f32_32x32x8f16(al[s], b_hi, c, 0, 0, 0);
		}
#pragma unroll
		for (int r = 0; r < 16; r++) zs[(8 * (r >> 2) + 4 * h + (r & 3)) * FAST_P + 32 * t + j] = ldexpf(c[r], -sh);
	}
}

// est[(row0+n)*rs + ro + p] = (w * win[p]) * sum_k Z[n][k] Bi[k][p] for n < N,
// Z from LDS rows of stride TR_ZS.  One wave; every lane must call it.
__device__ __forceinline__ void mfma_inverse(const float *__restrict__ zp, const uint16_t *__restrict__ bsp,
                                             int N, float w, const float *__restrict__ win, uint32_t row0,
                                             float *__restrict__ est, int rs, int ro)
{
	const int l = threadIdx.x, h = l >> 5, j = l & 31;
	float4 z[8];
#pragma unroll
	for (int s = 0; s < 8; s++) z[s] = *reinterpret_cast<const float4 *>(zp + j * TR_ZS + 8 * s + 4 * h);
	const int sh = tr_scale_exp(z);
	tr_f16x4 ah[8], al[8];
#pragma unroll
	for (int s = 0; s < 8; s++) tr_split_scaled(z[s], sh, ah[s], al[s]);
#pragma unroll
	for (int t = 0; t < 2; t++) {
		mlp_f32x16 c = {};
		const uint16_t *bh = bsp + 2 * FAST_P * FAST_P + (32 * t + j) * FAST_P + 4 * h;
		const uint16_t *bl = bh + FAST_P * FAST_P;
#pragma unroll
		for (int s = 0; s < 8; s++) {
			const tr_f16x4 b_hi = tr_load4(bh + 8 * s), b_lo = tr_load4(bl + 8 * s);
			c = __builtin_amdgcn_mfma_f32_32x32x8f16(ah[s], b_hi, c, 0, 0, 0);
			c = __builtin_amdgcn_mfma_f32_32x32x8f16(ah[s], b_lo, c, 0, 0, 0);
			c = __builtin_amdgcn_mfma_f32_32x32x8f16(al[s], b_hi, c, 0, 0, 0);
		}
		const float ww = w * win[32 * t + j];
#pragma unroll
		for (int r = 0; r < 16; r++) {
			const int n = 8 * (r >> 2) + 4 * h + (r & 3);
			if (n < N) est[(size_t)(row0 + n) * rs + ro + 32 * t + j] = ww * ldexpf(c[r], -sh);
		}
	}
}

// Haar-domain coefficients of a group into the lane registers z[n] (lane k),
// and each row's distance statistic: log1p of its energy over 64 sigma^2
// (for a detail row, the squared differences between the patch sub-means
// it compares; for the first row, the group mean's energy) into rstat[n].
template <int PX>
__device__ __forceinline__ void learned_coeffs(const float *__restrict__ img, const uint32_t *__restrict__ gidx,
                                               int N, const StepConst &sc, const uint16_t *__restrict__ bsp,
                                               float sigma, float *__restrict__ zs, float (&z)[FAST_NMAX],
                                               float *__restrict__ rstat)
{
	const int lane = threadIdx.x;
	mfma_forward<PX>(img, gidx, N, sc.W, sc.WH, bsp, zs);
	__syncthreads();
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) z[n] = n < N ? zs[n * FAST_P + lane] : 0.f;
	haar_lane(z, N, true);
	__syncthreads();
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) zs[n * FAST_P + lane] = z[n] * z[n];
	__syncthreads();
	if (lane < FAST_NMAX) {
		float e = 0.f;
		for (int k = 0; k < FAST_P; k++) e += zs[lane * FAST_P + k];
		rstat[lane] = lane < N ? log1pf(e / (64.f * sigma * sigma)) : 0.f;
	}
	__syncthreads();
}

// bf16 feature rows [32][MLP_XS] from the lane registers (rows >= N zero)
__device__ __forceinline__ void learned_features_lds(const float (&z)[FAST_NMAX], int N, float sigma,
                                                     const float *__restrict__ rstat, uint16_t *__restrict__ xb)
{
	const int lane = threadIdx.x;
	const float is = 1.f / sigma;
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) xb[n * MLP_XS + lane] = f32_to_bf16(n < N ? z[n] * is : 0.f);
	if (lane < 32) {
		const bool live = lane < N;
		xb[lane * MLP_XS + 64] = f32_to_bf16(live ? logf(sigma / 20.f) : 0.f);
		xb[lane * MLP_XS + 65] = f32_to_bf16(rstat[lane]);
		for (int k = 66; k < MLP_XS; k++) xb[lane * MLP_XS + k] = 0;
	}
}

template <int PX>
__global__ __launch_bounds__(64)
void filter_learned_kernel(const float *__restrict__ img, const uint32_t *__restrict__ idx,
                           const uint32_t *__restrict__ nsim, const uint32_t *__restrict__ base,
                           const uint32_t *__restrict__ gtile, StepConst sc, int g0, int Gc,
                           const uint16_t *__restrict__ bsp, const float *__restrict__ win, MlpInfer net,
                           float sigma, float gvar_thr,
                           float *__restrict__ est, float *__restrict__ entry_w, uint32_t *__restrict__ etile,
                           uint8_t *__restrict__ early, float *__restrict__ gweight,
                           uint8_t *__restrict__ gflag, uint32_t *__restrict__ n_flag, int C, int c)
{
	// coefficients, then the bf16 feature rows (read by the network's first
	// layer only), then gains (row stride 64), then the filtered rows (TR_ZS)
	__shared__ __align__(16) float zs[FAST_NMAX * TR_ZS];
	uint16_t *xb = reinterpret_cast<uint16_t *>(zs);
	static_assert(32 * MLP_XS * 2 <= FAST_NMAX * TR_ZS * 4, "feature rows fit the coefficient buffer");
	__shared__ __align__(16) uint16_t hb[32 * MLP_HS];
	const int gl = blockIdx.x;
	if (gl >= Gc) return;
	const int g = g0 + gl;
	const int lane = threadIdx.x;
	const int N = (int)nsim[g];
	const uint32_t row0 = base[gl];
	const uint32_t *gidx = idx + (size_t)g * sc.npmax;
	if (c == 0 && lane < N) etile[row0 + lane] = gtile[g];

	__shared__ float rstat[FAST_NMAX];
	float z[FAST_NMAX];
	learned_coeffs<PX>(img, gidx, N, sc, bsp, sigma, zs, z, rstat);
	learned_features_lds(z, N, sigma, rstat, xb);
	__syncthreads();
	mlp_rows32(net, xb, hb, zs, FAST_P);
	float s2 = 0.f, s1 = 0.f;
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) {
		if (n < N) {
			const float gn = zs[n * FAST_P + lane];
			z[n] = z[n] * gn;
			s1 += gn;
			s2 += gn * gn;
		}
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		s1 += __shfl_xor(s1, o, 64);
		s2 += __shfl_xor(s2, o, 64);
	}
	haar_lane(z, N, false);
	__syncthreads();
	// filtered rows for the inverse (rows n >= N are zero)
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) zs[n * TR_ZS + lane] = z[n];
	const float cnt = (float)(64 * N);
	const float gvar = s2 / cnt - (s1 / cnt) * (s1 / cnt);
	// one weight for all channels of a group, as the reference aggregates
	// colour groups (nldct.cpp:2293-2299), here from the first channel's
	// gains; a group is flagged when any channel's gain variance exceeds the
	// threshold
	const float w = c == 0 ? (float)(1. / ((double)sigma * sigma) / (double)s2) : gweight[g];
	const uint8_t f = gvar > gvar_thr ? 1 : 0;
	if (lane == 0) {
		if (c == 0) {
			gweight[g] = w;
			early[g] = 0;
			gflag[g] = f;
			if (f) atomicAdd(n_flag, 1u);
		} else if (f && !gflag[g]) {
			gflag[g] = 1;
			atomicAdd(n_flag, 1u);
		}
	}
	if (c == 0 && lane < N) entry_w[row0 + lane] = w;
	__syncthreads();
	mfma_inverse(zs, bsp, N, w, win, row0, est, C * FAST_P, c * FAST_P);
}

// Split tables of the matrix-core transforms from the fp32 bases
// (basis[k*64 + p], binv[k*64 + p]); 4096 threads.
__global__ void split_basis_kernel(const float *__restrict__ basis, const float *__restrict__ binv,
                                   uint16_t *__restrict__ bsp)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= FAST_P * FAST_P) return;
	const int k = i / FAST_P, p = i - k * FAST_P;
	const float b = basis[i], bi = binv[i];
	const _Float16 bh = (_Float16)b, bl = (_Float16)(b - (float)bh);
	const _Float16 ih = (_Float16)bi, il = (_Float16)(bi - (float)ih);
	__builtin_memcpy(bsp + i, &bh, 2);
	__builtin_memcpy(bsp + FAST_P * FAST_P + i, &bl, 2);
	__builtin_memcpy(bsp + 2 * FAST_P * FAST_P + p * FAST_P + k, &ih, 2);
	__builtin_memcpy(bsp + 3 * FAST_P * FAST_P + p * FAST_P + k, &il, 2);
}

// Training pairs: for every `stride`-th group, the features of its rows
// (noisy image, as filter_learned_kernel) and the step-2 Wiener gains that
// the basic estimate gives at the same neighbours:
// y^2 / (y^2 + (beta sigma)^2) (nldct.cpp:2131).  Rows n >= N are flagged 0.
template <int PX>
__global__ __launch_bounds__(64)
void learned_pairs_kernel(const float *__restrict__ noisy, const float *__restrict__ basic,
                          const uint32_t *__restrict__ idx, const uint32_t *__restrict__ nsim, StepConst sc,
                          int G, int stride, const uint16_t *__restrict__ bsp, float sigma, float beta_sigma2,
                          float *__restrict__ xout, float *__restrict__ tout, uint8_t *__restrict__ rflag)
{
	__shared__ __align__(16) float zs[FAST_NMAX * FAST_P];
	const int slot = blockIdx.x;
	const int g = slot * stride;
	if (g >= G) return;
	const int lane = threadIdx.x;
	const int N = (int)nsim[g];
	const uint32_t *gidx = idx + (size_t)g * sc.npmax;
	__shared__ float rstat[FAST_NMAX];
	float z[FAST_NMAX];
	learned_coeffs<PX>(noisy, gidx, N, sc, bsp, sigma, zs, z, rstat);
	const float is = 1.f / sigma, ls = logf(sigma / 20.f);
	const size_t r0 = (size_t)slot * FAST_NMAX;
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) xout[(r0 + n) * MLP_IN + lane] = n < N ? z[n] * is : 0.f;
	if (lane < 32) {
		xout[(r0 + lane) * MLP_IN + 64] = lane < N ? ls : 0.f;
		xout[(r0 + lane) * MLP_IN + 65] = rstat[lane];
		rflag[r0 + lane] = lane < N ? 1 : 0;
	}
	__syncthreads();
	float y[FAST_NMAX];
	learned_coeffs<PX>(basic, gidx, N, sc, bsp, sigma, zs, y, rstat);
#pragma unroll
	for (int n = 0; n < FAST_NMAX; n++) {
		const float wgt = y[n] * y[n] / (y[n] * y[n] + beta_sigma2);
		tout[(r0 + n) * MLP_OUT + lane] = n < N ? wgt : 0.f;
	}
}

// Pixels covered by the patches of flagged groups (plain two-step fallback
// of the learned pass); one block per group.
__global__ void fallback_mark_kernel(const uint32_t *__restrict__ idx, const uint32_t *__restrict__ nsim,
                                     const uint8_t *__restrict__ gflag, StepConst sc, int G,
                                     uint8_t *__restrict__ mask)
{
	const int g = blockIdx.x;
	if (g >= G || !gflag[g]) return;
	const int N = (int)nsim[g];
	const int sPx = sc.sPx, SL = sPx * sPx;
	for (int n = 0; n < N; n++) {
		const uint32_t pos = idx[(size_t)g * sc.npmax + n];
		for (int p = threadIdx.x; p < sc.P; p += blockDim.x) {
			const int ht = p / SL, r = p - ht * SL, hy = r / sPx, hx = r - hy * sPx;
			mask[pos + (uint32_t)ht * sc.WH + (uint32_t)hy * sc.W + hx] = 1;
		}
	}
}

// dst[c][t][y][x] = src[t][c][y][x]: the channel planes of a C-channel video
// as C one-channel videos
__global__ void planes_kernel(const float *__restrict__ src, int C, int WH, int64_t n, float *__restrict__ dst)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int64_t t = i / ((int64_t)WH * C), r = i - t * (int64_t)WH * C;
	const int c = (int)(r / WH), q = (int)(r - (int64_t)c * WH);
	dst[((int64_t)c * (n / C)) + t * WH + q] = src[i];
}

// out = plain where the mask of the sample's pixel is set
// (n samples of a [F][C][H][W] video; mask per pixel, W*H per frame)
__global__ void fallback_blend_kernel(const uint8_t *__restrict__ mask, const float *__restrict__ plain,
                                      int64_t n, int WH, int C, float *__restrict__ out)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int64_t t = i / ((int64_t)WH * C);
	const int64_t r = i % WH;
	if (mask[t * WH + r]) out[i] = plain[i];
}

#define LEARNED_ARGS const float *, const uint32_t *, const uint32_t *, const uint32_t *, const uint32_t *, \
	StepConst, int, int, const uint16_t *, const float *, MlpInfer, float, float, float *, \
	float *, uint32_t *, uint8_t *, float *, uint8_t *, uint32_t *, int, int
#define PAIRS_ARGS const float *, const float *, const uint32_t *, const uint32_t *, StepConst, int, int, \
	const uint16_t *, float, float, float *, float *, uint8_t *
template __global__ void filter_learned_kernel<4>(LEARNED_ARGS);
template __global__ void filter_learned_kernel<8>(LEARNED_ARGS);
template __global__ void learned_pairs_kernel<4>(PAIRS_ARGS);
template __global__ void learned_pairs_kernel<8>(PAIRS_ARGS);


#define FILTER1_ARGS const float *, const uint32_t *, const uint32_t *, const uint32_t *, \
	const uint32_t *, StepConst, int, int, const float *, const float *, const float *, float *, \
	float *, uint32_t *, uint8_t *, float *, uint32_t *, int, int, int64_t, float *, const float *
#define FILTER2_ARGS const float *, const float *, const uint32_t *, const uint32_t *, \
	const uint32_t *, const uint32_t *, StepConst, int, int, const float *, const float *, \
	const float *, float *, float *, uint32_t *, uint8_t *, float *, uint32_t *, int, int, int64_t, float *, \
	const float *, const float *
template __global__ void filter1_fast_kernel<4>(FILTER1_ARGS);
template __global__ void filter1_fast_kernel<8>(FILTER1_ARGS);
template __global__ void filter2_fast_kernel<4>(FILTER2_ARGS);
template __global__ void filter2_fast_kernel<8>(FILTER2_ARGS);

// ---------------------------------------------------------------------------
// Aggregation in the reference's order.
//
// computeAggregationStep{1,2} (nldct.cpp:2259-2425) adds every neighbour
// patch of every group into float numerator and weight images, group after
// group in processing order and neighbour after neighbour, so each pixel's
// float sums are folded in the order of the rows e of est.  The kernels below
// reproduce exactly that fold: for a block of pixels they gather the rows
// whose patches touch it (from per-position lists), sort them by e and add
// them one by one into float accumulators.  Only rows of the pixel's owner
// tile count (the tile whose interior holds it; subBuildTight copies only
// interiors, lib_videot.hpp:1128-1160).
// ---------------------------------------------------------------------------

// keys[e] = patch position of row e, for the groups
// [g0, g0+Gc) of a chunk; base = exclusive scan of nsim over the chunk.
__global__ void emit_entries_kernel(const uint32_t *__restrict__ idx, const uint32_t *__restrict__ nsim,
                                    const uint32_t *__restrict__ base, int npmax, int g0, int Gc,
                                    uint32_t *__restrict__ keys)
{
	// four entries per thread and round, their loads issued together
	const int64_t tot = (int64_t)Gc * npmax, gs = (int64_t)gridDim.x * blockDim.x;
	for (int64_t o0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o0 < tot; o0 += 4 * gs) {
		uint32_t ns[4], bs[4], ix[4];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int64_t o = o0 + u * gs;
			const int gl = o < tot ? (int)(o / npmax) : 0, n = (int)(o - (int64_t)gl * npmax);
			ns[u] = o < tot ? nsim[g0 + gl] : 0u;
			bs[u] = o < tot ? base[gl] : 0u;
			ix[u] = o < tot ? idx[(size_t)(g0 + gl) * npmax + n] : 0u;
		}
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int64_t o = o0 + u * gs;
			if (o >= tot) break;
			const int n = (int)(o % npmax);
			if (n < (int)ns[u]) keys[bs[u] + n] = ix[u];
		}
	}
}

// erec[e] = (position, tile, weight bits, row | frame << 16 of the
// position): the row metadata the chunked and binned aggregations read,
// one 16-byte record per row (the binned form takes row and frame from it
// instead of dividing the position once per bin).
__global__ void pack_rows_kernel(const uint32_t *__restrict__ epos, const uint32_t *__restrict__ etile,
                                 const float *__restrict__ entry_w, int64_t M, uint32_t W, uint32_t WH,
                                 uint4 *__restrict__ erec)
{
	const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (e < M) {
		const uint32_t q = epos[e];
		const uint32_t qt = q / WH, qy = (q - qt * WH) / W;
		erec[e] = make_uint4(q, etile[e], __float_as_uint(entry_w[e]), qy | (qt << 16));
	}
}

#define AGG_CAPB 1024   // rows of one band held in LDS
#define AGG_MAXPOS 512  // patch positions overlapping an 8x8 pixel block
#define AGG_STAGE 2048  // floats of staged row slices

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << (threadIdx.x & 63)) - 1ull; }

// Stable LSD radix sort of K keys in LDS by one wave, 8-bit digits.
// Returns the buffer that holds the result.
__device__ uint32_t *wave_radix_sort(uint32_t *a, uint32_t *b, int K, int nbits, uint32_t *hist)
{
	const int lane = threadIdx.x & 63;
	for (int shift = 0; shift < nbits; shift += 8) {
		for (int d = lane; d < 256; d += 64) hist[d] = 0;
		__syncthreads();
		for (int k = lane; k < K; k += 64) atomicAdd(&hist[(a[k] >> shift) & 255u], 1u);
		__syncthreads();
		{
			const uint32_t c0 = hist[4 * lane], c1 = hist[4 * lane + 1], c2 = hist[4 * lane + 2], c3 = hist[4 * lane + 3];
			const int local = (int)(c0 + c1 + c2 + c3);
			const int incl = wave_incl_scan(local);
			const uint32_t ex = (uint32_t)(incl - local);
			__syncthreads();
			hist[4 * lane] = ex;
			hist[4 * lane + 1] = ex + c0;
			hist[4 * lane + 2] = ex + c0 + c1;
			hist[4 * lane + 3] = ex + c0 + c1 + c2;
		}
		__syncthreads();
		for (int base = 0; base < K; base += 64) {
			const int k = base + lane;
			const bool valid = k < K;
			const uint32_t key = valid ? a[k] : 0u;
			const uint32_t d = (key >> shift) & 255u;
			uint64_t m = __ballot(valid);
#pragma unroll
			for (int bit = 0; bit < 8; bit++) {
				const bool bset = (d >> bit) & 1u;
				const uint64_t bb = __ballot(valid && bset);
				m &= bset ? bb : ~bb;
			}
			const int rank = __popcll(m & lanemask_lt());
			const int cnt = __popcll(m);
			const uint32_t pos = hist[d] + (uint32_t)rank;
			__syncthreads();
			if (valid) b[pos] = key;
			if (valid && rank == cnt - 1) hist[d] = pos + 1;
			__syncthreads();
		}
		uint32_t *t = a; a = b; b = t;
	}
	return a;
}

__device__ __forceinline__ int owner_tile(int x, int y, const StepConst &sc)
{
	return (y / sc.hTmp) * sc.nW + (x / sc.wTmp);
}

// First index in [lo, hi) of the sorted list v with v[i] >= key.
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *__restrict__ v, uint32_t lo,
                                                    uint32_t hi, uint32_t key)
{
	while (lo < hi) {
		const uint32_t mid = (lo + hi) >> 1;
		if (v[mid] < key) lo = mid + 1; else hi = mid;
	}
	return lo;
}

// Ordered aggregation of one 8 x 8 pixel block of one frame by one wave,
// lane = pixel.  The rows touching the block are those of the patch
// positions whose patch overlaps it; each position's rows are a sorted run of
// svals.  Rows are taken band by band (bands = consecutive row ranges, one
// per reference frame of a tile, so a band's rows are few), each band's rows
// are sorted in LDS, their slices at this frame staged in LDS, then added one
// by one into the lane's float accumulators: the order
// computeAggregationStep{1,2} adds them in (nldct.cpp:2284-2300,
// 2376-2393).  num/wsum hold the running sums and are updated in place.
template <int CC>
__global__ __launch_bounds__(64)
void agg_exact_kernel(const uint32_t *__restrict__ kstart, const uint32_t *__restrict__ svals,
                      const uint32_t *__restrict__ epos, const uint32_t *__restrict__ etile,
                      const float *__restrict__ entry_w, const float *__restrict__ est,
                      const float *__restrict__ win, StepConst sc, const uint32_t *__restrict__ bands,
                      int nbands, int tframe0, float *__restrict__ num, float *__restrict__ wsum,
                      uint32_t *__restrict__ nsplit)
{
	__shared__ uint32_t keysA[AGG_CAPB];
	__shared__ uint32_t keysB[AGG_CAPB];
	__shared__ uint32_t hist[256];
	__shared__ uint32_t pcur[AGG_MAXPOS], pend[AGG_MAXPOS], phead[AGG_MAXPOS], pub[AGG_MAXPOS];
	__shared__ uint32_t poff[AGG_MAXPOS + 1];
	__shared__ float stage[AGG_STAGE];
	__shared__ float swin[256];
	// per staged row: (qx - x0 + 16) | (qy - y0 + 16) << 8 | dt << 16 | tile << 24
	__shared__ uint32_t m_pos[64];
	__shared__ float m_w[64];
	__shared__ uint32_t m_e[64];
	const int lane = threadIdx.x;
	const int sPx = sc.sPx, sPt = sc.sPt, P = sc.P;
	const int SL = sPx * sPx;
	const int x0 = (int)blockIdx.x * 8, y0 = (int)blockIdx.y * 8, t0 = tframe0 + (int)blockIdx.z;
	const int qx0 = max(0, x0 - sPx + 1), qx1 = min(x0 + 7, sc.W - sPx);
	const int qy0 = max(0, y0 - sPx + 1), qy1 = min(y0 + 7, sc.H - sPx);
	const int qt0 = max(0, t0 - sPt + 1), qt1 = min(t0, sc.F - sPt);
	if (qx1 < qx0 || qy1 < qy0 || qt1 < qt0 || t0 >= sc.F) return;
	const int nqx = qx1 - qx0 + 1, nqy = qy1 - qy0 + 1;
	const int npos = nqx * nqy * (qt1 - qt0 + 1);

	for (int p = lane; p < npos; p += 64) {
		const int qx = qx0 + p % nqx, qy = qy0 + (p / nqx) % nqy, qt = qt0 + p / (nqx * nqy);
		const uint32_t key = (uint32_t)qt * sc.WH + (uint32_t)qy * sc.W + qx;
		const uint32_t a = kstart[key], b = kstart[key + 1];
		pcur[p] = a;
		pend[p] = b;
		phead[p] = a < b ? svals[a] : 0xffffffffu;
	}
	// this lane's pixel, its owner tile and running sums
	const int x = x0 + (lane & 7), y = y0 + (lane >> 3);
	const bool valid = x < sc.W && y < sc.H;
	const int owner = valid ? owner_tile(x, y, sc) : -1;
	const size_t pix = (size_t)t0 * sc.WH + (size_t)y * sc.W + x;
	float aw = 0.f, an[CC];
#pragma unroll
	for (int c = 0; c < CC; c++) an[c] = 0.f;
	if (valid) {
		aw = wsum[pix];
#pragma unroll
		for (int c = 0; c < CC; c++) an[c] = num[(size_t)t0 * sc.WHC + (size_t)c * sc.WH + (size_t)y * sc.W + x];
	}
	const int NB = min(64, AGG_STAGE / (SL * CC)); // rows per staged batch
	for (int i = lane; i < P; i += 64) swin[i] = win[i];
	// staging layout: a row slice is SL*CC floats; when it is at most 64
	// floats, 64/(SL*CC) rows are staged per pass (lane -> row jl, float rl)
	const int RS = SL * CC;
	const int rpp = RS <= 64 ? 64 / RS : 1;
	const int jl = RS <= 64 ? lane / RS : 0;
	const int rl = RS <= 64 ? lane - jl * RS : lane;
	const bool stage_lane = RS <= 64 ? jl < rpp : true;
	const int nslot = RS <= 64 ? 1 : (RS + 63) / 64;
	const int xo = x - x0 + 16, yo = y - y0 + 16;
	__syncthreads();
	for (int b = 0; b < nbands; b++) {
		uint32_t lo = bands[b];
		const uint32_t band_end = bands[b + 1];
		while (lo < band_end) {
			// runs of [lo, hi): one band, or a part of it when it is too large
			uint32_t hi = band_end;
			int K = 0;
			for (int attempt = 0;; attempt++) {
				// run of each position inside [lo, hi): [pcur, first row >= hi)
				for (int p = lane; p < npos; p += 64) {
					const uint32_t a = pcur[p];
					uint32_t ub = a;
					if (phead[p] < hi) ub = lower_bound_u32(svals, a + 1, pend[p], hi);
					pub[p] = ub;
				}
				__syncthreads();
				K = 0;
				for (int p0 = 0; p0 < npos; p0 += 64) {
					const int p = p0 + lane;
					const int len = p < npos ? (int)(pub[p] - pcur[p]) : 0;
					const int incl = wave_incl_scan(len);
					if (p < npos) poff[p] = (uint32_t)(K + incl - len);
					K += __shfl(incl, 63, 64);
				}
				if (K <= AGG_CAPB) break;
				if (lane == 0 && attempt == 0) atomicAdd(nsplit, 1u);
				hi = lo + (hi - lo + 1) / 2; // halve the row range and retry
				__syncthreads();
			}
			if (lane == 0) poff[npos] = (uint32_t)K;
			__syncthreads();
			if (K > 0) {
				// gather (entry k lies in the run p with poff[p] <= k < poff[p+1])
				for (int k = lane; k < K; k += 64) {
					int a = 0, bb = npos - 1;
					while (a < bb) {
						const int mid = (a + bb + 1) >> 1;
						if (poff[mid] <= (uint32_t)k) a = mid; else bb = mid - 1;
					}
					keysA[k] = svals[pcur[a] + ((uint32_t)k - poff[a])] - lo;
				}
				__syncthreads();
				int nbits = 1;
				while (nbits < 32 && (1ull << nbits) < (uint64_t)(hi - lo)) nbits++;
				const uint32_t *sorted = wave_radix_sort(keysA, keysB, K, nbits, hist);
				__syncthreads();
				for (int k0 = 0; k0 < K; k0 += NB) {
					const int nb = min(NB, K - k0);
					if (lane < nb) {
						const uint32_t e = sorted[k0 + lane] + lo;
						int qx, qy, qt;
						pos_decode(epos[e], sc.W, sc.WH, qx, qy, qt);
						m_pos[lane] = (uint32_t)(qx - x0 + 16) | ((uint32_t)(qy - y0 + 16) << 8) |
						              ((uint32_t)(t0 - qt) << 16) | (etile[e] << 24);
						m_w[lane] = entry_w[e];
						m_e[lane] = e;
					}
					__syncthreads();
					// stage the rows' slices at frame t0
					if (RS <= 64) {
						for (int j0 = 0; j0 < nb; j0 += rpp) {
							const int j = j0 + jl;
							if (stage_lane && j < nb) {
								const int c = rl / SL, sidx = rl - c * SL;
								const int dt = (int)((m_pos[j] >> 16) & 255u);
								stage[j * RS + rl] = est[((size_t)m_e[j] * CC + c) * P + dt * SL + sidx];
							}
						}
					} else {
						for (int j = 0; j < nb; j++)
							for (int k = 0; k < nslot; k++) {
								const int r = lane + 64 * k;
								if (r < RS) {
									const int c = r / SL, sidx = r - c * SL;
									const int dt = (int)((m_pos[j] >> 16) & 255u);
									stage[j * RS + r] = est[((size_t)m_e[j] * CC + c) * P + dt * SL + sidx];
								}
							}
					}
					__syncthreads();
					for (int j = 0; j < nb; j++) {
						const uint32_t mp = m_pos[j];
						const int dx = xo - (int)(mp & 255u);
						const int dy = yo - (int)((mp >> 8) & 255u);
						if ((unsigned)dx < (unsigned)sPx && (unsigned)dy < (unsigned)sPx && (int)(mp >> 24) == owner) {
							const int dt = (int)((mp >> 16) & 255u);
							const int sidx = dy * sPx + dx;
							aw = aw + m_w[j] * swin[dt * SL + sidx];
#pragma unroll
							for (int c = 0; c < CC; c++) an[c] = an[c] + stage[j * RS + c * SL + sidx];
						}
					}
					__syncthreads();
				}
			}
			// advance the position cursors past [lo, hi)
			if (K > 0)
				for (int p = lane; p < npos; p += 64) {
					const uint32_t ub = pub[p];
					if (ub != pcur[p]) {
						pcur[p] = ub;
						phead[p] = ub < pend[p] ? svals[ub] : 0xffffffffu;
					}
				}
			__syncthreads();
			lo = hi;
		}
	}
	if (valid) {
		wsum[pix] = aw;
#pragma unroll
		for (int c = 0; c < CC; c++) num[(size_t)t0 * sc.WHC + (size_t)c * sc.WH + (size_t)y * sc.W + x] = an[c];
	}
}

// ---------------------------------------------------------------------------
// Ordered aggregation, chunked form.  One wave per 8 x 8 pixel block of one
// frame, lane = pixel.  All rows whose patch overlaps the block (the runs of
// its patch positions in the position-sorted row list) are gathered into LDS
// and radix-sorted by row index e, in e-ranges of at most AGG2_CAP rows.
// Each chunk of up to 64 sorted rows then gets, per row, the 64-bit mask of
// the block pixels it covers inside its own tile; a 64 x 64 bit transpose
// across the wave turns the row masks into per-pixel masks, and every lane
// folds only its own rows, in increasing e: the order of
// computeAggregationStep{1,2} (nldct.cpp:2284-2300, 2376-2393).
// ---------------------------------------------------------------------------

#define AGG2_STAGE 1024  // floats of staged row slices per chunk
#define AGG2_SVN 16      // staged floats per lane in flight

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m)
{
	const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
	const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
	return ((uint64_t)hi << 32) | lo;
}

// Lane r holds row r of a 64 x 64 bit matrix (bit c = column c); returns
// column `lane` (bit r = row r's bit `lane`).  Recursive block swap.
__device__ __forceinline__ uint64_t transpose_bits64(uint64_t a)
{
	const int lane = threadIdx.x & 63;
	const uint64_t L[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
	                       0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
	for (int i = 0; i < 6; i++) {
		const int s = 32 >> i;
		const uint64_t p = shfl_xor_u64(a, s);
		a = (lane & s) ? ((a & ~L[i]) | ((p >> s) & L[i])) : ((a & L[i]) | ((p & L[i]) << s));
	}
	return a;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
	return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
	return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
	return v;
}

// n / d for n, d < 2^32 from a float reciprocal estimate and exact
// integer correction.
__device__ __forceinline__ uint32_t udiv_rcp(uint32_t n, uint32_t d, float rcp)
{
	uint32_t q = (uint32_t)fmaxf(0.f, (float)n * rcp);
	int64_t r = (int64_t)n - (int64_t)q * d;
	while (r < 0) { q--; r += d; }
	while (r >= (int64_t)d) { q++; r -= d; }
	return q;
}

// bits [a, b) of an 8-bit lane row, a and b clamped to [0, 8]
__device__ __forceinline__ uint32_t span8(int a, int b)
{
	a = max(a, 0); b = min(b, 8);
	return b > a ? ((1u << b) - 1u) & ~((1u << a) - 1u) : 0u;
}

// CAP: rows sorted per e-range; WP: window table size (>= patch samples).
// The LDS budget (about 15 KB at CAP 1024, WP 64) and the 168-register cap
// keep 10 waves per CU: the kernel waits on gathers, and fewer waves (the
// unconstrained 178-register form, or 27 KB of LDS) cost 45 % more time.
template <int CC, int CAP, int WP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
void agg_chunk_kernel(const uint32_t *__restrict__ kstart, const uint32_t *__restrict__ svals,
                      const uint4 *__restrict__ erec, const float *__restrict__ est,
                      const float *__restrict__ win, StepConst sc, int tframe0, int nframes,
                      float *__restrict__ num, float *__restrict__ wsum, uint32_t *__restrict__ nsplit)
{
	__shared__ uint32_t keysA[CAP];
	// the sort's second buffer and histogram share space with the staged slices
	constexpr int USZ = (CAP + 256) > AGG2_STAGE ? (CAP + 256) : AGG2_STAGE;
	__shared__ __align__(16) uint32_t ubuf[USZ];
	uint32_t *keysB = ubuf, *hist = ubuf + CAP;
	float *stage = reinterpret_cast<float *>(ubuf);
	__shared__ uint32_t pcur[AGG_MAXPOS];
	__shared__ uint16_t pcnt[AGG_MAXPOS];   // rows of the current e-range (<= CAP)
	__shared__ float swin[WP];
	__shared__ uint32_t m_pos2[2][64];   // (dx0 + 16) | (dy0 + 16) << 8 | dt << 16
	// per row of a chunk: stage index and window index of the block pixel
	// (0, 0) (offsets of the row's patch folded in), the weight bits and the
	// row index e
	__shared__ int4 m_rec2[2][64];
	// the unpipelined form (slices wider than the wave) keeps its row index
	// and weight in the record space
	uint32_t *m_pos = m_pos2[0], *m_e = reinterpret_cast<uint32_t *>(&m_rec2[0][0]);
	float *m_w = reinterpret_cast<float *>(&m_rec2[1][0]);
	const int lane = threadIdx.x;
	const int sPx = sc.sPx, sPt = sc.sPt, P = sc.P;
	const int SL = sPx * sPx;
	// One-dimensional grid.  Work items are ordered frame-fastest, then x,
	// then y, and each XCD (workgroups are dealt to the 8 XCDs in turn) takes
	// one contiguous eighth of them: the blocks that share rows (the frames
	// of one place, horizontal neighbours) run together on one L2.
	const int nbx = (sc.W + 7) / 8, nby = (sc.H + 7) / 8;
	const uint32_t nwork = (uint32_t)nframes * nbx * nby;
	const uint32_t per_xcd = (nwork + 7) / 8;
	const uint32_t wk = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
	if (wk >= nwork) return;
	const int bt = (int)(wk % (uint32_t)nframes);
	const uint32_t wr = wk / (uint32_t)nframes;
	const int x0 = (int)(wr % (uint32_t)nbx) * 8, y0 = (int)(wr / (uint32_t)nbx) * 8, t0 = tframe0 + bt;
	const int qx0 = max(0, x0 - sPx + 1), qx1 = min(x0 + 7, sc.W - sPx);
	const int qy0 = max(0, y0 - sPx + 1), qy1 = min(y0 + 7, sc.H - sPx);
	const int qt0 = max(0, t0 - sPt + 1), qt1 = min(t0, sc.F - sPt);
	if (qx1 < qx0 || qy1 < qy0 || qt1 < qt0 || t0 >= sc.F) return;
	const int nqx = qx1 - qx0 + 1, nqy = qy1 - qy0 + 1;
	const int npos = nqx * nqy * (qt1 - qt0 + 1);

	// position runs (ends in registers: position lane + 64 i), total row
	// count and e-range of the block
	constexpr int PPL = AGG_MAXPOS / 64;
	uint32_t pend[PPL];
	uint32_t ksum = 0, emin = 0xffffffffu, emax = 0u;
#pragma unroll
	for (int i = 0; i < PPL; i++) {
		const int p = lane + 64 * i;
		pend[i] = 0;
		if (p < npos) {
			const int qx = qx0 + p % nqx, qy = qy0 + (p / nqx) % nqy, qt = qt0 + p / (nqx * nqy);
			const uint32_t key = (uint32_t)qt * sc.WH + (uint32_t)qy * sc.W + qx;
			const uint32_t a = kstart[key], b = kstart[key + 1];
			pcur[p] = a;
			pend[i] = b;
			if (b > a) {
				ksum += b - a;
				emin = min(emin, svals[a]);
				emax = max(emax, svals[b - 1]);
			}
		}
	}
	ksum = wave_sum_u32(ksum);
	emin = wave_min_u32(emin);
	emax = wave_max_u32(emax);

	// this lane's pixel and the lanes owned by each tile present in the block
	const int lx = lane & 7, ly = lane >> 3;
	const int x = x0 + lx, y = y0 + ly;
	const bool valid = x < sc.W && y < sc.H;
	const int owner = valid ? owner_tile(x, y, sc) : -1;
	int own_t[4] = {-1, -1, -1, -1};
	uint64_t own_m[4] = {0, 0, 0, 0};
	{
		uint64_t rest = __ballot(valid);
#pragma unroll
		for (int i = 0; i < 4; i++) {
			if (rest) {
				const int lead = __ffsll((unsigned long long)rest) - 1;
				const int o = __shfl(owner, lead, 64);
				const uint64_t m = __ballot(owner == o);
				own_t[i] = o;
				own_m[i] = m;
				rest &= ~m;
			}
		}
	}
	const size_t pix = (size_t)t0 * sc.WH + (size_t)y * sc.W + x;
	float aw = 0.f, an[CC];
#pragma unroll
	for (int c = 0; c < CC; c++) an[c] = 0.f;
	if (valid) {
		aw = wsum[pix];
#pragma unroll
		for (int c = 0; c < CC; c++) an[c] = num[(size_t)t0 * sc.WHC + (size_t)c * sc.WH + (size_t)y * sc.W + x];
	}
	for (int i = lane; i < P; i += 64) swin[i] = win[i];
	const int RS = SL * CC;                       // floats of one staged row slice
	const int rpp = RS <= 64 ? 64 / RS : 1;       // slices staged per pass
	const int NBR = min(min(64, AGG2_STAGE / RS), AGG2_SVN * rpp);   // rows per chunk
	const int jl = RS <= 64 ? lane / RS : 0, rl = RS <= 64 ? lane - jl * RS : 0;
	const bool stage_lane = jl < rpp;
	const int cl = rl / SL, sl = rl - cl * SL;
	const float rcpW = 1.f / (float)sc.W, rcpWH = 1.f / (float)sc.WH;
	const int pix_l = ly * sPx + lx;   // this pixel's index in a patch slice anchored at the block origin
	__syncthreads();

	uint32_t left = ksum, lo = emin;
	int nrange = 0;
	while (left > 0) {
		// e-range [lo, hi) with at most CAP rows
		uint64_t hi64 = (uint64_t)emax + 1;
		if (left > CAP)
			hi64 = lo + max<uint64_t>(1, ((uint64_t)emax + 1 - lo) * (CAP * 3 / 4) / left);
		uint32_t hi = (uint32_t)hi64;
		uint32_t K = 0;
		for (;;) {
			uint32_t kl = 0;
#pragma unroll
			for (int i = 0; i < PPL; i++) {
				const int p = lane + 64 * i;
				if (p < npos) {
					const uint32_t a = pcur[p], b = pend[i];
					uint32_t l = a, r = b;   // first index in [a, b) with svals >= hi
					if (hi > emax) l = b;    // the last range takes every remaining row
					while (l < r) {
						const uint32_t mid = (l + r) >> 1;
						if (svals[mid] < hi) l = mid + 1; else r = mid;
					}
					pcnt[p] = (uint16_t)min<uint32_t>(l - a, 65535u);
					kl += l - a;
				}
			}
			K = wave_sum_u32(kl);
			if (K <= CAP) break;
			hi = lo + (hi - lo + 1) / 2;
		}
		nrange++;
		__syncthreads();
		if (K > 0) {
			// gather the range's rows (keys relative to lo)
			uint32_t base = 0;
			for (int p0 = 0; p0 < npos; p0 += 64) {
				const int p = p0 + lane;
				const uint32_t len = p < npos ? (uint32_t)pcnt[p] : 0u;
				const int incl = wave_incl_scan((int)len);
				uint32_t o = base + (uint32_t)incl - len;
				if (p < npos)
					for (uint32_t i = pcur[p], ie = pcur[p] + len; i < ie; i++) keysA[o++] = svals[i] - lo;
				base += (uint32_t)__shfl(incl, 63, 64);
			}
			__syncthreads();
			int nbits = 1;
			while (nbits < 32 && (1ull << nbits) < (uint64_t)(hi - lo)) nbits++;
			const uint32_t *sorted = wave_radix_sort(keysA, keysB, (int)K, nbits, hist);
			__syncthreads();
			if (sorted != keysA) {
				// keysB shares space with the staged slices: move the order back
				for (uint32_t k = lane; k < K; k += 64) keysA[k] = keysB[k];
				__syncthreads();
				sorted = keysA;
			}
			if (RS <= 64) {
				// Software-pipelined chunks: the row metadata of chunk c+2 and
				// the slices of chunk c+1 are in flight while chunk c is folded.
				const int nch = (int)((K + NBR - 1) / NBR);
				uint32_t eN = 0, qN = 0, tN = 0;
				float wN = 0.f;
				auto meta_load = [&](int c) {
					const int nbc = c < nch ? (int)min<uint32_t>(NBR, K - (uint32_t)c * NBR) : 0;
					if (lane < nbc) {
						eN = sorted[(uint32_t)c * NBR + lane] + lo;
						const uint4 rec = erec[eN];
						qN = rec.x;
						tN = rec.y;
						wN = __uint_as_float(rec.z);
					}
				};
				auto meta_commit = [&](int c) -> uint64_t {
					const int nbc = (int)min<uint32_t>(NBR, K - (uint32_t)c * NBR);
					const int buf = c & 1;
					uint64_t mrow = 0;
					if (lane < nbc) {
						const int qt = (int)udiv_rcp(qN, (uint32_t)sc.WH, rcpWH);
						const uint32_t rr = qN - (uint32_t)qt * (uint32_t)sc.WH;
						const int qy = (int)udiv_rcp(rr, (uint32_t)sc.W, rcpW);
						const int qx = (int)(rr - (uint32_t)qy * (uint32_t)sc.W);
						const int dx0 = qx - x0, dy0 = qy - y0, dt = t0 - qt;
						const uint32_t cols = span8(dx0, dx0 + sPx), rows = span8(dy0, dy0 + sPx);
						uint64_t rsel = 0;
#pragma unroll
						for (int r = 0; r < 8; r++) if ((rows >> r) & 1u) rsel |= 1ull << (8 * r);
						mrow = (uint64_t)cols * rsel;
						uint64_t om = 0;
#pragma unroll
						for (int i = 0; i < 4; i++) if (own_t[i] == (int)tN) om = own_m[i];
						mrow &= om;
						m_pos2[buf][lane] = (uint32_t)(dx0 + 16) | ((uint32_t)(dy0 + 16) << 8) | ((uint32_t)dt << 16);
						const int off = dy0 * sPx + dx0;
						m_rec2[buf][lane] = make_int4(lane * RS - off, dt * SL - off, (int)__float_as_uint(wN), (int)eN);
					}
					return mrow;
				};
				float sv[AGG2_SVN];
				auto stage_issue = [&](int c) {
					const int nbc = (int)min<uint32_t>(NBR, K - (uint32_t)c * NBR);
					const int buf = c & 1;
#pragma unroll
					for (int i = 0; i < AGG2_SVN; i++) {
						const int j = jl + i * rpp;
						if (stage_lane && j < nbc) {
							const int dt = (int)((m_pos2[buf][j] >> 16) & 255u);
							const uint32_t ej = reinterpret_cast<const uint32_t *>(&m_rec2[buf][j])[3];
							sv[i] = est[((size_t)ej * CC + cl) * P + dt * SL + sl];
						}
					}
				};
				auto stage_write = [&](int c) {
					const int nbc = (int)min<uint32_t>(NBR, K - (uint32_t)c * NBR);
#pragma unroll
					for (int i = 0; i < AGG2_SVN; i++) {
						const int j = jl + i * rpp;
						if (stage_lane && j < nbc) stage[j * RS + rl] = sv[i];
					}
				};
				meta_load(0);
				uint64_t mcur = meta_commit(0);
				meta_load(1);
				__syncthreads();
				stage_issue(0);
				stage_write(0);
				for (int c = 0; c < nch; c++) {
					uint64_t mnext = 0;
					if (c + 1 < nch) mnext = meta_commit(c + 1);
					meta_load(c + 2);
					__syncthreads();
					if (c + 1 < nch) stage_issue(c + 1);
					const int buf = c & 1;
					// each lane folds the rows that cover its pixel, in row
					// order, two rows per round (their loads overlap)
					uint64_t mine = transpose_bits64(mcur);
					while (mine) {
						const int j1 = __ffsll((unsigned long long)mine) - 1;
						mine &= mine - 1;
						const bool two = mine != 0;
						const int j2 = two ? __ffsll((unsigned long long)mine) - 1 : j1;
						if (two) mine &= mine - 1;
						const int4 r1 = m_rec2[buf][j1], r2 = m_rec2[buf][j2];
						const float v1 = swin[r1.y + pix_l], v2 = swin[r2.y + pix_l];
						float x1[CC], x2[CC];
#pragma unroll
						for (int cc = 0; cc < CC; cc++) {
							x1[cc] = stage[r1.x + cc * SL + pix_l];
							x2[cc] = stage[r2.x + cc * SL + pix_l];
						}
						aw = aw + __int_as_float(r1.z) * v1;
#pragma unroll
						for (int cc = 0; cc < CC; cc++) an[cc] = an[cc] + x1[cc];
						if (two) {
							aw = aw + __int_as_float(r2.z) * v2;
#pragma unroll
							for (int cc = 0; cc < CC; cc++) an[cc] = an[cc] + x2[cc];
						}
					}
					__syncthreads();
					if (c + 1 < nch) stage_write(c + 1);
					mcur = mnext;
				}
				__syncthreads();
			} else
			for (uint32_t k0 = 0; k0 < K; k0 += NBR) {
				const int nb = (int)min<uint32_t>(NBR, K - k0);
				uint64_t mrow = 0;
				if (lane < nb) {
					const uint32_t e = sorted[k0 + lane] + lo;
					const uint4 rec = erec[e];
					const uint32_t q = rec.x;
					const int qt = (int)udiv_rcp(q, (uint32_t)sc.WH, rcpWH);
					const uint32_t rr = q - (uint32_t)qt * (uint32_t)sc.WH;
					const int qy = (int)udiv_rcp(rr, (uint32_t)sc.W, rcpW);
					const int qx = (int)(rr - (uint32_t)qy * (uint32_t)sc.W);
					const int dx0 = qx - x0, dy0 = qy - y0, dt = t0 - qt;
					const uint32_t cols = span8(dx0, dx0 + sPx), rows = span8(dy0, dy0 + sPx);
					uint64_t rsel = 0;
#pragma unroll
					for (int r = 0; r < 8; r++) if ((rows >> r) & 1u) rsel |= 1ull << (8 * r);
					mrow = (uint64_t)cols * rsel;
					const int tl = (int)rec.y;
					uint64_t om = 0;
#pragma unroll
					for (int i = 0; i < 4; i++) if (own_t[i] == tl) om = own_m[i];
					mrow &= om;
					m_pos[lane] = (uint32_t)(dx0 + 16) | ((uint32_t)(dy0 + 16) << 8) | ((uint32_t)dt << 16);
					m_e[lane] = e;
					m_w[lane] = __uint_as_float(rec.z);
				}
				__syncthreads();
				// stage the rows' slices at frame t0 (lane -> fixed float of a
				// slice when a slice fits the wave)
				if (RS <= 64) {
					for (int j = jl; j < nb; j += rpp)
						if (stage_lane) {
							const int dt = (int)((m_pos[j] >> 16) & 255u);
							stage[j * RS + rl] = est[((size_t)m_e[j] * CC + cl) * P + dt * SL + sl];
						}
				} else {
					for (int i = lane; i < nb * RS; i += 64) {
						const int j = i / RS, r = i - j * RS;
						const int c = r / SL, sidx = r - c * SL;
						const int dt = (int)((m_pos[j] >> 16) & 255u);
						stage[i] = est[((size_t)m_e[j] * CC + c) * P + dt * SL + sidx];
					}
				}
				__syncthreads();
				uint64_t mine = transpose_bits64(mrow);
				while (mine) {
					const int j = __ffsll((unsigned long long)mine) - 1;
					mine &= mine - 1;
					const uint32_t mp = m_pos[j];
					const int dx = lx + 16 - (int)(mp & 255u);
					const int dy = ly + 16 - (int)((mp >> 8) & 255u);
					const int dt = (int)((mp >> 16) & 255u);
					const int sidx = dy * sPx + dx;
					aw = aw + m_w[j] * swin[dt * SL + sidx];
#pragma unroll
					for (int c = 0; c < CC; c++) an[c] = an[c] + stage[j * RS + c * SL + sidx];
				}
				__syncthreads();
			}
		}
		for (int p = lane; p < npos; p += 64) pcur[p] += pcnt[p];
		__syncthreads();
		left -= K;
		lo = hi;
	}
	if (lane == 0 && nrange > 1) atomicAdd(nsplit, 1u);
	if (valid) {
		wsum[pix] = aw;
#pragma unroll
		for (int c = 0; c < CC; c++) num[(size_t)t0 * sc.WHC + (size_t)c * sc.WH + (size_t)y * sc.W + x] = an[c];
	}
}

#define AGG2_ARGS const uint32_t *, const uint32_t *, const uint4 *, const float *, const float *, \
	StepConst, int, int, float *, float *, uint32_t *
template __global__ void agg_chunk_kernel<1, 512, 64>(AGG2_ARGS);
template __global__ void agg_chunk_kernel<1, 768, 64>(AGG2_ARGS);
template __global__ void agg_chunk_kernel<1, 1024, 64>(AGG2_ARGS);
template __global__ void agg_chunk_kernel<1, 1536, 64>(AGG2_ARGS);
template __global__ void agg_chunk_kernel<3, 1024, 64>(AGG2_ARGS);
template __global__ void agg_chunk_kernel<4, 1024, 64>(AGG2_ARGS);
template __global__ void agg_chunk_kernel<1, 1024, 256>(AGG2_ARGS);
template __global__ void agg_chunk_kernel<3, 1024, 256>(AGG2_ARGS);
template __global__ void agg_chunk_kernel<4, 1024, 256>(AGG2_ARGS);

#define AGG_ARGS const uint32_t *, const uint32_t *, const uint32_t *, const uint32_t *, \
	const float *, const float *, const float *, StepConst, const uint32_t *, int, int, float *, \
	float *, uint32_t *
template __global__ void agg_exact_kernel<1>(AGG_ARGS);
template __global__ void agg_exact_kernel<3>(AGG_ARGS);
template __global__ void agg_exact_kernel<4>(AGG_ARGS);

// Row index of the first row of each band: bands[b] = base[gfirst[b] - g0].
__global__ void band_bounds_kernel(const uint32_t *__restrict__ base, const uint32_t *__restrict__ gfirst,
                                   int nb, int g0, uint32_t *__restrict__ bands)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b < nb) bands[b] = base[gfirst[b] - g0];
}

// computeWeightedAggregation, nldct.cpp:2436-2448.
// ---------------------------------------------------------------------------
// Ordered aggregation by bins of 8 x 8 pixels x AGG_TB frames.
//
// bin_hist1_kernel / bin_scatter1_kernel (side stream, beside the filter):
// every row e gives one (bin, e) pair per bin its patch overlaps, in
// increasing e, sorted stably by the bin's low 8 bits straight from the
// rows; the further 8-bit passes (sortscan.hip) then leave each bin's rows
// in increasing e, and bin_bounds_kernel marks where each bin's list
// starts.  agg_bin_fold_kernel: one wave per bin, lane = pixel
// of the 8 x 8 block, AGG_TB running sums per lane (one per frame of the
// bin).  The bin's list is read in chunks of NBR rows; each row's slices at
// the bin's frames are staged once in LDS and added, frame by frame, into
// the lanes whose pixel the row covers, in increasing e: for every pixel the
// rows come in the order computeAggregationStep{1,2} adds them
// (nldct.cpp:2284-2300, 2376-2393), one float addition at a time, so the
// sums are those of the per-frame kernels above.
// ---------------------------------------------------------------------------

// bins of row e's patch (x blocks, y blocks, frame bins)
__device__ __forceinline__ void bin_span(uint32_t q, const StepConst &sc, int &bx0, int &bx1, int &by0, int &by1,
                                         int &bt0, int &bt1)
{
	int qx, qy, qt;
	pos_decode(q, sc.W, sc.WH, qx, qy, qt);
	bx0 = qx >> 3;
	bx1 = (qx + sc.sPx - 1) >> 3;
	by0 = qy >> 3;
	by1 = (qy + sc.sPx - 1) >> 3;
	bt0 = qt / AGG_TB;
	bt1 = (qt + sc.sPt - 1) / AGG_TB;
}

// bin key of the pairs of a row, in the order they are listed: frame bins,
// then y, then x
#define BIN_FOR_PAIRS(q, KEY, BODY)                                                                  \
	{                                                                                                \
		int bx0_, bx1_, by0_, by1_, bt0_, bt1_;                                                      \
		bin_span(q, sc, bx0_, bx1_, by0_, by1_, bt0_, bt1_);                                         \
		for (int bt_ = bt0_; bt_ <= bt1_; bt_++)                                                     \
			for (int by_ = by0_; by_ <= by1_; by_++)                                                 \
				for (int bx_ = bx0_; bx_ <= bx1_; bx_++) {                                           \
					const uint32_t KEY = (uint32_t)((bt_ * nby + by_) * nbx + bx_);                  \
					BODY;                                                                            \
				}                                                                                    \
	}

// First radix pass straight from the rows (no pair list in memory): tiles
// of BIN_RT rows, one wave each (grid-stride over the tiles).
// bin_hist1_kernel: the low-digit counts of each tile's pairs, digit-major
// (hist[d * ntiles + tile]).
__global__ __launch_bounds__(64)
void bin_hist1_kernel(const uint32_t *__restrict__ epos, const uint32_t *__restrict__ Mp, int64_t Mub,
                      StepConst sc, int nbx, int nby, int ntiles, uint32_t *__restrict__ hist)
{
	__shared__ uint32_t h[256];
	const int lane = threadIdx.x;
	const int64_t M = min<int64_t>(Mub, (int64_t)*Mp);
	for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
#pragma unroll
		for (int i = 0; i < 4; i++) h[lane + 64 * i] = 0;
		__syncthreads();
		const int64_t e0 = (int64_t)tile * BIN_RT;
		for (int r0 = 0; r0 < BIN_RT && e0 + r0 < M; r0 += 256) {   // four rows per lane, loads together
			uint32_t q[4];
#pragma unroll
			for (int u = 0; u < 4; u++) {
				const int64_t e = e0 + r0 + 64 * u + lane;
				q[u] = e < M ? epos[e] : 0u;
			}
#pragma unroll
			for (int u = 0; u < 4; u++)
				if (e0 + r0 + 64 * u + lane < M) BIN_FOR_PAIRS(q[u], key, atomicAdd(&h[key & 255u], 1u));
		}
		__syncthreads();
#pragma unroll
		for (int i = 0; i < 4; i++) hist[(size_t)(lane + 64 * i) * ntiles + tile] = h[lane + 64 * i];
		__syncthreads();
	}
}

// bin_scatter1_kernel: each tile's pairs to the slots the scan of the counts
// gives, in row order: a round's 64 rows list their pairs in LDS in (row,
// bin) order, then rounds of 64 pairs take consecutive slots per digit (the
// lanes of a round with the same digit by rank, the lowest advancing the
// digit's offset), as rs_scatter_kernel does for the later passes.
__global__ __launch_bounds__(64)
void bin_scatter1_kernel(const uint32_t *__restrict__ epos, const uint32_t *__restrict__ Mp, int64_t Mub,
                         StepConst sc, int nbx, int nby, int ntiles, const uint32_t *__restrict__ offs,
                         uint32_t *__restrict__ kout, uint32_t *__restrict__ vout)
{
	__shared__ uint32_t base[256];
	__shared__ uint32_t pk[8 * 64], pe[8 * 64];
	const int lane = threadIdx.x;
	const uint64_t lt = (1ull << lane) - 1ull;
	const int64_t M = min<int64_t>(Mub, (int64_t)*Mp);
	for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const int64_t e0 = (int64_t)tile * BIN_RT;
		if (e0 >= M) break;
#pragma unroll
		for (int i = 0; i < 4; i++) base[lane + 64 * i] = offs[(size_t)(lane + 64 * i) * ntiles + tile];
		uint32_t qn = e0 + lane < M ? epos[e0 + lane] : 0u;   // the next round's position
		__syncthreads();
		for (int r0 = 0; r0 < BIN_RT && e0 + r0 < M; r0 += 64) {
			const int64_t e = e0 + r0 + lane;
			const uint32_t q = qn;
			qn = r0 + 64 < BIN_RT && e + 64 < M ? epos[e + 64] : 0u;
			int c = 0;
			if (e < M) BIN_FOR_PAIRS(q, key, (void)key; c++);
			const int incl = wave_incl_scan(c);
			const int T = __shfl(incl, 63, 64);
			int o = incl - c;
			if (e < M) BIN_FOR_PAIRS(q, key, pk[o] = key; pe[o] = (uint32_t)e; o++);
			__syncthreads();
			for (int p0 = 0; p0 < T; p0 += 64) {
				const int i = p0 + lane;
				const bool valid = i < T;
				const uint32_t key = valid ? pk[i] : 0u, val = valid ? pe[i] : 0u;
				const uint32_t d = key & 255u;
				uint64_t m = __ballot(valid);
#pragma unroll
				for (int bit = 0; bit < 8; bit++) {
					const bool bs = (d >> bit) & 1u;
					const uint64_t bb = __ballot(valid && bs);
					m &= bs ? bb : ~bb;
				}
				const uint32_t rk = (uint32_t)__popcll(m & lt);
				const uint32_t bd = base[d];
				__syncthreads();   // every lane has read its digit's offset
				if (valid) {
					kout[bd + rk] = key;
					vout[bd + rk] = val;
					if (rk == 0) base[d] = bd + (uint32_t)__popcll(m);
				}
				__syncthreads();
			}
		}
		__syncthreads();   // the next tile rewrites base
	}
}
#undef BIN_FOR_PAIRS

// boff[b] = number of sorted keys below b, b = 0 .. nbins: pair index
// i <= n writes the starts of the bins from the key before it, exclusive, to
// its own key, inclusive.  One wave per segment of 1024 indices, 16
// consecutive ones per lane (grid-stride over the segments).
__global__ __launch_bounds__(64)
void bin_bounds_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ np, int64_t nub,
                       int nbins, uint32_t *__restrict__ boff)
{
	const int64_t n = min<int64_t>((int64_t)*np, nub);
	const int64_t nseg = (n + 1 + 1023) / 1024;
	for (int64_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
		const int64_t i0 = sg * 1024 + (int64_t)threadIdx.x * 16;
		uint32_t k[17];   // keys[i0 - 1 .. i0 + 15]
		k[0] = i0 > 0 && i0 - 1 < n ? keys[i0 - 1] : 0u;
		if (i0 + 16 <= n) {
			const uint4 *p = reinterpret_cast<const uint4 *>(keys + i0);
#pragma unroll
			for (int q = 0; q < 4; q++) {
				const uint4 v = p[q];
				k[4 * q + 1] = v.x;
				k[4 * q + 2] = v.y;
				k[4 * q + 3] = v.z;
				k[4 * q + 4] = v.w;
			}
		} else {
#pragma unroll
			for (int j = 0; j < 16; j++) k[j + 1] = i0 + j < n ? keys[i0 + j] : 0u;
		}
#pragma unroll
		for (int j = 0; j < 16; j++) {
			const int64_t i = i0 + j;
			if (i > n) break;
			const uint32_t lo = i == 0 ? 0u : k[j] + 1u;
			const uint32_t hi = i == n ? (uint32_t)nbins : k[j + 1];
			for (uint32_t b = lo; b <= hi; b++) boff[b] = (uint32_t)i;
		}
	}
}

// CC channels; NBR (<= 32) rows per chunk; patches of 4 PQ samples, frame
// slices of 4 SQ samples (PQ <= 16).
//
// LDS, in floats: a zero area Z, then NBR + 1 slots of CC P + 4 floats.
// Slot j >= 0 holds chunk row j: its samples (channel-major) and, after
// them, its record (byte offset wb4 of the window / sample index of the
// block pixel (0, 0) at the bin's first frame, weight).  Slot -1 holds the
// window (its first P floats) and the record (-4 |Z|, +0): a round's
// missing rows have index -1 (the lowest set bit of an empty mask), so
// their samples and window values are read from Z, and every round adds
// RPR terms without a select (adding +0 leaves a sum that is not -0
// unchanged, and these sums start at +0 and never become -0).
template <int CC, int NBR, int PQ, int SQ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CC == 1 ? 4 : 3)))
void agg_bin_fold_kernel(const uint32_t *__restrict__ boff, const uint32_t *__restrict__ blist,
                         const uint4 *__restrict__ erec, const float *__restrict__ est,
                         const float *__restrict__ win, StepConst sc, int nbx, int nby, int nbins,
                         float *__restrict__ num, float *__restrict__ wsum)
{
	static_assert(NBR <= 32, "row masks are 32 bits");
	constexpr int TB = AGG_TB;
	constexpr int P = 4 * PQ, SL = 4 * SQ, QR = CC * PQ;   // QR: float4 of one row's samples
	constexpr int QS = QR + 1;                             // float4 per slot (samples, record)
	constexpr int SPL = (NBR * QR + 63) / 64;              // float4 staged per lane per chunk
	constexpr int RPR = CC == 1 ? 4 : 2;                   // rows folded per round
	static_assert(SPL * 64 == NBR * QR, "the staging fills the chunk's slots exactly");
	constexpr int ZR = (TB - 1) * SL + 64;                 // span of the lane offsets f SL + pix_l
	constexpr int ZQ = ((CC - 1) * P + ZR + 3) / 4;        // float4 of Z
	__shared__ __align__(16) float4 lds4[ZQ + (NBR + 1) * QS];
	float *const slotm1 = reinterpret_cast<float *>(lds4 + ZQ);   // slot -1
	float *const slot0 = slotm1 + 4 * QS;                          // slot 0
	constexpr int SB = 16 * QS;                                    // bytes per slot
	constexpr int RECB = 4 * CC * P;                               // record offset in a slot
	// per row of a chunk (two chunks in flight): (row index e, bit dt set
	// for the slices at the bin's frames)
	__shared__ uint2 rowm[2][NBR];
	const int lane = threadIdx.x;
	// bins are dealt to the 8 XCDs in contiguous eighths (x fastest)
	const uint32_t per_xcd = ((uint32_t)nbins + 7) / 8;
	const uint32_t b = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
	if (b >= (uint32_t)nbins) return;
	const uint32_t k0 = boff[b], k1 = boff[b + 1];
	if (k0 >= k1) return;
	const int bx = (int)(b % (uint32_t)nbx);
	const uint32_t rb = b / (uint32_t)nbx;
	const int x0 = bx * 8, y0 = (int)(rb % (uint32_t)nby) * 8, t0 = (int)(rb / (uint32_t)nby) * TB;
	const int sPx = sc.sPx, sPt = sc.sPt;
	const int lx = lane & 7, ly = lane >> 3;
	const int x = x0 + lx, y = y0 + ly;
	const bool valid = x < sc.W && y < sc.H;
	// the lanes owned by each tile present in the block
	const int owner = valid ? owner_tile(x, y, sc) : -1;
	int own_t[4] = {-1, -1, -1, -1};
	uint64_t own_m[4] = {0, 0, 0, 0};
	{
		uint64_t rest = __ballot(valid);
#pragma unroll
		for (int i = 0; i < 4; i++) {
			if (rest) {
				const int lead = __ffsll((unsigned long long)rest) - 1;
				const int o = __shfl(owner, lead, 64);
				const uint64_t m = __ballot(owner == o);
				own_t[i] = o;
				own_m[i] = m;
				rest &= ~m;
			}
		}
	}
	float aw[TB], an[TB][CC];
#pragma unroll
	for (int f = 0; f < TB; f++) {
		aw[f] = 0.f;
#pragma unroll
		for (int c = 0; c < CC; c++) an[f][c] = 0.f;
		if (valid && t0 + f < sc.F) {
			aw[f] = wsum[(size_t)(t0 + f) * sc.WH + (size_t)y * sc.W + x];
#pragma unroll
			for (int c = 0; c < CC; c++)
				an[f][c] = num[(size_t)(t0 + f) * sc.WHC + (size_t)c * sc.WH + (size_t)y * sc.W + x];
		}
	}
	for (int i = lane; i < ZQ; i += 64) lds4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
	for (int i = lane; i < P; i += 64) slotm1[i] = win[i];
	if (lane == 0) *reinterpret_cast<int2 *>(slotm1 + CC * P) = make_int2(-16 * ZQ, 0);
	const int pix_l = ly * sPx + lx;
	const float4 *est4 = reinterpret_cast<const float4 *>(est);
	const int nch = (int)((k1 - k0 + NBR - 1) / NBR);

	// Software pipeline: while chunk c is folded, the slices of chunk c+1,
	// the row records of chunk c+2 and the list entries of chunk c+3 are in
	// flight (eA, rA: chunk c+1's entries and records, eB, rB: chunk c+2's,
	// eC: chunk c+3's entries), so no record load waits for its entry load.
	// The loads are unconditional (index clamped to the bin's last entry,
	// and the lanes past the chunk's NBR rows all at its first entry, one
	// line; meta_commit uses only the lanes of the chunk's rows), so their
	// results land in the registers that carry them and nothing waits for
	// them before they are used.
	auto ent_load = [&](int c) -> uint32_t {
		const uint32_t kc = k0 + (uint32_t)c * NBR;
		return blist[min(kc + (uint32_t)(lane < NBR ? lane : 0), k1 - 1u)];
	};
	uint32_t eA = ent_load(0);
	uint4 rA = erec[eA];
	uint32_t eB = 0, eC = 0;
	uint4 rB = make_uint4(0, 0, 0, 0);
	uint32_t &eN = eA;
	uint4 &recN = rA;
	// the chunk's row masks into rowm[c & 1] and its records (lane = row)
	// into rw, for the slot write with the chunk's samples; returns the
	// lane's coverage mask (lanes = rows) and the rows present at each frame
	int2 rw = make_int2(0, 0);
	auto meta_commit = [&](int c, uint32_t (&vm)[TB]) -> uint64_t {
		const uint32_t kc = k0 + (uint32_t)c * NBR;
		const bool rv = lane < NBR && kc + lane < k1;
		uint64_t msp = 0;
		int qt = 0;
		if (rv) {
			const uint32_t q = recN.x;
			qt = (int)(recN.w >> 16);
			const int qy = (int)(recN.w & 0xffffu);
			const int qx = (int)(q - (uint32_t)qt * (uint32_t)sc.WH - (uint32_t)qy * (uint32_t)sc.W);
			const int dx0 = qx - x0, dy0 = qy - y0;
			const uint32_t cols = span8(dx0, dx0 + sPx), rows = span8(dy0, dy0 + sPx);
			uint64_t rsel = 0;
#pragma unroll
			for (int r = 0; r < 8; r++) if ((rows >> r) & 1u) rsel |= 1ull << (8 * r);
			msp = (uint64_t)cols * rsel;
			uint64_t om = 0;
#pragma unroll
			for (int i = 0; i < 4; i++) if (own_t[i] == (int)recN.y) om = own_m[i];
			msp &= om;
			// slices dt of the bin's frames t0 .. t0 + TB - 1
			const int dlo = max(0, t0 - qt), dhi = min(sPt - 1, t0 + TB - 1 - qt);
			const uint32_t dm = (msp != 0 && dhi >= dlo) ? ((2u << dhi) - (1u << dlo)) : 0u;
			const int wb = (t0 - qt) * SL - (dy0 * sPx + dx0);
			rw = make_int2(4 * wb, (int)recN.z);
			rowm[c & 1][lane] = make_uint2(eN, dm);
		} else if (lane < NBR) {
			rowm[c & 1][lane] = make_uint2(0u, 0u);   // past the list's end: nothing staged
		}
#pragma unroll
		for (int f = 0; f < TB; f++) {
			const int dt = t0 + f - qt;
			vm[f] = (uint32_t)__ballot(rv && msp != 0 && dt >= 0 && dt < sPt);
		}
		return msp;
	};
	// staged slices of a chunk: float4 i = lane + 64 m of the chunk's rows
	// (row i / QR); only the slices at the bin's frames are loaded and
	// stored, the rest of a row's slot is never read
	float4 sv[SPL];
#pragma unroll
	for (int m = 0; m < SPL; m++) sv[m] = make_float4(0.f, 0.f, 0.f, 0.f);
	uint32_t sneed = 0;
#define AGG_BIN_STAGE_ISSUE(c)                                                                      \
	{                                                                                               \
		sneed = 0;                                                                                  \
		const int nbc_ = (int)min<uint32_t>(NBR, k1 - (k0 + (uint32_t)(c) * NBR));                 \
		_Pragma("unroll") for (int m = 0; m < SPL; m++) {                                         \
			const int i = lane + 64 * m;                                                            \
			const int j = i / QR, rq = i - j * QR;                                                  \
			if (j < nbc_) {                                                                         \
				const uint2 rm = rowm[(c) & 1][j];                                                  \
				if ((rm.y >> ((rq % PQ) / SQ)) & 1u) {                                              \
					sv[m] = est4[(size_t)rm.x * QR + rq];                                           \
					sneed |= 1u << m;                                                               \
				}                                                                                   \
			}                                                                                       \
		}                                                                                           \
	}
#define AGG_BIN_STAGE_WRITE()                                                                       \
	{                                                                                               \
		_Pragma("unroll") for (int m = 0; m < SPL; m++)                                           \
			if ((sneed >> m) & 1u) {                                                                \
				const int i = lane + 64 * m;                                                        \
				const int j = i / QR;                                                               \
				lds4[ZQ + QS + j * QS + (i - j * QR)] = sv[m];                                      \
			}                                                                                       \
		if (lane < NBR) *reinterpret_cast<int2 *>(slot0 + 4 * QS * lane + CC * P) = rw;            \
	}
	uint32_t vmC[TB], vmN[TB];
	uint64_t mC = meta_commit(0, vmC);
	eA = ent_load(1);
	rA = erec[eA];
	eB = ent_load(2);
	__syncthreads();
	AGG_BIN_STAGE_ISSUE(0);
	AGG_BIN_STAGE_WRITE();
	// byte offset of the lane's slice at each frame of the bin
	int plf4[TB];
#pragma unroll
	for (int f = 0; f < TB; f++) plf4[f] = 4 * (f * SL + pix_l);
	// the sums' first values are complete here (an empty asm that reads and
	// writes them), so the fold below waits for no global load of its own:
	// the slices of the next chunk stay in flight while a chunk is folded
#pragma unroll
	for (int f = 0; f < TB; f++) {
		asm volatile("" : "+v"(aw[f]));
#pragma unroll
		for (int cc = 0; cc < CC; cc++) asm volatile("" : "+v"(an[f][cc]));
	}
	for (int c = 0; c < nch; c++) {
		uint64_t mN = 0;
		if (c + 1 < nch) mN = meta_commit(c + 1, vmN);
		rB = erec[eB];
		eC = ent_load(c + 3);
		__syncthreads();
		if (c + 1 < nch) AGG_BIN_STAGE_ISSUE(c + 1);
		// each lane folds the rows that cover its pixel, frame by frame, in
		// row order, RPR rows per round (their LDS reads overlap)
		const uint32_t mine_sp = (uint32_t)transpose_bits64(mC);
#pragma unroll
		for (int f = 0; f < TB; f++) {
			uint32_t mine = mine_sp & vmC[f];
			while (mine) {
				int so[RPR];   // byte offset of the row's slot from slot 0 (-SB: no row)
#pragma unroll
				for (int u = 0; u < RPR; u++) {
					// the lowest set bit's index, -1 for an empty mask
					int jj;
					asm("v_ffbl_b32 %0, %1" : "=v"(jj) : "v"(mine));
					mine &= mine - 1;
					so[u] = __mul24(jj, SB);
				}
				const char *s0 = reinterpret_cast<const char *>(slot0);
				const char *wn = reinterpret_cast<const char *>(slotm1) + plf4[f];
				int2 rr[RPR];
				float vv[RPR], xx[RPR][CC];
#pragma unroll
				for (int u = 0; u < RPR; u++) rr[u] = *reinterpret_cast<const int2 *>(s0 + so[u] + RECB);
#pragma unroll
				for (int u = 0; u < RPR; u++) {
					vv[u] = *reinterpret_cast<const float *>(wn + rr[u].x);
					const char *sp = s0 + (so[u] + rr[u].x + plf4[f]);
#pragma unroll
					for (int cc = 0; cc < CC; cc++) xx[u][cc] = *reinterpret_cast<const float *>(sp + 4 * cc * P);
				}
#pragma unroll
				for (int u = 0; u < RPR; u++) {
					aw[f] = aw[f] + __int_as_float(rr[u].y) * vv[u];
#pragma unroll
					for (int cc = 0; cc < CC; cc++) an[f][cc] = an[f][cc] + xx[u][cc];
				}
			}
		}
		__syncthreads();
		if (c + 1 < nch) AGG_BIN_STAGE_WRITE();
		eA = eB;
		rA = rB;
		eB = eC;
		mC = mN;
#pragma unroll
		for (int f = 0; f < TB; f++) vmC[f] = vmN[f];
	}
#pragma unroll
	for (int f = 0; f < TB; f++)
		if (valid && t0 + f < sc.F) {
			wsum[(size_t)(t0 + f) * sc.WH + (size_t)y * sc.W + x] = aw[f];
#pragma unroll
			for (int c = 0; c < CC; c++)
				num[(size_t)(t0 + f) * sc.WHC + (size_t)c * sc.WH + (size_t)y * sc.W + x] = an[f][c];
		}
}
#undef AGG_BIN_STAGE_ISSUE
#undef AGG_BIN_STAGE_WRITE

#define AGGB_ARGS const uint32_t *, const uint32_t *, const uint4 *, const float *, const float *, StepConst, \
	int, int, int, float *, float *
template __global__ void agg_bin_fold_kernel<1, 32, 16, 4>(AGGB_ARGS);
template __global__ void agg_bin_fold_kernel<1, 32, 16, 16>(AGGB_ARGS);
template __global__ void agg_bin_fold_kernel<3, 16, 16, 4>(AGGB_ARGS);
template __global__ void agg_bin_fold_kernel<3, 16, 16, 16>(AGGB_ARGS);
template __global__ void agg_bin_fold_kernel<4, 16, 16, 4>(AGGB_ARGS);
template __global__ void agg_bin_fold_kernel<4, 16, 16, 16>(AGGB_ARGS);

__global__ void normalize_kernel(const float *__restrict__ num, const float *__restrict__ wsum,
                                 const float *__restrict__ noisy, StepConst sc,
                                 float *__restrict__ out)
{
	const int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	const int64_t n = (int64_t)sc.WHC * sc.F;
	if (o >= n) return;
	const int64_t t = o / sc.WHC, r = o - t * sc.WHC;
	const int64_t pix = t * sc.WH + r % sc.WH;
	const float w = wsum[pix];
	out[o] = w > 0.f ? num[o] / w : num[o] / noisy[o];
}

} // namespace bm4dof
