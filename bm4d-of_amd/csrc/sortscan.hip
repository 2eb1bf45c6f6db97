// Scans, the position-ordered row lists and the early-out weight fixup of
// the denoiser, as plain kernels: no host round trip and no memset or copy
// nodes, so a whole step can be captured into a hipGraph and replayed.
#include "common.hpp"

namespace bm4dof {

__global__ void fill_u32_kernel(uint32_t *__restrict__ p, uint32_t v, int64_t n)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

__global__ void fill_f32_kernel(float *__restrict__ p, float v, int64_t n)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

__global__ void fill_u8_kernel(uint8_t *__restrict__ p, uint8_t v, int64_t n)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ scans
// Three passes over blocks of SCAN_ITEMS elements: block scans with block
// totals, one block scanning the totals, then the block offsets added.
// OP 0: sum (exclusive), OP 1: maximum (inclusive).  Inputs at index
// >= n_in read as the identity (0).

#define SCAN_THREADS 256
#define SCAN_PER 4
#define SCAN_ITEMS (SCAN_THREADS * SCAN_PER)

template <int OP> __device__ __forceinline__ uint32_t scan_op(uint32_t a, uint32_t b)
{
	return OP == 0 ? a + b : (a > b ? a : b);
}

// inclusive block scan of one value per thread (256 threads)
template <int OP> __device__ uint32_t block_incl_scan(uint32_t v, uint32_t *sh)
{
	const int t = threadIdx.x;
	sh[t] = v;
	__syncthreads();
	for (int d = 1; d < SCAN_THREADS; d <<= 1) {
		const uint32_t o = t >= d ? sh[t - d] : 0u;
		__syncthreads();
		if (t >= d) sh[t] = scan_op<OP>(sh[t], o);
		__syncthreads();
	}
	return sh[t];
}

template <int OP>
__global__ __launch_bounds__(SCAN_THREADS)
void scan_block_kernel(const uint32_t *__restrict__ in, int64_t n_in, int64_t n, uint32_t *__restrict__ out,
                       uint32_t *__restrict__ bsum)
{
	__shared__ uint32_t sh[SCAN_THREADS];
	const int64_t b0 = (int64_t)blockIdx.x * SCAN_ITEMS + (int64_t)threadIdx.x * SCAN_PER;
	uint32_t v[SCAN_PER];
	uint32_t acc = 0;
#pragma unroll
	for (int i = 0; i < SCAN_PER; i++) {
		const int64_t j = b0 + i;
		v[i] = j < n_in ? in[j] : 0u;
		acc = scan_op<OP>(acc, v[i]);
	}
	const uint32_t incl = block_incl_scan<OP>(acc, sh);
	uint32_t run = OP == 0 ? incl - acc : (threadIdx.x ? sh[threadIdx.x - 1] : 0u);
#pragma unroll
	for (int i = 0; i < SCAN_PER; i++) {
		const int64_t j = b0 + i;
		if (OP == 0) {
			if (j < n) out[j] = run;   // exclusive
			run += v[i];
		} else {
			run = scan_op<OP>(run, v[i]);
			if (j < n) out[j] = run;   // inclusive
		}
	}
	if (threadIdx.x == SCAN_THREADS - 1) bsum[blockIdx.x] = incl;
}

// exclusive scan (OP) of nb block totals in place, one block
template <int OP>
__global__ __launch_bounds__(SCAN_THREADS)
void scan_top_kernel(uint32_t *__restrict__ bsum, int64_t nb)
{
	__shared__ uint32_t sh[SCAN_THREADS];
	uint32_t carry = 0;
	for (int64_t c0 = 0; c0 < nb; c0 += SCAN_THREADS) {
		const int64_t j = c0 + threadIdx.x;
		const uint32_t v = j < nb ? bsum[j] : 0u;
		const uint32_t incl = block_incl_scan<OP>(v, sh);
		const uint32_t ex = threadIdx.x ? sh[threadIdx.x - 1] : 0u;
		if (j < nb) bsum[j] = scan_op<OP>(carry, OP == 0 ? incl - v : ex);
		const uint32_t tot = sh[SCAN_THREADS - 1];
		__syncthreads();
		carry = scan_op<OP>(carry, tot);
	}
}

// One-wave forms of the two scan passes (1024 elements per block, 16 per
// lane), for the second stream: blocks of one wave get dispatched between
// the one-wave blocks of a kernel running beside them.
template <int OP> __device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v)
{
	const int lane = threadIdx.x & 63;
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		const uint32_t o = (uint32_t)__shfl_up((int)v, d, 64);
		if (lane >= d) v = scan_op<OP>(v, o);
	}
	return v;
}

template <int OP>
__global__ __launch_bounds__(64)
void scan_block_w_kernel(const uint32_t *__restrict__ in, int64_t n_in, int64_t n, uint32_t *__restrict__ out,
                         uint32_t *__restrict__ bsum)
{
	constexpr int PER = SCAN_ITEMS / 64;
	const int64_t nb = (n + SCAN_ITEMS - 1) / SCAN_ITEMS;
	for (int64_t blk = blockIdx.x; blk < nb; blk += gridDim.x) {
		const int64_t b0 = blk * SCAN_ITEMS + (int64_t)threadIdx.x * PER;
		uint32_t v[PER];
		uint32_t acc = 0;
#pragma unroll
		for (int i = 0; i < PER; i++) {
			const int64_t j = b0 + i;
			v[i] = j < n_in ? in[j] : 0u;
			acc = scan_op<OP>(acc, v[i]);
		}
		const uint32_t incl = wave_incl_scan_u32<OP>(acc);
		const uint32_t before = (uint32_t)__shfl_up((int)incl, 1, 64);
		uint32_t run = OP == 0 ? incl - acc : (threadIdx.x ? before : 0u);
#pragma unroll
		for (int i = 0; i < PER; i++) {
			const int64_t j = b0 + i;
			if (OP == 0) {
				if (j < n) out[j] = run;   // exclusive
				run += v[i];
			} else {
				run = scan_op<OP>(run, v[i]);
				if (j < n) out[j] = run;   // inclusive
			}
		}
		if (threadIdx.x == 63) bsum[blk] = incl;
	}
}

template <int OP>
__global__ __launch_bounds__(64)
void scan_top_w_kernel(uint32_t *__restrict__ bsum, int64_t nb)
{
	uint32_t carry = 0;
	for (int64_t c0 = 0; c0 < nb; c0 += 64) {
		const int64_t j = c0 + threadIdx.x;
		const uint32_t v = j < nb ? bsum[j] : 0u;
		const uint32_t incl = wave_incl_scan_u32<OP>(v);
		const uint32_t before = (uint32_t)__shfl_up((int)incl, 1, 64);
		const uint32_t ex = threadIdx.x ? before : 0u;
		if (j < nb) bsum[j] = scan_op<OP>(carry, OP == 0 ? incl - v : ex);
		carry = scan_op<OP>(carry, (uint32_t)__shfl((int)incl, 63, 64));
	}
}
template __global__ void scan_block_w_kernel<0>(const uint32_t *, int64_t, int64_t, uint32_t *, uint32_t *);
template __global__ void scan_block_w_kernel<1>(const uint32_t *, int64_t, int64_t, uint32_t *, uint32_t *);
template __global__ void scan_top_w_kernel<0>(uint32_t *, int64_t);
template __global__ void scan_top_w_kernel<1>(uint32_t *, int64_t);

template <int OP>
__global__ void scan_add_kernel(uint32_t *__restrict__ out, int64_t n, const uint32_t *__restrict__ bsum)
{
	const int64_t gs = (int64_t)gridDim.x * blockDim.x;
	for (int64_t j0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j0 < n; j0 += 4 * gs) {
		uint32_t a[4], o[4];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int64_t j = j0 + u * gs;
			a[u] = j < n ? bsum[j / SCAN_ITEMS] : 0u;
			o[u] = j < n ? out[j] : 0u;
		}
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int64_t j = j0 + u * gs;
			if (j < n) out[j] = scan_op<OP>(a[u], o[u]);
		}
	}
}

template __global__ void scan_block_kernel<0>(const uint32_t *, int64_t, int64_t, uint32_t *, uint32_t *);
template __global__ void scan_block_kernel<1>(const uint32_t *, int64_t, int64_t, uint32_t *, uint32_t *);
template __global__ void scan_top_kernel<0>(uint32_t *, int64_t);
template __global__ void scan_top_kernel<1>(uint32_t *, int64_t);
template __global__ void scan_add_kernel<0>(uint32_t *, int64_t, const uint32_t *);
template __global__ void scan_add_kernel<1>(uint32_t *, int64_t, const uint32_t *);

// ------------------------------------------------------ position-ordered rows
// The rows (group, neighbour) of a chunk, e = base[g] + n, listed by patch
// position: counts per position, exclusive scan -> kstart, scatter, then each
// position's run sorted by e (runs are short), which is the stable order.

// rows of the chunk: *M = base[Gc] (the scan's total)
__global__ void count_rows_kernel(const uint32_t *__restrict__ epos, const uint32_t *__restrict__ Mp,
                                  int64_t Mub, uint32_t *__restrict__ cnt)
{
	const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (e < Mub && e < (int64_t)*Mp) atomicAdd(&cnt[epos[e]], 1u);
}

__global__ void scatter_rows_kernel(const uint32_t *__restrict__ epos, const uint32_t *__restrict__ Mp,
                                    int64_t Mub, const uint32_t *__restrict__ kstart, uint32_t *__restrict__ fill,
                                    uint32_t *__restrict__ svals)
{
	const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (e >= Mub || e >= (int64_t)*Mp) return;
	const uint32_t k = epos[e];
	svals[kstart[k] + atomicAdd(&fill[k], 1u)] = (uint32_t)e;
}

__global__ void sort_runs_kernel(const uint32_t *__restrict__ kstart, int64_t npos, uint32_t *__restrict__ svals)
{
	const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (k >= npos) return;
	const uint32_t a = kstart[k], b = kstart[k + 1];
	for (uint32_t i = a + 1; i < b; i++) {
		const uint32_t v = svals[i];
		uint32_t j = i;
		while (j > a && svals[j - 1] > v) {
			svals[j] = svals[j - 1];
			j--;
		}
		svals[j] = v;
	}
}

// running totals of the run (statistics read once at its end)
__global__ void add_rows_kernel(const uint32_t *__restrict__ Mp, unsigned long long *__restrict__ total)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) atomicAdd(total, (unsigned long long)*Mp);
}

// ------------------------------------------------------------ early-out fixup
// A group whose patches are all equal is not filtered (nldct.cpp:1986-1999,
// 2222-2235) and is aggregated with the weights the previous groups left in
// aggreWeights: slot n holds the weight of the latest filtered group of the
// same tile with more than n neighbours, 1 at the tile's start
// (nldct.cpp:749, 818).  lastf[g] = 1 + index of the latest filtered group
// at or before g (inclusive max-scan), 0 when none.  One thread per group of
// the chunk; rows of early groups hold the raw patches and are scaled here.
__global__ void early_fix_kernel(const uint8_t *__restrict__ early, const uint32_t *__restrict__ nsim,
                                 const float *__restrict__ gweight, const uint32_t *__restrict__ lastf,
                                 const uint32_t *__restrict__ gtile, const uint32_t *__restrict__ tile_first,
                                 const uint32_t *__restrict__ base, int g0, int Gc, const float *__restrict__ win,
                                 int C, int P, float *__restrict__ est, float *__restrict__ entry_w)
{
	const int gl = blockIdx.x * blockDim.x + threadIdx.x;
	if (gl >= Gc) return;
	const int g = g0 + gl;
	if (!early[g]) return;
	const int N = (int)nsim[g];
	const int tf = (int)tile_first[gtile[g]];
	for (int n = 0; n < N; n++) {
		float w = 1.f;
		int gp = g - 1 >= tf ? (int)lastf[g - 1] - 1 : -1;
		while (gp >= tf && (int)nsim[gp] <= n) gp = gp - 1 >= tf ? (int)lastf[gp - 1] - 1 : -1;
		if (gp >= tf) w = gweight[gp];
		const uint32_t row = base[gl] + (uint32_t)n;
		float *r = est + (size_t)row * C * P;
		for (int i = 0; i < C * P; i++) r[i] = (w * win[i % P]) * r[i];
		entry_w[row] = w;
	}
}

// v[g] = 1 + g for filtered groups, 0 for early ones (input of the max-scan)
__global__ void filtered_index_kernel(const uint8_t *__restrict__ early, int n, uint32_t *__restrict__ v)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g < n) v[g] = early[g] ? 0u : (uint32_t)g + 1u;
}

// Counts the groups whose first neighbour is not the reference patch itself
// (equal distances in heap order): the reference then clears another
// position of its processing mask (nldct.cpp:2325-2329).
__global__ void first_moved_kernel(const uint32_t *__restrict__ idx, const uint32_t *__restrict__ gpos, int np,
                                   int G, uint32_t *__restrict__ count)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g < G && idx[(size_t)g * np] != gpos[g]) atomicAdd(count, 1u);
}

// First neighbour of every group.
__global__ void first_index_kernel(const uint32_t *__restrict__ idx, int np, int G, uint32_t *__restrict__ first)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g < G) first[g] = idx[(size_t)g * np];
}

// Groups the reference does not process: no rows (listed in drop[0..n)).
__global__ void drop_groups_kernel(const uint32_t *__restrict__ drop, int n, uint32_t *__restrict__ nsim)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) nsim[drop[i]] = 0u;
}

// ------------------------------------------------------- stable radix sort
// One LSD pass of 8-bit digits (bits [shift, shift + 8)) over (key, value)
// pairs; the pair count n is on the device, nblk tiles of RS_TILE pairs
// bound it.  rs_hist_kernel: digit counts per tile, digit-major
// (hist[d * nblk + tile]); an exclusive scan of them gives every tile's
// first output slot per digit; rs_scatter_kernel: each tile writes its
// pairs there in input order (equal digits keep their order: stable).
__global__ __launch_bounds__(RS_THREADS)
void rs_hist_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ np, int shift, int nblk,
                    uint32_t *__restrict__ hist)
{
	constexpr int DPT = 256 / RS_THREADS;   // digits per thread
	__shared__ uint32_t h[256];
	const uint64_t n = *np;
	for (int tile = blockIdx.x; tile < nblk; tile += gridDim.x) {
#pragma unroll
		for (int i = 0; i < DPT; i++) h[threadIdx.x + i * RS_THREADS] = 0;
		__syncthreads();
		const uint64_t b0 = (uint64_t)tile * RS_TILE;
		if (b0 < n)
			for (int i0 = 0; i0 < RS_IPT; i0 += 16) {   // sixteen loads in flight, then their counts
				uint32_t kv[16];
#pragma unroll
				for (int i = 0; i < 16; i++) {
					const uint64_t k = b0 + (uint64_t)(i0 + i) * RS_THREADS + threadIdx.x;
					kv[i] = k < n ? keys[k] : 0u;
				}
#pragma unroll
				for (int i = 0; i < 16; i++) {
					const uint64_t k = b0 + (uint64_t)(i0 + i) * RS_THREADS + threadIdx.x;
					if (k < n) atomicAdd(&h[(kv[i] >> shift) & 255u], 1u);
				}
			}
		__syncthreads();
#pragma unroll
		for (int i = 0; i < DPT; i++) {
			const int d = threadIdx.x + i * RS_THREADS;
			hist[(size_t)d * nblk + tile] = h[d];
		}
		__syncthreads();
	}
}

// One wave per tile: rounds of 64 pairs in input order; the lanes of a
// round with the same digit take consecutive slots after the digit's
// running offset, and the lowest of them advances it.
__global__ __launch_bounds__(RS_THREADS)
void rs_scatter_kernel(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                       const uint32_t *__restrict__ np, int shift, int nblk, const uint32_t *__restrict__ offs,
                       uint32_t *__restrict__ kout, uint32_t *__restrict__ vout)
{
	static_assert(RS_THREADS == 64, "one wave per tile");
	constexpr int DPT = 256 / RS_THREADS, BATCH = 16;
	__shared__ uint32_t base[256];
	const int lane = threadIdx.x;
	const uint64_t n = *np;
	for (int tile = blockIdx.x; tile < nblk; tile += gridDim.x) {
		const uint64_t b0 = (uint64_t)tile * RS_TILE;
		if (b0 >= n) break;
#pragma unroll
		for (int i = 0; i < DPT; i++) base[lane + i * 64] = offs[(size_t)(lane + i * 64) * nblk + tile];
		__syncthreads();
		const uint64_t lt = (1ull << lane) - 1ull;
		for (int r0 = 0; r0 < RS_IPT; r0 += BATCH) {
			if (b0 + (uint64_t)r0 * 64 >= n) break;
			uint32_t key[BATCH], val[BATCH];
#pragma unroll
			for (int r = 0; r < BATCH; r++) {
				const uint64_t k = b0 + (uint64_t)(r0 + r) * 64 + lane;
				key[r] = k < n ? kin[k] : 0u;
				val[r] = k < n ? vin[k] : 0u;
			}
#pragma unroll
			for (int r = 0; r < BATCH; r++) {
				const bool valid = b0 + (uint64_t)(r0 + r) * 64 + lane < n;
				const uint32_t d = valid ? (key[r] >> shift) & 255u : 0u;
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
					kout[bd + rk] = key[r];
					vout[bd + rk] = val[r];
					if (rk == 0) base[d] = bd + (uint32_t)__popcll(m);
				}
				__syncthreads();
			}
		}
		__syncthreads();   // the next tile rewrites base
	}
}

} // namespace bm4dof
