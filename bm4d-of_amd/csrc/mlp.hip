// Learned gain network: training (forward, backward, Adam) and standalone
// inference.  No counterpart in the reference: the network predicts, from
// the step-1 Haar-domain coefficients of one group row, the step-2 Wiener
// gains y^2 / (y^2 + (beta sigma)^2) (src/core/nldct.cpp:2131) so that one
// pass replaces the two steps.
//
// Parameters: a flat float32 master vector theta (W1 [66][128], b1, W2
// [128][128], b2, W3 [128][64], b3; W[k][j] input-major), Adam moments m, v,
// and two derived copies refreshed after every update:
//   * wq: W1..W3 rounded to bf16 (as float), used by the training forward;
//   * the inference blob: bf16 W1..W3 transposed for the matrix cores, and
//     the float32 biases.
// Training forward rounds the layer inputs to bf16 exactly as the matrix
// cores consume them; backward is straight-through (rounding has slope 1).
// The training products are plain float32 tiled GEMMs (a few hundred MFLOP
// per step); inference runs on the bf16 matrix cores (mlp_dev.hpp).
#include "common.hpp"
#include "ctx.hpp"
#include "mlp_dev.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace bm4dof {

// offsets in theta
constexpr int OW1 = 0, OB1 = OW1 + MLP_IN * MLP_H, OW2 = OB1 + MLP_H, OB2 = OW2 + MLP_H * MLP_H,
              OW3 = OB2 + MLP_H, OB3 = OW3 + MLP_H * MLP_OUT;
static_assert(OB3 + MLP_OUT == MLP_NPARAMS, "parameter layout");
static_assert(MLP_NPARAMS == BM4DOF_MLP_NPARAMS, "parameter count");
// inference blob (uint16 units), biases after it as float32
constexpr size_t IW1 = 0, IW2 = IW1 + (size_t)MLP_H * MLP_K1, IW3 = IW2 + (size_t)MLP_H * MLP_H,
                 IWEND = IW3 + (size_t)MLP_OUT * MLP_H;
constexpr size_t IB_BYTES = ((IWEND * 2 + 255) / 256) * 256;   // bias area offset (bytes)
constexpr size_t INF_BYTES = IB_BYTES + (size_t)(MLP_H + MLP_H + MLP_OUT) * 4;

MlpInfer mlp_infer_view(const void *blob)
{
	MlpInfer n;
	const uint16_t *w = (const uint16_t *)blob;
	const float *b = (const float *)((const char *)blob + IB_BYTES);
	n.w1t = w + IW1;
	n.w2t = w + IW2;
	n.w3t = w + IW3;
	n.b1 = b;
	n.b2 = b + MLP_H;
	n.b3 = b + 2 * MLP_H;
	return n;
}

__device__ __forceinline__ float bf16r(float f) { return bf16_to_f32(f32_to_bf16(f)); }

// theta -> rounded training copies (wq, same layout as theta's weights) and
// the inference blob.
__global__ void mlp_pack_kernel(const float *__restrict__ theta, float *__restrict__ wq,
                                uint16_t *__restrict__ blob, float *__restrict__ bias)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < MLP_NPARAMS) {
		const bool is_bias = (i >= OB1 && i < OW2) || (i >= OB2 && i < OW3) || i >= OB3;
		wq[i] = is_bias ? theta[i] : bf16r(theta[i]);
	}
	if (i < (int)IWEND) {
		float v = 0.f;
		if (i < (int)IW2) {
			const int j = i / MLP_K1, k = i - j * MLP_K1;
			v = k < MLP_IN ? theta[OW1 + k * MLP_H + j] : 0.f;
		} else if (i < (int)IW3) {
			const int r = i - (int)IW2, j = r / MLP_H, k = r - j * MLP_H;
			v = theta[OW2 + k * MLP_H + j];
		} else {
			const int r = i - (int)IW3, j = r / MLP_H, k = r - j * MLP_H;
			v = theta[OW3 + k * MLP_OUT + j];
		}
		blob[i] = f32_to_bf16(v);
	}
	if (i < MLP_H) bias[i] = theta[OB1 + i];
	if (i < MLP_H) bias[MLP_H + i] = theta[OB2 + i];
	if (i < MLP_OUT) bias[2 * MLP_H + i] = theta[OB3 + i];
}

// C[m][n] (=|+=) sum_k opA[m][k] opB[k][n]; opA[m][k] = ta ? A[k*lda+m] : A[m*lda+k].
__global__ __launch_bounds__(256)
void mlp_gemm_kernel(int M, int N, int K, const float *__restrict__ A, int lda, int ta,
                     const float *__restrict__ B, int ldb, int tb, float *__restrict__ C, int ldc, int acc)
{
	__shared__ float As[16][17], Bs[16][17];
	const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
	const int m = blockIdx.y * 16 + ty, n = blockIdx.x * 16 + tx;
	float s = 0.f;
	for (int k0 = 0; k0 < K; k0 += 16) {
		{
			const int mm = blockIdx.y * 16 + ty, kk = k0 + tx;
			As[ty][tx] = (mm < M && kk < K) ? (ta ? A[(size_t)kk * lda + mm] : A[(size_t)mm * lda + kk]) : 0.f;
			const int k2 = k0 + ty, nn = blockIdx.x * 16 + tx;
			Bs[ty][tx] = (k2 < K && nn < N) ? (tb ? B[(size_t)nn * ldb + k2] : B[(size_t)k2 * ldb + nn]) : 0.f;
		}
		__syncthreads();
#pragma unroll
		for (int k = 0; k < 16; k++) s = fmaf(As[ty][k], Bs[k][tx], s);
		__syncthreads();
	}
	if (m < M && n < N) {
		if (acc) C[(size_t)m * ldc + n] += s;
		else C[(size_t)m * ldc + n] = s;
	}
}

__global__ void mlp_round_kernel(const float *__restrict__ x, int64_t n, float *__restrict__ y)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i < n) y[i] = bf16r(x[i]);
}

// Z += bias (kept for the ReLU mask), H = bf16(relu(Z))
__global__ void mlp_bias_relu_kernel(float *__restrict__ Z, const float *__restrict__ bias, int B, int N,
                                     float *__restrict__ H)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i >= (int64_t)B * N) return;
	const float z = Z[i] + bias[i % N];
	Z[i] = z;
	H[i] = bf16r(fmaxf(z, 0.f));
}

// gains, d loss / d Z3, and the loss (double sum of squared errors)
__global__ void mlp_out_kernel(const float *__restrict__ Z3, const float *__restrict__ b3,
                               const float *__restrict__ T, int B, float *__restrict__ G,
                               float *__restrict__ dZ3, double *__restrict__ sse)
{
	__shared__ double part[256];
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	double e2 = 0.0;
	if (i < (int64_t)B * MLP_OUT) {
		const float z = Z3[i] + b3[i % MLP_OUT];
		const float g = 1.f / (1.f + expf(-z));
		const float d = g - T[i];
		if (G) G[i] = g;
		dZ3[i] = 2.f * d / (float)((int64_t)B * MLP_OUT) * g * (1.f - g);
		e2 = (double)d * d;
	}
	part[threadIdx.x] = e2;
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
		__syncthreads();
	}
	if (threadIdx.x == 0) atomicAdd(sse, part[0]);
}

__global__ void mlp_relu_mask_kernel(float *__restrict__ dH, const float *__restrict__ Z, int64_t n)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i < n && !(Z[i] > 0.f)) dH[i] = 0.f;
}

__global__ void mlp_colsum_kernel(const float *__restrict__ D, int B, int N, float *__restrict__ out)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= N) return;
	float s = 0.f;
	for (int m = 0; m < B; m++) s += D[(size_t)m * N + n];
	out[n] = s;
}

__global__ void mlp_adam_kernel(float *__restrict__ theta, float *__restrict__ m, float *__restrict__ v,
                                const float *__restrict__ g, int n, float lr, float b1, float b2, float eps,
                                float c1, float c2)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.f - b1) * gi;
	const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// batch rows: idx = seed-driven LCG per (step, row) unless the batch is the
// whole set (then rows in order)
__global__ void mlp_gather_kernel(const float *__restrict__ X, const float *__restrict__ T, int64_t nrows,
                                  int B, uint32_t seed, float *__restrict__ xb, float *__restrict__ tb)
{
	const int r = blockIdx.x;
	if (r >= B) return;
	int64_t src = r;
	if ((int64_t)B != nrows) {
		uint64_t s = ((uint64_t)seed << 32) ^ (uint64_t)(r * 2654435761u + 0x9E3779B9u);
		s = s * 6364136223846793005ull + 1442695040888963407ull;
		s ^= s >> 33;
		s *= 0xff51afd7ed558ccdull;
		s ^= s >> 33;
		src = (int64_t)(s % (uint64_t)nrows);
	}
	for (int k = threadIdx.x; k < MLP_IN; k += blockDim.x) xb[(size_t)r * MLP_IN + k] = X[(size_t)src * MLP_IN + k];
	for (int k = threadIdx.x; k < MLP_OUT; k += blockDim.x) tb[(size_t)r * MLP_OUT + k] = T[(size_t)src * MLP_OUT + k];
}

// Standalone inference of B rows (one wave per 32 rows).
__global__ __launch_bounds__(64)
void mlp_infer_kernel(const float *__restrict__ X, int B, MlpInfer net, float *__restrict__ G)
{
	__shared__ __align__(16) uint16_t xb[32 * MLP_XS];
	__shared__ __align__(16) uint16_t hb[32 * MLP_HS];
	__shared__ float gs[32 * MLP_OUT];
	const int r0 = blockIdx.x * 32;
	for (int i = threadIdx.x; i < 32 * MLP_XS; i += 64) {
		const int r = i / MLP_XS, k = i - r * MLP_XS;
		const float v = (r0 + r < B && k < MLP_IN) ? X[(size_t)(r0 + r) * MLP_IN + k] : 0.f;
		xb[i] = f32_to_bf16(v);
	}
	__syncthreads();
	mlp_rows32(net, xb, hb, gs, MLP_OUT);
	for (int i = threadIdx.x; i < 32 * MLP_OUT; i += 64) {
		const int r = i / MLP_OUT;
		if (r0 + r < B) G[(size_t)r0 * MLP_OUT + i] = gs[i];
	}
}

namespace {

inline int nb(int64_t n, int b) { return (int)((n + b - 1) / b); }

void gemm(bm4dof_ctx *ctx, int M, int N, int K, const float *A, int lda, int ta, const float *B, int ldb,
          int tb, float *C, int ldc, int acc)
{
	hipLaunchKernelGGL(mlp_gemm_kernel, dim3(nb(N, 16), nb(M, 16)), dim3(256), 0, ctx->stream, M, N, K, A, lda,
	                   ta, B, ldb, tb, C, ldc, acc);
	HIPCHK(hipGetLastError());
}

void repack(bm4dof_ctx *ctx)
{
	float *theta = ctx->mlp_theta.as<float>(MLP_NPARAMS);
	float *wq = ctx->mlp_wq.as<float>(MLP_NPARAMS);
	char *blob = ctx->mlp_inf.as<char>(INF_BYTES);
	hipLaunchKernelGGL(mlp_pack_kernel, dim3(nb(std::max<int64_t>(MLP_NPARAMS, IWEND), 256)), dim3(256), 0,
	                   ctx->stream, theta, wq, (uint16_t *)blob, (float *)(blob + IB_BYTES));
	HIPCHK(hipGetLastError());
}

void need_net(bm4dof_ctx *ctx)
{
	if (!ctx->mlp_ready) throw ArgError("no network loaded: call bm4dof_mlp_set first");
}

// Forward + backward of one batch already in the workspace (xb, tb), B rows.
// Leaves the gradient in ctx->mlp_grad and adds the squared error to *sse.
void fwd_bwd(bm4dof_ctx *ctx, const float *xb, const float *tb, int B, double *sse)
{
	const float *wq = ctx->mlp_wq.as<float>(MLP_NPARAMS);
	float *grad = ctx->mlp_grad.as<float>(MLP_NPARAMS);
	const size_t need = (size_t)B * (MLP_IN + 4 * MLP_H + 2 * MLP_H + 3 * MLP_OUT);
	float *ws = ctx->mlp_ws.as<float>(need);
	float *X0 = ws, *Z1 = X0 + (size_t)B * MLP_IN, *H1 = Z1 + (size_t)B * MLP_H, *Z2 = H1 + (size_t)B * MLP_H,
	      *H2 = Z2 + (size_t)B * MLP_H, *dH = H2 + (size_t)B * MLP_H, *dH2 = dH + (size_t)B * MLP_H,
	      *Z3 = dH2 + (size_t)B * MLP_H, *dZ3 = Z3 + (size_t)B * MLP_OUT;
	const int T = 256;
	hipLaunchKernelGGL(mlp_round_kernel, dim3(nb((int64_t)B * MLP_IN, T)), dim3(T), 0, ctx->stream, xb,
	                   (int64_t)B * MLP_IN, X0);
	gemm(ctx, B, MLP_H, MLP_IN, X0, MLP_IN, 0, wq + OW1, MLP_H, 0, Z1, MLP_H, 0);
	hipLaunchKernelGGL(mlp_bias_relu_kernel, dim3(nb((int64_t)B * MLP_H, T)), dim3(T), 0, ctx->stream, Z1,
	                   wq + OB1, B, MLP_H, H1);
	gemm(ctx, B, MLP_H, MLP_H, H1, MLP_H, 0, wq + OW2, MLP_H, 0, Z2, MLP_H, 0);
	hipLaunchKernelGGL(mlp_bias_relu_kernel, dim3(nb((int64_t)B * MLP_H, T)), dim3(T), 0, ctx->stream, Z2,
	                   wq + OB2, B, MLP_H, H2);
	gemm(ctx, B, MLP_OUT, MLP_H, H2, MLP_H, 0, wq + OW3, MLP_OUT, 0, Z3, MLP_OUT, 0);
	hipLaunchKernelGGL(mlp_out_kernel, dim3(nb((int64_t)B * MLP_OUT, T)), dim3(T), 0, ctx->stream, Z3,
	                   wq + OB3, tb, B, (float *)nullptr, dZ3, sse);
	HIPCHK(hipGetLastError());
	// layer 3
	gemm(ctx, MLP_H, MLP_OUT, B, H2, MLP_H, 1, dZ3, MLP_OUT, 0, grad + OW3, MLP_OUT, 0);
	hipLaunchKernelGGL(mlp_colsum_kernel, dim3(1), dim3(MLP_OUT), 0, ctx->stream, dZ3, B, MLP_OUT, grad + OB3);
	gemm(ctx, B, MLP_H, MLP_OUT, dZ3, MLP_OUT, 0, wq + OW3, MLP_OUT, 1, dH2, MLP_H, 0);
	hipLaunchKernelGGL(mlp_relu_mask_kernel, dim3(nb((int64_t)B * MLP_H, T)), dim3(T), 0, ctx->stream, dH2, Z2,
	                   (int64_t)B * MLP_H);
	// layer 2
	gemm(ctx, MLP_H, MLP_H, B, H1, MLP_H, 1, dH2, MLP_H, 0, grad + OW2, MLP_H, 0);
	hipLaunchKernelGGL(mlp_colsum_kernel, dim3(1), dim3(MLP_H), 0, ctx->stream, dH2, B, MLP_H, grad + OB2);
	gemm(ctx, B, MLP_H, MLP_H, dH2, MLP_H, 0, wq + OW2, MLP_H, 1, dH, MLP_H, 0);
	hipLaunchKernelGGL(mlp_relu_mask_kernel, dim3(nb((int64_t)B * MLP_H, T)), dim3(T), 0, ctx->stream, dH, Z1,
	                   (int64_t)B * MLP_H);
	// layer 1
	gemm(ctx, MLP_IN, MLP_H, B, X0, MLP_IN, 1, dH, MLP_H, 0, grad + OW1, MLP_H, 0);
	hipLaunchKernelGGL(mlp_colsum_kernel, dim3(1), dim3(MLP_H), 0, ctx->stream, dH, B, MLP_H, grad + OB1);
	HIPCHK(hipGetLastError());
}

} // namespace

MlpInfer mlp_net(bm4dof_ctx *ctx)
{
	need_net(ctx);
	return mlp_infer_view(ctx->mlp_inf.as<char>(INF_BYTES));
}

} // namespace bm4dof

using namespace bm4dof;

extern "C" {

__attribute__((visibility("default")))
int bm4dof_mlp_set(bm4dof_ctx *ctx, const float *theta)
{
	if (!ctx) return BM4DOF_EINVAL;
	return guarded(ctx, [&] {
		if (!theta) throw ArgError("theta is NULL");
		HIPCHK(hipSetDevice(ctx->device));
		float *d = ctx->mlp_theta.as<float>(MLP_NPARAMS);
		float *m = ctx->mlp_m.as<float>(MLP_NPARAMS);
		float *v = ctx->mlp_v.as<float>(MLP_NPARAMS);
		HIPCHK(hipMemcpyAsync(d, theta, (size_t)MLP_NPARAMS * 4, hipMemcpyHostToDevice, ctx->stream));
		HIPCHK(hipMemsetAsync(m, 0, (size_t)MLP_NPARAMS * 4, ctx->stream));
		HIPCHK(hipMemsetAsync(v, 0, (size_t)MLP_NPARAMS * 4, ctx->stream));
		ctx->mlp_step = 0;
		repack(ctx);
		HIPCHK(hipStreamSynchronize(ctx->stream));
		ctx->mlp_ready = true;
	});
}

__attribute__((visibility("default")))
int bm4dof_mlp_get(const bm4dof_ctx *cctx, float *theta)
{
	bm4dof_ctx *ctx = const_cast<bm4dof_ctx *>(cctx);
	if (!ctx) return BM4DOF_EINVAL;
	return guarded(ctx, [&] {
		need_net(ctx);
		if (!theta) throw ArgError("theta is NULL");
		HIPCHK(hipSetDevice(ctx->device));
		HIPCHK(hipMemcpyAsync(theta, ctx->mlp_theta.as<float>(MLP_NPARAMS), (size_t)MLP_NPARAMS * 4,
		                      hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
	});
}

__attribute__((visibility("default")))
int bm4dof_mlp_forward(bm4dof_ctx *ctx, const float *x, int64_t rows, float *gains, int on_device)
{
	if (!ctx) return BM4DOF_EINVAL;
	return guarded(ctx, [&] {
		need_net(ctx);
		if (rows < 0 || (rows > 0 && (!x || !gains))) throw ArgError("bad rows or NULL buffer");
		if (rows == 0) return;
		if (rows > ((int64_t)1 << 31) - 64) throw ArgError("too many rows", BM4DOF_EUNSUPP);
		HIPCHK(hipSetDevice(ctx->device));
		const float *dx = x;
		float *dg = gains;
		if (!on_device) {
			float *bx = ctx->io_noisy.as<float>((size_t)rows * MLP_IN);
			HIPCHK(hipMemcpyAsync(bx, x, (size_t)rows * MLP_IN * 4, hipMemcpyHostToDevice, ctx->stream));
			dx = bx;
			dg = ctx->io_final.as<float>((size_t)rows * MLP_OUT);
		}
		const MlpInfer net = mlp_infer_view(ctx->mlp_inf.as<char>(INF_BYTES));
		hipLaunchKernelGGL(mlp_infer_kernel, dim3(nb(rows, 32)), dim3(64), 0, ctx->stream, dx, (int)rows, net, dg);
		HIPCHK(hipGetLastError());
		if (!on_device)
			HIPCHK(hipMemcpyAsync(gains, dg, (size_t)rows * MLP_OUT * 4, hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
	});
}

__attribute__((visibility("default")))
int bm4dof_mlp_grad(bm4dof_ctx *ctx, const float *x, const float *t, int rows, float *loss, float *grad)
{
	if (!ctx) return BM4DOF_EINVAL;
	return guarded(ctx, [&] {
		need_net(ctx);
		if (rows <= 0 || !x || !t) throw ArgError("bad rows or NULL buffer");
		HIPCHK(hipSetDevice(ctx->device));
		float *bx = ctx->mlp_xb.as<float>((size_t)rows * MLP_IN);
		float *bt = ctx->mlp_tb.as<float>((size_t)rows * MLP_OUT);
		double *sse = ctx->mlp_sse.as<double>(1);
		HIPCHK(hipMemcpyAsync(bx, x, (size_t)rows * MLP_IN * 4, hipMemcpyHostToDevice, ctx->stream));
		HIPCHK(hipMemcpyAsync(bt, t, (size_t)rows * MLP_OUT * 4, hipMemcpyHostToDevice, ctx->stream));
		HIPCHK(hipMemsetAsync(sse, 0, 8, ctx->stream));
		fwd_bwd(ctx, bx, bt, rows, sse);
		double h_sse = 0.0;
		HIPCHK(hipMemcpyAsync(&h_sse, sse, 8, hipMemcpyDeviceToHost, ctx->stream));
		if (grad)
			HIPCHK(hipMemcpyAsync(grad, ctx->mlp_grad.as<float>(MLP_NPARAMS), (size_t)MLP_NPARAMS * 4,
			                      hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(hipStreamSynchronize(ctx->stream));
		if (loss) *loss = (float)(h_sse / ((double)rows * MLP_OUT));
	});
}

__attribute__((visibility("default")))
int bm4dof_mlp_train(bm4dof_ctx *ctx, const float *x, const float *t, int64_t rows, int batch, int steps,
                     float lr, unsigned seed, float *losses)
{
	if (!ctx) return BM4DOF_EINVAL;
	return guarded(ctx, [&] {
		need_net(ctx);
		if (rows <= 0 || batch <= 0 || steps < 0 || !x || !t) throw ArgError("bad training arguments");
		if (batch > rows) throw ArgError("batch larger than the data set");
		HIPCHK(hipSetDevice(ctx->device));
		float *theta = ctx->mlp_theta.as<float>(MLP_NPARAMS);
		float *m = ctx->mlp_m.as<float>(MLP_NPARAMS);
		float *v = ctx->mlp_v.as<float>(MLP_NPARAMS);
		float *bx = ctx->mlp_xb.as<float>((size_t)batch * MLP_IN);
		float *bt = ctx->mlp_tb.as<float>((size_t)batch * MLP_OUT);
		double *sse = ctx->mlp_sse.as<double>((size_t)std::max(steps, 1));
		HIPCHK(hipMemsetAsync(sse, 0, (size_t)std::max(steps, 1) * 8, ctx->stream));
		const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
		for (int s = 0; s < steps; s++) {
			hipLaunchKernelGGL(mlp_gather_kernel, dim3(batch), dim3(64), 0, ctx->stream, x, t, rows, batch,
			                   seed + (uint32_t)ctx->mlp_step, bx, bt);
			fwd_bwd(ctx, bx, bt, batch, sse + s);
			ctx->mlp_step++;
			const float c1 = 1.f - powf(b1, (float)ctx->mlp_step), c2 = 1.f - powf(b2, (float)ctx->mlp_step);
			hipLaunchKernelGGL(mlp_adam_kernel, dim3(nb(MLP_NPARAMS, 256)), dim3(256), 0, ctx->stream, theta, m, v,
			                   ctx->mlp_grad.as<float>(MLP_NPARAMS), MLP_NPARAMS, lr, b1, b2, eps, c1, c2);
			repack(ctx);
		}
		if (losses && steps > 0) {
			std::vector<double> h(steps);
			HIPCHK(hipMemcpyAsync(h.data(), sse, (size_t)steps * 8, hipMemcpyDeviceToHost, ctx->stream));
			HIPCHK(hipStreamSynchronize(ctx->stream));
			for (int s = 0; s < steps; s++) losses[s] = (float)(h[s] / ((double)batch * MLP_OUT));
		}
		HIPCHK(hipStreamSynchronize(ctx->stream));
	});
}

} // extern "C"
