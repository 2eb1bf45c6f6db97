// Device side of the learned gain network: 66 -> 128 -> 128 -> 64, ReLU,
// sigmoid, bf16 weights, float32 accumulation on the matrix cores
// (v_mfma_f32_32x32x8_bf16: 32 rows x 32 outputs x 8 inputs per step).
//
// MFMA lane layout (one wave, lane l, h = l / 32):
//   A (32 x 8):  lane l holds A[l % 32][4h + v], v = 0..3
//   B (8 x 32):  lane l holds B[4h + v][l % 32]
//   D (32 x 32): lane l, element r holds D[8 (r / 4) + 4h + r % 4][l % 32]
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/bm4dof.h"

namespace bm4dof {

constexpr int MLP_IN = 66, MLP_H = 128, MLP_OUT = 64;
constexpr int MLP_K1 = 72;                      // layer-1 inputs padded to 9 steps of 8
constexpr int MLP_XS = 76;                      // LDS row stride of the bf16 features
constexpr int MLP_HS = 132;                     // LDS row stride of the bf16 hidden rows
constexpr int MLP_NPARAMS = MLP_IN * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * MLP_OUT + MLP_OUT;

// Inference copy of the network: weights as bf16, transposed (output-major,
// inputs contiguous) so that one lane's 4 B-operand values are adjacent.
struct MlpInfer {
	const uint16_t *w1t;   // [128][72]
	const uint16_t *w2t;   // [128][128]
	const uint16_t *w3t;   // [64][128]
	const float *b1, *b2, *b3;
};

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f32_to_bf16(float f)
{
	uint32_t u;
	__builtin_memcpy(&u, &f, 4);
	return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);   // nearest, ties to even
}
__host__ __device__ inline float bf16_to_f32(uint16_t b)
{
	const uint32_t u = (uint32_t)b << 16;
	float f;
	__builtin_memcpy(&f, &u, 4);
	return f;
}

__device__ inline mlp_bf16x4 mlp_load4(const uint16_t *p)
{
	const uint2 v = *reinterpret_cast<const uint2 *>(p);
	mlp_bf16x4 r;
	r[0] = (short)(v.x & 0xffffu);
	r[1] = (short)(v.x >> 16);
	r[2] = (short)(v.y & 0xffffu);
	r[3] = (short)(v.y >> 16);
	return r;
}

// Gains of 32 feature rows, one wave.
//   xb:   LDS, bf16 features [32][MLP_XS] (inputs 66..71 zero)
//   hb:   LDS, bf16 scratch [32][MLP_HS]
// (row strides padded so that the 32 rows of one operand read fall into
// distinct LDS banks, up to two-way)
//   gout: LDS, float gains [32][gs]
// Every lane of the wave must call it; it synchronises the LDS itself.
__device__ inline void mlp_rows32(const MlpInfer &net, const uint16_t *xb, uint16_t *hb, float *gout, int gs)
{
	const int l = threadIdx.x & 63, h = l >> 5, j = l & 31;
	mlp_f32x16 acc[4];
	// layer 1
#pragma unroll
	for (int t = 0; t < 4; t++) {
		mlp_f32x16 c = {};
#pragma unroll
		for (int s = 0; s < MLP_K1 / 8; s++) {
			const mlp_bf16x4 a = mlp_load4(xb + j * MLP_XS + 8 * s + 4 * h);
			const mlp_bf16x4 b = mlp_load4(net.w1t + (32 * t + j) * MLP_K1 + 8 * s + 4 * h);
			c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
		}
		acc[t] = c;
	}
#pragma unroll
	for (int t = 0; t < 4; t++) {
		const float bias = net.b1[32 * t + j];
#pragma unroll
		for (int r = 0; r < 16; r++) {
			const int row = 8 * (r >> 2) + 4 * h + (r & 3);
			hb[row * MLP_HS + 32 * t + j] = f32_to_bf16(fmaxf(acc[t][r] + bias, 0.f));
		}
	}
	__syncthreads();
	// layer 2 (results stay in registers until every lane has read layer 1)
#pragma unroll
	for (int t = 0; t < 4; t++) {
		mlp_f32x16 c = {};
#pragma unroll
		for (int s = 0; s < MLP_H / 8; s++) {
			const mlp_bf16x4 a = mlp_load4(hb + j * MLP_HS + 8 * s + 4 * h);
			const mlp_bf16x4 b = mlp_load4(net.w2t + (32 * t + j) * MLP_H + 8 * s + 4 * h);
			c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
		}
		acc[t] = c;
	}
	__syncthreads();
#pragma unroll
	for (int t = 0; t < 4; t++) {
		const float bias = net.b2[32 * t + j];
#pragma unroll
		for (int r = 0; r < 16; r++) {
			const int row = 8 * (r >> 2) + 4 * h + (r & 3);
			hb[row * MLP_HS + 32 * t + j] = f32_to_bf16(fmaxf(acc[t][r] + bias, 0.f));
		}
	}
	__syncthreads();
	// layer 3 + sigmoid
#pragma unroll
	for (int t = 0; t < 2; t++) {
		mlp_f32x16 c = {};
#pragma unroll
		for (int s = 0; s < MLP_H / 8; s++) {
			const mlp_bf16x4 a = mlp_load4(hb + j * MLP_HS + 8 * s + 4 * h);
			const mlp_bf16x4 b = mlp_load4(net.w3t + (32 * t + j) * MLP_H + 8 * s + 4 * h);
			c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
		}
		const float bias = net.b3[32 * t + j];
#pragma unroll
		for (int r = 0; r < 16; r++) {
			const int row = 8 * (r >> 2) + 4 * h + (r & 3);
			gout[row * gs + 32 * t + j] = 1.f / (1.f + expf(-(c[r] + bias)));
		}
	}
	__syncthreads();
}

// Host: view of a context's inference blob (mlp.hip); throws when no network
// is loaded.
MlpInfer mlp_net(::bm4dof_ctx *ctx);

} // namespace bm4dof
