// Shared host/device definitions of the MI355X BM4D-OF denoiser.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace bm4dof {

// Crop box of one spatial tile, global pixel coordinates; every tile spans
// all frames (the reference tiles only in space, lib_videot.hpp:854-916).
struct TileBox {
	int ox, oy, ex, ey;   // crop [ox,ex) x [oy,ey)
};

// Constants of one step, passed by value to the kernels.
struct StepConst {
	int W, H, F, C;       // video size
	int WH;               // W*H
	int WHC;              // W*H*C (frame stride of a C-channel video)
	int sPx, sPt, P;      // patch: px, pt, px*px*pt
	int sWx;              // spatial search window
	int wtf, wtb;         // temporal search range forward / backward
	int npmax;            // nSimilarPatches (neighbour slots per group)
	float tau2;           // tau*tau*px*px*pt (nldct.cpp:1157)
	float beta_sigma2;    // (beta*sigma)^2      (nldct.cpp:1808-1809)
	float betaM_sigma2;   // (betaMean*sigma)^2  (nldct.cpp:1810-1811)
	float sigma2;         // sigma^2             (nldct.cpp:1813)
	int order_inv;        // orderInvariance
	int use_flow;         // flows present
	int nW, wTmp, hTmp;   // tile grid: columns, interior width / height
	// predictive search (a VBM3D_SEARCH build): window of the frames after
	// the first, matches kept per frame, distance bias dsub^2*255
	int pred_w, pred_np;
	float pred_dsub;
};

// Pixel block of the ordered aggregation: RX x RY x RT pixels, frame blocks
// counted from tb0.
struct AggRegion { int RX, RY, RT, tb0; };

// Binned ordered aggregation: bins of 8 x 8 pixels x AGG_TB frames; a bin
// gathers the rows of at most AGG_BPOS patch positions ((7 + sPx)^2 x
// (AGG_TB + sPt - 1)).
#define AGG_TB 4
#define AGG_BPOS 1024

// stable radix sort passes (sortscan.hip): threads per block, pairs per
// thread and per block (tile)
#define RS_THREADS 64
#define RS_IPT 256
#define RS_TILE (RS_THREADS * RS_IPT)
// rows per tile of the first bin-sort pass (at most 8 pairs per row)
#define BIN_RT 4096

// Decode a 3-index position t*WH + y*W + x.
__host__ __device__ inline void pos_decode(uint32_t pos, int W, int WH, int &x, int &y, int &t)
{
	t = (int)(pos / (uint32_t)WH);
	const uint32_t r = pos - (uint32_t)t * (uint32_t)WH;
	y = (int)(r / (uint32_t)W);
	x = (int)(r - (uint32_t)y * (uint32_t)W);
}

} // namespace bm4dof
