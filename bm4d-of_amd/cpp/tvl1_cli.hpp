// Scale-count cap shared by the tvl1flow and tvl1flow-seq tools.
#pragma once

#include <cmath>

// Number of pyramid scales so that the coarsest image stays at least 16
// pixels across (lib/tvl1flow/main.c:159-163): N = 1 + log(hypot(nx, ny) /
// 16) / log(1 / zfactor) in double precision (1 / zfactor in float), kept as
// a float, and nscales = N (truncated) when N < nscales.
inline int tvl1_cap_nscales(int nx, int ny, float zfactor, int nscales)
{
	const float inv = 1.0f / zfactor;
	const float N = (float)(1.0 + std::log(std::hypot((double)nx, (double)ny) / 16.0) / std::log((double)inv));
	if (N < (float)nscales) nscales = (int)N;
	return nscales;
}
