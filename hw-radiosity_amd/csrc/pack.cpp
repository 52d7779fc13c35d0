#include "pack.h"

#include <parallel/algorithm>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace rgpu {

namespace {
// the bits of x spread to every third position
inline uint32_t spread3(uint32_t x)
{
	x &= 0x3ffu;
	x = (x | (x << 16)) & 0x030000ffu;
	x = (x | (x << 8)) & 0x0300f00fu;
	x = (x | (x << 4)) & 0x030c30c3u;
	x = (x | (x << 2)) & 0x09249249u;
	return x;
}
}  // namespace

void spatial_order(const Patch *pool, int n, const float *G0, const float *G1, int *order)
{
	std::vector<uint64_t> key(n);
#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		uint32_t code = 0;
		for (int k = 0; k < 3; k++) {
			const double span = (double)G1[k] - (double)G0[k];
			double u = span > 0 ? ((double)pool[i].c[k] - (double)G0[k]) / span * 1024.0 : 0.0;
			if (!(u >= 0.0)) u = 0.0;  // NaN too
			if (u > 1023.0) u = 1023.0;
			code |= spread3((uint32_t)u) << k;
		}
		key[i] = ((uint64_t)code << 32) | (uint32_t)i;
	}
	__gnu_parallel::sort(key.begin(), key.end());
#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) order[i] = (int)(uint32_t)key[i];
}

}  // namespace rgpu
