#include "pack.h"

#include <parallel/algorithm>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace rgpu {

namespace {
inline void put4(float *o, float a, float b, float c, float d)
{
	o[0] = a;
	o[1] = b;
	o[2] = c;
	o[3] = d;
}
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
inline float bits(int x)
{
	float f;
	std::memcpy(&f, &x, 4);
	return f;
}
}  // namespace

void pack_leaf_items(const int *items, size_t ni, const Patch *pool, float *lnc, float *lse)
{
#pragma omp parallel for schedule(static)
	for (size_t k = 0; k < ni; k++) {
		const int id = items[k];
		const Patch &p = pool[id];
		put4(lnc + 8 * k, p.n[0], p.n[1], p.n[2], p.c[0]);
		put4(lnc + 8 * k + 4, p.c[1], p.c[2], bits(id), 0.f);
		for (int q = 0; q < 3; q++) put4(lse + 12 * k + 4 * q, p.se[q][0], p.se[q][1], p.se[q][2], p.se[q][3]);
	}
}

void pack_receivers(const Patch *pool, int n, float *vx, float *vy, float *vz, float *nrm, float *cA, float *rho,
                    float *acc, float *dacc, float *dB)
{
#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		const Patch &p = pool[i];
		for (int v = 0; v < 3; v++) {
			vx[3 * i + v] = p.p[v][0];
			vy[3 * i + v] = p.p[v][1];
			vz[3 * i + v] = p.p[v][2];
			put4(acc + 4 * (3 * i + v), p.accB[v][0], p.accB[v][1], p.accB[v][2], 0.f);
			put4(dacc + 4 * (3 * i + v), p.deltaaccB[v][0], p.deltaaccB[v][1], p.deltaaccB[v][2], 0.f);
		}
		put4(nrm + 4 * i, p.n[0], p.n[1], p.n[2], bits(p.parent));
		put4(cA + 4 * i, p.c[0], p.c[1], p.c[2], p.area);
		put4(rho + 4 * i, (float)p.Frgb[0], (float)p.Frgb[1], (float)p.Frgb[2], 0.f);
		put4(dB + 4 * i, p.deltaB[0], p.deltaB[1], p.deltaB[2], 0.f);
	}
}

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
