// Host packing of the prepared scene into the device layouts (float4
// groups, written as 4 consecutive floats), in parallel on the host.
#pragma once

#include <cstddef>

#include "patch.h"

namespace rgpu {

// Per list entry k of the octree (items[k] = an original's slot):
//   lnc[8k..8k+7]  = n.xyz, c.x | c.y, c.z, slot bits, 0
//   lse[12k..+11]  = the three side planes se[q] (xyz, w)
void pack_leaf_items(const int *items, size_t ni, const Patch *pool, float *lnc, float *lse);

// Per slot i: vertex coordinates vx/vy/vz[3i + v]; nrm = n.xyz, parent bits;
// cA = c.xyz, area; rho = Frgb; acc/dacc[3i + v] = accB/deltaaccB[v], 0;
// dB = deltaB, 0.
void pack_receivers(const Patch *pool, int n, float *vx, float *vy, float *vz, float *nrm, float *cA, float *rho,
                    float *acc, float *dacc, float *dB);

// order[0..n) = the slots sorted by the Z-order (Morton) code of their
// centroids on a 2^10 grid over [G0, G1], ties by slot; a permutation.
void spatial_order(const Patch *pool, int n, const float *G0, const float *G1, int *order);

}  // namespace rgpu
