// Host side of the upload: the order in which the device takes receivers.
#pragma once

#include "patch.h"

namespace rgpu {

// order[0..n) = the slots sorted by the Z-order (Morton) code of their
// centroids on a 2^10 grid over [G0, G1], ties by slot; a permutation.
void spatial_order(const Patch *pool, int n, const float *G0, const float *G1, int *order);

}  // namespace rgpu
