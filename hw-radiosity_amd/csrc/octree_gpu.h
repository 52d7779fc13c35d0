// The visibility octree built on the device (raycast.cpp:168-233): the same
// FlatOctree as build_octree, downloaded to the host.
#pragma once

#include <string>

#include "octree.h"

namespace rgpu {

// Octree over the originals [0, n) of `pool` inside [G0, G1], built on the
// current device on `stream` (a hipStream_t).  Returns 0, or -1 with err.
int build_octree_gpu(void *stream, const Patch *pool, int n, const float *G0, const float *G1, bool fma,
                     FlatOctree &T, std::string &err);

}  // namespace rgpu
