// The visibility octree, built on the host exactly as the reference builds
// it (raycast.cpp:168-233) and flattened for the device.
#pragma once

#include <vector>

#include "patch.h"

namespace rgpu {

struct FlatOctree {
	// reference-equivalent facts
	int n_nodes = 0;      // TreeNodeStorePtr
	int n_list = 0;       // TriListStorePtr (items plus one -1 per leaf)
	float min_grid = 0;   // MinGridLen
	int grid_num = 0;     // GridNum
	// device form
	int gt = 0;                   // cells per axis of the lookup table (grid_num + 2)
	std::vector<int> cell_leaf;   // gt^3: leaf reached by TreeNodeNum's descent from (gx, gy, gz)
	std::vector<float> leaf_box;  // 8 floats per leaf: g0.xyz, begin, g1.xyz, count|depth<<24
	std::vector<int> items;       // leaf triangle lists, ascending original index
	int n_leaves = 0;
};

// Octree over the originals [0, n) of `pool` inside [G0, G1].
FlatOctree build_octree(const Patch *pool, int n, const float *G0, const float *G1, bool fma, int threads);

}  // namespace rgpu
