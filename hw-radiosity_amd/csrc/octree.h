// The visibility octree, built on the host exactly as the reference builds
// it (raycast.cpp:168-233) and flattened for the device.
#pragma once

#include <string>
#include <vector>

#include "patch.h"

namespace rgpu {

// TreeNode (rad.h:41-45): 60 bytes, as TreeNodeStore holds it.
struct RefNode {
	float g0[3], g1[3];
	int sub[8];
	int list;
};
static_assert(sizeof(RefNode) == 60, "TreeNode layout");

struct FlatOctree {
	// the reference's own arrays (raycast.cpp:9-13) as BuildTree leaves them:
	// nodes in depth-first preorder, each leaf's list followed by -1 in
	// TriListStore in the same order
	std::vector<RefNode> ref_nodes;
	std::vector<int> ref_lists;
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
	// geometric point location (the explicit form-factor matrix's segment
	// visibility, made by build_segment_index): fine_r = 2^(deepest leaf
	// level) cells per axis over [G0, G1], fine_leaf[(a * fine_r + b) *
	// fine_r + c] = the leaf holding that cell
	int fine_r = 0;
	std::vector<int> fine_leaf;
	// complete leaf lists for that walk: seg_items[seg_begin[l] ..
	// seg_begin[l + 1]) = every original whose bounding box (grown by 1e-3
	// of a fine cell) touches a cell of leaf l, ascending
	std::vector<int> seg_begin, seg_items;
};

// Octree over the originals [0, n) of `pool` inside [G0, G1].
FlatOctree build_octree(const Patch *pool, int n, const float *G0, const float *G1, bool fma, int threads);

// The segment-visibility index (fine_r, fine_leaf, seg_begin, seg_items) of
// T over the originals [0, n); once per tree.
int build_segment_index(FlatOctree &T, const Patch *pool, int n, const float *G0, const float *G1, std::string &err);

}  // namespace rgpu
