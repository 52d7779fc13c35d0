// Octree build (raycast.cpp:24-233) and its flattening for the device.
//
// The tree is the reference's: a node is a leaf when it holds at most 10
// triangles or sits at depth 6; otherwise it has 8 children, child bit j set
// meaning the upper half in dimension j, split at the float midpoint; a
// triangle is listed in every child it overlaps by the CrossOver/Clip test.
// Subtrees are built in parallel (OpenMP tasks) and stitched; the node
// numbering does not matter to any result.
//
// The device never walks the tree.  The reference's TreeNodeNum turns a
// point into integer cell coordinates gnum = (int)((q - G0) / MinGridLen)
// and descends with gsize = GridNum >> level; the leaf it reaches depends on
// the three integers only, so the descent is tabulated once for every
// (gx, gy, gz) in [0, GridNum + 1]^3 and a device lookup is a single load.
#include "octree.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

namespace rgpu {
namespace {

// Clip (raycast.cpp:24-60): does segment p0-p1 meet the (x, y) rectangle?
bool segment_meets_rect(const float *p0, const float *p1, const float *g0, const float *g1, int x, int y)
{
	const int out0 = (p0[x] < g0[x]) | (p0[x] > g1[x]) << 1 | (p0[y] < g0[y]) << 2 | (p0[y] > g1[y]) << 3;
	const int out1 = (p1[x] < g0[x]) | (p1[x] > g1[x]) << 1 | (p1[y] < g0[y]) << 2 | (p1[y] > g1[y]) << 3;
	if (out0 & out1) return false;
	if (!(out0 | out1)) return true;
	const float a = p1[y] - p0[y], b = p1[x] - p0[x];
	const float side00 = a * (g0[x] - p0[x]) - b * (g0[y] - p0[y]);
	if (side00 == 0) return true;
	const bool neg = side00 < 0;
	if (((a * (g0[x] - p0[x]) - b * (g1[y] - p0[y])) < 0) != neg) return true;
	if (((a * (g1[x] - p0[x]) - b * (g0[y] - p0[y])) < 0) != neg) return true;
	if (((a * (g1[x] - p0[x]) - b * (g1[y] - p0[y])) < 0) != neg) return true;
	return false;
}

// InterpVector (raycast.cpp:91-100): the point of segment lo-hi where
// coordinate d equals g.
template <bool F> void point_at(const float *hi, const float *lo, float g, int d, float *out)
{
	const float span = hi[d] - lo[d];
	if (span == 0.0f) {
		std::memcpy(out, hi, 12);
		return;
	}
	const float wa = (g - lo[d]) / span, wb = (hi[d] - g) / span;
	for (int k = 0; k < 3; k++) out[k] = Arith<F>::lerp2(wa, hi[k], wb, lo[k]);
}

// CrossOver (raycast.cpp:109-155).  The three vertex pointers stay sorted
// across dimensions exactly as the reference's in-place sort leaves them.
template <bool F> bool overlaps(const Patch &t, const float *g0, const float *g1)
{
	const float *hi = t.p[0], *mid = t.p[1], *lo = t.p[2];
	float a[3], b[3];
	for (int x = 0; x < 3; x++) {
		if (hi[x] < mid[x]) std::swap(hi, mid);
		if (hi[x] < lo[x]) std::swap(hi, lo);
		if (mid[x] < lo[x]) std::swap(mid, lo);
		const int u = (x + 1) % 3, w = (x + 2) % 3;
		if (g1[x] < lo[x]) return false;
		point_at<F>(hi, lo, g1[x], x, a);
		if (g1[x] < mid[x]) {
			point_at<F>(mid, lo, g1[x], x, b);
			return segment_meets_rect(a, b, g0, g1, u, w);
		}
		if (g1[x] <= hi[x]) {
			point_at<F>(hi, mid, g1[x], x, b);
			return segment_meets_rect(a, b, g0, g1, u, w);
		}
		if (g0[x] > hi[x]) return false;
		point_at<F>(hi, lo, g0[x], x, a);
		if (g0[x] > mid[x]) {
			point_at<F>(hi, mid, g0[x], x, b);
			return segment_meets_rect(a, b, g0, g1, u, w);
		}
		if (g0[x] >= lo[x]) {
			point_at<F>(mid, lo, g0[x], x, b);
			return segment_meets_rect(a, b, g0, g1, u, w);
		}
	}
	return true;
}

// A sufficient condition for CrossOver to say no: the triangle's bounding
// box misses the node's box by more than rounding can move a point that
// InterpVector computes on one of its edges (such a point lies within the
// edge's range up to a few ulps; the margin is ~100 ulps of the larger
// coordinate plus an absolute floor).  Every test CrossOver makes compares
// such points, or vertices, with the box's planes, so beyond the margin each
// test lands on the same side it would for the bounding box.
struct TriBox {
	float lo[3], hi[3];
};
inline bool surely_apart(const TriBox &b, const float *g0, const float *g1)
{
	for (int k = 0; k < 3; k++) {
		const float m0 = 1e-5f * (std::fabs(g0[k]) + std::fabs(b.hi[k])) + 1e-30f;
		const float m1 = 1e-5f * (std::fabs(g1[k]) + std::fabs(b.lo[k])) + 1e-30f;
		if (b.hi[k] < g0[k] - m0 || b.lo[k] > g1[k] + m1) return true;
	}
	return false;
}

struct BNode {
	float g0[3], g1[3];
	int sub[8];
	int leaf;   // leaf number within its subtree, -1 for an inner node
	int depth;
};

struct Subtree {
	std::vector<BNode> nodes;
	std::vector<int> item_begin, item_count;  // per leaf
	std::vector<int> items;
	float min_grid;
};

template <bool F> struct Builder {
	const Patch *pool;
	const TriBox *box;
	int task_depth;

	int build(Subtree &st, const std::vector<int> &idx, int level, const float *g0, const float *g1)
	{
		const int id = (int)st.nodes.size();
		st.nodes.push_back(BNode());
		BNode nd;
		std::memcpy(nd.g0, g0, 12);
		std::memcpy(nd.g1, g1, 12);
		nd.depth = level;
		nd.leaf = -1;
		if ((int)idx.size() <= kMaxTriPerNode || level >= kMaxLevel) {
			if (g1[0] - g0[0] < st.min_grid) st.min_grid = g1[0] - g0[0];
			for (int i = 0; i < 8; i++) nd.sub[i] = -1;
			nd.leaf = (int)st.item_begin.size();
			st.item_begin.push_back((int)st.items.size());
			st.item_count.push_back((int)idx.size());
			st.items.insert(st.items.end(), idx.begin(), idx.end());
			st.nodes[id] = nd;
			return id;
		}
		float mid[3];
		for (int k = 0; k < 3; k++) mid[k] = (float)(0.5 * (double)g0[k] + 0.5 * (double)g1[k]);
		float q0[8][3], q1[8][3];
		std::vector<int> part[8];
		for (int c = 0; c < 8; c++) {
			for (int k = 0; k < 3; k++) {
				const bool upper = (c >> k) & 1;
				q0[c][k] = upper ? mid[k] : g0[k];
				q1[c][k] = upper ? g1[k] : mid[k];
			}
		}
		if (level < task_depth) {
			// children in parallel, each into its own subtree, then stitched
			std::unique_ptr<Subtree> kid[8];
			for (int c = 0; c < 8; c++) {
				kid[c].reset(new Subtree());
				kid[c]->min_grid = st.min_grid;
			}
			for (int c = 0; c < 8; c++) {
#pragma omp task shared(kid, q0, q1, idx) firstprivate(c)
				{
					std::vector<int> sel;
					sel.reserve(idx.size());
					for (int t : idx)
						if (!surely_apart(box[t], q0[c], q1[c]) && overlaps<F>(pool[t], q0[c], q1[c])) sel.push_back(t);
					build(*kid[c], sel, level + 1, q0[c], q1[c]);
				}
			}
#pragma omp taskwait
			for (int c = 0; c < 8; c++) {
				const int off = (int)st.nodes.size(), loff = (int)st.item_begin.size(), ioff = (int)st.items.size();
				for (BNode b : kid[c]->nodes) {
					for (int s = 0; s < 8; s++)
						if (b.sub[s] >= 0) b.sub[s] += off;
					if (b.leaf >= 0) b.leaf += loff;
					st.nodes.push_back(b);
				}
				for (size_t l = 0; l < kid[c]->item_begin.size(); l++) {
					st.item_begin.push_back(kid[c]->item_begin[l] + ioff);
					st.item_count.push_back(kid[c]->item_count[l]);
				}
				st.items.insert(st.items.end(), kid[c]->items.begin(), kid[c]->items.end());
				st.min_grid = std::min(st.min_grid, kid[c]->min_grid);
				nd.sub[c] = off;
			}
		} else {
			for (int c = 0; c < 8; c++) {
				std::vector<int> sel;
				sel.reserve(idx.size());
				for (int t : idx)
					if (!surely_apart(box[t], q0[c], q1[c]) && overlaps<F>(pool[t], q0[c], q1[c])) sel.push_back(t);
				nd.sub[c] = build(st, sel, level + 1, q0[c], q1[c]);
			}
		}
		st.nodes[id] = nd;
		return id;
	}
};

}  // namespace

FlatOctree build_octree(const Patch *pool, int n, const float *G0, const float *G1, bool fma, int threads)
{
	Subtree st;
	st.min_grid = G1[0] - G0[0];
	std::vector<int> all(n);
	for (int i = 0; i < n; i++) all[i] = i;
	const int nthreads = threads > 0 ? threads : omp_get_max_threads();
	std::vector<TriBox> box(n);
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int i = 0; i < n; i++)
		for (int k = 0; k < 3; k++) {
			const float a = pool[i].p[0][k], b = pool[i].p[1][k], c = pool[i].p[2][k];
			const bool nan = std::isnan(a) || std::isnan(b) || std::isnan(c);  // never rejected early
			box[i].lo[k] = nan ? -INFINITY : std::fmin(a, std::fmin(b, c));
			box[i].hi[k] = nan ? INFINITY : std::fmax(a, std::fmax(b, c));
		}
#pragma omp parallel num_threads(nthreads)
#pragma omp single
	{
		if (fma) Builder<true>{pool, box.data(), 3}.build(st, all, 0, G0, G1);
		else Builder<false>{pool, box.data(), 3}.build(st, all, 0, G0, G1);
	}
	FlatOctree T;
	T.n_nodes = (int)st.nodes.size();
	T.n_leaves = (int)st.item_begin.size();
	T.n_list = (int)st.items.size() + T.n_leaves;
	T.min_grid = st.min_grid;
	// GridNum; bounds that are not finite leave it -1 (no float-to-int cast
	// of a value out of range)
	const float grid_q = (G1[0] - G0[0]) / st.min_grid;
	T.grid_num = grid_q >= 0.0f && grid_q < 1023.0f ? (int)grid_q : -1;
	// the nodes are already in depth-first preorder (children stitched in
	// order after their parent), so the reference's arrays follow directly
	T.ref_nodes.resize(st.nodes.size());
	T.ref_lists.reserve((size_t)T.n_list);
	for (size_t i = 0; i < st.nodes.size(); i++) {
		const BNode &b = st.nodes[i];
		RefNode &r = T.ref_nodes[i];
		std::memcpy(r.g0, b.g0, 12);
		std::memcpy(r.g1, b.g1, 12);
		std::memcpy(r.sub, b.sub, sizeof r.sub);
		r.list = -1;
		if (b.leaf >= 0) {
			r.list = (int)T.ref_lists.size();
			const int beg = st.item_begin[b.leaf], cnt = st.item_count[b.leaf];
			T.ref_lists.insert(T.ref_lists.end(), st.items.begin() + beg, st.items.begin() + beg + cnt);
			T.ref_lists.push_back(-1);
		}
	}
	T.items = std::move(st.items);
	T.leaf_box.assign((size_t)T.n_leaves * 8, 0.0f);
	for (const BNode &b : st.nodes) {
		if (b.leaf < 0) continue;
		float *o = &T.leaf_box[(size_t)b.leaf * 8];
		const int begin = st.item_begin[b.leaf], cd = st.item_count[b.leaf] | (b.depth << 24);
		o[0] = b.g0[0]; o[1] = b.g0[1]; o[2] = b.g0[2];
		std::memcpy(&o[3], &begin, 4);
		o[4] = b.g1[0]; o[5] = b.g1[1]; o[6] = b.g1[2];
		std::memcpy(&o[7], &cd, 4);
	}
	// tabulate TreeNodeNum's descent (raycast.cpp:262-280); a grid the
	// table cannot hold (bounds that are not finite) is left empty, gt = 0
	if (T.grid_num < 0) {
		T.gt = 0;
		return T;
	}
	T.gt = T.grid_num + 2;
	const int gt = T.gt;
	T.cell_leaf.assign((size_t)gt * gt * gt, -1);
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int gx = 0; gx < gt; gx++)
		for (int gy = 0; gy < gt; gy++)
			for (int gz = 0; gz < gt; gz++) {
				int g[3] = {gx, gy, gz}, node = 0, gsize = T.grid_num;
				while (st.nodes[node].leaf < 0) {
					gsize >>= 1;
					int child = 0;
					for (int k = 0; k < 3; k++)
						if (g[k] >= gsize) {
							child |= 1 << k;
							g[k] -= gsize;
						}
					node = st.nodes[node].sub[child];
				}
				T.cell_leaf[((size_t)gx * gt + gy) * gt + gz] = st.nodes[node].leaf;
			}
	return T;
}

int build_segment_index(FlatOctree &T, const Patch *pool, int n, const float *G0, const float *G1, std::string &err)
{
	if (T.fine_r > 0) return 0;
	int D = 0;
	for (int l = 0; l < T.n_leaves; l++) {
		int cd;
		std::memcpy(&cd, &T.leaf_box[(size_t)l * 8 + 7], 4);
		D = std::max(D, cd >> 24);
	}
	const int R = 1 << D;
	double h[3];
	for (int k = 0; k < 3; k++) h[k] = ((double)G1[k] - (double)G0[k]) / R;
	// a leaf at depth d is a block of 2^(D - d) cells per axis starting at
	// its lower corner (the float midpoints sit within rounding of the
	// dyadic grid lines, so the nearest grid line is the corner)
	std::vector<int> fine((size_t)R * R * R, -1);
	for (int l = 0; l < T.n_leaves; l++) {
		const float *b = &T.leaf_box[(size_t)l * 8];
		int cd;
		std::memcpy(&cd, &b[7], 4);
		const int size = 1 << (D - (cd >> 24));
		int base[3];
		for (int k = 0; k < 3; k++) base[k] = (int)std::lround(((double)b[k] - G0[k]) / h[k]);
		for (int k = 0; k < 3; k++)
			if (base[k] < 0 || base[k] + size > R || base[k] % size) {
				err = "octree leaf off the dyadic grid";
				return -1;
			}
		for (int a = base[0]; a < base[0] + size; a++)
			for (int bb = base[1]; bb < base[1] + size; bb++)
				for (int c = base[2]; c < base[2] + size; c++) {
					int &cell = fine[((size_t)a * R + bb) * R + c];
					if (cell >= 0) {
						err = "octree leaves overlap";
						return -1;
					}
					cell = l;
				}
	}
	for (int x : fine)
		if (x < 0) {
			err = "octree leaves do not tile the root box";
			return -1;
		}
	// complete leaf lists: the reference's CrossOver/Clip test misses
	// triangles that cross a box through their interior, so the segment
	// walk lists a triangle in every leaf holding a fine cell that its
	// bounding box, grown by 1e-3 of a cell, touches
	std::vector<std::vector<int>> per(T.n_leaves);
	for (int i = 0; i < n; i++) {
		int lo[3], hi[3];
		for (int k = 0; k < 3; k++) {
			double mn = pool[i].p[0][k], mx = mn;
			for (int v = 1; v < 3; v++) {
				mn = std::min(mn, (double)pool[i].p[v][k]);
				mx = std::max(mx, (double)pool[i].p[v][k]);
			}
			lo[k] = std::max(0, (int)std::floor((mn - G0[k]) / h[k] - 1e-3));
			hi[k] = std::min(R - 1, (int)std::floor((mx - G0[k]) / h[k] + 1e-3));
		}
		for (int a = lo[0]; a <= hi[0]; a++)
			for (int b = lo[1]; b <= hi[1]; b++)
				for (int c = lo[2]; c <= hi[2]; c++) {
					std::vector<int> &L = per[fine[((size_t)a * R + b) * R + c]];
					if (L.empty() || L.back() != i) L.push_back(i);
				}
	}
	T.seg_begin.assign(T.n_leaves + 1, 0);
	for (int l = 0; l < T.n_leaves; l++) T.seg_begin[l + 1] = T.seg_begin[l] + (int)per[l].size();
	T.seg_items.resize(T.seg_begin[T.n_leaves]);
	for (int l = 0; l < T.n_leaves; l++) std::copy(per[l].begin(), per[l].end(), T.seg_items.begin() + T.seg_begin[l]);
	T.fine_leaf = std::move(fine);
	T.fine_r = R;
	return 0;
}

}  // namespace rgpu
