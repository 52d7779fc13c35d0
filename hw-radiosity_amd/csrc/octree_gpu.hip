// The visibility octree built on the device (BuildTree, raycast.cpp:168-233),
// level by level, into the same FlatOctree the host builder produces.
//
// The tree is the reference's: a node holding more than 10 triangles below
// depth 6 splits at its float midpoint into 8 children (child bit j = the
// upper half in dimension j), and a triangle is listed in every child that
// CrossOver/Clip (raycast.cpp:24-155) says it overlaps.  Per level:
//   k_oct_flags   one thread per (item of an inner node): the triangle
//                 against its node's 8 children -> a 0/1 flag at
//                 base8[node] + child * count + k (node-major, then child,
//                 then the item's position in the parent's list);
//   scan          exclusive prefix of the flags: the position of each kept
//                 item in the next level's items, so every child's list is
//                 contiguous and keeps the parent's (ascending) order;
//   k_oct_scatter the kept items and their child node;
//   k_oct_children the children's boxes and item ranges.
// Leaves (at most 10 items, or depth 6) keep their level's item range; at
// the end the leaf lists are gathered into one array and TreeNodeNum's
// descent (raycast.cpp:254-281) is tabulated for every integer cell, as the
// host builder does.  Node and leaf numbering is breadth-first instead of the
// reference's depth-first, which no result depends on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "octree_gpu.h"

namespace rgpu {
namespace {

constexpr int kB = 256;

struct ONode {
	float g0[3], g1[3];
	int begin, count;  // items in this level's item array
	int child;         // global index of child 0 (children contiguous), -1 for a leaf
	int leaf;          // leaf number, -1 for an inner node
	int depth;
	int pad;
};

struct OTri {
	float p[3][3];
	float lo[3], hi[3];  // bounding box (+-inf where a coordinate is NaN)
};

#define HIP_OK(expr, what)                                                        \
	do {                                                                          \
		hipError_t e_ = (expr);                                                   \
		if (e_ != hipSuccess) {                                                   \
			err = std::string(what) + ": " + hipGetErrorString(e_);               \
			return -1;                                                            \
		}                                                                         \
	} while (0)

// Clip (raycast.cpp:24-60): does segment p0-p1 meet the (x, y) rectangle?
__device__ bool seg_rect(const float *p0, const float *p1, const float *g0, const float *g1, int x, int y)
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

// InterpVector (raycast.cpp:91-100), with the compiler's contraction of
// AddVector (a * x + b * y) when F.
template <bool F> __device__ void point_at(const float *hi, const float *lo, float g, int d, float *out)
{
	const float span = hi[d] - lo[d];
	if (span == 0.0f) {
		out[0] = hi[0];
		out[1] = hi[1];
		out[2] = hi[2];
		return;
	}
	const float wa = (g - lo[d]) / span, wb = (hi[d] - g) / span;
	for (int k = 0; k < 3; k++) out[k] = F ? __builtin_fmaf(wa, hi[k], wb * lo[k]) : wa * hi[k] + wb * lo[k];
}

// CrossOver (raycast.cpp:109-155); the vertex pointers stay sorted across
// dimensions as the reference's in-place sort leaves them.
template <bool F> __device__ bool overlaps(const OTri &t, const float *g0, const float *g1)
{
	const float *hi = t.p[0], *mid = t.p[1], *lo = t.p[2];
	float a[3], b[3];
	for (int x = 0; x < 3; x++) {
		const float *s;
		if (hi[x] < mid[x]) { s = hi; hi = mid; mid = s; }
		if (hi[x] < lo[x]) { s = hi; hi = lo; lo = s; }
		if (mid[x] < lo[x]) { s = mid; mid = lo; lo = s; }
		const int u = (x + 1) % 3, w = (x + 2) % 3;
		if (g1[x] < lo[x]) return false;
		point_at<F>(hi, lo, g1[x], x, a);
		if (g1[x] < mid[x]) {
			point_at<F>(mid, lo, g1[x], x, b);
			return seg_rect(a, b, g0, g1, u, w);
		}
		if (g1[x] <= hi[x]) {
			point_at<F>(hi, mid, g1[x], x, b);
			return seg_rect(a, b, g0, g1, u, w);
		}
		if (g0[x] > hi[x]) return false;
		point_at<F>(hi, lo, g0[x], x, a);
		if (g0[x] > mid[x]) {
			point_at<F>(hi, mid, g0[x], x, b);
			return seg_rect(a, b, g0, g1, u, w);
		}
		if (g0[x] >= lo[x]) {
			point_at<F>(mid, lo, g0[x], x, b);
			return seg_rect(a, b, g0, g1, u, w);
		}
	}
	return true;
}

// A sufficient condition for CrossOver to say no (the host builder's
// early reject): the triangle's box misses the node's box by more than
// rounding can move a point InterpVector computes on one of its edges.
__device__ bool surely_apart(const OTri &t, const float *g0, const float *g1)
{
	for (int k = 0; k < 3; k++) {
		const float m0 = 1e-5f * (fabsf(g0[k]) + fabsf(t.hi[k])) + 1e-30f;
		const float m1 = 1e-5f * (fabsf(g1[k]) + fabsf(t.lo[k])) + 1e-30f;
		if (t.hi[k] < g0[k] - m0 || t.lo[k] > g1[k] + m1) return true;
	}
	return false;
}

// child c's box of a node (raycast.cpp:195-208): the midpoint in double
__device__ void child_box(const ONode &nd, int c, float *q0, float *q1)
{
	for (int k = 0; k < 3; k++) {
		const float mid = (float)(0.5 * (double)nd.g0[k] + 0.5 * (double)nd.g1[k]);
		const bool upper = (c >> k) & 1;
		q0[k] = upper ? mid : nd.g0[k];
		q1[k] = upper ? nd.g1[k] : mid;
	}
}

__global__ __launch_bounds__(kB) void k_oct_tris(const float *__restrict__ p9, int n, OTri *tri, int *items, int *item_node)
{
	const int i = blockIdx.x * kB + threadIdx.x;
	if (i >= n) return;
	OTri t;
	for (int v = 0; v < 3; v++)
		for (int k = 0; k < 3; k++) t.p[v][k] = p9[9 * (size_t)i + 3 * v + k];
	for (int k = 0; k < 3; k++) {
		const float a = t.p[0][k], b = t.p[1][k], c = t.p[2][k];
		const bool nan = isnan(a) || isnan(b) || isnan(c);  // never rejected early
		t.lo[k] = nan ? -INFINITY : fminf(a, fminf(b, c));
		t.hi[k] = nan ? INFINITY : fmaxf(a, fmaxf(b, c));
	}
	tri[i] = t;
	items[i] = i;
	item_node[i] = 0;
}

// inner[i] = 1 for a node that splits; cnt8[i] = 8 * its count (0 for a leaf)
__global__ __launch_bounds__(kB) void k_oct_classify(ONode *nodes, int m, int level, int *inner, int *cnt8)
{
	const int i = blockIdx.x * kB + threadIdx.x;
	if (i >= m) return;
	const bool split = nodes[i].count > kMaxTriPerNode && level < kMaxLevel;
	inner[i] = split;
	cnt8[i] = split ? 8 * nodes[i].count : 0;
}

template <bool F>
__global__ __launch_bounds__(kB) void k_oct_flags(const ONode *__restrict__ nodes, const OTri *__restrict__ tri,
                                                  const int *__restrict__ items, const int *__restrict__ item_node,
                                                  int nitems, const int *__restrict__ base8, int *flags)
{
	const int q = blockIdx.x * kB + threadIdx.x;
	if (q >= nitems) return;
	const int nd = item_node[q];
	const ONode &node = nodes[nd];
	if (!(node.count > kMaxTriPerNode && node.depth < kMaxLevel)) return;  // a leaf's items stay
	const OTri &t = tri[items[q]];
	const int k = q - node.begin;
	int *f = flags + base8[nd] + k;
	for (int c = 0; c < 8; c++) {
		float q0[3], q1[3];
		child_box(node, c, q0, q1);
		f[c * node.count] = !surely_apart(t, q0, q1) && overlaps<F>(t, q0, q1);
	}
}

__global__ __launch_bounds__(kB) void k_oct_scatter(const ONode *__restrict__ nodes, const int *__restrict__ items,
                                                    const int *__restrict__ item_node, int nitems,
                                                    const int *__restrict__ base8, const int *__restrict__ inner_idx,
                                                    const int *__restrict__ flags, const int *__restrict__ pos,
                                                    int child_base, int *items_next, int *node_next)
{
	const int q = blockIdx.x * kB + threadIdx.x;
	if (q >= nitems) return;
	const int nd = item_node[q];
	const ONode &node = nodes[nd];
	if (!(node.count > kMaxTriPerNode && node.depth < kMaxLevel)) return;
	const int k = q - node.begin;
	for (int c = 0; c < 8; c++) {
		const size_t f = (size_t)base8[nd] + (size_t)c * node.count + k;
		if (flags[f]) {
			items_next[pos[f]] = items[q];
			node_next[pos[f]] = child_base + 8 * inner_idx[nd] + c;
		}
	}
}

// children of inner node i (inner_idx = its rank among the level's inner
// nodes): boxes, depth and first item from the scanned flags (the counts
// follow in k_oct_counts)
__global__ __launch_bounds__(kB) void k_oct_children(ONode *nodes, int m, const int *__restrict__ inner,
                                                     const int *__restrict__ inner_idx, const int *__restrict__ base8,
                                                     const int *__restrict__ pos, int child_base, ONode *next)
{
	const int i = blockIdx.x * kB + threadIdx.x;
	if (i >= m || !inner[i]) return;
	ONode &node = nodes[i];
	const int first = 8 * inner_idx[i];
	node.child = child_base + first;
	for (int c = 0; c < 8; c++) {
		ONode ch;
		child_box(node, c, ch.g0, ch.g1);
		ch.begin = pos[(size_t)base8[i] + (size_t)c * node.count];
		ch.count = 0;
		ch.child = -1;
		ch.leaf = -1;
		ch.depth = node.depth + 1;
		ch.pad = 0;
		next[first + c] = ch;
	}
}

// child counts: the distance to the next child's begin (or the level total)
__global__ __launch_bounds__(kB) void k_oct_counts(ONode *next, int mn, int total_next)
{
	const int i = blockIdx.x * kB + threadIdx.x;
	if (i >= mn) return;
	const int end = i + 1 < mn ? next[i + 1].begin : total_next;
	next[i].count = end - next[i].begin;
}

__global__ __launch_bounds__(kB) void k_scan_block(const int *__restrict__ in, int *out, int *sums, size_t n)
{
	__shared__ int s[kB];
	const size_t i = blockIdx.x * (size_t)kB + threadIdx.x;
	const int x = i < n ? in[i] : 0;
	s[threadIdx.x] = x;
	__syncthreads();
	for (int off = 1; off < kB; off <<= 1) {
		const int y = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
		__syncthreads();
		s[threadIdx.x] += y;
		__syncthreads();
	}
	if (i < n) out[i] = s[threadIdx.x] - x;  // exclusive
	if (threadIdx.x == kB - 1) sums[blockIdx.x] = s[kB - 1];
}

__global__ __launch_bounds__(kB) void k_scan_add(int *out, const int *__restrict__ sums, size_t n)
{
	const size_t i = blockIdx.x * (size_t)kB + threadIdx.x;
	if (i < n) out[i] += sums[blockIdx.x];
}

// leaves of one level: leaf numbers (the leaves before node i number i
// minus the inner nodes before it) and MinGridLen (positive floats compare
// as their bit patterns)
__global__ __launch_bounds__(kB) void k_oct_leaves(ONode *nodes, int m, const int *__restrict__ inner,
                                                   const int *__restrict__ inner_idx, int leaf_base, unsigned *min_x)
{
	const int i = blockIdx.x * kB + threadIdx.x;
	if (i >= m || inner[i]) return;
	nodes[i].leaf = leaf_base + i - inner_idx[i];
	const float w = nodes[i].g1[0] - nodes[i].g0[0];
	atomicMin(min_x, __float_as_uint(w));
}

// leaf boxes and gathered lists
__global__ __launch_bounds__(kB) void k_oct_emit(const ONode *__restrict__ nodes, int m, const int *__restrict__ items,
                                                 const int *__restrict__ out_begin, float *leaf_box, int *out_items)
{
	const int i = blockIdx.x * kB + threadIdx.x;
	if (i >= m) return;
	const ONode &nd = nodes[i];
	if (nd.leaf < 0) return;
	const int b = out_begin[nd.leaf];
	for (int k = 0; k < nd.count; k++) out_items[b + k] = items[nd.begin + k];
	float *o = leaf_box + 8 * (size_t)nd.leaf;
	const int cd = nd.count | (nd.depth << 24);
	o[0] = nd.g0[0];
	o[1] = nd.g0[1];
	o[2] = nd.g0[2];
	o[3] = __int_as_float(b);
	o[4] = nd.g1[0];
	o[5] = nd.g1[1];
	o[6] = nd.g1[2];
	o[7] = __int_as_float(cd);
}

// TreeNodeNum's descent (raycast.cpp:262-280) for every integer cell
__global__ __launch_bounds__(kB) void k_oct_cells(const ONode *__restrict__ all, int grid_num, int gt, int *cell_leaf)
{
	const size_t q = blockIdx.x * (size_t)kB + threadIdx.x;
	const size_t total = (size_t)gt * gt * gt;
	if (q >= total) return;
	int g[3] = {(int)(q / ((size_t)gt * gt)), (int)((q / gt) % gt), (int)(q % gt)};
	int node = 0, gsize = grid_num;
	while (all[node].leaf < 0) {
		gsize >>= 1;
		int child = 0;
		for (int k = 0; k < 3; k++)
			if (g[k] >= gsize) {
				child |= 1 << k;
				g[k] -= gsize;
			}
		node = all[node].child + child;
	}
	cell_leaf[q] = all[node].leaf;
}

template <typename T> struct Dev {
	T *p = nullptr;
	~Dev() { if (p) (void)hipFree(p); }
	int alloc(size_t n, std::string &err)
	{
		if (p) (void)hipFree(p);
		p = nullptr;
		HIP_OK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "octree hipMalloc");
		return 0;
	}
};

// exclusive scan of n ints (in may equal out); the total into *total
int scan(const int *in, int *out, size_t n, int *total, hipStream_t st, std::string &err)
{
	const size_t nb = (n + kB - 1) / kB;
	Dev<int> sums, sums_x;
	if (sums.alloc(nb + 1, err) || sums_x.alloc(nb + 1, err)) return -1;
	if (n) hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(kB), 0, st, in, out, sums.p, n);
	int t = 0;
	if (nb > 1) {
		if (scan(sums.p, sums_x.p, nb, &t, st, err)) return -1;
		hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kB), 0, st, out, sums_x.p, n);
	} else if (nb == 1) {
		HIP_OK(hipMemcpyAsync(&t, sums.p, 4, hipMemcpyDeviceToHost, st), "scan total");
		HIP_OK(hipStreamSynchronize(st), "scan total");
	}
	HIP_OK(hipGetLastError(), "scan launch");
	*total = t;
	return 0;
}

}  // namespace

int build_octree_gpu(void *stream, const Patch *pool, int n, const float *G0, const float *G1, bool fma,
                     FlatOctree &T, std::string &err)
{
	hipStream_t st = static_cast<hipStream_t>(stream);
	if (n < 0) {
		err = "octree: negative triangle count";
		return -1;
	}
	std::vector<float> p9((size_t)n * 9);
	for (int i = 0; i < n; i++) std::memcpy(&p9[(size_t)i * 9], pool[i].p, 36);
	Dev<float> dp;
	Dev<OTri> tri;
	if (dp.alloc(p9.size(), err) || tri.alloc(n, err)) return -1;
	HIP_OK(hipMemcpyAsync(dp.p, p9.data(), p9.size() * 4, hipMemcpyHostToDevice, st), "octree upload");
	// level 0: the root with every triangle
	std::vector<Dev<ONode>> lv_nodes(kMaxLevel + 1);
	std::vector<Dev<int>> lv_items(kMaxLevel + 1), lv_inode(kMaxLevel + 1);
	std::vector<int> lv_m(kMaxLevel + 1, 0), lv_n(kMaxLevel + 1, 0), lv_base(kMaxLevel + 2, 0);
	if (lv_nodes[0].alloc(1, err) || lv_items[0].alloc(n, err) || lv_inode[0].alloc(n, err)) return -1;
	{
		ONode root;
		for (int k = 0; k < 3; k++) {
			root.g0[k] = G0[k];
			root.g1[k] = G1[k];
		}
		root.begin = 0;
		root.count = n;
		root.child = -1;
		root.leaf = -1;
		root.depth = 0;
		root.pad = 0;
		HIP_OK(hipMemcpyAsync(lv_nodes[0].p, &root, sizeof root, hipMemcpyHostToDevice, st), "octree root");
	}
	if (n) hipLaunchKernelGGL(k_oct_tris, dim3((n + kB - 1) / kB), dim3(kB), 0, st, dp.p, n, tri.p, lv_items[0].p, lv_inode[0].p);
	lv_m[0] = 1;
	lv_n[0] = n;
	Dev<unsigned> min_x;
	if (min_x.alloc(1, err)) return -1;
	{
		const float w = G1[0] - G0[0];
		unsigned bits;
		std::memcpy(&bits, &w, 4);
		HIP_OK(hipMemcpyAsync(min_x.p, &bits, 4, hipMemcpyHostToDevice, st), "octree min");
	}
	int levels = 0, n_leaves = 0;
	for (int L = 0; L <= kMaxLevel; L++) {
		levels = L + 1;
		const int m = lv_m[L], ni = lv_n[L];
		Dev<int> inner, inner_idx, cnt8, base8;
		if (inner.alloc(m, err) || inner_idx.alloc(m, err) || cnt8.alloc(m, err) || base8.alloc(m, err)) return -1;
		const unsigned gm = (unsigned)((m + kB - 1) / kB);
		hipLaunchKernelGGL(k_oct_classify, dim3(gm), dim3(kB), 0, st, lv_nodes[L].p, m, L, inner.p, cnt8.p);
		int n_inner = 0, n_flags = 0;
		if (scan(inner.p, inner_idx.p, m, &n_inner, st, err) || scan(cnt8.p, base8.p, m, &n_flags, st, err)) return -1;
		if (m) hipLaunchKernelGGL(k_oct_leaves, dim3(gm), dim3(kB), 0, st, lv_nodes[L].p, m, inner.p, inner_idx.p,
		                          n_leaves, min_x.p);
		n_leaves += m - n_inner;
		lv_base[L + 1] = lv_base[L] + m;
		if (n_inner == 0) break;
		// children of the inner nodes
		Dev<int> flags, pos;
		if (flags.alloc((size_t)n_flags, err) || pos.alloc((size_t)n_flags, err)) return -1;
		const unsigned gi = (unsigned)((ni + kB - 1) / kB);
		if (fma)
			hipLaunchKernelGGL(k_oct_flags<true>, dim3(gi), dim3(kB), 0, st, lv_nodes[L].p, tri.p, lv_items[L].p,
			                   lv_inode[L].p, ni, base8.p, flags.p);
		else
			hipLaunchKernelGGL(k_oct_flags<false>, dim3(gi), dim3(kB), 0, st, lv_nodes[L].p, tri.p, lv_items[L].p,
			                   lv_inode[L].p, ni, base8.p, flags.p);
		int n_next = 0;
		if (scan(flags.p, pos.p, (size_t)n_flags, &n_next, st, err)) return -1;
		const int mn = 8 * n_inner;
		if (lv_nodes[L + 1].alloc(mn, err) || lv_items[L + 1].alloc(n_next, err) || lv_inode[L + 1].alloc(n_next, err))
			return -1;
		const int child_base = lv_base[L + 1];
		hipLaunchKernelGGL(k_oct_scatter, dim3(gi), dim3(kB), 0, st, lv_nodes[L].p, lv_items[L].p, lv_inode[L].p, ni,
		                   base8.p, inner_idx.p, flags.p, pos.p, 0, lv_items[L + 1].p, lv_inode[L + 1].p);
		hipLaunchKernelGGL(k_oct_children, dim3(gm), dim3(kB), 0, st, lv_nodes[L].p, m, inner.p, inner_idx.p, base8.p,
		                   pos.p, child_base, lv_nodes[L + 1].p);
		hipLaunchKernelGGL(k_oct_counts, dim3((mn + kB - 1) / kB), dim3(kB), 0, st, lv_nodes[L + 1].p, mn, n_next);
		HIP_OK(hipGetLastError(), "octree level launch");
		lv_m[L + 1] = mn;
		lv_n[L + 1] = n_next;
	}
	// the nodes of all levels in one array (child indices are global), the
	// leaf lists gathered, leaf boxes
	const int n_nodes = lv_base[levels];
	Dev<ONode> all;
	if (all.alloc(n_nodes, err)) return -1;
	for (int L = 0; L < levels; L++)
		if (lv_m[L])
			HIP_OK(hipMemcpyAsync(all.p + lv_base[L], lv_nodes[L].p, (size_t)lv_m[L] * sizeof(ONode),
			                      hipMemcpyDeviceToDevice, st),
			       "octree nodes");
	std::vector<ONode> h_all(n_nodes);
	HIP_OK(hipMemcpyAsync(h_all.data(), all.p, (size_t)n_nodes * sizeof(ONode), hipMemcpyDeviceToHost, st), "octree nodes");
	unsigned mbits = 0;
	HIP_OK(hipMemcpyAsync(&mbits, min_x.p, 4, hipMemcpyDeviceToHost, st), "octree min");
	HIP_OK(hipStreamSynchronize(st), "octree nodes");
	std::vector<int> out_begin(n_leaves + 1, 0);
	for (const ONode &nd : h_all)
		if (nd.leaf >= 0) out_begin[nd.leaf + 1] = nd.count;
	for (int l = 0; l < n_leaves; l++) out_begin[l + 1] += out_begin[l];
	const int n_items = out_begin[n_leaves];
	Dev<int> d_begin, d_items;
	Dev<float> d_box;
	if (d_begin.alloc(n_leaves + 1, err) || d_items.alloc(n_items, err) || d_box.alloc((size_t)n_leaves * 8, err))
		return -1;
	HIP_OK(hipMemcpyAsync(d_begin.p, out_begin.data(), (size_t)(n_leaves + 1) * 4, hipMemcpyHostToDevice, st), "octree lists");
	for (int L = 0; L < levels; L++)
		if (lv_m[L])
			hipLaunchKernelGGL(k_oct_emit, dim3((lv_m[L] + kB - 1) / kB), dim3(kB), 0, st, all.p + lv_base[L], lv_m[L],
			                   lv_items[L].p, d_begin.p, d_box.p, d_items.p);
	float min_grid;
	std::memcpy(&min_grid, &mbits, 4);
	T = FlatOctree();
	T.n_nodes = n_nodes;
	T.n_leaves = n_leaves;
	T.n_list = n_items + n_leaves;
	T.min_grid = min_grid;
	const float grid_q = (G1[0] - G0[0]) / min_grid;  // GridNum, range-checked before the cast
	if (!(grid_q >= 0.0f && grid_q < 1023.0f)) {
		err = "octree grid of " + std::to_string(grid_q) + " cells per axis: too large to tabulate";
		return -1;
	}
	T.grid_num = (int)grid_q;
	T.gt = T.grid_num + 2;
	const size_t cells = (size_t)T.gt * T.gt * T.gt;
	Dev<int> d_cells;
	if (d_cells.alloc(cells, err)) return -1;
	hipLaunchKernelGGL(k_oct_cells, dim3((unsigned)((cells + kB - 1) / kB)), dim3(kB), 0, st, all.p, T.grid_num, T.gt,
	                   d_cells.p);
	HIP_OK(hipGetLastError(), "octree cells launch");
	T.items.resize(n_items);
	T.leaf_box.resize((size_t)n_leaves * 8);
	T.cell_leaf.resize(cells);
	HIP_OK(hipMemcpyAsync(T.items.data(), d_items.p, (size_t)n_items * 4, hipMemcpyDeviceToHost, st), "octree download");
	HIP_OK(hipMemcpyAsync(T.leaf_box.data(), d_box.p, T.leaf_box.size() * 4, hipMemcpyDeviceToHost, st), "octree download");
	HIP_OK(hipMemcpyAsync(T.cell_leaf.data(), d_cells.p, cells * 4, hipMemcpyDeviceToHost, st), "octree download");
	HIP_OK(hipStreamSynchronize(st), "octree build");
	// the reference's arrays: the breadth-first nodes renumbered in
	// depth-first preorder, the lists in that order
	std::vector<int> dfs(n_nodes, -1), order;
	order.reserve(n_nodes);
	std::vector<int> stack = {0};
	while (!stack.empty()) {
		const int b = stack.back();
		stack.pop_back();
		dfs[b] = (int)order.size();
		order.push_back(b);
		if (h_all[b].child >= 0)
			for (int c = 7; c >= 0; c--) stack.push_back(h_all[b].child + c);
	}
	T.ref_nodes.resize(n_nodes);
	T.ref_lists.clear();
	T.ref_lists.reserve((size_t)T.n_list);
	for (int i = 0; i < n_nodes; i++) {
		const ONode &b = h_all[order[i]];
		RefNode &r = T.ref_nodes[i];
		std::memcpy(r.g0, b.g0, 12);
		std::memcpy(r.g1, b.g1, 12);
		r.list = -1;
		for (int c = 0; c < 8; c++) r.sub[c] = b.child >= 0 ? dfs[b.child + c] : -1;
		if (b.leaf >= 0) {
			r.list = (int)T.ref_lists.size();
			T.ref_lists.insert(T.ref_lists.end(), T.items.begin() + out_begin[b.leaf],
			                   T.items.begin() + out_begin[b.leaf + 1]);
			T.ref_lists.push_back(-1);
		}
	}
	return 0;
}

}  // namespace rgpu
