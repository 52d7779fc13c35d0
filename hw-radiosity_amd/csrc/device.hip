// Device kernels and buffers of the shooting loop (gfx950).
//
// Layout in HBM
//   octree: cell_leaf[gt^3] (TreeNodeNum tabulated), leaf_box[2*l] =
//       {g0.xyz, first item}, {g1.xyz, count | depth<<24}
//   leaf lists, leaf-major with the triangle inline (the list order is the
//       reference's ascending original index): lnc[2*k] = {n.xyz, c.x},
//       lnc[2*k+1] = {c.y, c.z, original id, cull radius}, lse[3*k+s] = side plane s.
//       The contents are the originals' slots after pre-partition, which is
//       what RayHitted tests (raycast.cpp:329).
//   receivers (every pool slot), SoA by vertex: vx/vy/vz[3*r+v],
//       acc/dacc[3*r+v] = accB/deltaaccB of vertex v (float4),
//       nrm[r] = {n.xyz, parent}, cA[r] = {c.xyz, area}, rho[r] = Frgb as
//       floats, dB[r] = deltaB
//
// One shooting iteration (rad.cpp:406-451) is four launches on one stream:
//   k_gen     every block turns the best candidate key into the source
//             descriptor -- the whole source and the halves AdaptCalFF
//             shoots from (shade.cpp:141-168, 76-99) -- and one lane per
//             (receiver, vertex), 84 or 168 receivers per block, computes
//             CalFF up to the ray (shade.cpp:107-123) for each half into a
//             per-(receiver, vertex, half) form-factor buffer, with the
//             block's ray count;
//   k_compact the rays as a dense list in the receivers' spatial order,
//             grouped by origin key within a block; it also clears the
//             source's deltaB and the key (at most 8192 slots: no list,
//             k_trace_direct traces each ray position in place);
//   k_trace   two lanes per ray (four and eight over the end of each XCD's
//             share, or for few rays): RayHitted (raycast.cpp:335-360); a
//             ray that misses the source zeroes its form factor;
//   k_update  CalRadiosity (shade.cpp:244-259), the three vertex lanes of a
//             receiver combined by shuffles in the reference's order, and
//             the next iteration's candidate key (listCanditate,
//             rad.cpp:387-404) by wave, block and one 64-bit atomicMax per
//             block.
// With k sources per iteration (-parallel_src) k_collect/k_topk select them
// and gen/compact/trace/update repeat per source in the reference's order.
// The host polls a done flag every -batch iterations.
//
// Every float operation is IEEE single or double in source order
// (-ffp-contract=off, correctly rounded division and square root); with
// F=true an explicit fma sits at each of the reference compiler's
// contraction sites, so results are bit-identical to the reference.  Four
// shortcuts are exact by construction: a triangle test is rejected from
// a2 * rcp(a1) only when the exact quotient is certainly on the rejected
// side (a 2^-19 relative guard band around a <= 2^-21 error), or when the
// point that estimate gives is farther from the triangle than any point its
// side-plane tests can accept (test_entry); TreeNodeNum's
// double divide is a multiply by 1/MinGridLen unless the product is within
// 1e-9 of a cell edge; and the double compares against 1e-4 and 1e-6 are
// float compares against the nearest floats on the right side.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "device.h"
#include "../../include/radiosity_gpu.h"
#include "octree_gpu.h"
#include "pack.h"

namespace rgpu {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// k_trace's block: one wave, so that a wave whose rays walk far holds only
// its own slot (rays' costs are heavy-tailed)
constexpr int kTraceBlock = 64;
// k_trace lanes the machine runs at once (256 CUs x 4 SIMDs x 6 waves, its
// occupancy, x 64)
constexpr long long kTraceSlots = 256LL * 4 * 6 * 64;
constexpr int kRecvPerWave = 21;  // 3 lanes per receiver, lane 63 idle
// k_trace with many rays: kTailPct percent of each XCD's share of the list
// in blocks of 64 / kTailLanes rays at kTailLanes lanes per ray, then the
// last kEndPct percent in blocks of 8 rays at eight lanes (hall, per launch:
// 214 us without; 209, 205.7, 204.0, 204.3, 205.7 us at 10, 15, 20, 25, 30
// percent at four lanes; eight lanes alone 217 us at 20, 232 at 30; 20 at
// four then 3 at eight 200.8 us, then 5 at eight 201.4)
constexpr int kTailPct = 20;
constexpr int kTailLanes = 4;
constexpr int kEndPct = 3;
// receiver groups per block of k_gen, k_compact and k_update: 2, or 1 below
// kOneGroupSlots slots (two 2-group blocks per CU), where halving a block's
// sequential work pays (the 2120-slot room: 48.6 -> 42.0 us per iteration;
// the 253,728-slot hall: 0.4475 s with 2, 0.4530 s with 1)
constexpr int kOneGroupSlots = 2 * 256 * 2 * 4 * 21;
constexpr int kMinQueue = 8;      // iterations queued per host check at least (up to -batch)
constexpr int kGraphIters = 8;    // shooting iterations per captured graph
constexpr int kV3Batch = 32;      // -variant v3: iteration slots queued per host check
template <int G> constexpr int kRecvPerBlock = G * kWavesPerBlock * kRecvPerWave;
constexpr int kMaxSrc = 256;
constexpr int kRayBuckets = 16;   // ray keys k_compact groups by
constexpr int kMaxWalk = 1 << 24;
constexpr double kPiD = 3.1415926535897932;
constexpr float kGuardUp = 1.0f + 0x1p-19f;
constexpr float kGuardDown = 1.0f - 0x1p-19f;


struct DScene {
	const float4 *__restrict__ lnc;
	const float4 *__restrict__ lse;
	const int *__restrict__ cell_leaf;
	const float4 *__restrict__ leaf_box;
	float g0x, g0y, g0z, g1x, g1y, g1z;
	double mgl, inv_mgl;
	int gt;
	float thr_t;     // smallest float f with (double)f >= 1e-4
	float thr_t_lo;  // thr_t * (1 - 2^-19)
	float thr_v;     // largest float f with (double)f <= 1e-6
	float cull_cl;   // bound on |c|_1 over the list records with a finite cull radius
};

struct DRecv {
	const float *__restrict__ vx;
	const float *__restrict__ vy;
	const float *__restrict__ vz;
	const float4 *__restrict__ nrm;
	const float4 *__restrict__ cA;
	const float4 *__restrict__ rho;
	float4 *acc;
	float4 *dacc;
	float4 *dB;
	int nslots;
	// receiver processing order: order[p] = the slot at position p (slots
	// along a Z-order curve of their centroids, so that a block's, a wave's
	// and an XCD's rays start close together); results do not depend on it
	const int *__restrict__ order;
};

struct SrcPatch {
	float c[3], n[3], area, nlen;
};

struct SrcDesc {
	int phys, log, split, pad;
	float dB[3], pad2;
	SrcPatch whole, half[2];
};

struct Loop {
	unsigned long long key;  // best candidate of the coming iteration
	int count;               // candidates (multi-source selection)
	int done;
	int shots;
	int nsrc;
	int ncand;  // candidates after the last update: a lower bound on the shots left
	int nrays;  // rays of the current source (length of the compacted ray list)
	// V3 splitting: slots in use, an error raised by the commit, and glibc
	// rand()'s state (34 words, the oldest at rng[rng_k])
	int trinum;
	int v3_error;
	int rng_k;
	int v3_pending;  // the last pass made new slots: another pass before the selection
	unsigned rng[34];
	SrcDesc src[kMaxSrc];
	int v3_range[2][2];  // V3: the slots [r0, r1) of the pass of each parity
};

struct Cnt {
	int rays, ref_rays, cells, tests, levels, precision, stalled;
};

// ---- arithmetic -----------------------------------------------------------
template <bool F> __device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
	if (F) return __builtin_fmaf(az, bz, __builtin_fmaf(ax, bx, ay * by));
	return ax * bx + ay * by + az * bz;
}
template <bool F> __device__ __forceinline__ float axpy(float t, float v, float p)
{
	if (F) return __builtin_fmaf(t, v, p);
	return p + t * v;
}

// Candidate key: unshot energy (order-preserving bits, -0 folded to +0) in
// the high word; the slot in the low word so that the maximum takes the
// highest slot on ties (V0's sort, rad.cpp:398-402), or with LOWEST the
// lowest slot (MaxDeltaRad's strict >, rad.cpp:373).
template <bool LOWEST> __device__ __forceinline__ unsigned long long cand_key(float sum, int slot)
{
	const float s = sum + 0.0f;
	const uint32_t b = __float_as_uint(s);
	const uint32_t ob = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
	const uint32_t lo = LOWEST ? (0xffffffffu - (uint32_t)slot) : (uint32_t)slot;
	return ((unsigned long long)ob << 32) | lo;
}
template <bool LOWEST> __device__ __forceinline__ int key_slot(unsigned long long k)
{
	const uint32_t lo = (uint32_t)k;
	return LOWEST ? (int)(0xffffffffu - lo) : (int)lo;
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long k)
{
	for (int off = 32; off > 0; off >>= 1) {
		const unsigned long long o = __shfl_xor(k, off);
		k = o > k ? o : k;
	}
	return k;
}
__device__ __forceinline__ int wave_sum(int x)
{
	for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
	return x;
}
// (int)(((double)q - G0) / MinGridLen) of TreeNodeNum (raycast.cpp:263),
// from the product with the reciprocal unless its fraction is within 1e-9
// of an integer; there the exact quotient decides.  The division sits
// behind an asm barrier so that the compiler keeps it in a branch (it is
// rare) instead of computing it for every lane and selecting.
__device__ __forceinline__ int cell_coord(float q, float g0, const DScene &S)
{
	const double x = (double)q - (double)g0;
	const double qd = x * S.inv_mgl;
	const double fl = floor(qd);
	const double fr = qd - fl;
	int g = (int)fl;
	if (!(fr > 1e-9 && fr < 1.0 - 1e-9)) {
		asm volatile("" ::: "memory");
		g = (int)(x / S.mgl);
	}
	return g;
}

// ---- RayHitted ------------------------------------------------------------
// One list entry of TriListHitted (raycast.cpp:316-333): TriHitted
// (raycast.cpp:283-313) of the ray against the record {a, b} = {n.xyz, c.x},
// {c.y, c.z, id, cull radius} with side planes se[0..2], keeping the nearest
// hit.  Exact shortcuts before the correctly rounded quotient t = a2 / a1:
//  - t from the reciprocal estimate (relative error <= 2^-21) decides when it
//    is outside [1e-4, t_min) by more than a 2^-19 guard band;
//  - the estimate's point q_a = p + t v rejects when it lies farther from the
//    centroid than the record's cull radius r plus a margin.  r (finite only
//    for records whose centroid, normal and side planes are exactly those of
//    their own vertices and whose smallest angle keeps 1 / sin(angle / 2) <=
//    16) bounds every point of the plane that passes the three side tests
//    exactly; the margin 2^-12 (2 |t| |v|_1 + |p|_1 + max |c|_1 + 2 r) covers
//    the side tests' rounding (at most 2^-21 of those magnitudes, moved by up
//    to 16x at a corner), the distance of the reference's rounded hit point
//    from the plane and |q_a - q|, with a factor of 4 to spare.
template <bool F, bool SHARED = false>
__device__ __forceinline__ void test_entry(const DScene &S, const float4 &a, const float4 &b, const float4 *se,
                                           float px, float py, float pz, float vx, float vy, float vz, int ban,
                                           float ck, float cb, float &tmin, float &tmin_hi, int &best, int e,
                                           int &best_e, bool &nan_hit, Cnt &c)
{
	// The shortcuts are evaluated for every lane without branches (most
	// waves hold lanes on every path, so branches would only add mask
	// updates); a lane leaves at one branch: banned or rejected.
	const int id = __float_as_int(b.z);
	const bool own = id == ban;
	c.tests += own ? 0 : 1;
	const float a1 = dot3<F>(vx, vy, vz, a.x, a.y, a.z);
	const float a2 = dot3<F>(a.w - px, b.x - py, b.y - pz, a.x, a.y, a.z);
	const float fa1 = fabsf(a1);
	// & and | rather than && and ||: no branch per condition
	const bool est = (fa1 >= 0x1p-100f) & (fa1 <= 0x1p100f) & (fabsf(a2) <= 0x1p100f);
	const float ta = a2 * __builtin_amdgcn_rcpf(a1);
	const float qx = __builtin_fmaf(ta, vx, px) - a.w, qy = __builtin_fmaf(ta, vy, py) - b.x,
	            qz = __builtin_fmaf(ta, vz, pz) - b.y;
	const float d2 = __builtin_fmaf(qz, qz, __builtin_fmaf(qy, qy, qx * qx));
	const float lim = __builtin_fmaf(b.w, 1.0f + 0x1p-11f, __builtin_fmaf(fabsf(ta), ck, cb));
	const bool rej = est & ((ta < S.thr_t_lo) | (ta > tmin_hi) | (d2 > lim * lim));
	if (own | rej) return;
	const float tt = a2 / a1;
	// TriHitted: t < 0.0001 (in double) or t >= t_min rejects
	if (tt < S.thr_t || tt >= tmin) return;
	const float hx = axpy<F>(tt, vx, px), hy = axpy<F>(tt, vy, py), hz = axpy<F>(tt, vz, pz);
	const float4 s0 = se[0];
	if (dot3<F>(hx, hy, hz, s0.x, s0.y, s0.z) + s0.w < 0.0f) return;
	const float4 s1 = se[1];
	if (dot3<F>(hx, hy, hz, s1.x, s1.y, s1.z) + s1.w < 0.0f) return;
	const float4 s2 = se[2];
	if (dot3<F>(hx, hy, hz, s2.x, s2.y, s2.z) + s2.w < 0.0f) return;
	tmin = tt;
	tmin_hi = tt * kGuardUp;
	best = id;
	if (SHARED) {  // what the lanes sharing a list need to combine
		best_e = e;
		nan_hit |= tt != tt;  // a NaN quotient passes every compare: accepted
	}
}

// TriListHitted over entries sub, sub + L, sub + 2 L, ... of a leaf's list
// (one lane's share when L lanes test the list of one ray; L = 1: the whole
// list).  Two entries per trip, the next one loaded while the current one
// is tested (no register copies between trips); the loads are not
// conditional (the index is clamped to the share), so that the wait before
// a test covers only the loads that entry needs.
template <bool F, int L>
__device__ __forceinline__ void test_list(const DScene &S, int begin, int cnt, int sub, float px, float py, float pz,
                                          float vx, float vy, float vz, int ban, float ck, float cb, float &tmin,
                                          float &tmin_hi, int &best, int &best_e, bool &nan_hit, Cnt &c)
{
	const int my = cnt > sub ? (cnt - sub + L - 1) / L : 0;
	if (my <= 0) return;
	const float4 *nc = S.lnc + 2 * (size_t)begin;
	const float4 *se = S.lse + 3 * (size_t)begin;
	float4 ea0 = nc[2 * sub], eb0 = nc[2 * sub + 1];
	for (int m = 0; m < my; m += 2) {
		const int e0 = sub + L * m, e1 = sub + L * min(m + 1, my - 1);
		const float4 ea1 = nc[2 * e1], eb1 = nc[2 * e1 + 1];
		test_entry<F, (L > 1)>(S, ea0, eb0, se + 3 * e0, px, py, pz, vx, vy, vz, ban, ck, cb, tmin, tmin_hi, best, e0,
		                       best_e, nan_hit, c);
		if (m + 1 >= my) break;
		const int e2 = sub + L * min(m + 2, my - 1);
		ea0 = nc[2 * e2];
		eb0 = nc[2 * e2 + 1];
		test_entry<F, (L > 1)>(S, ea1, eb1, se + 3 * e1, px, py, pz, vx, vy, vz, ban, ck, cb, tmin, tmin_hi, best, e1,
		                       best_e, nan_hit, c);
	}
}

// The walk of raycast.cpp:335-360 over the tabulated octree: find the leaf
// of q (TreeNodeNum), test its list (TriListHitted/TriHitted,
// raycast.cpp:283-333) and return the nearest hit of the first leaf with any
// hit; else step q to just past the leaf's nearest exit plane (NextPoint,
// raycast.cpp:235-251).
// With L > 1 the L lanes sub = 0 .. L-1 of an aligned group walk the same
// ray together: each tests every L-th entry of a leaf's list, and the group
// takes the nearest hit, the earliest entry on ties, as the sequential scan
// does; where a lane accepted a NaN quotient (after which the sequential
// scan accepts whatever follows) the group scans the leaf again in order.
// The walk's other steps run in every lane of the group alike (so every
// lane counts them; each counts only its own triangle tests).
template <bool F, int L = 1>
__device__ int walk(const DScene &S, float px, float py, float pz, float vx, float vy, float vz, int ban, Cnt &c,
                    int sub = 0)
{
	float t = 0.f, qx = px, qy = py, qz = pz;
	int prev = -1;
	// Brent's cycle detection over the leaves searched: the walk's next leaf
	// is a function of the leaf it leaves (the hit test and the exit point
	// use only the ray and that leaf's box), so a leaf searched twice starts
	// a cycle the reference never leaves.
	int tort = -1, power = 1, lam = 0;
	const bool use_x = fabsf(vx) > S.thr_v, use_y = fabsf(vy) > S.thr_v, use_z = fabsf(vz) > S.thr_v;
	// cull margin terms of test_entry: |ta| |v|_1 2^-11 + (|p|_1 + max |c|_1) 2^-12
	const float ck = (fabsf(vx) + fabsf(vy) + fabsf(vz)) * 0x1p-11f;
	const float cb = (fabsf(px) + fabsf(py) + fabsf(pz) + S.cull_cl) * 0x1p-12f;
	for (int steps = 0;;) {
		if (qx < S.g0x || qx > S.g1x || qy < S.g0y || qy > S.g1y || qz < S.g0z || qz > S.g1z) return -1;
		const int gx = cell_coord(qx, S.g0x, S), gy = cell_coord(qy, S.g0y, S), gz = cell_coord(qz, S.g0z, S);
		if ((unsigned)gx >= (unsigned)S.gt || (unsigned)gy >= (unsigned)S.gt || (unsigned)gz >= (unsigned)S.gt) {
			c.stalled++;  // outside the tabulated descent: cannot happen for q inside [G0, G1]
			return -2;
		}
		const int leaf = S.cell_leaf[(gx * S.gt + gy) * S.gt + gz];
		if (++steps > kMaxWalk) { c.stalled++; return -2; }
		if (leaf == prev) {
			c.precision++;
			const float nt = (float)((double)t + 1.0);
			if (nt == t) { c.stalled++; return -2; }  // the reference loops forever here
			t = nt;
			qx = axpy<F>(t, vx, px);
			qy = axpy<F>(t, vy, py);
			qz = axpy<F>(t, vz, pz);
			continue;
		}
		const float4 b0 = S.leaf_box[2 * leaf], b1 = S.leaf_box[2 * leaf + 1];
		const int begin = __float_as_int(b0.w), cd = __float_as_int(b1.w), cnt = cd & 0xffffff;
		// the exit plane of each axis, picked before the list's tests so that
		// only three of the box's six bounds stay live across them
		const float ex = vx > 0 ? b1.x : b0.x, ey = vy > 0 ? b1.y : b0.y, ez = vz > 0 ? b1.z : b0.z;
		c.cells++;
		c.levels += cd >> 24;
		float tmin = 1000000.0f, tmin_hi = 1000000.0f * kGuardUp;
		int best = -1, best_e = 0;
		bool nan_hit = false;
		if (L == 1) {
			// two entries per trip, the next one loaded while the current one
			// is tested (no register copies between trips); the loads are not
			// conditional (the index is clamped to the list), so that the wait
			// before a test covers only the loads that entry needs
			const float4 *nc = S.lnc + 2 * (size_t)begin;
			const float4 *se = S.lse + 3 * (size_t)begin;
			if (cnt > 0) {
				float4 ea0 = nc[0], eb0 = nc[1];
				for (int e = 0; e < cnt; e += 2) {
					const int e1 = min(e + 1, cnt - 1);
					const float4 ea1 = nc[2 * e1], eb1 = nc[2 * e1 + 1];
					test_entry<F>(S, ea0, eb0, se + 3 * e, px, py, pz, vx, vy, vz, ban, ck, cb, tmin, tmin_hi, best, e,
					              best_e, nan_hit, c);
					if (e + 1 >= cnt) break;
					const int e2 = min(e + 2, cnt - 1);
					ea0 = nc[2 * e2];
					eb0 = nc[2 * e2 + 1];
					test_entry<F>(S, ea1, eb1, se + 3 * e + 3, px, py, pz, vx, vy, vz, ban, ck, cb, tmin, tmin_hi, best,
					              e + 1, best_e, nan_hit, c);
				}
			}
		} else {
			test_list<F, L>(S, begin, cnt, sub, px, py, pz, vx, vy, vz, ban, ck, cb, tmin, tmin_hi, best, best_e, nan_hit,
			                c);
			// butterfly over the group: after the step with offset o each
			// lane holds the combination of its aligned 2o lanes
			bool redo = nan_hit;
			for (int o = 1; o < L; o <<= 1) {
				const float t2 = __shfl_xor(tmin, o);
				const int b2 = __shfl_xor(best, o), e2 = __shfl_xor(best_e, o);
				redo |= __shfl_xor((int)redo, o) != 0;
				if (b2 >= 0 && (best < 0 || t2 < tmin || (t2 == tmin && e2 < best_e))) {
					tmin = t2;
					best = b2;
					best_e = e2;
				}
			}
			if (redo) {
				// the whole list in order, in every lane of the group (its
				// tests were counted by the shares)
				Cnt tmp = c;
				tmin = 1000000.0f;
				tmin_hi = 1000000.0f * kGuardUp;
				best = -1;
				test_list<F, 1>(S, begin, cnt, 0, px, py, pz, vx, vy, vz, ban, ck, cb, tmin, tmin_hi, best, best_e,
				                nan_hit, tmp);
			}
		}
		if (best >= 0) return best;
		if (leaf == tort) { c.stalled++; return -2; }
		if (++lam == power) { tort = leaf; power <<= 1; lam = 0; }
		prev = leaf;
		float tm = 1000000.0f;
		if (use_x) {
			const float s = (ex - px) / vx;
			tm = s < tm ? s : tm;
		}
		if (use_y) {
			const float s = (ey - py) / vy;
			tm = s < tm ? s : tm;
		}
		if (use_z) {
			const float s = (ez - pz) / vz;
			tm = s < tm ? s : tm;
		}
		tm = (float)((double)tm + 0.1);
		qx = axpy<F>(tm, vx, px);
		qy = axpy<F>(tm, vy, py);
		qz = axpy<F>(tm, vz, pz);
		t = tm;
	}
}

// CalFF (shade.cpp:107-123) up to the ray: the point-to-patch form factor,
// or 0 where the reference returns 0 before casting.
template <bool F>
__device__ __forceinline__ float ff_prefix(const SrcPatch &s, float nx, float ny, float nz, float nlen, float px,
                                           float py, float pz)
{
	const float dx = s.c[0] - px, dy = s.c[1] - py, dz = s.c[2] - pz;
	const float ds = dot3<F>(dx, dy, dz, s.n[0], s.n[1], s.n[2]);
	const float dd = -dot3<F>(dx, dy, dz, nx, ny, nz);
	if ((ds < 0) != (dd < 0)) return 0.0f;
	const float r2 = dot3<F>(dx, dy, dz, dx, dy, dz);
	const float dl = __builtin_sqrtf(r2);
	const float c1 = ds / dl / s.nlen;
	const float c2 = dd / dl / nlen;
	float ff = c1 * c2;
	if (ff <= 0.0f) return 0.0f;
	ff = (float)((double)ff * ((double)s.area / ((double)r2 * kPiD)));
	return ff <= 0.0f ? 0.0f : ff;
}

// ff_prefix(...) != 0, for counting the full-source ray AdaptCalFF casts and
// discards.  After ff_prefix's own sign test, estimates from hardware
// reciprocals (relative error below 2^-18 in all) of |ds| / dl, |dd| / dl,
// both cosines, their product and the whole prefix: when every one lies in
// [2^-100, 2^100], each exact intermediate is a positive normal float (or a
// double of that size), so the prefix is non-zero; anything else (zero,
// underflow, overflow, NaN) takes ff_prefix itself.
template <bool F>
__device__ __forceinline__ bool ff_prefix_nonzero(const SrcPatch &s, float nx, float ny, float nz, float nlen,
                                                  float px, float py, float pz)
{
	const float dx = s.c[0] - px, dy = s.c[1] - py, dz = s.c[2] - pz;
	const float ds = dot3<F>(dx, dy, dz, s.n[0], s.n[1], s.n[2]);
	const float dd = -dot3<F>(dx, dy, dz, nx, ny, nz);
	if ((ds < 0) != (dd < 0)) return false;
	const float r2 = dot3<F>(dx, dy, dz, dx, dy, dz);
	const float ir = __builtin_amdgcn_rsqf(r2);
	const float q1 = fabsf(ds) * ir, q2 = fabsf(dd) * ir;
	const float c1 = q1 * __builtin_amdgcn_rcpf(s.nlen), c2 = q2 * __builtin_amdgcn_rcpf(nlen);
	const float cc = c1 * c2;
	const float e = cc * (s.area * __builtin_amdgcn_rcpf(r2 * (float)kPiD));
	const float lo = fminf(fminf(fminf(q1, q2), fminf(c1, c2)), fminf(cc, e));
	const float hi = fmaxf(fmaxf(fmaxf(q1, q2), fmaxf(c1, c2)), fmaxf(cc, e));
	// fminf/fmaxf drop a NaN operand: test each estimate for NaN through e and lo/hi's own
	const bool sure = lo >= 0x1p-100f && hi <= 0x1p100f && e == e && q1 == q1 && q2 == q2;
	if (sure) return true;
	return ff_prefix<F>(s, nx, ny, nz, nlen, px, py, pz) != 0.0f;
}

// Whole CalFF: prefix, then the ray (shade.cpp:124).
template <bool F>
__device__ float point_ff(const DScene &S, const SrcPatch &s, int logsrc, float nx, float ny, float nz, float nlen,
                          int logdst, float px, float py, float pz, Cnt &c)
{
	const float ff = ff_prefix<F>(s, nx, ny, nz, nlen, px, py, pz);
	if (ff == 0.0f) return 0.0f;
	c.rays++;
	c.ref_rays++;
	return walk<F>(S, px, py, pz, s.c[0] - px, s.c[1] - py, s.c[2] - pz, logdst, c) == logsrc ? ff : 0.0f;
}

// Source descriptor of slot `idx`: the whole patch and, when its longest
// squared edge is at least 0.1, the two halves AdaptCalFF shoots from.
template <bool F> __device__ void make_source(const DRecv &R, int idx, SrcDesc &d)
{
	float p[3][3];
	for (int v = 0; v < 3; v++) {
		p[v][0] = R.vx[3 * idx + v];
		p[v][1] = R.vy[3 * idx + v];
		p[v][2] = R.vz[3 * idx + v];
	}
	const float4 nr = R.nrm[idx], ca = R.cA[idx], db = R.dB[idx];
	d.phys = idx;
	d.log = __float_as_int(nr.w);
	d.dB[0] = db.x;
	d.dB[1] = db.y;
	d.dB[2] = db.z;
	const float nlen = __builtin_sqrtf(dot3<F>(nr.x, nr.y, nr.z, nr.x, nr.y, nr.z));
	d.whole.c[0] = ca.x; d.whole.c[1] = ca.y; d.whole.c[2] = ca.z;
	d.whole.n[0] = nr.x; d.whole.n[1] = nr.y; d.whole.n[2] = nr.z;
	d.whole.area = ca.w;
	d.whole.nlen = nlen;
	int edge = 0;
	float maxlen = 0.0f;
	for (int v = 0; v < 3; v++) {
		const int w = (v + 1) % 3;
		const float ex = p[w][0] - p[v][0], ey = p[w][1] - p[v][1], ez = p[w][2] - p[v][2];
		const float l = dot3<F>(ex, ey, ez, ex, ey, ez);
		if (maxlen < l) { maxlen = l; edge = v; }
	}
	d.split = !((double)maxlen < 0.1);
	if (!d.split) return;
	const int i0 = edge, i1 = (edge + 1) % 3, i2 = (edge + 2) % 3;
	float m[3];
	for (int k = 0; k < 3; k++) m[k] = (float)((double)(p[i0][k] + p[i1][k]) / 2.0);
	for (int k = 0; k < 3; k++) {
		d.half[0].c[k] = (float)((double)(m[k] + p[i2][k] + p[i0][k]) / 3.0);
		d.half[1].c[k] = (float)((double)(m[k] + p[i1][k] + p[i2][k]) / 3.0);
	}
	for (int h = 0; h < 2; h++) {
		d.half[h].n[0] = nr.x; d.half[h].n[1] = nr.y; d.half[h].n[2] = nr.z;
		d.half[h].area = (float)((double)ca.w / 2.0);
		d.half[h].nlen = nlen;
	}
}

// ---- kernels --------------------------------------------------------------
template <bool LOWEST>
__global__ __launch_bounds__(kBlock) void k_select(DRecv R, Loop *L, float converge)
{
	__shared__ unsigned long long wmax[kWavesPerBlock];
	__shared__ int wn[kWavesPerBlock];
	const int r = blockIdx.x * kBlock + threadIdx.x;
	unsigned long long key = 0;
	int n = 0;
	if (r < R.nslots && __float_as_int(R.nrm[r].w) >= 0) {
		const float4 db = R.dB[r];
		const float sum = db.x + db.y + db.z;
		if (sum > converge) {
			key = cand_key<LOWEST>(sum, r);
			n = 1;
		}
	}
	key = wave_max(key);
	n = wave_sum(n);
	if ((threadIdx.x & 63) == 0) {
		wmax[threadIdx.x >> 6] = key;
		wn[threadIdx.x >> 6] = n;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < kWavesPerBlock; w++) {
			key = wmax[w] > key ? wmax[w] : key;
			n += wn[w];
		}
		if (key) atomicMax(&L->key, key);
		if (n) atomicAdd(&L->ncand, n);
	}
}

template <bool F, bool LOWEST> __global__ void k_prep(DRecv R, Loop *L)
{
	if (threadIdx.x != 0 || L->done) return;
	const unsigned long long key = L->key;
	L->key = 0;
	if (key == 0) {
		L->done = 1;
		return;
	}
	const int idx = key_slot<LOWEST>(key);
	make_source<F>(R, idx, L->src[0]);
	R.dB[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
	L->nsrc = 1;
	L->shots++;
	L->ncand = 0;
}

// k_prep for -variant v3: nothing while a split is still being shaded
// (v3_pending: the pass of parity par continues it); else the next source
// and a first pass over every slot in use
template <bool F> __global__ void k_v3_prep(DRecv R, Loop *L, int par)
{
	if (threadIdx.x != 0 || L->done || L->v3_error || L->v3_pending) return;
	const unsigned long long key = L->key;
	L->key = 0;
	if (key == 0) {
		L->done = 1;
		return;
	}
	const int idx = key_slot<false>(key);
	make_source<F>(R, idx, L->src[0]);
	R.dB[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
	L->nsrc = 1;
	L->shots++;
	L->ncand = 0;
	L->v3_range[par][0] = 0;
	L->v3_range[par][1] = L->trinum;
}

// Multi-source selection (-parallel_src k > 1): collect candidate keys, then
// one block takes the k = min(ceil(|C|/5), parallel_src) largest in
// descending order (listCanditate, rad.cpp:398-402).
__global__ __launch_bounds__(kBlock) void k_collect(DRecv R, Loop *L, unsigned long long *cand, float converge)
{
	if (L->done) return;
	const int r = blockIdx.x * kBlock + threadIdx.x;
	if (r >= R.nslots || __float_as_int(R.nrm[r].w) < 0) return;
	const float4 db = R.dB[r];
	const float sum = db.x + db.y + db.z;
	if (sum > converge) cand[atomicAdd(&L->count, 1)] = cand_key<false>(sum, r);
}

template <bool F> __global__ __launch_bounds__(1024) void k_topk(DRecv R, Loop *L, const unsigned long long *cand, int parallel_src)
{
	__shared__ unsigned long long wmax[1024 / kWave];
	__shared__ int picked[kMaxSrc];
	if (L->done) return;
	const int count = L->count;
	const int k = min((count + 4) / 5, parallel_src);
	if (k <= 0) {
		__syncthreads();
		if (threadIdx.x == 0) { L->done = 1; L->count = 0; }
		return;
	}
	unsigned long long bound = ~0ull;
	for (int round = 0; round < k; round++) {
		unsigned long long best = 0;
		for (int i = threadIdx.x; i < count; i += blockDim.x) {
			const unsigned long long c = cand[i];
			if (c < bound && c > best) best = c;
		}
		best = wave_max(best);
		if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = best;
		__syncthreads();
		if (threadIdx.x == 0) {
			for (int w = 1; w < (int)(blockDim.x / kWave); w++) best = wmax[w] > best ? wmax[w] : best;
			wmax[0] = best;
			picked[round] = key_slot<false>(best);
		}
		__syncthreads();
		bound = wmax[0];
		__syncthreads();
	}
	if (threadIdx.x < k) make_source<F>(R, picked[threadIdx.x], L->src[threadIdx.x]);
	__syncthreads();
	if (threadIdx.x < k) R.dB[picked[threadIdx.x]] = make_float4(0.f, 0.f, 0.f, 0.f);
	if (threadIdx.x == 0) {
		L->nsrc = k;
		L->shots++;
		L->count = 0;
		L->ncand = 0;
	}
}

// Receiver-side mapping shared by k_gen and k_update: 168 receivers per
// block, 3 lanes (vertices) per receiver, 21 receivers per wave.  Blocks
// are XCD-aware: the hardware deals consecutive blocks to different XCDs,
// so each XCD is given a contiguous run of receivers.
template <int G> struct RecvMap {
	int vb, lane, wid, v, base;
	__device__ RecvMap(int nblocks)
	{
		const int xcd = blockIdx.x & 7, per = (nblocks + 7) >> 3;
		vb = xcd * per + (blockIdx.x >> 3);
		lane = threadIdx.x & 63;
		wid = threadIdx.x >> 6;
		const int rl = lane / 3;
		v = lane - 3 * rl;
		base = lane - v;
	}
	__device__ int slot(int g) const
	{
		return ((vb * G + g) * kWavesPerBlock + wid) * kRecvPerWave + lane / 3;
	}
	__device__ bool valid(int g, int nslots, int nblocks) const
	{
		return vb < nblocks && lane < 63 && slot(g) < nslots;
	}
};

template <int W = kWavesPerBlock> __device__ void add_block_counters(const int (&vals_in)[8], long long *bcnt)
{
	__shared__ int wcnt[W][8];
	int vals[8];
	for (int k = 0; k < 8; k++) vals[k] = wave_sum(vals_in[k]);
	if ((threadIdx.x & 63) == 0)
		for (int k = 0; k < 8; k++) wcnt[threadIdx.x >> 6][k] = vals[k];
	__syncthreads();
	if (threadIdx.x == 0)
		for (int k = 0; k < 8; k++) {
			long long sum = 0;
			for (int w = 0; w < W; w++) sum += wcnt[w][k];
			if (sum) bcnt[(size_t)blockIdx.x * 8 + k] += sum;
		}
}

// Source k's rays, part 1: which receivers take part (fixed at the first
// source: parent >= 0, not a source, not all nine accB >= 255,
// rad.cpp:420-428; bit 0 of live[slot], bit 1 holding parent >= 0 alone,
// which k_update's candidate test needs) and CalFF up to the ray (shade.cpp:107-123) for each
// (receiver, vertex, half) into ffbuf[2 * (3 * slot + vertex) + half]; 0
// where no ray is cast.  For k_compact, per receiver vertex in position
// order (rcode[3 p + vertex], p the receiver's position in R.order): bit h
// set where half h casts a ray, and the vertex's key (rkey, same order)
// above.
// PREP 1 or 2 (one source per iteration): k_prep's selection here instead,
// from the key the last k_update left (2: the lowest slot on ties).  Every
// block builds the source descriptor in LDS; block 0 publishes it and counts
// the shot; k_compact, after every block has read the source's unshot
// energy, zeroes it and the key.
template <bool F, int G, int PREP>
__global__ __launch_bounds__(kBlock) void k_gen(DRecv R, Loop *L, int k, unsigned char *live, float *ffbuf,
                                                const unsigned char *rkey, unsigned char *rcode, long long *bcnt,
                                                int nblocks, int *raycnt)
{
	__shared__ SrcDesc sd;
	const SrcDesc *sp;
	if (PREP) {
		if (L->done) return;
		const unsigned long long key = L->key;
		if (key == 0) {
			if (blockIdx.x == 0 && threadIdx.x == 0) L->done = 1;
			return;
		}
		if (threadIdx.x == 0) make_source<F>(R, key_slot<PREP == 2>(key), sd);
		__syncthreads();
		if (blockIdx.x == 0 && threadIdx.x == 0) {
			L->src[0] = sd;
			L->nsrc = 1;
			L->shots++;
			L->ncand = 0;
		}
		sp = &sd;
	} else {
		if (L->done || k >= L->nsrc) return;
		sp = &L->src[k];
	}
	const SrcDesc &s = *sp;
	const RecvMap<G> m(nblocks);
	int ref_rays = 0, nray = 0;
#pragma unroll 1
	for (int g = 0; g < G; g++) {
		const bool valid = m.valid(g, R.nslots, nblocks);
		const int pos = m.slot(g);
		const int r = valid ? R.order[pos] : -1;
		// loaded before the stores below, which a byte pointer may alias
		const int key = valid ? rkey[3 * pos + m.v] : 0;
		bool lv;
		if (k == 0) {
			int parent = -1;
			float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
			if (valid) {
				parent = __float_as_int(R.nrm[r].w);
				acc = R.acc[3 * r + m.v];
			}
			const int vfull = !(acc.x < 255.0f || acc.y < 255.0f || acc.z < 255.0f);
			const bool full = __shfl(vfull, m.base) && __shfl(vfull, m.base + 1) && __shfl(vfull, m.base + 2);
			bool chosen = false;
			if (PREP) chosen = sd.phys == r;
			else
				for (int j = 0; j < L->nsrc; j++) chosen |= (L->src[j].phys == r);
			lv = valid && parent >= 0 && !chosen && !full;
			if (valid && m.v == 0) live[r] = (unsigned char)(lv | (parent >= 0) << 1);
		} else {
			lv = valid && (live[r] & 1);
		}
		if (!valid) continue;
		float f0 = 0.f, f1 = 0.f;
		const int sv = 3 * r + m.v;
		if (lv) {
			const float4 nr = R.nrm[r];
			const float nlen = __builtin_sqrtf(dot3<F>(nr.x, nr.y, nr.z, nr.x, nr.y, nr.z));
			const float px = R.vx[sv], py = R.vy[sv], pz = R.vz[sv];
			if (s.split) {
				// the full-source ray AdaptCalFF computes and then discards
				if (ff_prefix_nonzero<F>(s.whole, nr.x, nr.y, nr.z, nlen, px, py, pz)) ref_rays++;
				f0 = ff_prefix<F>(s.half[0], nr.x, nr.y, nr.z, nlen, px, py, pz);
				f1 = ff_prefix<F>(s.half[1], nr.x, nr.y, nr.z, nlen, px, py, pz);
			} else {
				f0 = ff_prefix<F>(s.whole, nr.x, nr.y, nr.z, nlen, px, py, pz);
			}
		}
		ffbuf[2 * sv] = f0;
		ffbuf[2 * sv + 1] = f1;
		const int bits = (f0 != 0.0f) | (f1 != 0.0f) << 1;
		rcode[3 * pos + m.v] = (unsigned char)(bits | key << 2);
		nray += (bits & 1) + (bits >> 1);
	}
	// rays of this block's contiguous lane range [vb * 6 * kRecvPerBlock<G>, +6 * kRecvPerBlock<G>)
	__shared__ int wray[kWavesPerBlock];
	for (int off = 32; off > 0; off >>= 1) nray += __shfl_xor(nray, off);
	if ((threadIdx.x & 63) == 0) wray[threadIdx.x >> 6] = nray;
	__syncthreads();
	if (threadIdx.x == 0 && m.vb < nblocks) {
		int t = 0;
		for (int w = 0; w < kWavesPerBlock; w++) t += wray[w];
		raycnt[m.vb] = t;
	}
	const int vals[8] = {0, ref_rays, 0, 0, 0, 0, 0, 0};
	add_block_counters(vals, bcnt);
}

// The rays of source k as a dense list: block b (one per k_gen range) takes
// positions q of [b * 6 * kRecvPerBlock<G>, +6 * kRecvPerBlock<G>), lane i = 6 *
// order[q / 6] + q % 6, and places those whose form-factor prefix is
// non-zero (bit q % 2 of rcode[q / 2]; both it and order are read in q's
// order) at offset sum(raycnt[0..b)), grouped by their origin's key (the
// list length of the leaf holding the receiver vertex, in kRayBuckets
// steps), smallest key first, in ascending q within a key; the last block
// stores the total.  A ray's cost follows mostly the lists it tests, so
// grouping the rays of one neighbourhood by that key makes each wave's 64
// rays alike (less divergence: k_trace 272 -> 255 us on the hall); which
// rays there are, and each one's result, do not depend on the order.
template <int G>
__global__ __launch_bounds__(kBlock) void k_compact(Loop *L, int k, const unsigned char *rcode, const int *order,
                                                    int nlanes, const int *raycnt, int nblocks, int *rays,
                                                    float4 *dB, int prep)
{
	if (L->done || k >= L->nsrc) return;
	constexpr int kRounds = (6 * kRecvPerBlock<G> + kBlock - 1) / kBlock;
	__shared__ int wsum[kWavesPerBlock];
	__shared__ int slot[kRounds][kWavesPerBlock][kRayBuckets];  // rays per (round, wave, key), then their first slot
	__shared__ int ktot[kRayBuckets], kfirst[kRayBuckets];
	__shared__ int total;
	const int b = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	int off = 0;
	for (int j = threadIdx.x; j < b; j += kBlock) off += raycnt[j];
	for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o);
	if (lane == 0) wsum[w] = off;
	constexpr int chunk = 6 * kRecvPerBlock<G>;
	const int i0 = b * chunk, i1 = min(nlanes, i0 + chunk);
	const unsigned long long below = (1ull << lane) - 1;
	int ri[kRounds], rk[kRounds];
	unsigned long long rm[kRounds];
#pragma unroll
	for (int r = 0; r < kRounds; r++) {
		const int q = i0 + r * kBlock + threadIdx.x;
		int i = 0, key = -1;  // -1: no ray
		if (q < i1) {
			const int cd = rcode[q >> 1];
			i = 6 * order[q / 6] + q % 6;
			key = (cd >> (q & 1)) & 1 ? cd >> 2 : -1;
		}
		unsigned long long mine = 0;
#pragma unroll
		for (int kb = 0; kb < kRayBuckets; kb++) {
			const unsigned long long m = __ballot(key == kb);
			if (key == kb) mine = m;
			if (lane == kb) slot[r][w][kb] = __popcll(m);
		}
		ri[r] = i;
		rk[r] = key;
		rm[r] = mine;
	}
	__syncthreads();
	if (threadIdx.x < kRayBuckets) {
		int t = 0;
		for (int r = 0; r < kRounds; r++)
			for (int v = 0; v < kWavesPerBlock; v++) t += slot[r][v][threadIdx.x];
		ktot[threadIdx.x] = t;
	}
	__syncthreads();
	if (threadIdx.x == 0) {  // smallest key first, after the rays of the blocks before
		int run = wsum[0] + wsum[1] + wsum[2] + wsum[3];
		for (int kb = 0; kb < kRayBuckets; kb++) {
			kfirst[kb] = run;
			run += ktot[kb];
		}
		total = run;
	}
	__syncthreads();
	if (threadIdx.x < kRayBuckets) {  // within a key: rounds, then waves, in order
		int run = kfirst[threadIdx.x];
		for (int r = 0; r < kRounds; r++)
			for (int v = 0; v < kWavesPerBlock; v++) {
				const int c = slot[r][v][threadIdx.x];
				slot[r][v][threadIdx.x] = run;
				run += c;
			}
	}
	__syncthreads();
#pragma unroll
	for (int r = 0; r < kRounds; r++)
		if (rk[r] >= 0) rays[slot[r][w][rk[r]] + __popcll(rm[r] & below)] = ri[r];
	if (b == nblocks - 1 && threadIdx.x == 0) {
		L->nrays = total;
		if (prep) {  // what k_prep does after the selection (k_gen with PREP)
			L->key = 0;
			dB[L->src[0].phys] = make_float4(0.f, 0.f, 0.f, 0.f);
		}
	}
}

// Key of each receiver vertex for k_compact's grouping: the list length of
// the leaf that holds the vertex (the first leaf every ray from it tests),
// in steps of 8, capped at kRayBuckets - 1; 0 outside the grid.  Stored in
// position order: rkey[3 p + v] for vertex v of slot R.order[p].
__global__ __launch_bounds__(kBlock) void k_ray_keys(DScene S, DRecv R, int nv, unsigned char *rkey)
{
	const int t = blockIdx.x * kBlock + threadIdx.x;
	if (t >= nv) return;
	const int sv = 3 * R.order[t / 3] + t % 3;
	const float px = R.vx[sv], py = R.vy[sv], pz = R.vz[sv];
	int key = 0;
	if (!(px < S.g0x || px > S.g1x || py < S.g0y || py > S.g1y || pz < S.g0z || pz > S.g1z)) {
		const int gx = cell_coord(px, S.g0x, S), gy = cell_coord(py, S.g0y, S), gz = cell_coord(pz, S.g0z, S);
		if ((unsigned)gx < (unsigned)S.gt && (unsigned)gy < (unsigned)S.gt && (unsigned)gz < (unsigned)S.gt) {
			const int leaf = S.cell_leaf[(gx * S.gt + gy) * S.gt + gz];
			const int cnt = __float_as_int(S.leaf_box[2 * leaf + 1].w) & 0xffffff;
			key = min(kRayBuckets - 1, cnt >> 3);
		}
	}
	rkey[t] = (unsigned char)key;
}

// Source k's rays, part 2: the compacted list, which keeps the receivers'
// spatial order, so a wave's rays start from neighbouring receivers toward
// the same point and every lane has a ray.  RayHitted (raycast.cpp:335-360);
// a ray that does not reach the source zeroes its form factor
// (shade.cpp:124-126).  XCD-contiguous block order.  Two lanes share each
// ray's leaf lists (each tests every second entry), or four or eight where
// fewer would leave most of the machine's wave slots idle: measured on the
// hall against one lane per ray (242 us) and four (239 us), two take 227 us,
// the list loop's lanes being busier than the walk's doubled steps cost; on
// the 2120-slot room eight lanes take an iteration from 49.4 to 48.4 us.
__host__ __device__ __forceinline__ int lanes_per_ray(int nr)
{
	return 8LL * nr <= kTraceSlots ? 8 : 4LL * nr <= kTraceSlots ? 4 : 2;
}

// Blocks of k_trace with two lanes per ray (nr rays, nr > kTraceSlots / 4):
// XCD x's segment [nr x / 8, nr (x + 1) / 8) of the list (XCD-contiguous, as
// the receivers' spatial order) is split at sc = end - len * kEndPct / 100
// and sb = sc - len * kTailPct / 100.  Its head [start, sb) goes to blocks
// 8 b + x, b < head, 32 rays each at two lanes per ray; its tail [sb, sc)
// to blocks 8 (head + b) + x, b < tail, 64 / kTailLanes rays each at
// kTailLanes lanes per ray; its end [sc, end) to the next `end` blocks of
// the XCD, 8 rays each at eight lanes.  Dispatched last, the tail's waves
// are short, so the launch does not end on long waves
// started just before the list ran out (timed per wave on the hall without
// the split: every slot busy until 62 % of the launch, then 40 % of it
// draining waves of 47 us on average, up to 130 us).  The host sizes the
// grid for 6 n rays; blocks past the split of the actual count exit.
struct TraceSplit {
	int head, tail, end;  // blocks per XCD
};
__host__ __device__ __forceinline__ TraceSplit trace_split(long long nr)
{
	const long long sm = nr >> 3;                                            // shortest segment
	const long long hmax = (sm + 1) - sm * kTailPct / 100 - sm * kEndPct / 100;  // longest head
	const long long tmax = (sm + 1) * kTailPct / 100;                       // longest tail
	const long long emax = (sm + 1) * kEndPct / 100;
	const int rt = kTraceBlock / kTailLanes;
	return TraceSplit{(int)((hmax + 31) / 32), (int)((tmax + rt - 1) / rt), (int)((emax + 7) / 8)};
}

// Six waves per SIMD (80 VGPRs, a 28-byte spill per leaf step) against five
// (82 VGPRs, none): 223.6 -> 213.4 us per launch on the hall; on the 2120-slot
// room, where latency rules, 48.2 -> 48.5 us per iteration.
template <bool F>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_trace(
	DScene S, DRecv R, Loop *L, int k, float *ffbuf, long long *bcnt, const int *rays)
{
	if (L->done || k >= L->nsrc) return;
	const int nr = L->nrays;
	const int xcd = blockIdx.x & 7;
	int b = blockIdx.x >> 3;
	// this lane group's ray j (-1: none) and its lanes per ray
	int lanes = lanes_per_ray(nr), j = -1;
	if (lanes > 2) {
		// few rays: 64 / lanes rays per block, XCD-contiguous
		const int rpb = kTraceBlock / lanes;
		const int nb = (nr + rpb - 1) / rpb, per = (nb + 7) >> 3;
		const int jj = (xcd * per + b) * rpb + (int)threadIdx.x / lanes;
		if (b < per && jj < nr) j = jj;
	} else {
		const TraceSplit ts = trace_split(nr);
		const int s0 = (int)((long long)nr * xcd >> 3), s1 = (int)((long long)nr * (xcd + 1) >> 3);
		const int sc = s1 - (int)((long long)(s1 - s0) * kEndPct / 100);
		const int sb = sc - (int)((long long)(s1 - s0) * kTailPct / 100);
		if (b < ts.head) {
			const int jj = s0 + b * (kTraceBlock / 2) + (int)threadIdx.x / 2;
			if (jj < sb) j = jj;
		} else if ((b -= ts.head) < ts.tail) {
			lanes = kTailLanes;
			const int jj = sb + b * (kTraceBlock / kTailLanes) + (int)threadIdx.x / kTailLanes;
			if (jj < sc) j = jj;
		} else if ((b -= ts.tail) < ts.end) {
			lanes = 8;
			const int jj = sc + b * (kTraceBlock / 8) + (int)threadIdx.x / 8;
			if (jj < s1) j = jj;
		}
	}
	Cnt c = {0, 0, 0, 0, 0, 0, 0};
	if (j >= 0) {
		const int i = rays[j], sub = (int)threadIdx.x & (lanes - 1);
		const SrcDesc &s = L->src[k];
		const int sv = i >> 1, h = i & 1, r = sv / 3;
		const float px = R.vx[sv], py = R.vy[sv], pz = R.vz[sv];
		const int ban = __float_as_int(R.nrm[r].w);
		const SrcPatch &sp = s.split ? s.half[h] : s.whole;
		const float vx = sp.c[0] - px, vy = sp.c[1] - py, vz = sp.c[2] - pz;
		Cnt w = {0, 0, 0, 0, 0, 0, 0};
		const int hit = lanes == 8   ? walk<F, 8>(S, px, py, pz, vx, vy, vz, ban, w, sub)
		                : lanes == 4 ? walk<F, 4>(S, px, py, pz, vx, vy, vz, ban, w, sub)
		                             : walk<F, 2>(S, px, py, pz, vx, vy, vz, ban, w, sub);
		// every lane of a group counts its own tests; the rest once
		c.tests = w.tests;
		if (sub == 0) {
			c.rays = c.ref_rays = 1;
			c.cells = w.cells;
			c.levels = w.levels;
			c.precision = w.precision;
			c.stalled = w.stalled;
			if (hit != s.log) ffbuf[i] = 0.0f;
		}
	}
	const int vals[8] = {c.rays, c.ref_rays, c.cells, c.tests, c.levels, c.precision, c.stalled, 0};
	add_block_counters<kTraceBlock / kWave>(vals, bcnt);
}

// Source k's rays with no dense list, where every one of the 6 n ray
// positions fits the machine at eight lanes per ray (48 n <= kTraceSlots):
// lane group q (eight lanes) takes position q of k_compact's order, lane
// i = 6 order[q / 6] + q % 6, and traces it if rcode marks a ray there.
// There are lane groups to spare, so the list and its launch buy nothing;
// what k_compact does besides (the key and the source's deltaB with PREP)
// block 0 does here.
template <bool F>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_trace_direct(
	DScene S, DRecv R, Loop *L, int k, float *ffbuf, long long *bcnt, const unsigned char *rcode, int nlanes,
	float4 *dB, int prep)
{
	if (L->done || k >= L->nsrc) return;
	const SrcDesc &s = L->src[k];
	if (prep && blockIdx.x == 0 && threadIdx.x == 0) {
		L->key = 0;
		dB[s.phys] = make_float4(0.f, 0.f, 0.f, 0.f);
	}
	const int q = blockIdx.x * (kTraceBlock / 8) + (int)threadIdx.x / 8, sub = (int)threadIdx.x & 7;
	Cnt c = {0, 0, 0, 0, 0, 0, 0};
	if (q < nlanes && ((rcode[q >> 1] >> (q & 1)) & 1)) {
		const int i = 6 * R.order[q / 6] + q % 6;
		const int sv = i >> 1, h = i & 1, r = sv / 3;
		const float px = R.vx[sv], py = R.vy[sv], pz = R.vz[sv];
		const int ban = __float_as_int(R.nrm[r].w);
		const SrcPatch &sp = s.split ? s.half[h] : s.whole;
		Cnt w = {0, 0, 0, 0, 0, 0, 0};
		const int hit = walk<F, 8>(S, px, py, pz, sp.c[0] - px, sp.c[1] - py, sp.c[2] - pz, ban, w, sub);
		c.tests = w.tests;
		if (sub == 0) {
			c.rays = c.ref_rays = 1;
			c.cells = w.cells;
			c.levels = w.levels;
			c.precision = w.precision;
			c.stalled = w.stalled;
			if (hit != s.log) ffbuf[i] = 0.0f;
		}
	}
	const int vals[8] = {c.rays, c.ref_rays, c.cells, c.tests, c.levels, c.precision, c.stalled, 0};
	add_block_counters<kTraceBlock / kWave>(vals, bcnt);
}

// Source k, part 3: CalRadiosity (shade.cpp:244-259) per receiver, its
// three vertex lanes combined by shuffles in the reference's order; after
// the last source the next iteration's candidate key (listCanditate,
// rad.cpp:387-404) by wave, block and one 64-bit atomicMax per block.
template <bool F, bool FUSED_SELECT, bool LOWEST, int G>
__global__ __launch_bounds__(kBlock) void k_update(DRecv R, Loop *L, int k, const unsigned char *live,
                                                   const float *ffbuf, float converge, long long *bcnt, int nblocks)
{
	if (L->done || k >= L->nsrc) return;
	__shared__ unsigned long long wmax[kWavesPerBlock];
	__shared__ int wn[kWavesPerBlock];
	const RecvMap<G> m(nblocks);
	const SrcDesc &s = L->src[k];
	const bool last = k == L->nsrc - 1;
	unsigned long long key = 0;
	int recv = 0, ncand = 0;
#pragma unroll
	for (int g = 0; g < G; g++) {
		const int r = m.slot(g);
		const bool valid = m.valid(g, R.nslots, nblocks);
		const int lb = valid ? live[r] : 0;  // bit 0: takes part; bit 1: parent >= 0
		const bool lv = lb & 1;
		const int sv = 3 * r + m.v;
		float ff = 0.f;
		float4 rho = make_float4(0.f, 0.f, 0.f, 0.f);
		if (lv) {
			ff = s.split ? ffbuf[2 * sv] + ffbuf[2 * sv + 1] : ffbuf[2 * sv];
			rho = R.rho[r];
		}
		const float kr = (float)(0.5 / 255.0);
		const float dr = rho.x * kr * s.dB[0] / 3 * ff;
		const float dg = rho.y * kr * s.dB[1] / 3 * ff;
		const float dbb = rho.z * kr * s.dB[2] / 3 * ff;
		const float r1 = __shfl(dr, m.base + 1), g1 = __shfl(dg, m.base + 1), b1 = __shfl(dbb, m.base + 1);
		const float r2 = __shfl(dr, m.base + 2), g2 = __shfl(dg, m.base + 2), b2 = __shfl(dbb, m.base + 2);
		float4 db = make_float4(0.f, 0.f, 0.f, 0.f);
		if (valid && m.v == 0) db = R.dB[r];
		if (lv) {
			float4 acc = R.acc[sv];
			acc.x += dr;
			acc.y += dg;
			acc.z += dbb;
			R.acc[sv] = acc;
			R.dacc[sv] = make_float4(dr, dg, dbb, 0.f);
			if (m.v == 0) {
				db.x += dr; db.x += r1; db.x += r2;
				db.y += dg; db.y += g1; db.y += g2;
				db.z += dbb; db.z += b1; db.z += b2;
				R.dB[r] = db;
				recv++;
			}
		}
		if (last && valid && m.v == 0 && (lb & 2)) {
			const float sum = db.x + db.y + db.z;
			if (sum > converge) {
				ncand++;
				if (FUSED_SELECT) {
					const unsigned long long kk = cand_key<LOWEST>(sum, r);
					key = kk > key ? kk : key;
				}
			}
		}
	}
	if (FUSED_SELECT) key = wave_max(key);
	ncand = wave_sum(ncand);
	if (m.lane == 0) {
		wmax[m.wid] = key;
		wn[m.wid] = ncand;
	}
	const int vals[8] = {0, 0, 0, 0, 0, 0, 0, recv};
	add_block_counters(vals, bcnt);  // synchronises the block
	if (threadIdx.x == 0) {
		for (int w = 1; w < kWavesPerBlock; w++) {
			key = wmax[w] > key ? wmax[w] : key;
			ncand += wn[w];
		}
		if (FUSED_SELECT && key) atomicMax(&L->key, key);
		if (ncand) atomicAdd(&L->ncand, ncand);
	}
}


// Explicit form factors (the linear system of the gathering solvers):
// F[i][j] = CalFF(source = whole patch j, receiver = patch i, point = the
// centroid of i) with its visibility ray (shade.cpp:107-127), 0 on the
// diagonal and for logical slots; a row whose sum exceeds 1 is scaled to 1.
// One block per row: the row's rays share their origin.  F is row-major
// with leading dimension ld (>= n); columns [n, ld) are left untouched.
template <bool F>
__global__ __launch_bounds__(kBlock) void k_ff_rows(DScene S, DRecv R, int n, int ld, float *Fm, long long *bcnt)
{
	__shared__ double wsum[kWavesPerBlock];
	const int i = blockIdx.x;
	if (i >= n) return;
	const float4 ni = R.nrm[i], ci = R.cA[i];
	const int pi = __float_as_int(ni.w);
	const float nlen_i = __builtin_sqrtf(dot3<F>(ni.x, ni.y, ni.z, ni.x, ni.y, ni.z));
	float *row = Fm + (size_t)i * ld;
	Cnt c = {0, 0, 0, 0, 0, 0, 0};
	double sum = 0.0;
	for (int j = threadIdx.x; j < n; j += kBlock) {
		float ff = 0.0f;
		const float4 nj = R.nrm[j], cj = R.cA[j];
		const int pj = __float_as_int(nj.w);
		if (j != i && pi >= 0 && pj >= 0) {
			SrcPatch sp;
			sp.c[0] = cj.x; sp.c[1] = cj.y; sp.c[2] = cj.z;
			sp.n[0] = nj.x; sp.n[1] = nj.y; sp.n[2] = nj.z;
			sp.area = cj.w;
			sp.nlen = __builtin_sqrtf(dot3<F>(nj.x, nj.y, nj.z, nj.x, nj.y, nj.z));
			ff = point_ff<F>(S, sp, pj, ni.x, ni.y, ni.z, nlen_i, pi, ci.x, ci.y, ci.z, c);
		}
		row[j] = ff;
		sum += (double)ff;
	}
	for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
	if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = sum;
	__syncthreads();
	const double total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
	if (total > 1.0) {
		const float scale = (float)(1.0 / total);
		for (int j = threadIdx.x; j < n; j += kBlock) row[j] *= scale;
	}
	const int vals[8] = {c.rays, c.ref_rays, c.cells, c.tests, c.levels, c.precision, c.stalled, 0};
	add_block_counters(vals, bcnt);
}

// ---- explicit matrix with segment visibility -------------------------------
// The gathering matrix asks whether anything blocks the open segment between
// two centroids.  Point location is geometric (the 2^D grid of the deepest
// octree level, each cell mapped to its leaf), the walk is Amanatides-Woo over
// that grid (at most 3 * 2^D steps), and each leaf's complete list is tested
// once (a leaf is a box, so its cells are consecutive along a segment).  The
// test of one original is TriHitted's plane and side-plane arithmetic with the
// segment's parameter range (1e-4, 1 - 1e-4).
constexpr float kSegLo = 0.0001f, kSegHi = 0.9999f;

struct DFine {
	const int *__restrict__ leaf;        // R^3 cell -> leaf
	const int *__restrict__ begin;       // leaf -> first record, n_leaves + 1
	const float4 *__restrict__ nc;       // 2 per record: {n, c.x}, {c.y, c.z, id, -}
	const float4 *__restrict__ se;       // 3 per record: side planes
	int r;
	float g0[3], h[3], inv[3];
};

template <bool F>
__device__ bool seg_occluded(const DFine &G, float px, float py, float pz, float vx, float vy, float vz, int ban0,
                             int ban1, Cnt &c)
{
	const int R = G.r;
	const float p[3] = {px, py, pz}, v[3] = {vx, vy, vz};
	int cell[3], step[3];
	float tmax[3], tdelta[3];
#pragma unroll
	for (int k = 0; k < 3; k++) {
		const float rel = p[k] - G.g0[k];
		int ck = (int)floorf(rel * G.inv[k]);
		ck = ck < 0 ? 0 : (ck >= R ? R - 1 : ck);
		cell[k] = ck;
		if (v[k] > 0.0f) {
			step[k] = 1;
			tmax[k] = ((float)(ck + 1) * G.h[k] - rel) / v[k];
			tdelta[k] = G.h[k] / v[k];
		} else if (v[k] < 0.0f) {
			step[k] = -1;
			tmax[k] = ((float)ck * G.h[k] - rel) / v[k];
			tdelta[k] = -G.h[k] / v[k];
		} else {
			step[k] = 0;
			tmax[k] = tdelta[k] = __builtin_inff();
		}
	}
	c.rays++;
	int last = -1;
	for (;;) {
		const int leaf = G.leaf[(cell[0] * R + cell[1]) * R + cell[2]];  // R^3 < 2^31
		if (leaf != last) {
			last = leaf;
			c.cells++;
			const int e0 = G.begin[leaf], e1 = G.begin[leaf + 1];
			// the whole 32-byte record at once, the next one in flight; the
			// load is not conditional (the index is clamped to the list), so
			// that the wait before a test covers only the loads it needs
			float4 na = make_float4(0.f, 0.f, 0.f, 0.f), nb = na;
			if (e0 < e1) {
				na = G.nc[2 * e0];
				nb = G.nc[2 * e0 + 1];
			}
			for (int e = e0; e < e1; e++) {
				const float4 a = na, b = nb;
				const int en = min(e + 1, e1 - 1);
				na = G.nc[2 * en];
				nb = G.nc[2 * en + 1];
				const int id = __float_as_int(b.z);
				if (id == ban0 || id == ban1) continue;
				c.tests++;
				const float a1 = dot3<F>(vx, vy, vz, a.x, a.y, a.z);
				const float a2 = dot3<F>(a.w - px, b.x - py, b.y - pz, a.x, a.y, a.z);
				const float tt = a2 / a1;
				if (!(tt > kSegLo && tt < kSegHi)) continue;
				const float hx = axpy<F>(tt, vx, px), hy = axpy<F>(tt, vy, py), hz = axpy<F>(tt, vz, pz);
				const float4 *se = G.se + 3 * e;
				const float4 s0 = se[0], s1 = se[1], s2 = se[2];
				if (!(dot3<F>(hx, hy, hz, s0.x, s0.y, s0.z) + s0.w < 0.0f) &&
				    !(dot3<F>(hx, hy, hz, s1.x, s1.y, s1.z) + s1.w < 0.0f) &&
				    !(dot3<F>(hx, hy, hz, s2.x, s2.y, s2.z) + s2.w < 0.0f))
					return true;
			}
		}
		const int ax = tmax[0] <= tmax[1] ? (tmax[0] <= tmax[2] ? 0 : 2) : (tmax[1] <= tmax[2] ? 1 : 2);
		if (tmax[ax] >= 1.0f) return false;
		cell[ax] += step[ax];
		if ((unsigned)cell[ax] >= (unsigned)R) return false;
		tmax[ax] += tdelta[ax];
	}
}

// Visibility of every unordered pair, once: tile (ti <= tj) of 64 x 64 pairs
// per block; wave w takes rows ti*64 + 16w .. +15, lane l column tj*64 + l,
// so a wave's 64 rays at a time leave one centroid for 64 neighbouring ones.
// The segment runs from the lower slot index to the higher; pairs that do
// not face each other (CalFF's sign test, which is symmetric) are left 0.
// vis[i * nt + j/64] bit j%64 = 1 when i and j see each other; the block
// writes its tile and, transposed through LDS, the mirror tile.
__device__ __forceinline__ int tri_tile(int t, int nt, int &tj)
{
	// t -> (ti, tj) over ti <= tj, row by row: row r starts at
	// o(r) = r (2 nt - r + 1) / 2; the root of o(r) = t, then exact fix-up
	const double b = 2.0 * nt + 1.0;
	int ti = (int)((b - __builtin_sqrt(b * b - 8.0 * t)) * 0.5);
	auto o = [nt](long long r) { return r * (2LL * nt - r + 1) / 2; };
	while (ti > 0 && o(ti) > t) ti--;
	while (ti + 1 < nt && o(ti + 1) <= t) ti++;
	tj = ti + (int)(t - o(ti));
	return ti;
}

template <bool F>
__global__ __launch_bounds__(kBlock) void k_vis_tiles(DFine G, DRecv R, int n, int nt,
                                                      unsigned long long *vis, long long *bcnt)
{
	__shared__ unsigned long long rowbits[64];
	int tj;
	const int ti = tri_tile(blockIdx.x, nt, tj);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int j = tj * 64 + lane;
	Cnt c = {0, 0, 0, 0, 0, 0, 0};
	float4 nj = make_float4(0.f, 0.f, 0.f, 0.f), cj = nj;
	int pj = -1;
	if (j < n) {
		nj = R.nrm[j];
		cj = R.cA[j];
		pj = __float_as_int(nj.w);
	}
	for (int q = 0; q < 16; q++) {
		const int i = ti * 64 + w * 16 + q;  // wave-uniform
		bool see = false;
		if (i < n && j < n && i < j && pj >= 0) {
			const float4 ni = R.nrm[i], ci = R.cA[i];
			const int pi = __float_as_int(ni.w);
			const float dx = cj.x - ci.x, dy = cj.y - ci.y, dz = cj.z - ci.z;
			const float ds = dot3<F>(dx, dy, dz, nj.x, nj.y, nj.z);
			const float dd = -dot3<F>(dx, dy, dz, ni.x, ni.y, ni.z);
			if (pi >= 0 && (ds < 0) == (dd < 0) && ds != 0.0f && dd != 0.0f)
				see = !seg_occluded<F>(G, ci.x, ci.y, ci.z, dx, dy, dz, pi, pj, c);
		}
		const unsigned long long bits = __ballot(see);
		if (lane == 0) rowbits[w * 16 + q] = bits;
	}
	__syncthreads();
	if (threadIdx.x < 64) {
		const int r = threadIdx.x;
		if (ti * 64 + r < n) vis[(size_t)(ti * 64 + r) * nt + tj] = rowbits[r];
		// mirror: row tj*64 + r, bits over the tile's rows
		unsigned long long m = 0;
		for (int q = 0; q < 64; q++) m |= ((rowbits[q] >> r) & 1ull) << q;
		if (tj * 64 + r < n) {
			if (ti == tj) m |= rowbits[r];  // diagonal tile: both halves in one word
			vis[(size_t)(tj * 64 + r) * nt + ti] = m;
		}
	}
	const int vals[8] = {c.rays, c.ref_rays, c.cells, c.tests, c.levels, c.precision, c.stalled, 0};
	add_block_counters(vals, bcnt);
}

// Row i (one block): F[i][j] = CalFF's value from the centroid of i to the
// whole patch j (shade.cpp:107-123) when i and j see each other (vis); 0 on
// the diagonal and for logical slots; the row scaled by (float)(1 / sum)
// when its double sum (fixed reduction order) exceeds 1.
template <bool F>
__global__ __launch_bounds__(kBlock) void k_ff_seg(DRecv R, int n, int ld, int nt, const unsigned long long *vis,
                                                   float *Fm)
{
	__shared__ double wsum[kWavesPerBlock];
	const int i = blockIdx.x;
	if (i >= n) return;
	const float4 ni = R.nrm[i], ci = R.cA[i];
	const int pi = __float_as_int(ni.w);
	const float nlen_i = __builtin_sqrtf(dot3<F>(ni.x, ni.y, ni.z, ni.x, ni.y, ni.z));
	float *row = Fm + (size_t)i * ld;
	const unsigned long long *vrow = vis + (size_t)i * nt;
	double sum = 0.0;
	for (int j = threadIdx.x; j < n; j += kBlock) {
		float ff = 0.0f;
		const float4 nj = R.nrm[j], cj = R.cA[j];
		const int pj = __float_as_int(nj.w);
		if (j != i && pi >= 0 && pj >= 0 && ((vrow[j >> 6] >> (j & 63)) & 1ull)) {
			SrcPatch sp;
			sp.c[0] = cj.x; sp.c[1] = cj.y; sp.c[2] = cj.z;
			sp.n[0] = nj.x; sp.n[1] = nj.y; sp.n[2] = nj.z;
			sp.area = cj.w;
			sp.nlen = __builtin_sqrtf(dot3<F>(nj.x, nj.y, nj.z, nj.x, nj.y, nj.z));
			ff = ff_prefix<F>(sp, ni.x, ni.y, ni.z, nlen_i, ci.x, ci.y, ci.z);
		}
		row[j] = ff;
		sum += (double)ff;
	}
	for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
	if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = sum;
	__syncthreads();
	const double total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
	if (total > 1.0) {
		const float scale = (float)(1.0 / total);
		for (int j = threadIdx.x; j < n; j += kBlock) row[j] *= scale;
	}
}

// The prepared scene's slot records (the 236-byte reference layout, copied
// to the device as they are) gathered into the device layouts: k_pack_items
// fills the octree's list entries with the originals inline, k_pack_recv the
// receivers' arrays.  Pure data movement.
__global__ __launch_bounds__(kBlock) void k_pack_items(const Patch *__restrict__ pool, const int *__restrict__ items,
                                                       const float *__restrict__ radius, int ni, float4 *lnc,
                                                       float4 *lse)
{
	const int k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= ni) return;
	const int id = items[k];
	const Patch &p = pool[id];
	lnc[2 * (size_t)k] = make_float4(p.n[0], p.n[1], p.n[2], p.c[0]);
	lnc[2 * (size_t)k + 1] = make_float4(p.c[1], p.c[2], __int_as_float(id), radius[id]);
	for (int q = 0; q < 3; q++) lse[3 * (size_t)k + q] = make_float4(p.se[q][0], p.se[q][1], p.se[q][2], p.se[q][3]);
}

__global__ __launch_bounds__(kBlock) void k_pack_recv(const Patch *__restrict__ pool, int n, float *vx, float *vy,
                                                      float *vz, float4 *nrm, float4 *cA, float4 *rho, float4 *acc,
                                                      float4 *dacc, float4 *dB)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) return;
	const Patch &p = pool[i];
	for (int v = 0; v < 3; v++) {
		vx[3 * i + v] = p.p[v][0];
		vy[3 * i + v] = p.p[v][1];
		vz[3 * i + v] = p.p[v][2];
		acc[3 * i + v] = make_float4(p.accB[v][0], p.accB[v][1], p.accB[v][2], 0.f);
		dacc[3 * i + v] = make_float4(p.deltaaccB[v][0], p.deltaaccB[v][1], p.deltaaccB[v][2], 0.f);
	}
	nrm[i] = make_float4(p.n[0], p.n[1], p.n[2], __int_as_float(p.parent));
	cA[i] = make_float4(p.c[0], p.c[1], p.c[2], p.area);
	rho[i] = make_float4((float)p.Frgb[0], (float)p.Frgb[1], (float)p.Frgb[2], 0.f);
	dB[i] = make_float4(p.deltaB[0], p.deltaB[1], p.deltaB[2], 0.f);
}

// RayHitted for n rays, L lanes per ray (as k_trace: each tests every L-th
// entry of a leaf's list; cells from the group's first lane, tests summed)
template <bool F, int L>
__global__ __launch_bounds__(kBlock) void k_cast(DScene S, const RayIn *rays, int n, int *hit, int *cells, int *tests)
{
	const int g = blockIdx.x * kBlock + threadIdx.x;
	const int i = g / L, sub = g % L;
	if (i >= n) return;  // a whole group: its lanes share i
	const RayIn r = rays[i];
	Cnt c = {0, 0, 0, 0, 0, 0, 0};
	const int h = walk<F, L>(S, r.p[0], r.p[1], r.p[2], r.v[0], r.v[1], r.v[2], r.ban, c, sub);
	int t = c.tests;
	for (int o = 1; o < L; o <<= 1) t += __shfl_xor(t, o);
	if (sub == 0) {
		hit[i] = h;
		cells[i] = c.cells;
		tests[i] = t;
	}
}

// k_cast over n rays with lanes per ray as k_trace would choose them
template <bool F>
void launch_cast(hipStream_t st, const DScene &S, const RayIn *rays, int n, int *hit, int *cells, int *tests)
{
	const int lpr = lanes_per_ray(n);
	const dim3 g((unsigned)(((size_t)n * lpr + kBlock - 1) / kBlock)), b(kBlock);
	if (lpr == 8) hipLaunchKernelGGL((k_cast<F, 8>), g, b, 0, st, S, rays, n, hit, cells, tests);
	else if (lpr == 4) hipLaunchKernelGGL((k_cast<F, 4>), g, b, 0, st, S, rays, n, hit, cells, tests);
	else hipLaunchKernelGGL((k_cast<F, 2>), g, b, 0, st, S, rays, n, hit, cells, tests);
}

// Scratch of rad_cast_static, the variant of rad_cast that allocates nothing:
// the rays and their three results for up to kStaticRays rays per call.  One
// array per GPU, held until the process ends; every call on a GPU shares it.
constexpr int kStaticRays = RAD_CAST_STATIC_MAX;
constexpr size_t kStaticRayBytes = (size_t)kStaticRays * sizeof(RayIn);
constexpr size_t kStaticBytes = kStaticRayBytes + (size_t)3 * kStaticRays * sizeof(int);
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

template <bool F>
__global__ __launch_bounds__(kBlock) void k_ff_pairs(DScene S, DRecv R, const int *pairs, int n, float *ff3)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) return;
	SrcDesc s;
	make_source<F>(R, pairs[2 * i], s);
	const int d = pairs[2 * i + 1];
	const float4 nr = R.nrm[d];
	const float nlen = __builtin_sqrtf(dot3<F>(nr.x, nr.y, nr.z, nr.x, nr.y, nr.z));
	Cnt c = {0, 0, 0, 0, 0, 0, 0};
	for (int v = 0; v < 3; v++) {
		const float px = R.vx[3 * d + v], py = R.vy[3 * d + v], pz = R.vz[3 * d + v];
		const int logdst = __float_as_int(nr.w);
		float ff;
		if (s.split)
			ff = point_ff<F>(S, s.half[0], s.log, nr.x, nr.y, nr.z, nlen, logdst, px, py, pz, c) +
			     point_ff<F>(S, s.half[1], s.log, nr.x, nr.y, nr.z, nlen, logdst, px, py, pz, c);
		else
			ff = point_ff<F>(S, s.whole, s.log, nr.x, nr.y, nr.z, nlen, logdst, px, py, pz, c);
		ff3[3 * i + v] = ff;
	}
}

// ---- V3: Delta-ff adaptive splitting (shade.cpp:463-525 with PARALLEL) -------
// The reference's V3 build shades receivers in slot order, and a receiver of
// area >= -adapt_area whose vertex form factors do not predict the one at its
// centroid (|mean - centre| >= -delta_ff), or at random points drawn with
// rand() (GetPointInTriangle, shade.cpp:49-68), is split at its centroid into
// three (PartitionTriangleWithCenterAndStore, shade.cpp:261-306); the new
// slots are shaded in further passes (rad.cpp:406-451).  On the device:
// k_v3_gen computes every receiver's form factors (its vertices and centroid)
// in parallel; k_v3_commit walks the receivers in order on one block, keeping
// the pool count and the rand() state, drawing each receiver's samples 64 at
// a time across the lanes, and writing the splits; k_v3_apply then runs
// CalRadiosity (shade.cpp:244-259) for the receivers that keep their form
// factors.  A split only changes its own slot and new ones, so the form
// factors of the later receivers of a pass are those the reference computes.
enum { V3_NONE = 0, V3_APPLY = 1 };

struct V3Dev {
	Patch *pool;         // the slot records (static fields; the SoA arrays hold accB, deltaaccB, deltaB)
	float *ff;           // 4 per slot: the vertices' form factors, then the centroid's
	unsigned char *act;  // per slot: V3_APPLY when CalRadiosity follows
	int cap;             // slots available (min(MaxTri, -triangle))
	int n_orig;
	float adapt_area, sample_area, delta_ff;
};

// AdaptCalFF(CalFF(..)) at a point of a receiver with normal n
// (shade.cpp:141-168): the whole source's form factor, or, for a split
// source, the discarded whole one and then the sum of the two halves'.  On
// the four lanes q = 0..3 of an aligned group: lanes 0-1 walk the first half
// (or the whole source), lanes 2-3 the second half, two lanes sharing each
// ray's lists (walk<F, 2>); every lane of the group returns the form
// factor, the halves summed in AdaptCalFF's order.  Tests
// are counted by every lane, the rest once per ray (its first lane).  The
// four lanes call it together.
template <bool F>
__device__ float adapt_ff4(const DScene &S, const SrcDesc &s, float nx, float ny, float nz, float nlen, int logdst,
                           float px, float py, float pz, int q, Cnt &c)
{
	const int h = q >> 1, sub = q & 1;
	float f = 0.0f;
	if (s.split || h == 0) {
		const SrcPatch &sp = s.split ? s.half[h] : s.whole;
		if (s.split && q == 0 && ff_prefix<F>(s.whole, nx, ny, nz, nlen, px, py, pz) != 0.0f) c.ref_rays++;
		const float ff = ff_prefix<F>(sp, nx, ny, nz, nlen, px, py, pz);
		if (ff != 0.0f) {
			Cnt w = {0, 0, 0, 0, 0, 0, 0};
			const int hit = walk<F, 2>(S, px, py, pz, sp.c[0] - px, sp.c[1] - py, sp.c[2] - pz, logdst, w, sub);
			c.tests += w.tests;
			if (sub == 0) {
				c.rays++;
				c.ref_rays++;
				c.cells += w.cells;
				c.levels += w.levels;
				c.precision += w.precision;
				c.stalled += w.stalled;
			}
			f = hit == s.log ? ff : 0.0f;
		}
	}
	const int g0 = (int)(threadIdx.x & 63) & ~3;
	const float fa = __shfl(f, g0), fb = __shfl(f, g0 + 2);
	return s.split ? fa + fb : fa;
}

__device__ __forceinline__ float ffabs(float x) { return (float)fabs((double)x); }

// Slot r as the reference holds it: static fields from the record, the rest
// from the SoA arrays
__device__ void v3_load(const V3Dev &V, const DRecv &R, int r, Patch &t)
{
	t = V.pool[r];
	for (int v = 0; v < 3; v++) {
		const float4 a = R.acc[3 * r + v], d = R.dacc[3 * r + v];
		t.accB[v][0] = a.x; t.accB[v][1] = a.y; t.accB[v][2] = a.z;
		t.deltaaccB[v][0] = d.x; t.deltaaccB[v][1] = d.y; t.deltaaccB[v][2] = d.z;
	}
	const float4 b = R.dB[r];
	t.deltaB[0] = b.x; t.deltaB[1] = b.y; t.deltaB[2] = b.z;
}
__device__ void v3_store(const V3Dev &V, const DRecv &R, int r, const Patch &t)
{
	V.pool[r] = t;
	// the receiver arrays are read-only to the shooting kernels; the V3 commit
	// writes only slots no kernel of its pass reads afterwards
	for (int v = 0; v < 3; v++) {
		const_cast<float *>(R.vx)[3 * r + v] = t.p[v][0];
		const_cast<float *>(R.vy)[3 * r + v] = t.p[v][1];
		const_cast<float *>(R.vz)[3 * r + v] = t.p[v][2];
		R.acc[3 * r + v] = make_float4(t.accB[v][0], t.accB[v][1], t.accB[v][2], 0.f);
		R.dacc[3 * r + v] = make_float4(t.deltaaccB[v][0], t.deltaaccB[v][1], t.deltaaccB[v][2], 0.f);
	}
	const_cast<float4 *>(R.nrm)[r] = make_float4(t.n[0], t.n[1], t.n[2], __int_as_float(t.parent));
	const_cast<float4 *>(R.cA)[r] = make_float4(t.c[0], t.c[1], t.c[2], t.area);
	const_cast<float4 *>(R.rho)[r] = make_float4((float)t.Frgb[0], (float)t.Frgb[1], (float)t.Frgb[2], 0.f);
	R.dB[r] = make_float4(t.deltaB[0], t.deltaB[1], t.deltaB[2], 0.f);
}

// Form factors of receivers [r0, r1): 4 lanes per receiver, the vertices'
// (shade.cpp:442-450) and, where area >= -adapt_area, the centroid's
// (shade.cpp:469-470); act = V3_APPLY for each receiver the reference
// shades (parent >= 0, not the source, not saturated: rad.cpp:419-428).
template <bool F>
__global__ __launch_bounds__(kBlock) void k_v3_gen(DScene S, DRecv R, V3Dev V, Loop *L, int par, long long *bcnt)
{
	Cnt c = {0, 0, 0, 0, 0, 0, 0};
	const int r0 = L->v3_range[par][0], r1 = L->v3_range[par][1];
	// four lanes per form factor (adapt_ff4), four form factors per receiver;
	// the grid strides over the pass's slots
	for (int t = blockIdx.x * kBlock + threadIdx.x; !L->done && t < 16 * (r1 - r0); t += gridDim.x * kBlock) {
		const int r = r0 + t / 16, w = (t / 4) % 4, g4 = t % 4;
		const SrcDesc &s = L->src[0];
		const float4 nr = R.nrm[r];
		const int parent = __float_as_int(nr.w);
		bool sat = true;
		for (int v = 0; v < 3; v++) {
			const float4 a = R.acc[3 * r + v];
			sat = sat && !(a.x < 255.0f || a.y < 255.0f || a.z < 255.0f);
		}
		const bool live = parent >= 0 && r != s.phys && !sat;
		if (w == 0 && g4 == 0) V.act[r] = live ? V3_APPLY : V3_NONE;
		float f = 0.0f;
		const float4 ca = R.cA[r];
		if (live && (w < 3 || ca.w >= V.adapt_area)) {
			const float nlen = __builtin_sqrtf(dot3<F>(nr.x, nr.y, nr.z, nr.x, nr.y, nr.z));
			float px, py, pz;
			if (w < 3) {
				px = R.vx[3 * r + w];
				py = R.vy[3 * r + w];
				pz = R.vz[3 * r + w];
			} else {
				px = ca.x;
				py = ca.y;
				pz = ca.z;
			}
			f = adapt_ff4<F>(S, s, nr.x, nr.y, nr.z, nlen, parent, px, py, pz, g4, c);
		}
		if (g4 == 0) V.ff[4 * (size_t)r + w] = f;
	}
	const int vals[8] = {c.rays, c.ref_rays, c.cells, c.tests, c.levels, c.precision, c.stalled, 0};
	add_block_counters(vals, bcnt);
}

// glibc rand() (TYPE_3): r[i] = r[i - 3] + r[i - 31] (mod 2^32), output
// r[i] >> 1; the state is the last 34 values, the oldest at st[k].
//
// v3_draw: the next n (<= 3 kWave) values, drawn by the block's first wave
// into seq[34 .. 34 + n), after the state's 34 in order in seq[0, 34).  In
// a block of 31 consecutive values every r[i - 31] is already known, so
// within the block r[b + t] = r[b + t % 3 - 3] + (the sum of known_u over
// u <= t, u = t mod 3): three interleaved prefix sums, four shuffle steps.
constexpr int kV3Seq = 34 + 3 * kWave + 31;
__device__ __forceinline__ void v3_draw(const unsigned *st, int sk, unsigned *seq, int n, int tid)
{
	if (tid < kWave) {  // one wave: its LDS accesses stay in order
		const int lane = tid;
		if (lane < 34) seq[lane] = st[(sk + lane) % 34];
		for (int b = 34; b < 34 + n; b += 31) {
			unsigned x = lane < 31 ? seq[b + lane - 31] : 0u;
			for (int d = 3; d < 31; d <<= 1) {  // 3, 6, 12, 24: within a residue class
				const unsigned y = __shfl_up(x, d);
				if (lane >= d) x += y;
			}
			if (lane < 31) seq[b + lane] = x + seq[b + lane % 3 - 3];
		}
	}
	__syncthreads();
}
// the state after the first n values of seq's draws (k = 0: oldest first)
__device__ __forceinline__ void v3_advance(unsigned *st, int &sk, const unsigned *seq, int n, int tid)
{
	__syncthreads();
	if (tid < 34) st[tid] = seq[n + tid];
	if (tid == 0) sk = 0;
	__syncthreads();
}

template <bool F> __device__ void v3_center(float p[3][3], float *c)
{
	for (int k = 0; k < 3; k++) c[k] = (float)((double)(p[0][k] + p[1][k] + p[2][k]) / 3.0);
}
// CalArea (vector.h:58-62): n = (p1 - p0) x (p2 - p0), not normalised
template <bool F> __device__ void v3_area(Patch &t)
{
	float a[3], b[3];
	for (int k = 0; k < 3; k++) {
		a[k] = t.p[1][k] - t.p[0][k];
		b[k] = t.p[2][k] - t.p[0][k];
	}
	auto det2 = [](float x, float y, float z, float w) { return F ? __builtin_fmaf(x, y, -(z * w)) : x * y - z * w; };
	t.n[0] = det2(a[1], b[2], b[1], a[2]);
	t.n[1] = det2(a[2], b[0], b[2], a[0]);
	t.n[2] = det2(a[0], b[1], b[0], a[1]);
	t.area = __builtin_sqrtf(dot3<F>(t.n[0], t.n[1], t.n[2], t.n[0], t.n[1], t.n[2])) / 2;
}

// The receivers [r0, r1) in order (see above) on one block of four waves:
// every wave follows the same decisions; a chunk's 64 sample points take
// four lanes each (adapt_ff4), the first wave draws the random numbers,
// thread 0 writes the splits.
template <bool F>
__global__ __launch_bounds__(4 * kWave) void k_v3_commit(DScene S, DRecv R, V3Dev V, Loop *L, int par,
                                                         long long *bcnt)
{
	__shared__ unsigned st[34], seq[kV3Seq];
	__shared__ int sk;
	__shared__ unsigned long long pmask[4];  // each wave's departing points (bit 4k: its point k)
	__shared__ float pdff[kWave];            // each point's |g - fs|
	const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
	const int pt = tid >> 2, q = tid & 3;  // a chunk's point and the lane in its group of four
	Cnt c = {0, 0, 0, 0, 0, 0, 0};
	if (L->done || L->v3_error) return;
	const int r0 = L->v3_range[par][0], r1 = L->v3_range[par][1];
	if (tid < 34) st[tid] = L->rng[tid];
	if (tid == 0) sk = L->rng_k;
	__syncthreads();
	const SrcDesc &s = L->src[0];
	const int limit = V.cap;
	int trinum = L->trinum;
	bool stop = false;
	for (int base = r0; base < r1 && !stop; base += kWave) {
		// Once the pool is full every later receiver keeps its form factors
		// (the first test below), so nothing is left to do in order.
		if (trinum + 4 >= limit) break;
		// This chunk's receivers that need the ordered steps, found in
		// parallel: live, of area >= -adapt_area (shade.cpp:455-458); the
		// others keep their vertices' form factors.
		const int rl = base + lane;
		const bool cand = rl < r1 && V.act[rl] == V3_APPLY && R.cA[rl].w >= V.adapt_area;
		unsigned long long todo = __ballot(cand);
		while (todo) {
			const int r = base + (__ffsll((long long)todo) - 1);
			todo &= todo - 1;
			const float4 ca = R.cA[r];
			const float area = ca.w;
			if (trinum + 4 >= limit || (double)area < 1e-8) continue;   // shade.cpp:463-466
			const float f0 = V.ff[4 * (size_t)r], f1 = V.ff[4 * (size_t)r + 1], f2 = V.ff[4 * (size_t)r + 2];
			float groud = (float)((double)(f0 + f1 + f2) / 3.0);
			float dff = ffabs(groud - V.ff[4 * (size_t)r + 3]);
			if (dff < V.delta_ff) {
				// GetPointInTriangle samples until one departs by more than -delta_ff
				const int samples = (int)(area / V.sample_area);
				const float4 nr = R.nrm[r];
				const float nlen = __builtin_sqrtf(dot3<F>(nr.x, nr.y, nr.z, nr.x, nr.y, nr.z));
				const int logdst = __float_as_int(nr.w);
				float P[3][3];
				for (int v = 0; v < 3; v++) {
					P[v][0] = R.vx[3 * r + v];
					P[v][1] = R.vy[3 * r + v];
					P[v][2] = R.vz[3 * r + v];
				}
				for (int base = 0; base < samples; base += kWave) {
					const int m = min(kWave, samples - base);
					v3_draw(st, sk, seq, 3 * m, tid);
					bool brk = false;
					if (pt < m) {
						const unsigned *o = seq + 34 + 3 * pt;
						float w1 = (float)(int)(o[0] >> 1), w2 = (float)(int)(o[1] >> 1), w3 = (float)(int)(o[2] >> 1);
						const float tot = w1 + w2 + w3;
						w1 /= tot;
						w2 /= tot;
						w3 /= tot;
						const float qx = dot3<F>(w1, w2, w3, P[0][0], P[1][0], P[2][0]);
						const float qy = dot3<F>(w1, w2, w3, P[0][1], P[1][1], P[2][1]);
						const float qz = dot3<F>(w1, w2, w3, P[0][2], P[1][2], P[2][2]);
						const float g = dot3<F>(w1, w2, w3, f0, f1, f2);
						const float fs = adapt_ff4<F>(S, s, nr.x, nr.y, nr.z, nlen, logdst, qx, qy, qz, q, c);
						const float mydff = ffabs(g - fs);
						brk = mydff > V.delta_ff;
						if (q == 0) pdff[pt] = mydff;
					}
					const unsigned long long bw = __ballot(brk && q == 0);
					if (lane == 0) pmask[wv] = bw;
					__syncthreads();
					unsigned long long bal = 0;  // bit p: point p departs
					for (int w = 0; w < 4; w++) {
						const unsigned long long x = pmask[w];
						for (int k = 0; k < 16; k++) bal |= ((x >> (4 * k)) & 1ull) << (16 * w + k);
					}
					if (bal) {
						const int j = __ffsll((long long)bal) - 1;
						dff = pdff[j];
						// the state after the 3 (j + 1) draws the reference made
						v3_advance(st, sk, seq, 3 * (j + 1), tid);
						break;
					}
					dff = pdff[m - 1];
					v3_advance(st, sk, seq, 3 * m, tid);
				}
				if (dff < V.delta_ff) continue;  // CalRadiosity with the vertices' form factors
			}
			// PartitionTriangleWithCenterAndStore (shade.cpp:261-306)
			if (trinum + 4 >= limit || (double)area < 1e-8) continue;
			Patch src;
			v3_load(V, R, r, src);
			float maxl = 0.0f, minl = 1e+30f;
			for (int v = 0; v < 3; v++) {
				const int w = (v + 1) % 3;
				const float lx = src.p[w][0] - src.p[v][0], ly = src.p[w][1] - src.p[v][1], lz = src.p[w][2] - src.p[v][2];
				const float len = dot3<F>(lx, ly, lz, lx, ly, lz);
				maxl = (maxl < len) ? len : maxl;
				minl = (len < minl) ? len : minl;
			}
			if ((double)(minl / maxl) < 1e-1 || minl < 1) continue;
			if (src.parent != r && r < V.n_orig) {
				// the split would rewrite an original the octree lists (a slot
				// pre-partition overwrote): the rays of the later receivers
				// would change mid-pass; not supported
				if (tid == 0) L->v3_error = r + 1;
				stop = true;
				break;
			}
			const int id[3] = {trinum, trinum + 1, trinum + 2};
			Patch t[3];
			for (int k = 0; k < 3; k++) {
				t[k] = src;
				for (int x = 0; x < 3; x++) t[k].p[k][x] = src.c[x];
				v3_center<F>(t[k].p, t[k].c);
				v3_area<F>(t[k]);
			}
			t[0].neighbor[0] = id[2], t[0].neighbor[2] = id[1];
			t[1].neighbor[1] = id[0], t[1].neighbor[0] = id[2];
			t[2].neighbor[2] = id[1], t[2].neighbor[1] = id[0];
			if (tid == 0) {
				v3_store(V, R, id[0], t[0]);
				v3_store(V, R, id[1], t[1]);
			}
			if (src.parent == r) {
				// an original becomes logical; its three parts are shaded next pass
				if (tid == 0) {
					v3_store(V, R, id[2], t[2]);
					src.parent = -1;
					v3_store(V, R, r, src);
				}
				trinum += 3;
				V.act[r] = V3_NONE;
			} else {
				// slot r holds the third part (the last slot is given back), its
				// centroid corner shaded anew (shade.cpp:508-520)
				if (tid == 0) v3_store(V, R, r, t[2]);
				trinum += 2;
				const Patch &t3 = t[2];
				const float nlen = __builtin_sqrtf(dot3<F>(t3.n[0], t3.n[1], t3.n[2], t3.n[0], t3.n[1], t3.n[2]));
				float f3 = 0.0f;
				if (tid < 4)
					f3 = adapt_ff4<F>(S, s, t3.n[0], t3.n[1], t3.n[2], nlen, src.parent, t3.p[2][0], t3.p[2][1], t3.p[2][2],
					                  q, c);
				if (tid == 0) V.ff[4 * (size_t)r + 2] = f3;
			}
			__syncthreads();
		}
	}
	__syncthreads();
	if (tid < 34) L->rng[tid] = st[tid];
	if (tid == 0) {
		L->rng_k = sk;
		L->trinum = trinum;
		// the next pass: the slots this one made
		L->v3_range[par ^ 1][0] = r1;
		L->v3_range[par ^ 1][1] = trinum;
	}
	const int vals[8] = {c.rays, c.ref_rays, c.cells, c.tests, c.levels, c.precision, c.stalled, 0};
	add_block_counters<4>(vals, bcnt);
}

// CalRadiosity (shade.cpp:244-259) of receiver r with its form factors V.ff
__device__ __forceinline__ void v3_receive(const DRecv &R, const V3Dev &V, const SrcDesc &s, int r)
{
	const float4 rho = R.rho[r];
	const float kr = (float)(0.5 / 255.0);
	const float lo[3] = {rho.x * kr * s.dB[0] / 3, rho.y * kr * s.dB[1] / 3, rho.z * kr * s.dB[2] / 3};
	float4 db = R.dB[r];
	float dbv[3] = {db.x, db.y, db.z};
	for (int v = 0; v < 3; v++) {
		const float ff = V.ff[4 * (size_t)r + v];
		float4 a = R.acc[3 * r + v];
		float av[3] = {a.x, a.y, a.z}, dv[3];
		for (int c = 0; c < 3; c++) {
			const float d = lo[c] * ff;
			dbv[c] += d;
			av[c] += d;
			dv[c] = d;
		}
		R.acc[3 * r + v] = make_float4(av[0], av[1], av[2], 0.f);
		R.dacc[3 * r + v] = make_float4(dv[0], dv[1], dv[2], 0.f);
	}
	R.dB[r] = make_float4(dbv[0], dbv[1], dbv[2], 0.f);
}

// ... for the receivers of the pass that keep their form factors, one lane
// each
__global__ __launch_bounds__(kBlock) void k_v3_apply(DRecv R, V3Dev V, Loop *L, int par)
{
	if (L->done || L->v3_error) return;
	const int r0 = L->v3_range[par][0], r1 = L->v3_range[par][1];
	for (int r = r0 + blockIdx.x * kBlock + threadIdx.x; r < r1; r += gridDim.x * kBlock)
		if (V.act[r] == V3_APPLY) v3_receive(R, V, L->src[0], r);
}

// the best candidate over the slots in use (listCanditate with one source,
// rad.cpp:387-404: the highest slot on ties), once the passes are over: the
// next pass (parity par) has no slot; else v3_pending makes the next
// iteration's k_v3_prep leave the source as it is, for that pass
__global__ __launch_bounds__(kBlock) void k_v3_select(DRecv R, Loop *L, float converge, int par)
{
	__shared__ unsigned long long wmax[kWavesPerBlock];
	__shared__ int wn[kWavesPerBlock];
	if (L->done || L->v3_error) return;
	const bool pending = L->v3_range[par][0] < L->v3_range[par][1];
	if (blockIdx.x == 0 && threadIdx.x == 0) L->v3_pending = pending;
	if (pending) return;
	unsigned long long key = 0;
	int n = 0;
	for (int r = blockIdx.x * kBlock + threadIdx.x; r < L->trinum; r += gridDim.x * kBlock) {
		if (__float_as_int(R.nrm[r].w) < 0) continue;
		const float4 db = R.dB[r];
		const float sum = db.x + db.y + db.z;
		if (sum > converge) {
			const unsigned long long k = cand_key<false>(sum, r);
			key = k > key ? k : key;
			n++;
		}
	}
	key = wave_max(key);
	n = wave_sum(n);
	if ((threadIdx.x & 63) == 0) {
		wmax[threadIdx.x >> 6] = key;
		wn[threadIdx.x >> 6] = n;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < kWavesPerBlock; w++) {
			key = wmax[w] > key ? wmax[w] : key;
			n += wn[w];
		}
		if (key) atomicMax(&L->key, key);
		if (n) atomicAdd(&L->ncand, n);
	}
}

// ---- InitRadiosity's neighbour search (rad.cpp:295-334) ----------------------
// The reference compares every corner of triangle i with every corner of
// every j (O(n^2)); here corners are hashed by their exact coordinates into
// chained buckets and only corners in one bucket are compared.  For each
// ordered pair (i, v), (j, c) with i != j and p_i[v] == p_j[c]: when
// p_i[v+1] == p_j[c+1] ("same direction") or p_i[v+1] == p_j[c+2]
// ("opposite") and CosTheta(n_i, n_j) > 0.85, the reference writes
// i.neighbor[v] = j and j.neighbor[c or c+2] = i.  Each write carries its
// position in the reference's loop order, (i, j, 3v + c) packed into 64
// bits, and a slot edge keeps the greatest: the last write wins, as in the
// reference's sequential order.
struct NbPoint {
	float p[3][3];
	float n[3];
};

__device__ __forceinline__ unsigned nb_hash(float x, float y, float z, unsigned mask)
{
	// -0 folded onto +0: equal coordinates have equal bits
	unsigned long long h = __float_as_uint(x + 0.0f) * 0x9E3779B97F4A7C15ull;
	h ^= (h >> 29) + __float_as_uint(y + 0.0f) * 0xC2B2AE3D27D4EB4Full;
	h ^= (h >> 31) + __float_as_uint(z + 0.0f) * 0x165667B19E3779F9ull;
	h ^= h >> 32;
	return (unsigned)h & mask;
}

__device__ __forceinline__ bool nb_same(const float *a, const float *b)
{
	return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

__global__ __launch_bounds__(kBlock) void k_nb_insert(const NbPoint *__restrict__ T, int n, unsigned mask, int *heads,
                                                      int *next)
{
	const int r = blockIdx.x * kBlock + threadIdx.x;
	if (r >= 3 * n) return;
	const float *p = T[r / 3].p[r % 3];
	if (isnan(p[0]) || isnan(p[1]) || isnan(p[2])) {  // never equal to anything
		next[r] = -2;
		return;
	}
	next[r] = atomicExch(&heads[nb_hash(p[0], p[1], p[2], mask)], r);
}

template <bool F>
__global__ __launch_bounds__(kBlock) void k_nb_pairs(const NbPoint *__restrict__ T, int n, unsigned mask,
                                                     const int *__restrict__ heads, const int *__restrict__ next,
                                                     unsigned long long *best)
{
	const int u = blockIdx.x * kBlock + threadIdx.x;
	if (u >= 3 * n || next[u] == -2) return;
	const int i = u / 3, v = u % 3;
	const NbPoint &tp = T[i];
	const float *pv = tp.p[v], *pv1 = tp.p[(v + 1) % 3];
	for (int w = heads[nb_hash(pv[0], pv[1], pv[2], mask)]; w >= 0; w = next[w]) {
		const int j = w / 3, c = w % 3;
		if (j == i) continue;
		const NbPoint &tn = T[j];
		if (!nb_same(pv, tn.p[c])) continue;  // another point in the same bucket
		int other;
		if (nb_same(pv1, tn.p[(c + 1) % 3])) other = c;
		else if (nb_same(pv1, tn.p[(c + 2) % 3])) other = (c + 2) % 3;
		else continue;
		const float li = __builtin_sqrtf(dot3<F>(tp.n[0], tp.n[1], tp.n[2], tp.n[0], tp.n[1], tp.n[2]));
		const float lj = __builtin_sqrtf(dot3<F>(tn.n[0], tn.n[1], tn.n[2], tn.n[0], tn.n[1], tn.n[2]));
		const float cosine = dot3<F>(tp.n[0], tp.n[1], tp.n[2], tn.n[0], tn.n[1], tn.n[2]) / li / lj;
		if (!((double)cosine > 0.85)) continue;
		const unsigned long long order =
		    ((unsigned long long)i << 34) | ((unsigned long long)j << 4) | (unsigned long long)(3 * v + c);
		atomicMax(&best[3 * (size_t)i + v], order);      // i.neighbor[v] = j
		atomicMax(&best[3 * (size_t)j + other], order);  // j.neighbor[c or c+2] = i
	}
}

__global__ __launch_bounds__(kBlock) void k_nb_final(const unsigned long long *__restrict__ best, int n, int *nb)
{
	const int k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= 3 * n) return;
	const unsigned long long order = best[k];
	const int s = k / 3, i = (int)(order >> 34), j = (int)((order >> 4) & 0x3fffffffu);
	nb[k] = order == 0 ? -1 : (s == i ? j : i);
}

template <typename T> __global__ void k_fill_n(T *p, size_t n, T v)
{
	for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}

// ---- host side --------------------------------------------------------------
#define HIP_OK(expr, what)                                                        \
	do {                                                                          \
		hipError_t e_ = (expr);                                                   \
		if (e_ != hipSuccess) {                                                   \
			err = std::string(what) + ": " + hipGetErrorString(e_);               \
			return -1;                                                            \
		}                                                                         \
	} while (0)

// A device array of n elements; the allocation is kept and reused while
// later sizes fit in it.
template <typename T> struct DevBuf {
	T *p = nullptr;
	size_t n = 0, cap = 0;
	~DevBuf() { if (p) (void)hipFree(p); }
	int alloc(size_t count, std::string &err)
	{
		n = count;
		if (count == 0 || (p && count <= cap)) return 0;
		if (p) (void)hipFree(p);
		p = nullptr;
		cap = 0;
		HIP_OK(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
		cap = count;
		return 0;
	}
	int put(const T *src, size_t count, hipStream_t s, std::string &err)
	{
		if (alloc(count, err)) return -1;
		if (count) HIP_OK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
		return 0;
	}
};

float smallest_float_at_least(double x)
{
	float f = (float)x;
	while ((double)f < x) f = std::nextafter(f, INFINITY);
	while ((double)std::nextafter(f, -INFINITY) >= x) f = std::nextafter(f, -INFINITY);
	return f;
}
float largest_float_at_most(double x)
{
	float f = (float)x;
	while ((double)f > x) f = std::nextafter(f, -INFINITY);
	while ((double)std::nextafter(f, INFINITY) <= x) f = std::nextafter(f, INFINITY);
	return f;
}

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy at its own address); host code uses
// this address, never the array's name.
int scratch_address(void **out, std::string &err)
{
	static std::mutex mu;
	static void *addr[64] = {};
	int dev = -1;
	HIP_OK(hipGetDevice(&dev), "hipGetDevice");
	if (dev < 0 || dev >= 64) {
		err = "device ordinal out of range for the static scratch table";
		return -1;
	}
	std::lock_guard<std::mutex> lock(mu);
	if (!addr[dev]) {
		void *p = nullptr;
		HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
		addr[dev] = p;
	}
	*out = addr[dev];
	return 0;
}

// glibc rand()'s state at the start of a process (srandom(1): the TYPE_3
// generator of degree 31 and separation 3, its first 310 outputs
// discarded), as the 34 most recent words with the oldest first.
void glibc_rand_start(unsigned st[34])
{
	unsigned w[344];
	int x = 1;
	w[0] = 1;
	for (int i = 1; i < 31; i++) {
		const int hi = x / 127773, lo = x % 127773;
		x = 16807 * lo - 2836 * hi;
		if (x < 0) x += 2147483647;
		w[i] = (unsigned)x;
	}
	for (int i = 31; i < 34; i++) w[i] = w[i - 31];
	for (int i = 34; i < 344; i++) w[i] = w[i - 31] + w[i - 3];
	for (int i = 0; i < 34; i++) st[i] = w[310 + i];
}

}  // namespace

struct DeviceEngine::Impl {
	hipStream_t stream = nullptr;
	DevBuf<float4> lnc, lse, leaf_box, nrm, cA, rho, acc, dacc, dB;
	DevBuf<int> cell_leaf;
	DevBuf<float> vx, vy, vz;
	DevBuf<int> fine_leaf, seg_begin;  // segment-visibility index (explicit matrix)
	DevBuf<unsigned long long> vis;    // pair visibility bits of the explicit matrix
	DevBuf<float4> seg_nc, seg_se;
	bool have_seg = false;
	DevBuf<Loop> loop;
	DevBuf<long long> bcnt;
	DevBuf<unsigned long long> cand;
	DevBuf<unsigned char> live;
	DevBuf<float> ffbuf;
	DevBuf<int> raycnt, rays;  // per-k_gen-block ray counts, compacted ray list
	DevBuf<unsigned char> rkey;  // per receiver vertex in position order: k_compact's grouping key
	DevBuf<unsigned char> rcode; // the same order: which of the vertex's rays exist, and its key (k_gen)
	DevBuf<int> order;         // receiver processing order (DRecv::order)
	DevBuf<Patch> pool;        // the uploaded slot records (packed by k_pack_items / k_pack_recv)
	DevBuf<int> items;         // the octree's list entries (original slots)
	DevBuf<float> radius;      // per original: cull radius of its list records (cull_radius)
	DScene S;
	DRecv R;
	Loop *host_loop = nullptr;  // pinned mirror of the loop header
	int nblocks = 0, n_orig = 0, trace_nblocks = 0;
	int groups = 2;  // receiver groups per block (kOneGroupSlots)
	int v3_par = 0;  // parity of the next V3 pass
	bool fma = true, lowest = false;
	int parallel_src = 1, batch = 32;
	float converge = 1200.f;
	// V3 splitting
	bool v3 = false;
	V3Dev V = {};
	DevBuf<float> v3ff;
	DevBuf<unsigned char> v3act;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	void *scratch = nullptr;  // this device's address of the static scratch (rad_cast_static)
	// kGraphIters shooting iterations captured once per upload and replayed
	hipGraphExec_t graph_exec = nullptr;
	hipGraph_t graph = nullptr;
	void drop_graph()
	{
		if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
		if (graph) (void)hipGraphDestroy(graph);
		graph_exec = nullptr;
		graph = nullptr;
	}
	static constexpr int kEv = 16;
	hipEvent_t kev[2 * kEv] = {};
};

DeviceEngine::~DeviceEngine()
{
	if (!d_) return;
	if (d_->stream) (void)hipStreamSynchronize(d_->stream);
	d_->drop_graph();
	if (d_->host_loop) (void)hipHostFree(d_->host_loop);
	if (d_->ev0) (void)hipEventDestroy(d_->ev0);
	if (d_->ev1) (void)hipEventDestroy(d_->ev1);
	for (hipEvent_t &e : d_->kev)
		if (e) (void)hipEventDestroy(e);
	Impl *d = d_;
	d_ = nullptr;
	hipStream_t s = d->stream;
	delete d;  // frees the device buffers
	if (s) (void)hipStreamDestroy(s);
}

DeviceEngine *DeviceEngine::create(std::string &err)
{
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
		err = "no HIP device available: the engine runs on the GPU only";
		return nullptr;
	}
	DeviceEngine *e = new DeviceEngine();
	e->d_ = new Impl();
	if (hipStreamCreateWithFlags(&e->d_->stream, hipStreamNonBlocking) != hipSuccess ||
	    hipHostMalloc((void **)&e->d_->host_loop, sizeof(Loop), hipHostMallocDefault) != hipSuccess ||
	    hipEventCreate(&e->d_->ev0) != hipSuccess || hipEventCreate(&e->d_->ev1) != hipSuccess) {
		err = "HIP stream/event/pinned allocation failed";
		delete e;
		return nullptr;
	}
	for (hipEvent_t &ev : e->d_->kev)
		if (hipEventCreate(&ev) != hipSuccess) {
			err = "HIP event creation failed";
			delete e;
			return nullptr;
		}
	if (scratch_address(&e->d_->scratch, err)) {
		delete e;
		return nullptr;
	}
	return e;
}

int DeviceEngine::current_device()
{
	int d = -1;
	return hipGetDevice(&d) == hipSuccess ? d : -1;
}

DeviceEngine *DeviceEngine::create_on(int device, std::string &err)
{
	if (hipSetDevice(device) != hipSuccess) {
		err = "HIP device " + std::to_string(device) + " cannot be selected";
		return nullptr;
	}
	return create(err);
}

__global__ void k_warm(int *p)
{
	if (p && blockIdx.x == 0 && threadIdx.x == 0) p[0] = 0;
}

int DeviceEngine::warm(std::string &err)
{
	// the first launch loads the device code; the first copies each way set
	// up the runtime's staging for pageable host memory (milliseconds each,
	// otherwise paid by the first upload and the first download)
	int *p = nullptr;
	HIP_OK(hipMalloc(&p, 4096), "warm");
	std::vector<int> h(1024, 0);
	hipLaunchKernelGGL(k_warm, dim3(1), dim3(64), 0, d_->stream, p);
	hipError_t e = hipMemcpyAsync(p, h.data(), 4096, hipMemcpyHostToDevice, d_->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(h.data(), p, 4096, hipMemcpyDeviceToHost, d_->stream);
	// k_trace keeps a few values in scratch: its first dispatch in a process
	// has the runtime set up scratch memory (about 3 ms), paid here.  A loop
	// state marked done makes it return at once.
	Loop *lp = nullptr;
	std::vector<unsigned char> zero(sizeof(Loop), 0);
	reinterpret_cast<Loop *>(zero.data())->done = 1;
	if (e == hipSuccess) e = hipMalloc(&lp, sizeof(Loop));
	if (e == hipSuccess) e = hipMemcpyAsync(lp, zero.data(), sizeof(Loop), hipMemcpyHostToDevice, d_->stream);
	if (e == hipSuccess) {
		const DScene S{};
		const DRecv R{};
		hipLaunchKernelGGL(k_trace<true>, dim3(1), dim3(kTraceBlock), 0, d_->stream, S, R, lp, 0, nullptr, nullptr, nullptr);
		hipLaunchKernelGGL(k_trace<false>, dim3(1), dim3(kTraceBlock), 0, d_->stream, S, R, lp, 0, nullptr, nullptr, nullptr);
		e = hipGetLastError();
	}
	if (e == hipSuccess) e = hipStreamSynchronize(d_->stream);
	(void)hipFree(lp);
	(void)hipFree(p);
	HIP_OK(e, "warm");
	return 0;
}

int DeviceEngine::upload(const HostScene &s, const FlatOctree &t, std::string &err)
{
	d_->have_seg = false;
	d_->drop_graph();  // its launches name the buffers of the previous upload
	Impl &d = *d_;
	const int n0 = s.n_orig, n = s.trinum;
	d.n_orig = n0;
	d.fma = s.o.fma != 0;
	d.lowest = s.o.lowest != 0;
	d.parallel_src = s.o.parallel_src;
	d.batch = s.o.batch > 0 ? s.o.batch : 32;
	d.converge = s.o.converge;
	d.v3 = s.o.v3 != 0;
	if (d.v3 && (d.lowest || d.parallel_src != 1)) {
		err = "-variant v3 runs with one source per iteration (-parallel_src 1, -select high)";
		return -1;
	}
	if (!d.lowest && d.parallel_src > kMaxSrc) {
		err = "-parallel_src above " + std::to_string(kMaxSrc) + " is not supported";
		return -1;
	}
	if ((size_t)t.gt * t.gt * t.gt > (size_t)1 << 30) {
		err = "octree cell table too large";
		return -1;
	}
	// leaf lists with the originals inline, as the octree sees them: the
	// post-pre-partition contents of slots [0, n_orig)
	const size_t ni = t.items.size();
	for (size_t k = 0; k < ni; k++)
		if (t.items[k] < 0 || t.items[k] >= n0) {
			err = "octree item out of range";
			return -1;
		}
	for (int l = 0; l < t.n_leaves; l++) {
		int begin, cd;
		std::memcpy(&begin, &t.leaf_box[(size_t)l * 8 + 3], 4);
		std::memcpy(&cd, &t.leaf_box[(size_t)l * 8 + 7], 4);
		if (begin < 0 || (size_t)begin + (size_t)(cd & 0xffffff) > ni) {
			err = "octree leaf list out of range";
			return -1;
		}
	}
	for (int leaf : t.cell_leaf)
		if (leaf < 0 || leaf >= t.n_leaves) {
			err = "octree cell table entry out of range";
			return -1;
		}
	// slots the SoA arrays and the record pool hold: V3 splits add slots up
	// to the pool limit min(MaxTri, -triangle) (shade.cpp:16-30)
	const int cap = d.v3 ? std::max(n, std::min(s.o.max_tri, s.o.triangle)) : n;
	std::vector<int> order(n);
	spatial_order(s.pool.data(), n, s.G0, s.G1, order.data());
	// cull radii of the originals as the octree sees them, and the bound on
	// |c|_1 over those with a finite radius
	std::vector<float> rad((size_t)n0 + 1);
	const float cull_cl = cull_radii(s, rad.data());
	hipStream_t st = d.stream;
	if (d.order.put(order.data(), order.size(), st, err)) return -1;
	// the slot records as they are, then gathered on the device
	if (d.pool.alloc((size_t)cap, err) ||
	    (n && hipMemcpyAsync(d.pool.p, s.pool.data(), (size_t)n * sizeof(Patch), hipMemcpyHostToDevice, st) != hipSuccess) ||
	    d.items.put(t.items.data(), ni, st, err) ||
	    d.radius.put(rad.data(), rad.size(), st, err) ||
	    d.lnc.alloc(ni * 2, err) || d.lse.alloc(ni * 3, err) ||
	    d.cell_leaf.put(t.cell_leaf.data(), t.cell_leaf.size(), st, err) ||
	    d.leaf_box.put((const float4 *)t.leaf_box.data(), t.leaf_box.size() / 4, st, err) ||
	    d.vx.alloc((size_t)cap * 3, err) || d.vy.alloc((size_t)cap * 3, err) || d.vz.alloc((size_t)cap * 3, err) ||
	    d.nrm.alloc(cap, err) || d.cA.alloc(cap, err) || d.rho.alloc(cap, err) || d.acc.alloc((size_t)cap * 3, err) ||
	    d.dacc.alloc((size_t)cap * 3, err) || d.dB.alloc(cap, err)) {
		if (err.empty()) err = "upload of the slot records failed";
		return -1;
	}
	if (ni > 0)
		hipLaunchKernelGGL(k_pack_items, dim3((unsigned)((ni + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, d.pool.p,
		                   d.items.p, d.radius.p, (int)ni, d.lnc.p, d.lse.p);
	if (n > 0)
		hipLaunchKernelGGL(k_pack_recv, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, d.pool.p, n, d.vx.p,
		                   d.vy.p, d.vz.p, d.nrm.p, d.cA.p, d.rho.p, d.acc.p, d.dacc.p, d.dB.p);
	HIP_OK(hipGetLastError(), "pack launch");
	// grid sizes first: every kernel that adds into bcnt indexes it by its
	// hardware block index, so bcnt must cover the largest grid
	d.groups = n <= kOneGroupSlots ? 1 : 2;
	const int rpb = d.groups == 1 ? kRecvPerBlock<1> : kRecvPerBlock<2>;
	d.nblocks = (n + rpb - 1) / rpb;
	// k_trace: two lanes for each of up to 6 rays per slot, or four lanes on
	// up to kTraceSlots lanes
	{
		// k_trace's grid: every split of up to 6 n rays (two per receiver
		// vertex), or the machine's slots for the few-ray mapping
		const TraceSplit ts = trace_split((long long)n * 6);
		d.trace_nblocks = (int)std::max(8LL * (ts.head + ts.tail + ts.end), (long long)(kTraceSlots / kTraceBlock));
	}
	const int grid = ((d.nblocks + 7) / 8) * 8;  // a whole number of blocks per XCD
	const int tgrid = ((d.trace_nblocks + 7) / 8) * 8;
	const int v3grid = d.v3 ? (int)(((size_t)cap * 4 + kBlock - 1) / kBlock) : 0;
	std::vector<long long> zeros((size_t)std::max(std::max(grid, tgrid), v3grid) * 8, 0);
	if (d.bcnt.put(zeros.data(), zeros.size(), st, err)) return -1;
	if (!d.lowest && d.parallel_src > 1 && d.cand.alloc((size_t)n + 1, err)) return -1;
	if (d.live.alloc((size_t)n + 1, err) || d.ffbuf.alloc((size_t)n * 6 + 2, err) ||
	    d.raycnt.alloc((size_t)d.nblocks + 1, err) || d.rays.alloc((size_t)n * 6 + 1, err))
		return -1;
	std::memset(d.host_loop, 0, sizeof(Loop));
	d.host_loop->trinum = n;
	glibc_rand_start(d.host_loop->rng);  // rand() is first called in the solve (shade.cpp:53)
	if (d.v3) {
		if (d.v3ff.alloc((size_t)cap * 4 + 4, err) || d.v3act.alloc((size_t)cap + 1, err)) return -1;
		d.V = V3Dev{d.pool.p, d.v3ff.p, d.v3act.p, std::min(s.o.max_tri, s.o.triangle), n0, s.o.adapt_area,
		            s.o.sample_area, s.o.delta_ff};
	}
	// -parallel_src 0: listCanditate returns no source (rad.cpp:400), so the
	// first iteration finds nothing to shoot
	if (!d.lowest && d.parallel_src <= 0) d.host_loop->done = 1;
	if (d.loop.put(d.host_loop, 1, st, err)) return -1;
	const float thr_t = smallest_float_at_least(0.0001);
	d.S = DScene{d.lnc.p, d.lse.p, d.cell_leaf.p, d.leaf_box.p, s.G0[0], s.G0[1], s.G0[2], s.G1[0], s.G1[1], s.G1[2],
	             (double)t.min_grid, 1.0 / (double)t.min_grid, t.gt, thr_t, thr_t * kGuardDown,
	             largest_float_at_most(1e-6), cull_cl};
	d.R = DRecv{d.vx.p, d.vy.p, d.vz.p, d.nrm.p, d.cA.p, d.rho.p, d.acc.p, d.dacc.p, d.dB.p, n, d.order.p};
	if (d.rkey.alloc((size_t)n * 3 + 1, err) || d.rcode.alloc((size_t)n * 3 + 1, err)) return -1;
	if (n > 0)
		hipLaunchKernelGGL(k_ray_keys, dim3((unsigned)(((size_t)n * 3 + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, d.S,
		                   d.R, n * 3, d.rkey.p);
	if (d.lowest || d.parallel_src <= 1) {
		const int sb = (n + kBlock - 1) / kBlock;
		if (sb > 0) {
			if (d.lowest) hipLaunchKernelGGL(k_select<true>, dim3(sb), dim3(kBlock), 0, st, d.R, d.loop.p, d.converge);
			else hipLaunchKernelGGL(k_select<false>, dim3(sb), dim3(kBlock), 0, st, d.R, d.loop.p, d.converge);
		}
	}
	HIP_OK(hipGetLastError(), "select launch");
	HIP_OK(hipStreamSynchronize(st), "upload");
	shoot_ms_ = kernel_ms_ = 0;
	kernel_launches_ = kernel_timed_ = 0;
	return 0;
}

int DeviceEngine::shoot(int max_shots, int *shots, int *converged, std::string &err)
{
	Impl &d = *d_;
	hipStream_t st = d.stream;
	*shots = 0;
	*converged = 0;
	if (d.R.nslots == 0) {  // no triangles: no candidate, as listCanditate finds none
		*converged = 1;
		return 0;
	}
	HIP_OK(hipMemcpyAsync(d.host_loop, d.loop.p, 64, hipMemcpyDeviceToHost, st), "loop read");
	HIP_OK(hipStreamSynchronize(st), "loop read");
	if (d.host_loop->done) {
		*converged = 1;
		return 0;
	}
	const int start = d.host_loop->shots;
	if (d.v3) {
		// An iteration is the source, then passes over the receivers until a
		// pass makes no new slot (rad.cpp:414-444), then the next source.
		// The device runs it from queued slots of five launches: the source
		// (kept while a split is still being shaded), one pass, and the
		// selection once no pass is left (v3_pending).  Each pass's slots
		// are in Loop by parity; the grids stride over them.  Batches of
		// kV3Batch slots between host checks; a call ends between iterations.
		float ms_total = 0.f;
		auto slot = [&]() {
			const int p = d.v3_par;
			const size_t nt = (size_t)std::max(d.host_loop->trinum, 1);  // as last read back
			const dim3 g((unsigned)std::min<size_t>((nt * 16 + kBlock - 1) / kBlock, 4096));
			const dim3 ga((unsigned)std::min<size_t>((nt + kBlock - 1) / kBlock, 1024)), blk(kBlock);
			if (d.fma) {
				hipLaunchKernelGGL(k_v3_prep<true>, dim3(1), dim3(64), 0, st, d.R, d.loop.p, p);
				hipLaunchKernelGGL(k_v3_gen<true>, g, blk, 0, st, d.S, d.R, d.V, d.loop.p, p, d.bcnt.p);
				hipLaunchKernelGGL(k_v3_commit<true>, dim3(1), dim3(4 * kWave), 0, st, d.S, d.R, d.V, d.loop.p, p, d.bcnt.p);
			} else {
				hipLaunchKernelGGL(k_v3_prep<false>, dim3(1), dim3(64), 0, st, d.R, d.loop.p, p);
				hipLaunchKernelGGL(k_v3_gen<false>, g, blk, 0, st, d.S, d.R, d.V, d.loop.p, p, d.bcnt.p);
				hipLaunchKernelGGL(k_v3_commit<false>, dim3(1), dim3(4 * kWave), 0, st, d.S, d.R, d.V, d.loop.p, p, d.bcnt.p);
			}
			hipLaunchKernelGGL(k_v3_apply, ga, blk, 0, st, d.R, d.V, d.loop.p, p);
			hipLaunchKernelGGL(k_v3_select, ga, blk, 0, st, d.R, d.loop.p, d.converge, p ^ 1);
			d.v3_par = p ^ 1;
		};
		auto read_back = [&]() -> int {
			HIP_OK(hipGetLastError(), "v3 launch");
			HIP_OK(hipMemcpyAsync(d.host_loop, d.loop.p, 64, hipMemcpyDeviceToHost, st), "loop read");
			HIP_OK(hipStreamSynchronize(st), "v3 batch");
			if (d.host_loop->v3_error) {
				err = "-variant v3: the split of slot " + std::to_string(d.host_loop->v3_error - 1) +
				      " would rewrite an original the octree lists (pre-partition overwrote it); not supported";
				return -1;
			}
			return 0;
		};
		for (;;) {
			const int left = max_shots - (d.host_loop->shots - start);
			if (left <= 0 || d.host_loop->done) break;
			const int b = std::min(kV3Batch, left);  // slots that pick a source: at most b shots
			HIP_OK(hipEventRecord(d.ev0, st), "event");
			for (int i = 0; i < b; i++) slot();
			HIP_OK(hipEventRecord(d.ev1, st), "event");
			if (read_back()) return -1;
			float ms = 0;
			(void)hipEventElapsedTime(&ms, d.ev0, d.ev1);
			ms_total += ms;
		}
		while (!d.host_loop->done && d.host_loop->v3_pending) {  // finish the iteration under way
			slot();
			if (read_back()) return -1;
		}
		kernel_launches_ += d.host_loop->shots - start;
		HIP_OK(hipMemcpyAsync(d.host_loop, d.loop.p, 64, hipMemcpyDeviceToHost, st), "loop read");
		HIP_OK(hipStreamSynchronize(st), "v3 shoot");
		shoot_ms_ += ms_total;
		*shots = d.host_loop->shots - start;
		*converged = d.host_loop->done;
		return 0;
	}
	const int grid = ((d.nblocks + 7) / 8) * 8;
	const int tgrid = ((d.trace_nblocks + 7) / 8) * 8;
	if (d.bcnt.n < (size_t)std::max(grid, tgrid) * 8 ||
	    d.ffbuf.n < (size_t)d.R.nslots * 6 || d.live.n < (size_t)d.R.nslots ||
	    d.raycnt.n < (size_t)d.nblocks || d.rays.n < (size_t)d.R.nslots * 6) {
		err = "device buffers smaller than the launch grids";
		return -1;
	}
	const bool multi = !d.lowest && d.parallel_src > 1;
	const int cb = (d.R.nslots + kBlock - 1) / kBlock;
	// one shooting iteration's launches; with events, the ray-tracing kernel
	// of the first source is bracketed by event pair ev
	// one source per iteration: k_gen selects it (k_prep only without receivers)
	const int prep = multi || grid <= 0 ? 0 : d.lowest ? 2 : 1;
	// every ray position fits the machine at eight lanes per ray: no dense list
	const bool direct = 48LL * d.R.nslots <= kTraceSlots;
	const int dgrid = (int)((6LL * d.R.nslots + kTraceBlock / 8 - 1) / (kTraceBlock / 8));
	auto iteration = [&](int ev) {
		if (multi) {
			hipLaunchKernelGGL(k_collect, dim3(cb > 0 ? cb : 1), dim3(kBlock), 0, st, d.R, d.loop.p, d.cand.p, d.converge);
			if (d.fma) hipLaunchKernelGGL(k_topk<true>, dim3(1), dim3(1024), 0, st, d.R, d.loop.p, d.cand.p, d.parallel_src);
			else hipLaunchKernelGGL(k_topk<false>, dim3(1), dim3(1024), 0, st, d.R, d.loop.p, d.cand.p, d.parallel_src);
		} else if (!prep) {  // no receivers: k_prep still ends the loop when no candidate is left
			if (d.fma) {
				if (d.lowest) hipLaunchKernelGGL((k_prep<true, true>), dim3(1), dim3(64), 0, st, d.R, d.loop.p);
				else hipLaunchKernelGGL((k_prep<true, false>), dim3(1), dim3(64), 0, st, d.R, d.loop.p);
			} else {
				if (d.lowest) hipLaunchKernelGGL((k_prep<false, true>), dim3(1), dim3(64), 0, st, d.R, d.loop.p);
				else hipLaunchKernelGGL((k_prep<false, false>), dim3(1), dim3(64), 0, st, d.R, d.loop.p);
			}
		}
		const int nk = multi ? d.parallel_src : 1;
		for (int k = 0; grid > 0 && k < nk; k++) {
			const dim3 g(grid), blk(kBlock), tg(tgrid);
			// k_gen and k_compact for G receiver groups per block
			auto gen = [&](auto gc) {
				constexpr int G = decltype(gc)::value;
#define GEN(F, P) hipLaunchKernelGGL((k_gen<F, G, P>), g, blk, 0, st, d.R, d.loop.p, k, d.live.p, d.ffbuf.p, d.rkey.p, d.rcode.p, d.bcnt.p, d.nblocks, d.raycnt.p)
				if (d.fma) { if (prep == 2) GEN(true, 2); else if (prep) GEN(true, 1); else GEN(true, 0); }
				else { if (prep == 2) GEN(false, 2); else if (prep) GEN(false, 1); else GEN(false, 0); }
#undef GEN
				if (!direct)
					hipLaunchKernelGGL(k_compact<G>, dim3(d.nblocks), blk, 0, st, d.loop.p, k, d.rcode.p, d.order.p,
					                   6 * d.R.nslots, d.raycnt.p, d.nblocks, d.rays.p, d.dB.p, prep);
			};
			if (d.groups == 1) gen(std::integral_constant<int, 1>{});
			else gen(std::integral_constant<int, 2>{});
			if (ev >= 0 && k == 0) (void)hipEventRecord(d.kev[2 * ev], st);
			if (direct) {
				const dim3 dg((unsigned)dgrid);
				if (d.fma) hipLaunchKernelGGL(k_trace_direct<true>, dg, dim3(kTraceBlock), 0, st, d.S, d.R, d.loop.p, k, d.ffbuf.p, d.bcnt.p, d.rcode.p, 6 * d.R.nslots, d.dB.p, prep);
				else hipLaunchKernelGGL(k_trace_direct<false>, dg, dim3(kTraceBlock), 0, st, d.S, d.R, d.loop.p, k, d.ffbuf.p, d.bcnt.p, d.rcode.p, 6 * d.R.nslots, d.dB.p, prep);
			} else if (d.fma) {
				hipLaunchKernelGGL(k_trace<true>, tg, dim3(kTraceBlock), 0, st, d.S, d.R, d.loop.p, k, d.ffbuf.p, d.bcnt.p, d.rays.p);
			} else {
				hipLaunchKernelGGL(k_trace<false>, tg, dim3(kTraceBlock), 0, st, d.S, d.R, d.loop.p, k, d.ffbuf.p, d.bcnt.p, d.rays.p);
			}
			if (ev >= 0 && k == 0) (void)hipEventRecord(d.kev[2 * ev + 1], st);
			auto update = [&](auto gc) {
				constexpr int G = decltype(gc)::value;
#define UPD(F, FS, LO) hipLaunchKernelGGL((k_update<F, FS, LO, G>), g, blk, 0, st, d.R, d.loop.p, k, d.live.p, d.ffbuf.p, d.converge, d.bcnt.p, d.nblocks)
				if (multi) { if (d.fma) UPD(true, false, false); else UPD(false, false, false); }
				else if (d.lowest) { if (d.fma) UPD(true, true, true); else UPD(false, true, true); }
				else { if (d.fma) UPD(true, true, false); else UPD(false, true, false); }
#undef UPD
			};
			if (d.groups == 1) update(std::integral_constant<int, 1>{});
			else update(std::integral_constant<int, 2>{});
		}
	};
	// With -graph 1 and without kernel timing, runs of kGraphIters
	// iterations replay one captured graph (the kernels exit at once after
	// convergence, and none of the launches is a memory set).
	const bool replay = graphs && !timing;
	if (replay && !d.graph_exec) {
		HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
		for (int i = 0; i < kGraphIters; i++) iteration(-1);
		hipGraph_t gph = nullptr;
		const hipError_t ce = hipStreamEndCapture(st, &gph);
		if (ce != hipSuccess || !gph) {
			err = std::string("graph capture: ") + hipGetErrorString(ce);
			return -1;
		}
		d.graph = gph;
		HIP_OK(hipGraphInstantiate(&d.graph_exec, d.graph, nullptr, nullptr, 0), "graph instantiate");
	}
	int queued = 0;
	while (queued < max_shots) {
		// A candidate's unshot energy only grows until it is shot, so the
		// candidate count bounds the shots left from below: queue no more
		// iterations than that, and none run idle after convergence.
		// Below kMinQueue the bound is not used: a host round trip costs more
		// than the few launches that may find nothing left to do.
		const int per_iter = multi ? d.parallel_src : 1;
		// No candidate left: the next iteration is the one that finds none.
		const int left = d.host_loop->ncand == 0 && queued > 0
		                     ? 1
		                     : std::max(kMinQueue, (d.host_loop->ncand + per_iter - 1) / per_iter);
		const int b = std::min(std::min(d.batch, left), max_shots - queued);
		HIP_OK(hipEventRecord(d.ev0, st), "event");
		if (replay) {
			for (int i = 0; i + kGraphIters <= b; i += kGraphIters) HIP_OK(hipGraphLaunch(d.graph_exec, st), "graph launch");
			for (int i = 0; i < b % kGraphIters; i++) iteration(-1);
		} else if (!timing) {
			for (int i = 0; i < b; i++) iteration(-1);
		}
		for (int i = 0; timing && i < b; i++) {
			iteration(i % Impl::kEv);
			if (i % Impl::kEv == Impl::kEv - 1 || i == b - 1) {
				// collect this group of kernel timings before the events are reused
				HIP_OK(hipEventSynchronize(d.kev[2 * (i % Impl::kEv) + 1]), "event");
				for (int j = 0; j <= i % Impl::kEv; j++) {
					float ms = 0;
					if (hipEventElapsedTime(&ms, d.kev[2 * j], d.kev[2 * j + 1]) == hipSuccess) {
						kernel_ms_ += ms;
						kernel_timed_++;
					}
				}
			}
		}
		HIP_OK(hipGetLastError(), "shoot launch");
		HIP_OK(hipEventRecord(d.ev1, st), "event");
		HIP_OK(hipMemcpyAsync(d.host_loop, d.loop.p, 64, hipMemcpyDeviceToHost, st), "loop read");
		HIP_OK(hipStreamSynchronize(st), "shoot");
		float ms = 0;
		(void)hipEventElapsedTime(&ms, d.ev0, d.ev1);
		shoot_ms_ += ms;
		queued += b;
		kernel_launches_ += b;
		if (d.host_loop->done) break;
	}
	*shots = d.host_loop->shots - start;
	*converged = d.host_loop->done;
	return 0;
}

int DeviceEngine::download(HostScene &s, std::string &err)
{
	Impl &d = *d_;
	if (d.v3) {
		// the slots V3 splits made or rewrote: the records, then the counts
		HIP_OK(hipMemcpyAsync(d.host_loop, d.loop.p, 64, hipMemcpyDeviceToHost, d.stream), "loop read");
		HIP_OK(hipStreamSynchronize(d.stream), "download");
		const int tn = d.host_loop->trinum;
		if (tn < 0 || (size_t)tn > s.pool.size()) {
			err = "V3 slot count beyond the host pool";
			return -1;
		}
		if (tn) HIP_OK(hipMemcpy(s.pool.data(), d.pool.p, (size_t)tn * sizeof(Patch), hipMemcpyDeviceToHost), "download");
		s.trinum = tn;
		d.R.nslots = tn;
	}
	const int n = d.R.nslots;
	std::vector<float4> acc((size_t)n * 3), dacc((size_t)n * 3), dB(n);
	hipStream_t st = d.stream;
	if (n) {
		HIP_OK(hipMemcpyAsync(acc.data(), d.acc.p, acc.size() * sizeof(float4), hipMemcpyDeviceToHost, st), "download");
		HIP_OK(hipMemcpyAsync(dacc.data(), d.dacc.p, dacc.size() * sizeof(float4), hipMemcpyDeviceToHost, st), "download");
		HIP_OK(hipMemcpyAsync(dB.data(), d.dB.p, dB.size() * sizeof(float4), hipMemcpyDeviceToHost, st), "download");
	}
	HIP_OK(hipStreamSynchronize(st), "download");
	for (int i = 0; i < n; i++) {
		Patch &p = s.pool[i];
		for (int v = 0; v < 3; v++) {
			const float4 a = acc[3 * i + v], b = dacc[3 * i + v];
			p.accB[v][0] = a.x; p.accB[v][1] = a.y; p.accB[v][2] = a.z;
			p.deltaaccB[v][0] = b.x; p.deltaaccB[v][1] = b.y; p.deltaaccB[v][2] = b.z;
		}
		p.deltaB[0] = dB[i].x;
		p.deltaB[1] = dB[i].y;
		p.deltaB[2] = dB[i].z;
	}
	return 0;
}

int DeviceEngine::cast(int n, const RayIn *rays, int *hit, int *cells, int *tests, std::string &err)
{
	Impl &d = *d_;
	if (n <= 0) return 0;
	DevBuf<RayIn> dr;
	DevBuf<int> dh, dc, dt;
	hipStream_t st = d.stream;
	if (dr.put(rays, n, st, err) || dh.alloc(n, err) || dc.alloc(n, err) || dt.alloc(n, err)) return -1;
	HIP_OK(hipEventRecord(d.ev0, st), "event");
	if (d.fma) launch_cast<true>(st, d.S, dr.p, n, dh.p, dc.p, dt.p);
	else launch_cast<false>(st, d.S, dr.p, n, dh.p, dc.p, dt.p);
	HIP_OK(hipGetLastError(), "cast launch");
	HIP_OK(hipEventRecord(d.ev1, st), "event");
	HIP_OK(hipMemcpyAsync(hit, dh.p, n * sizeof(int), hipMemcpyDeviceToHost, st), "cast");
	if (cells) HIP_OK(hipMemcpyAsync(cells, dc.p, n * sizeof(int), hipMemcpyDeviceToHost, st), "cast");
	if (tests) HIP_OK(hipMemcpyAsync(tests, dt.p, n * sizeof(int), hipMemcpyDeviceToHost, st), "cast");
	HIP_OK(hipStreamSynchronize(st), "cast");
	float ms = 0;
	if (hipEventElapsedTime(&ms, d.ev0, d.ev1) == hipSuccess) cast_ms_ = ms;
	return 0;
}

int DeviceEngine::build_octree(const HostScene &s, FlatOctree &t, std::string &err)
{
	return build_octree_gpu(d_->stream, s.pool.data(), s.n_orig, s.G0, s.G1, s.o.fma != 0, t, err);
}

int DeviceEngine::link_neighbors(HostScene &s, std::string &err)
{
	Impl &d = *d_;
	const int n = s.trinum;
	if (n <= 0) return 0;
	if ((long long)n >= (1ll << 30)) {
		err = "neighbour search: too many triangles for 30-bit slot indices";
		return -1;
	}
	static_assert(sizeof(NbPoint) == 12 * sizeof(float), "NbPoint is 12 floats");
	std::vector<NbPoint> pts(n);
	pack_points(s, reinterpret_cast<float *>(pts.data()));
	const size_t refs = (size_t)3 * n;
	size_t cap = 1;
	while (cap < 2 * refs) cap <<= 1;
	const unsigned mask = (unsigned)(cap - 1);
	DevBuf<NbPoint> dT;
	DevBuf<int> heads, next, nb;
	DevBuf<unsigned long long> best;
	hipStream_t st = d.stream;
	if (dT.put(pts.data(), n, st, err) || heads.alloc(cap, err) || next.alloc(refs, err) || nb.alloc(refs, err) ||
	    best.alloc(refs, err))
		return -1;
	hipLaunchKernelGGL(k_fill_n<int>, dim3(1024), dim3(kBlock), 0, st, heads.p, cap, -1);
	hipLaunchKernelGGL(k_fill_n<unsigned long long>, dim3(1024), dim3(kBlock), 0, st, best.p, refs, 0ull);
	const unsigned g = (unsigned)((refs + kBlock - 1) / kBlock);
	hipLaunchKernelGGL(k_nb_insert, dim3(g), dim3(kBlock), 0, st, dT.p, n, mask, heads.p, next.p);
	if (s.o.fma) hipLaunchKernelGGL(k_nb_pairs<true>, dim3(g), dim3(kBlock), 0, st, dT.p, n, mask, heads.p, next.p, best.p);
	else hipLaunchKernelGGL(k_nb_pairs<false>, dim3(g), dim3(kBlock), 0, st, dT.p, n, mask, heads.p, next.p, best.p);
	hipLaunchKernelGGL(k_nb_final, dim3(g), dim3(kBlock), 0, st, best.p, n, nb.p);
	HIP_OK(hipGetLastError(), "neighbour search launch");
	std::vector<int> h(refs);
	HIP_OK(hipMemcpyAsync(h.data(), nb.p, refs * sizeof(int), hipMemcpyDeviceToHost, st), "neighbour download");
	HIP_OK(hipStreamSynchronize(st), "neighbour search");
	unpack_neighbors(s, h.data());
	return 0;
}

int DeviceEngine::cast_static(int n, const RayIn *rays, int *hit, int *cells, int *tests, std::string &err)
{
	Impl &d = *d_;
	if (n > kStaticRays) {
		err = "rad_cast_static: " + std::to_string(n) + " rays exceed its limit of " + std::to_string(kStaticRays) +
		      " per call";
		return -1;
	}
	if (n <= 0) return 0;
	if (!d.scratch) {
		err = "rad_cast_static: no static scratch on this device";
		return -1;
	}
	// staging in this device's scratch: rays, then hit, cells and tests;
	// k_cast writes every result, so nothing needs clearing
	unsigned char *base = static_cast<unsigned char *>(d.scratch);
	RayIn *dr = reinterpret_cast<RayIn *>(base);
	int *dh = reinterpret_cast<int *>(base + kStaticRayBytes);
	int *dc = dh + kStaticRays, *dt = dc + kStaticRays;
	hipStream_t st = d.stream;
	HIP_OK(hipMemcpyAsync(dr, rays, (size_t)n * sizeof(RayIn), hipMemcpyHostToDevice, st), "upload");
	HIP_OK(hipEventRecord(d.ev0, st), "event");
	if (d.fma) launch_cast<true>(st, d.S, dr, n, dh, dc, dt);
	else launch_cast<false>(st, d.S, dr, n, dh, dc, dt);
	HIP_OK(hipGetLastError(), "cast launch");
	HIP_OK(hipEventRecord(d.ev1, st), "event");
	HIP_OK(hipMemcpyAsync(hit, dh, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, st), "cast");
	if (cells) HIP_OK(hipMemcpyAsync(cells, dc, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, st), "cast");
	if (tests) HIP_OK(hipMemcpyAsync(tests, dt, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, st), "cast");
	HIP_OK(hipStreamSynchronize(st), "cast");
	float ms = 0;
	if (hipEventElapsedTime(&ms, d.ev0, d.ev1) == hipSuccess) cast_ms_ = ms;
	return 0;
}

int DeviceEngine::ff_pairs(int n, const int *pairs, float *ff3, std::string &err)
{
	Impl &d = *d_;
	if (n <= 0) return 0;
	for (int i = 0; i < 2 * n; i++)
		if (pairs[i] < 0 || pairs[i] >= d.R.nslots) {
			err = "pair index out of range";
			return -1;
		}
	DevBuf<int> dp;
	DevBuf<float> df;
	hipStream_t st = d.stream;
	if (dp.put(pairs, (size_t)2 * n, st, err) || df.alloc((size_t)3 * n, err)) return -1;
	const int blocks = (n + kBlock - 1) / kBlock;
	if (d.fma) hipLaunchKernelGGL(k_ff_pairs<true>, dim3(blocks), dim3(kBlock), 0, st, d.S, d.R, dp.p, n, df.p);
	else hipLaunchKernelGGL(k_ff_pairs<false>, dim3(blocks), dim3(kBlock), 0, st, d.S, d.R, dp.p, n, df.p);
	HIP_OK(hipGetLastError(), "ff launch");
	HIP_OK(hipMemcpyAsync(ff3, df.p, (size_t)3 * n * sizeof(float), hipMemcpyDeviceToHost, st), "ff");
	HIP_OK(hipStreamSynchronize(st), "ff");
	return 0;
}


int DeviceEngine::form_factors(float *dF, int n, int ld, int vis, FlatOctree &t, const HostScene &s, std::string &err)
{
	Impl &d = *d_;
	if (n <= 0 || n > d.R.nslots || ld < n) {
		err = "form_factors: bad size";
		return -1;
	}
	if (vis != VIS_SEGMENT && vis != VIS_REFERENCE) {
		err = "form_factors: unknown visibility mode";
		return -1;
	}
	hipStream_t st = d.stream;
	if (vis == VIS_SEGMENT && !d.have_seg) {
		// the segment test reads the originals' planes; pre-partition
		// rewrites originals in place when it splits, so it needs none
		if (s.trinum != s.n_orig) {
			err = "segment visibility needs a scene that pre-partition did not split (" + std::to_string(s.trinum) +
			      " slots, " + std::to_string(s.n_orig) + " originals)";
			return -1;
		}
		if (build_segment_index(t, s.pool.data(), s.n_orig, s.G0, s.G1, err)) return -1;
		const size_t ne = t.seg_items.size();
		if (ne >= ((size_t)1 << 29) || (size_t)t.fine_r * t.fine_r * t.fine_r >= ((size_t)1 << 31)) {
			err = "segment index too large for 32-bit record and cell indices";  // the walk indexes in int
			return -1;
		}
		std::vector<float4> nc(2 * ne), se(3 * ne);
		for (size_t e = 0; e < ne; e++) {
			const Patch &p = s.pool[t.seg_items[e]];
			float idf;
			std::memcpy(&idf, &t.seg_items[e], 4);
			nc[2 * e] = make_float4(p.n[0], p.n[1], p.n[2], p.c[0]);
			nc[2 * e + 1] = make_float4(p.c[1], p.c[2], idf, 0.f);
			for (int q = 0; q < 3; q++) se[3 * e + q] = make_float4(p.se[q][0], p.se[q][1], p.se[q][2], p.se[q][3]);
		}
		if (d.fine_leaf.put(t.fine_leaf.data(), t.fine_leaf.size(), st, err) ||
		    d.seg_begin.put(t.seg_begin.data(), t.seg_begin.size(), st, err) ||
		    d.seg_nc.put(nc.data(), nc.size(), st, err) || d.seg_se.put(se.data(), se.size(), st, err))
			return -1;
		HIP_OK(hipStreamSynchronize(st), "segment index upload");
		d.have_seg = true;
	}
	const size_t nblk_ff = std::max<size_t>((size_t)n, (size_t)((n + 63) / 64) * ((n + 63) / 64 + 1) / 2);
	if (d.bcnt.n < nblk_ff * 8) {  // the largest grid of the two kernels indexes bcnt
		std::vector<long long> zeros(nblk_ff * 8, 0);
		if (d.bcnt.put(zeros.data(), zeros.size(), st, err)) return -1;
		HIP_OK(hipStreamSynchronize(st), "counter upload");
	}
	HIP_OK(hipEventRecord(d.ev0, st), "event");
	if (vis == VIS_SEGMENT) {
		DFine G;
		G.leaf = d.fine_leaf.p;
		G.begin = d.seg_begin.p;
		G.nc = d.seg_nc.p;
		G.se = d.seg_se.p;
		G.r = t.fine_r;
		for (int k = 0; k < 3; k++) {
			G.g0[k] = s.G0[k];
			G.h[k] = (s.G1[k] - s.G0[k]) / (float)t.fine_r;
			G.inv[k] = (float)t.fine_r / (s.G1[k] - s.G0[k]);
		}
		const int nt = (n + 63) / 64;
		const long long tiles = (long long)nt * (nt + 1) / 2;
		if (d.vis.alloc((size_t)n * nt, err)) return -1;
		if (d.fma) hipLaunchKernelGGL(k_vis_tiles<true>, dim3((unsigned)tiles), dim3(kBlock), 0, st, G, d.R, n, nt, d.vis.p, d.bcnt.p);
		else hipLaunchKernelGGL(k_vis_tiles<false>, dim3((unsigned)tiles), dim3(kBlock), 0, st, G, d.R, n, nt, d.vis.p, d.bcnt.p);
		if (d.fma) hipLaunchKernelGGL(k_ff_seg<true>, dim3(n), dim3(kBlock), 0, st, d.R, n, ld, nt, d.vis.p, dF);
		else hipLaunchKernelGGL(k_ff_seg<false>, dim3(n), dim3(kBlock), 0, st, d.R, n, ld, nt, d.vis.p, dF);
	} else {
		if (d.fma) hipLaunchKernelGGL(k_ff_rows<true>, dim3(n), dim3(kBlock), 0, st, d.S, d.R, n, ld, dF, d.bcnt.p);
		else hipLaunchKernelGGL(k_ff_rows<false>, dim3(n), dim3(kBlock), 0, st, d.S, d.R, n, ld, dF, d.bcnt.p);
	}
	HIP_OK(hipGetLastError(), "form factor launch");
	HIP_OK(hipEventRecord(d.ev1, st), "event");
	HIP_OK(hipStreamSynchronize(st), "form factors");
	float ms = 0;
	if (hipEventElapsedTime(&ms, d.ev0, d.ev1) == hipSuccess) cast_ms_ = ms;
	return 0;
}

void DeviceEngine::counters(long long *out)
{
	Impl &d = *d_;
	for (int k = 0; k < CNT_N; k++) out[k] = 0;
	if (!d.bcnt.p) return;
	std::vector<long long> h(d.bcnt.n);
	if (hipMemcpy(h.data(), d.bcnt.p, h.size() * sizeof(long long), hipMemcpyDeviceToHost) != hipSuccess) return;
	for (size_t i = 0; i < h.size(); i++) out[i % 8] += h[i];
}

}  // namespace rgpu
