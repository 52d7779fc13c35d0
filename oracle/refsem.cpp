// refsem -- CPU restatement (the oracle) of the reference's radiosity path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the
// benchmark's CPU-baseline leg as the checker.  The engine never uses it.
//
// Semantics follow the reference as shipped ("V0": ADAPT ROLLBACK
// DONT_PARTITION PRE_PARTITION PARALLEL, include/config.h:5-13).  Arithmetic is
// compiled with -ffp-contract=off; with fma=1 an explicit fmaf sits exactly
// where GCC 11.4 (-O2 -march=native, the reference Makefile:7) contracts:
//   dot3(x,y)   = fma(x2,y2, fma(x0,y0, x1*y1))           vector.h:36,66,69
//   a*b - c*d   = fma(a,b, -(c*d))                          vector.h:72-74
//   p + t*v     = fma(t,v,p)                                raycast.cpp:249,300,349
//   InterpVector a*x + b*y = fma(a,x, b*y)                  raycast.cpp:97
// Everything else runs in source order; double steps run in double.
#include "refsem.h"

#include <omp.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

const double kPi = 3.1415926535897932;  // rad.h:5
const int kMaxWalk = 1 << 24;            // walk steps before a ray is reported stalled

// One pool slot, byte-compatible with the reference Triangle (rad.h:25-39).
struct Rec {
	float p[3][3];
	int Frgb[3];
	int Brgb[3];
	float accB[3][3];
	float deltaaccB[3][3];
	float deltaB[3];
	float se[3][4];
	float c[3];
	float n[3];
	float area;
	int parent;
	int neighbor[3];
};
static_assert(sizeof(Rec) == 236, "record layout");

// Octree node (rad.h:41-45).
struct Node {
	float g0[3], g1[3];
	int sub[8];
	int list;
};
static_assert(sizeof(Node) == 60, "node layout");

// ---- arithmetic with or without the contraction sites --------------------
template <bool F> inline float dot3(const float *x, const float *y)
{
	if (F) return fmaf(x[2], y[2], fmaf(x[0], y[0], x[1] * y[1]));
	return x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
}
template <bool F> inline float mulsub(float a, float b, float c, float d)
{
	if (F) return fmaf(a, b, -(c * d));
	return a * b - c * d;
}
template <bool F> inline float axpy(float t, float v, float p)  // p + t*v
{
	if (F) return fmaf(t, v, p);
	return p + t * v;
}
template <bool F> inline void cross(const float *a, const float *b, float *v)  // vector.h:71-75
{
	v[0] = mulsub<F>(a[1], b[2], b[1], a[2]);
	v[1] = mulsub<F>(a[2], b[0], b[2], a[0]);
	v[2] = mulsub<F>(a[0], b[1], b[0], a[1]);
}
template <bool F> inline float norm(const float *v) { return sqrtf(dot3<F>(v, v)); }
template <bool F> inline float norm2(const float *v) { return dot3<F>(v, v); }

// CalCenter (vector.h:53-56): float sum, divided by 3.0 in double.
inline void center(const float p[3][3], float *c)
{
	for (int k = 0; k < 3; k++)
		c[k] = (float)((double)(p[0][k] + p[1][k] + p[2][k]) / 3.0);
}
// AverageVector (vector.h:27-30): float sum, halved in double.
inline void average(const float *a, const float *b, float *m)
{
	for (int k = 0; k < 3; k++) m[k] = (float)((double)(a[k] + b[k]) / 2.0);
}
inline bool is_light(const Rec &t)  // rad.h:47-49
{
	return t.Brgb[0] == 255 || t.Brgb[1] == 255 || t.Brgb[2] == 255;
}

// ---- fp::tostring<float> semantics (float2string.h:103-271) ---------------
// Truncated 8-digit decimal of the exact binary value, trailing zeros
// dropped.  Returns false where the reference writes nothing (subnormal, inf,
// NaN, |x| >= 2^64), so the caller's buffer keeps its previous text.
bool format_float(float x, char *out)
{
	uint32_t raw;
	memcpy(&raw, &x, 4);
	const bool neg = raw >> 31;
	const int ef = (raw >> 23) & 0xff;
	const uint32_t frac = raw & 0x7fffff;
	if (ef == 0xff) return false;
	if (ef == 0) {
		if (frac) return false;
		strcpy(out, neg ? "-0.0" : "0.0");
		return true;
	}
	const int e = ef - 127 - 23;
	if (e >= 41) return false;
	const uint64_t m = frac | (1u << 23);
	uint64_t ip;
	if (e > 0) ip = m << e;
	else if (e < 0) ip = (-e > 23) ? 0 : (m >> -e);
	else ip = m;
	char *d = out;
	if (neg) *d++ = '-';
	char tmp[24];
	int k = 0;
	while (ip) { tmp[k++] = (char)('0' + ip % 10); ip /= 10; }
	if (k == 0) *d++ = '0';
	while (k) *d++ = tmp[--k];
	*d++ = '.';
	const int shift = 60 + e;
	int ndig = 0;
	char fd[8];
	if (shift < 60 && shift >= 0) {
		uint64_t f = m << shift;
		int j = 0;
		while (f && j < 8) {
			f &= 0x0fffffffffffffffull;
			f *= 10;
			int digit = (int)(f >> 60);
			fd[j++] = (char)('0' + digit);
			if (digit) ndig = j;
		}
	}
	if (ndig == 0) *d++ = '0';
	for (int j = 0; j < ndig; j++) *d++ = fd[j];
	*d = 0;
	return true;
}

struct Options {
	int debug = 0;
	float adapt_area = 5.0f, sample_area = -1.0f, converge = 1200.0f, delta_ff = 0.000005f;
	int write_cycle = 10, zip = 0, triangle = 1000000;
	float light = 1.0f;
	std::string out = "./test";
	int interactive = 0, jobs = 2, parallel_src = 1;
	int fma = 1, lowest = 0, hash = 1, skip_discarded = 0, max_tri = 1000000, old_format = 0, quiet = 0;
	int v3 = 0;  // the reference's V3 build: Delta-ff splitting live (config.h without DONT_PARTITION)
};

// glibc's rand() with its default seed (srand is never called): the TYPE_3
// additive feedback generator of degree 31 and separation 3.  srandom(1)
// fills r[0..30] with r[0] = 1, r[i] = 16807 r[i-1] mod (2^31 - 1), then
// r[i] = r[i-31] + r[i-3] (mod 2^32) and output k is r[k + 344] >> 1
// (the first 310 values are discarded).  Used only by V3's sampling
// (GetPointInTriangle, shade.cpp:49-68).
struct GlibcRand {
	uint32_t r[34];
	int k = 0;  // r[k % 34] is the oldest of the last 34 values
	GlibcRand()
	{
		int32_t w = 1;
		uint32_t s[344];
		s[0] = 1;
		for (int i = 1; i < 31; i++) {
			const int32_t hi = w / 127773, lo = w % 127773;
			w = 16807 * lo - 2836 * hi;
			if (w < 0) w += 2147483647;
			s[i] = (uint32_t)w;
		}
		for (int i = 31; i < 34; i++) s[i] = s[i - 31];
		for (int i = 34; i < 344; i++) s[i] = s[i - 31] + s[i - 3];
		for (int i = 0; i < 34; i++) r[i] = s[310 + i];
		k = 0;  // r holds s[310 .. 343]; the next value is s[344]
	}
	int next()
	{
		// s[n] = s[n - 31] + s[n - 3] with n = 344 + k: r holds s[n-34 .. n-1]
		const uint32_t v = r[(k + 3) % 34] + r[(k + 31) % 34];
		r[k % 34] = v;
		k = (k + 1) % 34;
		return (int)(v >> 1);
	}
};

struct Counters {
	long long v[RS_CNT_N];
	void clear() { memset(v, 0, sizeof v); }
};

// A half of the source (PartitionSource, shade.cpp:76-99) or the whole source.
constexpr float kSegLo = 0.0001f, kSegHi = 0.9999f;  // open segment of the matrix visibility

struct SrcPatch {
	float c[3], n[3], area;
};

struct Source {
	int phys, log;
	float deltaB[3];
	SrcPatch whole;
	bool split;
	SrcPatch half[2];
};

}  // namespace

struct rs_ctx {
	Options o;
	std::string err;
	std::vector<Rec> T;
	int trinum = 0;
	float G0[3] = {0, 0, 0}, G1[3] = {0, 0, 0};
	std::vector<Node> nodes;
	std::vector<int> lists;
	float min_grid = 0;
	int grid_num = 0;
	int n_orig = 0;  // originals the octree was built over
	int loop = 0;
	bool warned_full = false;
	char fmt_buf[64] = {0};  // the reference's static buffer (rad.cpp:179)
	GlibcRand rng;           // rand()'s state, from the process start
	Counters cnt;

	void log(const char *fmt, ...)
	{
		if (o.quiet) return;
		va_list ap;
		va_start(ap, fmt);
		vfprintf(stderr, fmt, ap);
		va_end(ap);
	}

	// ---------------- input (ReadTriangle, rad.cpp:47-111) ----------------
	template <bool F> void derive(Rec &t)
	{
		center(t.p, t.c);
		float v0[3], v1[3];
		for (int k = 0; k < 3; k++) { v0[k] = t.p[1][k] - t.p[0][k]; v1[k] = t.p[2][k] - t.p[0][k]; }
		cross<F>(v0, v1, t.n);
		t.area = norm<F>(t.n) / 2;
		float len = norm<F>(t.n);
		for (int k = 0; k < 3; k++) t.n[k] /= len;
		for (int i = 0; i < 3; i++) {
			const float *a = t.p[i], *b = t.p[(i + 1) % 3];
			float e[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, s[3];
			cross<F>(t.n, e, s);
			t.se[i][0] = s[0]; t.se[i][1] = s[1]; t.se[i][2] = s[2];
			t.se[i][3] = -dot3<F>(a, s);
		}
	}

	void reserve_pool(int n0)
	{
		long cap = std::max<long>(n0, std::min(o.max_tri, o.triangle)) + 16;
		T.assign(cap, Rec());
		memset(T.data(), 0, sizeof(Rec) * T.size());
	}

	int load(int n, const float *p9, const int *f3, const int *b3)
	{
		reserve_pool(n);
		float lim0[3] = {1000000.0f, 1000000.0f, 1000000.0f};
		float lim1[3] = {-1000000.0f, -1000000.0f, -1000000.0f};
		for (int i = 0; i < n; i++) {
			Rec &t = T[i];
			memcpy(t.p, p9 + 9 * i, 36);
			memcpy(t.Frgb, f3 + 3 * i, 12);
			memcpy(t.Brgb, b3 + 3 * i, 12);
			if (o.fma) derive<true>(t); else derive<false>(t);
			for (int v = 0; v < 3; v++)
				for (int k = 0; k < 3; k++) {
					lim0[k] = t.p[v][k] < lim0[k] ? t.p[v][k] : lim0[k];
					lim1[k] = lim1[k] < t.p[v][k] ? t.p[v][k] : lim1[k];
				}
			t.parent = i;
			t.neighbor[0] = t.neighbor[1] = t.neighbor[2] = -1;
		}
		trinum = n;
		float m = 0;
		for (int k = 0; k < 3; k++) { float d = lim1[k] - lim0[k]; m = m < d ? d : m; }
		for (int k = 0; k < 3; k++) G0[k] = (float)((double)lim0[k] - kPi);
		for (int k = 0; k < 3; k++) G1[k] = (float)((double)(G0[k] + m) + 2 * kPi);
		log("[INFO] Read #Triangle %d from input file\n", trinum);
		return 0;
	}

	int load_file(const char *path)
	{
		FILE *f = fopen(path, "r");
		if (!f) { err = std::string("File open error::") + path; return -1; }
		std::vector<float> p;
		std::vector<int> fr, br;
		char word[256];
		while (fscanf(f, "%255s", word) == 1) {
			int iv[6];
			float fv[9];
			if (fscanf(f, "%d%d%d%d%d%d", iv, iv + 1, iv + 2, iv + 3, iv + 4, iv + 5) != 6) break;
			int got = 0;
			for (int k = 0; k < 9; k++) got += fscanf(f, "%f", fv + k) == 1;
			if (got != 9) { fclose(f); err = "malformed .tri"; return -1; }
			fr.insert(fr.end(), iv, iv + 3);
			br.insert(br.end(), iv + 3, iv + 6);
			p.insert(p.end(), fv, fv + 9);
		}
		fclose(f);
		return load((int)fr.size() / 3, p.data(), fr.data(), br.data());
	}

	// ---------------- octree (raycast.cpp:24-233) ----------------
	static bool clip2(const float *p0, const float *p1, const float *g0, const float *g1, int x, int y)
	{
		int m1 = (p0[x] < g0[x]) | (p0[x] > g1[x]) << 1 | (p0[y] < g0[y]) << 2 | (p0[y] > g1[y]) << 3;
		int m2 = (p1[x] < g0[x]) | (p1[x] > g1[x]) << 1 | (p1[y] < g0[y]) << 2 | (p1[y] > g1[y]) << 3;
		if (m1 & m2) return false;
		if (!(m1 | m2)) return true;
		float a = p1[y] - p0[y], b = p1[x] - p0[x];
		float t = a * (g0[x] - p0[x]) - b * (g0[y] - p0[y]);
		if (t == 0) return true;
		bool tn = t < 0;
		if (((a * (g0[x] - p0[x]) - b * (g1[y] - p0[y])) < 0) != tn) return true;
		if (((a * (g1[x] - p0[x]) - b * (g0[y] - p0[y])) < 0) != tn) return true;
		if (((a * (g1[x] - p0[x]) - b * (g1[y] - p0[y])) < 0) != tn) return true;
		return false;
	}
	template <bool F> static void interp(const float *hi, const float *lo, float g, int d, float *p)
	{
		float dd = hi[d] - lo[d];
		if (dd == 0.0f) { p[0] = hi[0]; p[1] = hi[1]; p[2] = hi[2]; return; }
		float a = (g - lo[d]) / dd, b = (hi[d] - g) / dd;
		for (int k = 0; k < 3; k++) p[k] = F ? fmaf(a, hi[k], b * lo[k]) : a * hi[k] + b * lo[k];
	}
	template <bool F> static bool overlaps(const Rec &t, const float *g0, const float *g1)
	{
		const float *mx = t.p[0], *md = t.p[1], *mn = t.p[2];
		float p0[3], p1[3];
		for (int x = 0; x < 3; x++) {
			if (mx[x] < md[x]) std::swap(mx, md);
			if (mx[x] < mn[x]) std::swap(mx, mn);
			if (md[x] < mn[x]) std::swap(md, mn);
			const int y1 = (x + 1) % 3, y2 = (x + 2) % 3;
			if (g1[x] < mn[x]) return false;
			interp<F>(mx, mn, g1[x], x, p0);
			if (g1[x] < md[x]) { interp<F>(md, mn, g1[x], x, p1); return clip2(p0, p1, g0, g1, y1, y2); }
			if (g1[x] <= mx[x]) { interp<F>(mx, md, g1[x], x, p1); return clip2(p0, p1, g0, g1, y1, y2); }
			if (g0[x] > mx[x]) return false;
			interp<F>(mx, mn, g0[x], x, p0);
			if (g0[x] > md[x]) { interp<F>(mx, md, g0[x], x, p1); return clip2(p0, p1, g0, g1, y1, y2); }
			if (g0[x] >= mn[x]) { interp<F>(md, mn, g0[x], x, p1); return clip2(p0, p1, g0, g1, y1, y2); }
		}
		return true;
	}
	template <bool F> int build_node(const std::vector<int> &idx, int level, const float *g0, const float *g1)
	{
		int id = (int)nodes.size();
		nodes.push_back(Node());
		Node nd;
		memcpy(nd.g0, g0, 12);
		memcpy(nd.g1, g1, 12);
		nd.list = -1;
		if (idx.size() <= 10 || level >= 6) {  // MaxTriPerNode, MaxLevel (rad.h:7,12)
			if (g1[0] - g0[0] < min_grid) min_grid = g1[0] - g0[0];
			for (int i = 0; i < 8; i++) nd.sub[i] = -1;
			nd.list = (int)lists.size();
			lists.insert(lists.end(), idx.begin(), idx.end());
			lists.push_back(-1);
			nodes[id] = nd;
			return id;
		}
		float mid[3];
		for (int k = 0; k < 3; k++) mid[k] = (float)(0.5 * (double)g0[k] + 0.5 * (double)g1[k]);
		for (int i = 0; i < 8; i++) {
			float q0[3] = {g0[0], g0[1], g0[2]}, q1[3] = {g1[0], g1[1], g1[2]};
			for (int j = 0; j < 3; j++) {
				if ((i >> j) & 1) q0[j] = mid[j];
				else q1[j] = mid[j];
			}
			std::vector<int> sub;
			sub.reserve(idx.size());
			for (int t : idx)
				if (overlaps<F>(T[t], q0, q1)) sub.push_back(t);
			nd.sub[i] = build_node<F>(sub, level + 1, q0, q1);
		}
		nodes[id] = nd;
		return id;
	}
	int build_tree()
	{
		nodes.clear();
		lists.clear();
		fine.clear();
		n_orig = trinum;
		std::vector<int> all(trinum);
		for (int i = 0; i < trinum; i++) all[i] = i;
		min_grid = G1[0] - G0[0];
		if (o.fma) build_node<true>(all, 0, G0, G1); else build_node<false>(all, 0, G0, G1);
		grid_num = (int)((G1[0] - G0[0]) / min_grid);
		log("[INFO] 3D-DDA #TreeNode = %d, #Trilist = %d\n", (int)nodes.size(), (int)lists.size());
		return 0;
	}

	// ---------------- visibility (raycast.cpp:235-360) ----------------
	int descend(const int *g_in, long long *levels) const
	{
		int g[3] = {g_in[0], g_in[1], g_in[2]};
		int ctn = 0, gsize = grid_num;
		while (nodes[ctn].list < 0) {
			gsize >>= 1;
			int index = 0;
			for (int i = 0; i < 3; i++)
				if (g[i] >= gsize) { index |= 1 << i; g[i] -= gsize; }
			ctn = nodes[ctn].sub[index];
			if (levels) ++*levels;
		}
		return ctn;
	}
	int cell_of(const float *q, long long *levels) const
	{
		if (q[0] < G0[0] || q[0] > G1[0] || q[1] < G0[1] || q[1] > G1[1] || q[2] < G0[2] || q[2] > G1[2])
			return -1;
		int g[3];
		for (int i = 0; i < 3; i++) g[i] = (int)(((double)q[i] - (double)G0[i]) / (double)min_grid);
		return descend(g, levels);
	}
	template <bool F> static bool tri_hit(const float *p, const float *v, const Rec &t, float *tmin)
	{
		float a1 = dot3<F>(v, t.n);
		float vv[3] = {t.c[0] - p[0], t.c[1] - p[1], t.c[2] - p[2]};
		float a2 = dot3<F>(vv, t.n);
		float tt = a2 / a1;
		if ((double)tt < 0.0001) return false;
		if (tt >= *tmin) return false;
		float q[3];
		for (int k = 0; k < 3; k++) q[k] = axpy<F>(tt, v[k], p[k]);
		for (int s = 0; s < 3; s++)
			if (dot3<F>(q, t.se[s]) + t.se[s][3] < 0.0f) return false;
		*tmin = tt;
		return true;
	}
	template <bool F> int ray_hit(const float *p, const float *v, int ban, long long *c) const
	{
		float t = 0.f, q[3] = {p[0], p[1], p[2]};
		int prev = -1, cur, steps = 0;
		// The leaf searched next is a function of the leaf searched last
		// (hits and the exit point depend on the ray and that leaf only), so
		// a leaf searched twice means the reference cycles forever: Brent's
		// cycle detection reports such a ray at once.
		int tort = -1, power = 1, lam = 0;
		c[RS_CNT_RAYS]++;
		while ((cur = cell_of(q, &c[RS_CNT_LEVELS])) >= 0) {
			// The reference has no bound on this walk.  Report a ray instead
			// of looping where the reference never returns: when t += 1.0
			// no longer changes t, q repeats forever; kMaxWalk is a backstop.
			if (++steps > kMaxWalk) { c[RS_CNT_CAPPED]++; return -2; }
			if (prev == cur) {
				c[RS_CNT_PRECISION]++;
				float nt = (float)((double)t + 1.0);
				if (nt == t) { c[RS_CNT_CAPPED]++; return -2; }
				t = nt;
				for (int k = 0; k < 3; k++) q[k] = axpy<F>(t, v[k], p[k]);
				continue;
			}
			const Node &nd = nodes[cur];
			c[RS_CNT_CELLS]++;
			int best = -1;
			float tmin = 1000000.0f;
			for (int e = nd.list; lists[e] >= 0; e++) {
				int id = lists[e];
				if (id == ban) continue;
				c[RS_CNT_TESTS]++;
				if (tri_hit<F>(p, v, T[id], &tmin)) best = id;
			}
			if (best >= 0) return best;
			if (cur == tort) { c[RS_CNT_CAPPED]++; return -2; }
			if (++lam == power) { tort = cur; power <<= 1; lam = 0; }
			prev = cur;
			// NextPoint (raycast.cpp:235-251)
			float tm = 1000000.0f;
			for (int i = 0; i < 3; i++) {
				if ((double)fabsf(v[i]) > 1e-6) {
					float s = v[i] > 0 ? (nd.g1[i] - p[i]) / v[i] : (nd.g0[i] - p[i]) / v[i];
					tm = s < tm ? s : tm;
				}
			}
			tm = (float)((double)tm + 0.1);
			for (int k = 0; k < 3; k++) q[k] = axpy<F>(tm, v[k], p[k]);
			t = tm;
		}
		return -1;
	}

	// ---------------- segment visibility of the explicit matrix ----------------
	// Not a reference function: the gathering matrix asks whether the open
	// segment between two centroids is blocked.  Point location is
	// geometric (the 2^D grid of the deepest level; a cell descends by the
	// bits of its coordinates) and the walk is Amanatides-Woo over that grid;
	// a leaf's list is tested once (its cells are consecutive on a segment).
	std::vector<int> fine;
	int fine_r = 0;
	void build_fine()
	{
		if (!fine.empty() || nodes.empty()) return;
		int D = 0;
		std::vector<std::pair<int, int>> stack{{0, 0}};
		while (!stack.empty()) {
			auto [nd, d] = stack.back();
			stack.pop_back();
			if (nodes[nd].list >= 0) { D = std::max(D, d); continue; }
			for (int k = 0; k < 8; k++) stack.push_back({nodes[nd].sub[k], d + 1});
		}
		fine_r = 1 << D;
		const int R = fine_r;
		fine.assign((size_t)R * R * R, -1);
		for (int a = 0; a < R; a++)
			for (int b = 0; b < R; b++)
				for (int c = 0; c < R; c++) {
					int nd = 0;
					for (int bit = D - 1; nodes[nd].list < 0; bit--)
						nd = nodes[nd].sub[((a >> bit) & 1) | (((b >> bit) & 1) << 1) | (((c >> bit) & 1) << 2)];
					fine[((size_t)a * R + b) * R + c] = nd;
				}
		// complete lists (the reference's CrossOver/Clip listing misses
		// triangles that cross a box through their interior): an original
		// goes to every leaf holding a fine cell its bounding box, grown by
		// 1e-3 of a cell, touches
		seg.assign(nodes.size(), {});
		for (int i = 0; i < n_orig; i++) {
			int lo[3], hi[3];
			for (int k = 0; k < 3; k++) {
				const double h = ((double)G1[k] - (double)G0[k]) / R;
				double mn = T[i].p[0][k], mx = mn;
				for (int v = 1; v < 3; v++) {
					mn = std::min(mn, (double)T[i].p[v][k]);
					mx = std::max(mx, (double)T[i].p[v][k]);
				}
				lo[k] = std::max(0, (int)std::floor((mn - G0[k]) / h - 1e-3));
				hi[k] = std::min(R - 1, (int)std::floor((mx - G0[k]) / h + 1e-3));
			}
			for (int a = lo[0]; a <= hi[0]; a++)
				for (int b = lo[1]; b <= hi[1]; b++)
					for (int c = lo[2]; c <= hi[2]; c++) {
						std::vector<int> &L = seg[fine[((size_t)a * R + b) * R + c]];
						if (L.empty() || L.back() != i) L.push_back(i);
					}
		}
	}
	std::vector<std::vector<int>> seg;  // complete list per leaf node
	// one original against the open segment (the walk's and the brute test)
	template <bool F> bool seg_hits(const float *p, const float *v, const Rec &t) const
	{
		const float a1 = dot3<F>(v, t.n);
		const float vv[3] = {t.c[0] - p[0], t.c[1] - p[1], t.c[2] - p[2]};
		const float a2 = dot3<F>(vv, t.n);
		const float tt = a2 / a1;
		if (!(tt > kSegLo && tt < kSegHi)) return false;
		float q[3];
		for (int k = 0; k < 3; k++) q[k] = axpy<F>(tt, v[k], p[k]);
		for (int s2 = 0; s2 < 3; s2++)
			if (dot3<F>(q, t.se[s2]) + t.se[s2][3] < 0.0f) return false;
		return true;
	}
	template <bool F> bool seg_brute(const float *p, const float *v, int ban0, int ban1) const
	{
		for (int id = 0; id < n_orig; id++)
			if (id != ban0 && id != ban1 && seg_hits<F>(p, v, T[id])) return true;
		return false;
	}
	template <bool F>
	bool seg_occluded(const float *p, const float *v, int ban0, int ban1, long long *c) const
	{
		const int R = fine_r;
		int cell[3], step[3];
		float tmax[3], tdelta[3];
		for (int k = 0; k < 3; k++) {
			const float h = (G1[k] - G0[k]) / (float)R, inv = (float)R / (G1[k] - G0[k]);
			const float rel = p[k] - G0[k];
			int ck = (int)floorf(rel * inv);
			ck = ck < 0 ? 0 : (ck >= R ? R - 1 : ck);
			cell[k] = ck;
			if (v[k] > 0.0f) {
				step[k] = 1;
				tmax[k] = ((float)(ck + 1) * h - rel) / v[k];
				tdelta[k] = h / v[k];
			} else if (v[k] < 0.0f) {
				step[k] = -1;
				tmax[k] = ((float)ck * h - rel) / v[k];
				tdelta[k] = -h / v[k];
			} else {
				step[k] = 0;
				tmax[k] = tdelta[k] = INFINITY;
			}
		}
		c[RS_CNT_RAYS]++;
		int last = -1;
		for (;;) {
			const int nd = fine[((size_t)cell[0] * R + cell[1]) * R + cell[2]];
			if (nd != last) {
				last = nd;
				c[RS_CNT_CELLS]++;
				for (int id : seg[nd]) {
					if (id == ban0 || id == ban1) continue;
					c[RS_CNT_TESTS]++;
					if (seg_hits<F>(p, v, T[id])) return true;
				}
			}
			const int ax = tmax[0] <= tmax[1] ? (tmax[0] <= tmax[2] ? 0 : 2) : (tmax[1] <= tmax[2] ? 1 : 2);
			if (tmax[ax] >= 1.0f) return false;
			cell[ax] += step[ax];
			if (cell[ax] < 0 || cell[ax] >= R) return false;
			tmax[ax] += tdelta[ax];
		}
	}
	// CalFF's value without the ray (shade.cpp:107-123)
	template <bool F> float ff_value(const SrcPatch &s, const float *dn, const float *p) const
	{
		float dir[3] = {s.c[0] - p[0], s.c[1] - p[1], s.c[2] - p[2]};
		float ds = dot3<F>(dir, s.n);
		float dd = -dot3<F>(dir, dn);
		if ((ds < 0) != (dd < 0)) return 0.0f;
		float dl = norm<F>(dir), sl = norm<F>(s.n), el = norm<F>(dn);
		float c1 = ds / dl / sl;
		float c2 = dd / dl / el;
		float ff = c1 * c2;
		if (ff <= 0.0f) return 0.0f;
		ff = (float)((double)ff * ((double)s.area / ((double)norm2<F>(dir) * kPi)));
		return ff <= 0.0f ? 0.0f : ff;
	}

	// ---------------- form factor (shade.cpp:76-168) ----------------
	template <bool F> float cal_ff(const SrcPatch &s, int logsrc, const float *dn, int logdst, const float *p, long long *c) const
	{
		float dir[3] = {s.c[0] - p[0], s.c[1] - p[1], s.c[2] - p[2]};
		float ds = dot3<F>(dir, s.n);
		float dd = -dot3<F>(dir, dn);
		if ((ds < 0) != (dd < 0)) return 0.0f;
		float dl = norm<F>(dir), sl = norm<F>(s.n), el = norm<F>(dn);
		float c1 = ds / dl / sl;
		float c2 = dd / dl / el;
		float ff = c1 * c2;
		if (ff <= 0.0f) return 0.0f;
		ff = (float)((double)ff * ((double)s.area / ((double)norm2<F>(dir) * kPi)));
		if (ff <= 0.0f) return 0.0f;
		if (ray_hit<F>(p, dir, logdst, c) == logsrc) return ff;
		return 0.0f;
	}
	template <bool F> Source make_source(int phys) const
	{
		const Rec &s = T[phys];
		Source src;
		src.phys = phys;
		src.log = s.parent;
		memcpy(src.deltaB, s.deltaB, 12);
		memcpy(src.whole.c, s.c, 12);
		memcpy(src.whole.n, s.n, 12);
		src.whole.area = s.area;
		int edge = 0;
		float maxlen = 0.0f;
		for (int v = 0; v < 3; v++) {
			float l[3];
			for (int k = 0; k < 3; k++) l[k] = s.p[(v + 1) % 3][k] - s.p[v][k];
			float len = norm2<F>(l);
			if (maxlen < len) { maxlen = len; edge = v; }
		}
		src.split = !((double)maxlen < 0.1);
		if (src.split) {
			int i0 = edge, i1 = (edge + 1) % 3, i2 = (edge + 2) % 3;
			float m[3];
			average(s.p[i0], s.p[i1], m);
			float h1[3][3], h2[3][3];
			memcpy(h1[0], m, 12); memcpy(h1[1], s.p[i2], 12); memcpy(h1[2], s.p[i0], 12);
			memcpy(h2[0], m, 12); memcpy(h2[1], s.p[i1], 12); memcpy(h2[2], s.p[i2], 12);
			center(h1, src.half[0].c);
			center(h2, src.half[1].c);
			memcpy(src.half[0].n, s.n, 12);
			memcpy(src.half[1].n, s.n, 12);
			src.half[0].area = src.half[1].area = (float)((double)s.area / 2.0);
		}
		return src;
	}
	// Shade's form factors for one receiver (shade.cpp:442-450, V0)
	template <bool F> void shade_ff(const Source &s, const Rec &d, float ff[3], long long *c) const
	{
		for (int v = 0; v < 3; v++) {
			float f = 0.0f;
			if (!s.split || !o.skip_discarded) f = cal_ff<F>(s.whole, s.log, d.n, d.parent, d.p[v], c);
			if (s.split) {
				float f1 = cal_ff<F>(s.half[0], s.log, d.n, d.parent, d.p[v], c);
				float f2 = cal_ff<F>(s.half[1], s.log, d.n, d.parent, d.p[v], c);
				f = f1 + f2;
			}
			ff[v] = f;
		}
	}
	// AdaptCalFF(CalFF(..)) at one point p of the receiver with normal dn
	// (shade.cpp:141-168, 445-446)
	template <bool F> float point_ff(const Source &s, const float *dn, int logdst, const float *p, long long *c) const
	{
		float f = 0.0f;
		if (!s.split || !o.skip_discarded) f = cal_ff<F>(s.whole, s.log, dn, logdst, p, c);
		if (s.split) {
			float f1 = cal_ff<F>(s.half[0], s.log, dn, logdst, p, c);
			float f2 = cal_ff<F>(s.half[1], s.log, dn, logdst, p, c);
			f = f1 + f2;
		}
		return f;
	}
	static float ffabs(float x) { return (float)fabs((double)x); }  // rad.h:16
	// GetPointInTriangle (shade.cpp:49-68): a point from three rand() weights
	// and the form factor interpolated there; GCC contracts both sums as dot
	// products (fma(w3, x3, fma(w1, x1, w2 * x2))).
	template <bool F> float random_point(const Rec &t, float *p, const float ff[3])
	{
		float r1 = (float)rng.next(), r2 = (float)rng.next(), r3 = (float)rng.next();
		const float tot = r1 + r2 + r3;
		r1 /= tot;
		r2 /= tot;
		r3 /= tot;
		const float w[3] = {r1, r2, r3};
		for (int v = 0; v < 3; v++) {
			const float x[3] = {t.p[0][v], t.p[1][v], t.p[2][v]};
			p[v] = dot3<F>(w, x);
		}
		return dot3<F>(w, ff);
	}
	// CalArea (vector.h:58-62): n = (p1 - p0) x (p2 - p0), not normalized;
	// area = |n| / 2
	template <bool F> static void cal_area(Rec &t)
	{
		float v0[3], v1[3];
		for (int k = 0; k < 3; k++) {
			v0[k] = t.p[1][k] - t.p[0][k];
			v1[k] = t.p[2][k] - t.p[0][k];
		}
		cross<F>(v0, v1, t.n);
		t.area = norm<F>(t.n) / 2;
	}
	// PartitionTriangleWithCenterAndStore (shade.cpp:261-306) with its 3-way
	// PartitionDestination (shade.cpp:213-241): -1 when slot i cannot split,
	// 0 when it split and, being its own parent, became logical, 1 when its
	// slot now holds the third part (the last slot is given back).
	template <bool F> int split_center(int i)
	{
		if (full_pool(4) || (double)T[i].area < 1e-8) return -1;
		float maxl = 0.0f, minl = 1e+30f;
		for (int v = 0; v < 3; v++) {
			float l[3];
			for (int k = 0; k < 3; k++) l[k] = T[i].p[(v + 1) % 3][k] - T[i].p[v][k];
			const float len = norm2<F>(l);
			maxl = (maxl < len) ? len : maxl;  // std::max
			minl = (len < minl) ? len : minl;  // std::min
		}
		if ((double)(minl / maxl) < 1e-1 || minl < 1) return -1;
		const int id[3] = {alloc_slot(), alloc_slot(), alloc_slot()};
		const Rec src = T[i];
		for (int k = 0; k < 3; k++) {
			Rec &t = T[id[k]];
			t = src;
			memcpy(t.p[k], src.c, 12);
			memcpy(t.accB[k], src.accB[k], 12);
			center(t.p, t.c);
			cal_area<F>(t);
			t.parent = src.parent;
		}
		T[id[0]].neighbor[0] = id[2], T[id[0]].neighbor[2] = id[1];
		T[id[1]].neighbor[1] = id[0], T[id[1]].neighbor[0] = id[2];
		T[id[2]].neighbor[2] = id[1], T[id[2]].neighbor[1] = id[0];
		if (src.parent == i) {
			T[i].parent = -1;
			return 0;
		}
		T[i] = T[id[2]];
		trinum--;  // FreeTriangle
		return 1;
	}
	// Shade in the V3 build (shade.cpp:435-525, PARALLEL without
	// DONT_PARTITION): returns the split flag
	template <bool F> int shade_v3(const Source &s, int i, int logdst, long long *c)
	{
		Rec &d = T[i];
		float ff[3];
		for (int v = 0; v < 3; v++) ff[v] = point_ff<F>(s, d.n, logdst, d.p[v], c);
		if (d.area < o.adapt_area) {
			receive(s.deltaB, d, ff);
			return 0;
		}
		if (full_pool(4) || (double)d.area < 1e-8) {
			receive(s.deltaB, d, ff);
			return 0;
		}
		float groud = (float)((double)(ff[0] + ff[1] + ff[2]) / 3.0);
		float ffs = point_ff<F>(s, d.n, logdst, d.c, c);
		float dff = ffabs(groud - ffs);
		if (dff < o.delta_ff) {
			const int samples = (int)(d.area / o.sample_area);
			for (int k = 0; k < samples; k++) {
				float q[3];
				groud = random_point<F>(d, q, ff);
				ffs = point_ff<F>(s, d.n, logdst, q, c);
				dff = ffabs(groud - ffs);
				if (dff > o.delta_ff) break;
			}
			if (dff < o.delta_ff) {
				receive(s.deltaB, d, ff);
				return 0;
			}
		}
		const int r = split_center<F>(i);
		if (r == 1) {  // slot i holds the third part: its centroid corner shaded anew
			Rec &t3 = T[i];
			const float ff2[3] = {ff[0], ff[1], point_ff<F>(s, t3.n, logdst, t3.p[2], c)};
			receive(s.deltaB, t3, ff2);
		}
		if (r < 0) receive(s.deltaB, d, ff);
		return r;
	}
	// parallelRadiosity in the V3 build, run in receiver order (its -jobs 1
	// order; more threads race on the slot pool): new slots made by splits
	// are shaded in further passes over the slots they occupy
	// (rad.cpp:406-451).
	template <bool F> int shoot_once_v3()
	{
		std::vector<int> C = candidates();
		if (C.empty()) return 0;
		std::vector<Source> src;
		for (int s : C) src.push_back(make_source<F>(s));
		std::set<int> chosen(C.begin(), C.end());
		long long c[RS_CNT_N] = {0};
		int prev = 0;
		for (;;) {
			const int n = trinum;
			for (int i = prev; i < n; i++) {
				if (T[i].parent < 0 || chosen.count(i) || saturated(T[i])) continue;
				const int logdst = T[i].parent;
				for (const Source &s : src) {
					if (s.phys == i) continue;
					shade_v3<F>(s, i, logdst, c);
					c[RS_CNT_SHADES]++;
				}
			}
			prev = n;
			if (n == trinum) break;
		}
		for (int k = 0; k < RS_CNT_N; k++) cnt.v[k] += c[k];
		for (int s : C) T[s].deltaB[0] = T[s].deltaB[1] = T[s].deltaB[2] = 0.0f;
		cnt.v[RS_CNT_ITERS]++;
		return 1;
	}

	// CalRadiosity (shade.cpp:244-259)
	static void receive(const float *src_dB, Rec &d, const float ff[3])
	{
		const float k = (float)(0.5 / 255.0);
		float lo[3];
		for (int c = 0; c < 3; c++) lo[c] = (float)d.Frgb[c] * k * src_dB[c] / 3;
		for (int v = 0; v < 3; v++)
			for (int c = 0; c < 3; c++) {
				float db = lo[c] * ff[v];
				d.deltaB[c] += db;
				d.accB[v][c] += db;
				d.deltaaccB[v][c] = db;
			}
	}

	// ---------------- InitRadiosity (rad.cpp:259-356) ----------------
	template <bool F> float cos_theta(const float *a, const float *b) const
	{
		return dot3<F>(a, b) / norm<F>(a) / norm<F>(b);
	}
	static bool veq(const float *a, const float *b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }

	// The neighbour pass visits (i, j, v, n) in that order and writes
	// T[i].neighbor[v] = j and T[j].neighbor[n or n+2] = i.  The last write to
	// each (slot, edge) wins.
	template <bool F> void neighbors_exact()
	{
		for (int i = 0; i < trinum; i++) {
			Rec &tp = T[i];
			for (int j = 0; j < trinum; j++) {
				if (j == i) continue;
				Rec &tn = T[j];
				for (int v = 0; v < 3; v++)
					for (int n = 0; n < 3; n++) {
						if (!veq(tp.p[v], tn.p[n])) continue;
						if (veq(tp.p[(v + 1) % 3], tn.p[(n + 1) % 3])) {
							if ((double)cos_theta<F>(tp.n, tn.n) > 0.85) { tp.neighbor[v] = j; tn.neighbor[n] = i; }
						} else if (veq(tp.p[(v + 1) % 3], tn.p[(n + 2) % 3])) {
							if ((double)cos_theta<F>(tp.n, tn.n) > 0.85) { tp.neighbor[v] = j; tn.neighbor[(n + 2) % 3] = i; }
						}
					}
			}
		}
	}
	struct VKey {
		uint32_t x, y, z;
		bool operator==(const VKey &o) const { return x == o.x && y == o.y && z == o.z; }
	};
	struct VHash {
		size_t operator()(const VKey &k) const
		{
			uint64_t h = k.x * 0x9E3779B97F4A7C15ull ^ (uint64_t)k.y * 0xC2B2AE3D27D4EB4Full ^ (uint64_t)k.z * 0x165667B19E3779F9ull;
			return (size_t)(h ^ (h >> 29));
		}
	};
	static bool vkey(const float *p, VKey &k)
	{
		if (p[0] != p[0] || p[1] != p[1] || p[2] != p[2]) return false;  // NaN never compares equal
		float q[3] = {p[0] + 0.0f, p[1] + 0.0f, p[2] + 0.0f};           // -0 == +0
		memcpy(&k, q, 12);
		return true;
	}
	template <bool F> void neighbors_hash()
	{
		std::unordered_map<VKey, std::vector<int>, VHash> at;  // vertex -> 3*j+n
		at.reserve(trinum * 2);
		for (int j = 0; j < trinum; j++)
			for (int n = 0; n < 3; n++) {
				VKey k;
				if (vkey(T[j].p[n], k)) at[k].push_back(3 * j + n);
			}
		// Collect writes as (slot*3+edge) -> (key, value); keep the max key.
		struct W { uint64_t key; int val; };
		std::vector<W> best((size_t)trinum * 3, W{0, 0});
		std::vector<char> has((size_t)trinum * 3, 0);
		auto put = [&](int slot, int edge, uint64_t key, int val) {
			size_t s = (size_t)slot * 3 + edge;
			if (!has[s] || key > best[s].key) { best[s] = W{key, val}; has[s] = 1; }
		};
		for (int i = 0; i < trinum; i++) {
			const Rec &tp = T[i];
			for (int v = 0; v < 3; v++) {
				VKey k;
				if (!vkey(tp.p[v], k)) continue;
				auto it = at.find(k);
				if (it == at.end()) continue;
				for (int code : it->second) {
					int j = code / 3, n = code % 3;
					if (j == i) continue;
					const Rec &tn = T[j];
					// loop order key (i, j, v, n)
					uint64_t key = ((uint64_t)i << 34) | ((uint64_t)j << 4) | (uint64_t)(v * 3 + n);
					if (veq(tp.p[(v + 1) % 3], tn.p[(n + 1) % 3])) {
						if ((double)cos_theta<F>(tp.n, tn.n) > 0.85) { put(i, v, key, j); put(j, n, key, i); }
					} else if (veq(tp.p[(v + 1) % 3], tn.p[(n + 2) % 3])) {
						if ((double)cos_theta<F>(tp.n, tn.n) > 0.85) { put(i, v, key, j); put(j, (n + 2) % 3, key, i); }
					}
				}
			}
		}
		for (int s = 0; s < trinum; s++)
			for (int e = 0; e < 3; e++)
				if (has[(size_t)s * 3 + e]) T[s].neighbor[e] = best[(size_t)s * 3 + e].val;
	}
	int init_radiosity()
	{
		float minarea = 100000000.0f, maxarea = 0.0f;
		for (int i = 0; i < trinum; i++) {
			Rec &t = T[i];
			minarea = t.area < minarea ? t.area : minarea;
			maxarea = maxarea < t.area ? t.area : maxarea;
			if (is_light(t)) {
				t.deltaB[0] = t.deltaB[1] = t.deltaB[2] = 0.0f;
				for (int v = 0; v < 3; v++) {
					for (int c = 0; c < 3; c++) t.accB[v][c] = (float)t.Brgb[c] * o.light;
					for (int c = 0; c < 3; c++) t.deltaB[c] += 100 * t.accB[v][c] / 3;
				}
			} else {
				memset(t.accB, 0, sizeof t.accB);
				t.deltaB[0] = t.deltaB[1] = t.deltaB[2] = 0.0f;
			}
		}
		if (o.hash) { if (o.fma) neighbors_hash<true>(); else neighbors_hash<false>(); }
		else { if (o.fma) neighbors_exact<true>(); else neighbors_exact<false>(); }
		if (o.sample_area < 0.0f) o.sample_area = (float)((double)maxarea / 100.0);
		if (!o.quiet) {
			printf("[INFO] Max Patch Area = %f\n", maxarea);
			printf("[INFO] Min Patch Area = %f\n", minarea);
			printf("[INFO] Config Table\n");
			printf("       -debug          %i\n", o.debug);
			printf("       -adapt_area     %f\n", o.adapt_area);
			printf("       -sample_area    %f\n", o.sample_area);
			printf("       -converge       %f\n", o.converge);
			printf("       -delta_ff       %f\n", o.delta_ff);
			printf("       -write_cycle    %d\n", o.write_cycle);
			printf("       -triangle       %d\n", o.triangle);
			printf("       -light          %f\n", o.light);
			printf("       -interactive    %d\n", o.interactive);
			printf("       -jobs           %d\n", o.jobs);
			printf("       -parallel_src   %d\n", o.parallel_src);
			printf("[INFO] Complete initialization\n");
		}
		return 0;
	}

	// ---------------- pre-partition (shade.cpp:11-43, 171-211, 311-431, 556-598) ----
	bool full_pool(int need) const { return trinum + need >= std::min(o.max_tri, o.triangle); }
	int alloc_slot()
	{
		if (trinum == o.max_tri || trinum == o.triangle) {
			if (!warned_full) { log("\n**** Max Triangle exceed : %i ***\n", trinum); warned_full = true; }
			return -1;
		}
		return trinum++;
	}
	void set_neighbor_to(int nb, int from, int to)
	{
		if (nb < 0) return;
		for (int v = 0; v < 3; v++)
			if (T[nb].neighbor[v] == from) T[nb].neighbor[v] = to;
	}
	// PartitionDestination, 2-way: halves s at edge into slots a and b.
	void halve(int s_idx, int a_idx, int b_idx, int edge)
	{
		const Rec s = T[s_idx];
		Rec &a = T[a_idx], &b = T[b_idx];
		int i0 = edge, i1 = (edge + 1) % 3, i2 = (edge + 2) % 3;
		float m[3];
		average(s.p[i0], s.p[i1], m);
		memcpy(a.p[0], m, 12); memcpy(a.p[1], s.p[i2], 12); memcpy(a.p[2], s.p[i0], 12);
		memcpy(a.n, s.n, 12); memcpy(a.Frgb, s.Frgb, 12); memcpy(a.deltaB, s.deltaB, 12);
		center(a.p, a.c);
		memcpy(b.p[0], m, 12); memcpy(b.p[1], s.p[i1], 12); memcpy(b.p[2], s.p[i2], 12);
		memcpy(b.n, s.n, 12); memcpy(b.Frgb, s.Frgb, 12); memcpy(b.deltaB, s.deltaB, 12);
		center(b.p, b.c);
		float ma[3];
		average(s.accB[i0], s.accB[i1], ma);
		memcpy(a.accB[0], ma, 12); memcpy(a.accB[1], s.accB[i2], 12); memcpy(a.accB[2], s.accB[i0], 12);
		memcpy(b.accB[0], ma, 12); memcpy(b.accB[1], s.accB[i1], 12); memcpy(b.accB[2], s.accB[i2], 12);
		a.area = b.area = (float)((double)s.area / 2.0);
		a.parent = b.parent = s.parent;
	}
	// PartitionTriangleAndStore: returns -1 (refused), 0 or 1 (s replaced).
	template <bool F> int split_pair(int s, int tv[4])
	{
		tv[0] = tv[1] = tv[2] = tv[3] = -1;
		if (full_pool(4) || (double)T[s].area < 1e-8) return -1;
		int edge = 0;
		float maxlen = 0.0f;
		for (int v = 0; v < 3; v++) {
			float l[3];
			for (int k = 0; k < 3; k++) l[k] = T[s].p[(v + 1) % 3][k] - T[s].p[v][k];
			float len = norm2<F>(l);
			if (maxlen < len) { maxlen = len; edge = v; }
		}
		if (maxlen < 1) return -1;
		int nb = T[s].neighbor[edge];
		if (nb >= 0 && (double)T[nb].area < 1e-1) return -1;
		int t1 = alloc_slot(), t2 = alloc_slot();
		halve(s, t1, t2, edge);
		int n1 = T[s].neighbor[(edge + 2) % 3], n2 = T[s].neighbor[(edge + 1) % 3];
		T[t1].neighbor[1] = n1;
		T[t2].neighbor[1] = n2;
		T[t1].neighbor[2] = T[t2].neighbor[0] = -1;
		int replaced = 0;
		if (T[s].parent == s) T[s].parent = -1;
		else { T[s] = T[t2]; t2 = s; trinum--; replaced = 1; }
		set_neighbor_to(n1, s, t1);
		set_neighbor_to(n2, s, t2);
		T[t1].neighbor[0] = t2;
		T[t2].neighbor[2] = t1;
		tv[0] = t1; tv[1] = t2;
		if (nb < 0) return replaced;
		int t3 = alloc_slot(), t4 = alloc_slot();
		int ne = 0;
		for (int v = 0; v < 3; v++)
			if (T[nb].neighbor[v] == s) ne = v;
		if (nb < s)
			for (int v = 0; v < 3; v++)
				for (int c = 0; c < 3; c++) T[nb].accB[v][c] -= T[nb].deltaaccB[v][c];
		halve(nb, t3, t4, ne);
		n1 = T[nb].neighbor[(ne + 2) % 3];
		n2 = T[nb].neighbor[(ne + 1) % 3];
		T[t3].neighbor[1] = n1;
		T[t4].neighbor[1] = n2;
		if (T[nb].parent == nb) T[nb].parent = -1;
		else { T[nb] = T[t4]; t4 = nb; trinum--; }
		set_neighbor_to(n1, nb, t3);
		set_neighbor_to(n2, nb, t4);
		T[t3].neighbor[0] = t4;
		T[t4].neighbor[2] = t3;
		// shade.cpp:416 compares the two vertex arrays' addresses
		if (nb == s && edge == ne) {
			T[t1].neighbor[2] = t3; T[t3].neighbor[2] = t1;
			T[t2].neighbor[0] = t4; T[t4].neighbor[0] = t2;
		} else {
			T[t1].neighbor[2] = t4; T[t4].neighbor[0] = t1;
			T[t2].neighbor[0] = t3; T[t3].neighbor[2] = t2;
		}
		tv[2] = t3; tv[3] = t4;
		return replaced;
	}
	template <bool F> int prepartition_impl()
	{
		log("[DEBUG] Before PrePartitionTriangle %d\n", trinum);
		std::set<std::pair<float, int>> S;
		for (int i = 0; i < trinum; i++) {
			if (T[i].area < 1.f || is_light(T[i])) continue;
			S.insert(std::make_pair(T[i].area, i));
		}
		int status = 0;
		while (true) {
			if (S.empty()) { status = 1; break; }  // the reference dereferences an empty set here
			std::pair<float, int> top = *S.rbegin();
			S.erase(top);
			int tp = top.second;
			if (T[tp].area < 1.0f) break;
			int tv[4];
			if (split_pair<F>(tp, tv) < 0) break;
			if (is_light(T[tp]))
				for (int i = 0; i < 4; i++) {
					if (tv[i] < 0) continue;
					memcpy(T[tv[i]].Brgb, T[tp].Brgb, 12);
					for (int v = 0; v < 3; v++) {
						memcpy(T[tv[i]].accB[v], T[tp].accB[v], 12);
						T[tv[i]].deltaB[v] = T[tp].deltaB[v];
					}
				}
			for (int i = 0; i < 4; i++)
				if (tv[i] >= 0) S.insert(std::make_pair(T[tv[i]].area, tv[i]));
		}
		log("[DEBUG] After PrePartitionTriangle %d\n", trinum);
		return status;
	}
	int prepartition() { return o.fma ? prepartition_impl<true>() : prepartition_impl<false>(); }

	// ---------------- shooting (rad.cpp:359-451) ----------------
	std::vector<int> candidates() const
	{
		std::vector<std::pair<float, int>> C;
		for (int i = 0; i < trinum; i++) {
			const Rec &t = T[i];
			if (t.parent < 0) continue;
			float d = t.deltaB[0] + t.deltaB[1] + t.deltaB[2];
			if (d > o.converge) C.push_back(std::make_pair(d, i));
		}
		std::vector<int> ret;
		if (o.lowest) {  // MaxDeltaRad (rad.cpp:359-386): strict >, first index wins
			int best = -1;
			float bd = 0.f;
			for (auto &e : C)
				if (e.first > bd) { bd = e.first; best = e.second; }
			if (best >= 0) ret.push_back(best);
			return ret;
		}
		std::sort(C.begin(), C.end());
		int k = std::min(((int)C.size() + 4) / 5, o.parallel_src);
		for (int i = 0; i < k; i++) ret.push_back(C[C.size() - 1 - i].second);
		return ret;
	}
	static bool saturated(const Rec &d)
	{
		for (int v = 0; v < 3; v++)
			for (int c = 0; c < 3; c++)
				if (d.accB[v][c] < 255.0f) return false;
		return true;
	}
	template <bool F> int shoot_once()
	{
		std::vector<int> C = candidates();
		if (C.empty()) return 0;
		std::vector<Source> src;
		for (int s : C) src.push_back(make_source<F>(s));
		std::vector<char> chosen(trinum, 0);
		for (int s : C) chosen[s] = 1;
		const int n = trinum;
		long long tot[RS_CNT_N] = {0};
#pragma omp parallel
		{
			long long c[RS_CNT_N] = {0};
#pragma omp for schedule(dynamic, 64)
			for (int i = 0; i < n; i++) {
				Rec &d = T[i];
				if (d.parent < 0 || chosen[i] || saturated(d)) continue;
				for (const Source &s : src) {
					if (s.phys == i) continue;
					float ff[3];
					shade_ff<F>(s, d, ff, c);
					receive(s.deltaB, d, ff);
					c[RS_CNT_SHADES]++;
				}
			}
#pragma omp critical
			for (int k = 0; k < RS_CNT_N; k++) tot[k] += c[k];
		}
		for (int k = 0; k < RS_CNT_N; k++) cnt.v[k] += tot[k];
		for (int s : C) T[s].deltaB[0] = T[s].deltaB[1] = T[s].deltaB[2] = 0.0f;
		cnt.v[RS_CNT_ITERS]++;
		return 1;
	}
	int shoot()
	{
		if (o.v3) return o.fma ? shoot_once_v3<true>() : shoot_once_v3<false>();
		return o.fma ? shoot_once<true>() : shoot_once<false>();
	}

	// ---------------- output (rad.cpp:117-242) ----------------
	const char *fmt(float x)
	{
		char tmp[64];
		if (format_float(x, tmp)) strcpy(fmt_buf, tmp);
		return fmt_buf;
	}
	int write_out(const char *path)
	{
		FILE *f = fopen(path, "w");
		if (!f) { err = std::string("Error: output file ") + path + " open failed"; return -1; }
		for (int i = 0; i < trinum; i++) {
			const Rec &t = T[i];
			if (t.parent < 0) continue;
			if (o.old_format) {
				fputs("Triangle \n", f);
				for (int v = 0; v < 3; v++) {
					fprintf(f, "%f %f %f ", t.p[v][0], t.p[v][1], t.p[v][2]);
					fprintf(f, "%i %i %i 0 0 0 \n", (int)t.accB[v][0], (int)t.accB[v][1], (int)t.accB[v][2]);
				}
				continue;
			}
			fputs("Triangle \n", f);
			int col[3];
			for (int c = 0; c < 3; c++) {
				float s = 0;
				for (int v = 0; v < 3; v++) s += t.accB[v][c];
				s /= 3;
				s = (255.0f < s) ? 255.0f : s;
				s = (s < 0.0f) ? 0.0f : s;
				if (std::isnan(s)) s = 0;
				col[c] = (int)s;
			}
			fprintf(f, "%d %d %d 0 0 0\n", col[0], col[1], col[2]);
			std::string nstr;
			for (int k = 0; k < 3; k++) { nstr += fmt(t.n[k]); nstr += k < 2 ? " " : "\n"; }
			for (int v = 0; v < 3; v++) {
				for (int k = 0; k < 3; k++) { fputs(fmt(t.p[v][k]), f); fputs(" ", f); }
				fputs(nstr.c_str(), f);
			}
		}
		fclose(f);
		return 0;
	}

	int run(const char *prefix)
	{
		init_radiosity();
		prepartition();
		int lp;
		for (lp = 1;; lp++) {
			if (lp % o.write_cycle == 0) log("[INFO] Loop %d BEGIN\n", lp);
			if (shoot() == 0) break;
			if (lp % o.write_cycle == 0) {
				std::string p = std::string(prefix) + "." + std::to_string(lp);
				write_out(p.c_str());
				log("[INFO] END Loop. There are %i triangles.\n", trinum);
			}
		}
		log("[INFO] Complete in #Iterations %d\n", lp);
		std::string fin = std::string(prefix) + ".fin";
		if (write_out(fin.c_str()) != 0) return -1;
		loop = lp;
		return lp;
	}
};

namespace {
bool parse_opts(Options &o, int argc, const char **argv, std::string &err)
{
	for (int i = 0; i < argc; i++) {
		std::string a = argv[i];
		auto need = [&](const char *what) -> const char * {
			if (i + 1 >= argc) { err = std::string("missing value for ") + what; return nullptr; }
			return argv[++i];
		};
		const char *v = nullptr;
#define OPT_F(name, field) else if (a == name) { if (!(v = need(name)) || sscanf(v, "%f", &o.field) != 1) { err = name; return false; } }
#define OPT_I(name, field) else if (a == name) { if (!(v = need(name)) || sscanf(v, "%d", &o.field) != 1) { err = name; return false; } }
		if (false) {}
		OPT_I("-debug", debug)
		OPT_F("-adapt_area", adapt_area)
		OPT_F("-sample_area", sample_area)
		OPT_F("-converge", converge)
		OPT_F("-delta_ff", delta_ff)
		OPT_I("-write_cycle", write_cycle)
		else if (a == "-zip") o.zip = 1;
		OPT_I("-triangle", triangle)
		OPT_F("-light", light)
		else if (a == "-o") { if (!(v = need("-o"))) return false; o.out = v; }
		else if (a == "-interactive") o.interactive = 1;
		OPT_I("-jobs", jobs)
		OPT_I("-parallel_src", parallel_src)
		OPT_I("-fma", fma)
		OPT_I("-skip_discarded", skip_discarded)
		OPT_I("-max_tri", max_tri)
		OPT_I("-old_format", old_format)
		OPT_I("-quiet", quiet)
		else if (a == "-select") { if (!(v = need("-select"))) return false; o.lowest = std::string(v) == "low"; }
		else if (a == "-neighbors") { if (!(v = need("-neighbors"))) return false; o.hash = std::string(v) != "exact"; }
		else if (a == "-variant") { if (!(v = need("-variant"))) return false; o.v3 = std::string(v) == "v3"; }
		else { err = "unknown option " + a; return false; }
#undef OPT_F
#undef OPT_I
	}
	return true;
}
}  // namespace

extern "C" {

rs_ctx *rs_new(int argc, const char **argv)
{
	rs_ctx *c = new rs_ctx();
	if (!parse_opts(c->o, argc, argv, c->err)) {
		fprintf(stderr, "refsem: %s\n", c->err.c_str());
		delete c;
		return nullptr;
	}
	if (c->o.jobs > 0) omp_set_num_threads(c->o.jobs);
	c->cnt.clear();
	return c;
}
void rs_free(rs_ctx *c) { delete c; }
const char *rs_error(rs_ctx *c) { return c->err.c_str(); }
int rs_load_file(rs_ctx *c, const char *path) { return c->load_file(path); }
int rs_load_arrays(rs_ctx *c, int n, const float *p9, const int *f3, const int *b3) { return c->load(n, p9, f3, b3); }
int rs_build_tree(rs_ctx *c) { return c->build_tree(); }
int rs_init_radiosity(rs_ctx *c) { return c->init_radiosity(); }
int rs_prepartition(rs_ctx *c) { return c->prepartition(); }
int rs_iterate(rs_ctx *c, int max_shots, int *converged)
{
	int done = 0;
	*converged = 0;
	while (done < max_shots) {
		if (c->shoot() == 0) { *converged = 1; break; }
		done++;
	}
	return done;
}
int rs_run(rs_ctx *c, const char *prefix) { return c->run(prefix); }
int rs_write_fin(rs_ctx *c, const char *path) { return c->write_out(path); }
int rs_trinum(rs_ctx *c) { return c->trinum; }
int rs_record_size(void) { return (int)sizeof(Rec); }
int rs_dump(rs_ctx *c, void *rec, float *g)
{
	if (rec) memcpy(rec, c->T.data(), sizeof(Rec) * c->trinum);
	if (g) { memcpy(g, c->G0, 12); memcpy(g + 3, c->G1, 12); }
	return c->trinum;
}
int rs_tree_info(rs_ctx *c, int *out3, float *mgl)
{
	out3[0] = (int)c->nodes.size();
	out3[1] = (int)c->lists.size();
	out3[2] = c->grid_num;
	*mgl = c->min_grid;
	return 0;
}
int rs_rand(int n, int *out)
{
	GlibcRand r;
	for (int i = 0; i < n; i++) out[i] = r.next();
	return 0;
}
int rs_descend(rs_ctx *c, int gx, int gy, int gz)
{
	int g[3] = {gx, gy, gz};
	return c->descend(g, nullptr);
}
int rs_tree_arrays(rs_ctx *c, void *nodes, int *lists)
{
	if (nodes) memcpy(nodes, c->nodes.data(), sizeof(Node) * c->nodes.size());
	if (lists) memcpy(lists, c->lists.data(), sizeof(int) * c->lists.size());
	return 0;
}
int rs_raycast(rs_ctx *c, int n, const void *rays, int *hit, int *cells, int *tests)
{
	struct R { float p[3], v[3]; int ban, target; };
	const R *r = (const R *)rays;
#pragma omp parallel for schedule(dynamic, 256)
	for (int i = 0; i < n; i++) {
		long long k[RS_CNT_N] = {0};
		int h = c->o.fma ? c->ray_hit<true>(r[i].p, r[i].v, r[i].ban, k) : c->ray_hit<false>(r[i].p, r[i].v, r[i].ban, k);
		if (hit) hit[i] = h;
		if (cells) cells[i] = (int)k[RS_CNT_CELLS];
		if (tests) tests[i] = (int)k[RS_CNT_TESTS];
	}
	return 0;
}
int rs_ff_pairs(rs_ctx *c, int n, const int *pairs, float *out3)
{
	for (int i = 0; i < n; i++) {
		long long k[RS_CNT_N] = {0};
		const Rec &d = c->T[pairs[2 * i + 1]];
		if (c->o.fma) c->shade_ff<true>(c->make_source<true>(pairs[2 * i]), d, out3 + 3 * i, k);
		else c->shade_ff<false>(c->make_source<false>(pairs[2 * i]), d, out3 + 3 * i, k);
	}
	return 0;
}
// Gathering form factor of the explicit matrix: F[i][j] = CalFF from the
// centroid of receiver i (normal n_i, ban = its logical index) to the whole
// patch j (no AdaptCalFF halving), visible iff RayHitted returns j's logical
// index (shade.cpp:107-127, raycast.cpp:335-360); 0 on the diagonal and for
// logical slots.
static float centroid_ff(rs_ctx *c, int i, int j, int vis, long long *k)
{
	const Rec &d = c->T[i], &s = c->T[j];
	if (i == j || d.parent < 0 || s.parent < 0) return 0.0f;
	SrcPatch w;
	memcpy(w.c, s.c, 12);
	memcpy(w.n, s.n, 12);
	w.area = s.area;
	if (vis == RS_VIS_REFERENCE)
		return c->o.fma ? c->cal_ff<true>(w, s.parent, d.n, d.parent, d.c, k)
		                : c->cal_ff<false>(w, s.parent, d.n, d.parent, d.c, k);
	const float ff = c->o.fma ? c->ff_value<true>(w, d.n, d.c) : c->ff_value<false>(w, d.n, d.c);
	if (ff == 0.0f || vis == RS_VIS_NONE) return ff;
	// one test per unordered pair: the segment runs from the lower slot
	// index to the higher, so visibility is symmetric
	const Rec &lo = i < j ? d : s, &hi = i < j ? s : d;
	const float v[3] = {hi.c[0] - lo.c[0], hi.c[1] - lo.c[1], hi.c[2] - lo.c[2]};
	bool occ;
	if (vis == RS_VIS_BRUTE)
		occ = c->o.fma ? c->seg_brute<true>(lo.c, v, lo.parent, hi.parent) : c->seg_brute<false>(lo.c, v, lo.parent, hi.parent);
	else
		occ = c->o.fma ? c->seg_occluded<true>(lo.c, v, lo.parent, hi.parent, k)
		               : c->seg_occluded<false>(lo.c, v, lo.parent, hi.parent, k);
	return occ ? 0.0f : ff;
}

// Cost sample of InitRadiosity's exact neighbour pass as the reference runs
// it (rad.cpp:295-334: i serial, j in an OpenMP parallel for, every pair's
// vertex comparisons): rows i in [0, rows), matches counted, nothing
// written.  Returns the matches; the caller times the call.
long long rs_neighbor_rows(rs_ctx *c, int rows)
{
	const int n = c->trinum;
	long long matches = 0;
	for (int i = 0; i < rows && i < n; i++) {
		const Rec &tp = c->T[i];
#pragma omp parallel for reduction(+ : matches)
		for (int j = 0; j < n; j++) {
			if (j == i) continue;
			const Rec &tn = c->T[j];
			for (int v = 0; v < 3; v++)
				for (int m = 0; m < 3; m++) {
					if (!rs_ctx::veq(tp.p[v], tn.p[m])) continue;
					if (rs_ctx::veq(tp.p[(v + 1) % 3], tn.p[(m + 1) % 3]) ||
					    rs_ctx::veq(tp.p[(v + 1) % 3], tn.p[(m + 2) % 3]))
						matches++;
				}
		}
	}
	return matches;
}

int rs_ff_centroid(rs_ctx *c, int vis, int n, const int *pairs, float *out)
{
	c->build_fine();
#pragma omp parallel for schedule(dynamic, 256)
	for (int q = 0; q < n; q++) {
		long long k[RS_CNT_N] = {0};
		out[q] = centroid_ff(c, pairs[2 * q], pairs[2 * q + 1], vis, k);
	}
	return 0;
}

// Rows of the matrix for receivers rows[0..nr): out[r * n + j].  A row whose
// sum exceeds 1 is scaled by (float)(1 / sum); the sum is formed in double in
// the engine kernel's fixed order: 256 lane sums over j = lane (mod 256) in
// ascending j, an xor butterfly (offsets 32, 16, .., 1) inside each group of
// 64 lanes, then the four group sums left to right.
int rs_ff_rows(rs_ctx *c, int vis, int nr, const int *rows, float *out)
{
	const int n = c->trinum;
	c->build_fine();
#pragma omp parallel for schedule(dynamic, 1)
	for (int r = 0; r < nr; r++) {
		const int i = rows[r];
		float *row = out + (size_t)r * n;
		long long k[RS_CNT_N] = {0};
		double lane[256] = {0};
		for (int j = 0; j < n; j++) {
			row[j] = centroid_ff(c, i, j, vis, k);
			lane[j & 255] += (double)row[j];
		}
		double wsum[4];
		for (int w = 0; w < 4; w++) {
			double v[64];
			for (int l = 0; l < 64; l++) v[l] = lane[64 * w + l];
			for (int off = 32; off > 0; off >>= 1) {
				double nv[64];
				for (int l = 0; l < 64; l++) nv[l] = v[l] + v[l ^ off];
				memcpy(v, nv, sizeof v);
			}
			wsum[w] = v[0];
		}
		const double total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
		if (total > 1.0) {
			const float scale = (float)(1.0 / total);
			for (int j = 0; j < n; j++) row[j] *= scale;
		}
	}
	return 0;
}

int rs_shade_pairs(rs_ctx *c, int n, const int *pairs, float *out9)
{
	for (int i = 0; i < n; i++) {
		Rec d = c->T[pairs[2 * i + 1]];
		memset(d.accB, 0, sizeof d.accB);
		memset(d.deltaaccB, 0, sizeof d.deltaaccB);
		memset(d.deltaB, 0, sizeof d.deltaB);
		d.Frgb[0] = d.Frgb[1] = d.Frgb[2] = 255;
		float ff[3];
		rs_ff_pairs(c, 1, pairs + 2 * i, ff);
		const float dB[3] = {1000.f, 2000.f, 3000.f};
		rs_ctx::receive(dB, d, ff);
		memcpy(out9 + 9 * i, d.deltaaccB, 36);
	}
	return 0;
}
void rs_counters(rs_ctx *c, long long *out) { memcpy(out, c->cnt.v, sizeof c->cnt.v); }
void rs_reset_counters(rs_ctx *c) { c->cnt.clear(); }

}  // extern "C"
