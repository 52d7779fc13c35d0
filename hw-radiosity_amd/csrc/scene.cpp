// Host stages: ReadTriangle, InitRadiosity (with a sort-based neighbour
// search) and PrePartitionTriangles.  These are index-order
// dependent steps of the reference; they run once per solve on the host and
// the result is uploaded to the device.
#include "scene.h"

#include <parallel/algorithm>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <set>
#include <utility>

namespace rgpu {

void HostScene::log(const char *fmt, ...) const
{
	if (o.quiet) return;
	va_list ap;
	va_start(ap, fmt);
	vfprintf(stderr, fmt, ap);
	va_end(ap);
}

namespace {

template <bool F> void derive_fields(Patch &t)
{
	using A = Arith<F>;
	centroid(t.p, t.c);
	float e1[3], e2[3];
	for (int k = 0; k < 3; k++) {
		e1[k] = t.p[1][k] - t.p[0][k];
		e2[k] = t.p[2][k] - t.p[0][k];
	}
	A::cross(e1, e2, t.n);
	t.area = A::len(t.n) / 2;  // CalArea, vector.h:58-62
	const float l = A::len(t.n);
	t.n[0] /= l;  // normalize, vector.h:77-80
	t.n[1] /= l;
	t.n[2] /= l;
	// side planes with inward normals (rad.cpp:81-88)
	for (int i = 0; i < 3; i++) {
		const float *a = t.p[i], *b = t.p[(i + 1) % 3];
		const float e[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
		float s[3];
		A::cross(t.n, e, s);
		t.se[i][0] = s[0];
		t.se[i][1] = s[1];
		t.se[i][2] = s[2];
		t.se[i][3] = -A::dot(a, s);
	}
}

}  // namespace

float cull_radius(const Patch &t, bool fma, float max_amp)
{
	constexpr float kInf = std::numeric_limits<float>::infinity();
	Patch d;
	std::memcpy(d.p, t.p, sizeof d.p);
	if (fma) derive_fields<true>(d);
	else derive_fields<false>(d);
	// the record's planes must be exactly those of its own vertices
	if (std::memcmp(d.c, t.c, sizeof d.c) || std::memcmp(d.n, t.n, sizeof d.n) || std::memcmp(d.se, t.se, sizeof d.se))
		return kInf;
	double e[3][3], len[3];
	for (int i = 0; i < 3; i++) {
		for (int k = 0; k < 3; k++) e[i][k] = (double)t.p[(i + 1) % 3][k] - (double)t.p[i][k];
		len[i] = std::sqrt(e[i][0] * e[i][0] + e[i][1] * e[i][1] + e[i][2] * e[i][2]);
		if (!(len[i] > 0.0) || !std::isfinite(len[i])) return kInf;
	}
	// smallest angle: at vertex i between -e[i-1] and e[i]
	double min_angle = 4.0;
	for (int i = 0; i < 3; i++) {
		const double *a = e[(i + 2) % 3], *b = e[i];
		double c = -(a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / (len[(i + 2) % 3] * len[i]);
		c = std::max(-1.0, std::min(1.0, c));
		min_angle = std::min(min_angle, std::acos(c));
	}
	if (!(min_angle > 0.0) || 1.0 / std::sin(min_angle / 2.0) > (double)max_amp) return kInf;
	double r = 0.0;
	for (int v = 0; v < 3; v++) {
		double q = 0.0;
		for (int k = 0; k < 3; k++) q += ((double)t.p[v][k] - t.c[k]) * ((double)t.p[v][k] - t.c[k]);
		r = std::max(r, std::sqrt(q));
	}
	const float rf = (float)r;
	return std::isfinite(rf) ? std::nextafter(rf, kInf) : kInf;
}

void pack_points(const HostScene &s, float *out12)
{
#pragma omp parallel for schedule(static)
	for (int i = 0; i < s.trinum; i++) {
		std::memcpy(out12 + 12 * (size_t)i, s.pool[i].p, 9 * sizeof(float));
		std::memcpy(out12 + 12 * (size_t)i + 9, s.pool[i].n, 3 * sizeof(float));
	}
}

void unpack_neighbors(HostScene &s, const int *nb3)
{
#pragma omp parallel for schedule(static)
	for (int i = 0; i < s.trinum; i++) std::memcpy(s.pool[i].neighbor, nb3 + 3 * (size_t)i, 3 * sizeof(int));
}

float cull_radii(const HostScene &s, float *radius)
{
	double cl = 0.0;
#pragma omp parallel for schedule(static) reduction(max : cl)
	for (int i = 0; i < s.n_orig; i++) {
		const Patch &p = s.pool[i];
		radius[i] = cull_radius(p, s.o.fma != 0, 16.0f);
		if (std::isfinite(radius[i]))
			cl = std::max(cl, std::fabs((double)p.c[0]) + std::fabs((double)p.c[1]) + std::fabs((double)p.c[2]));
	}
	return std::nextafter((float)cl, std::numeric_limits<float>::infinity());
}

int HostScene::set_triangles(int n, const float *p9, const int *f3, const int *b3, std::string &err)
{
	if (n < 0) { err = "negative triangle count"; return -1; }
	const long cap = std::max<long>(n, std::min(o.max_tri, o.triangle)) + 16;
	if (!pool.reset((size_t)cap)) { err = "cannot allocate the triangle pool"; return -1; }
	float lo[3] = {1000000.0f, 1000000.0f, 1000000.0f};
	float hi[3] = {-1000000.0f, -1000000.0f, -1000000.0f};
#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		Patch &t = pool[i];
		std::memcpy(t.p, p9 + 9 * (size_t)i, sizeof t.p);
		std::memcpy(t.Frgb, f3 + 3 * (size_t)i, sizeof t.Frgb);
		std::memcpy(t.Brgb, b3 + 3 * (size_t)i, sizeof t.Brgb);
		if (o.fma) derive_fields<true>(t);
		else derive_fields<false>(t);
		t.parent = i;
		t.neighbor[0] = t.neighbor[1] = t.neighbor[2] = -1;
	}
	for (int i = 0; i < n; i++) {  // in index order, as the reference's min/max
		const Patch &t = pool[i];
		for (int v = 0; v < 3; v++)
			for (int k = 0; k < 3; k++) {
				const float x = t.p[v][k];
				if (x < lo[k]) lo[k] = x;   // std::min(lo, x)
				if (hi[k] < x) hi[k] = x;   // std::max(hi, x)
			}
	}
	trinum = n_orig = n;
	float m = 0;
	for (int k = 0; k < 3; k++) {
		const float d = hi[k] - lo[k];
		if (m < d) m = d;
	}
	// grid corners padded by pi (rad.cpp:105-110); G0 + m is a float sum
	for (int k = 0; k < 3; k++) G0[k] = (float)((double)lo[k] - kPi);
	for (int k = 0; k < 3; k++) G1[k] = (float)((double)(G0[k] + m) + 2 * kPi);
	log("[INFO] Read #Triangle %d from input file\n", trinum);
	return 0;
}

int HostScene::load_tri(const char *path, std::string &err)
{
	FILE *f = std::fopen(path, "r");
	if (!f) { err = std::string("File open error::") + path; return -1; }
	std::vector<float> p;
	std::vector<int> fr, br;
	char word[256];
	// keyword, 3 Frgb ints, 3 Brgb ints, 3 x 3 floats (rad.cpp:59-75)
	while (std::fscanf(f, "%255s", word) == 1) {
		int iv[6];
		float fv[9];
		if (std::fscanf(f, "%d%d%d%d%d%d", iv, iv + 1, iv + 2, iv + 3, iv + 4, iv + 5) != 6) break;
		for (int k = 0; k < 9; k++)
			if (std::fscanf(f, "%f", fv + k) != 1) {
				std::fclose(f);
				err = std::string("malformed triangle record in ") + path;
				return -1;
			}
		fr.insert(fr.end(), iv, iv + 3);
		br.insert(br.end(), iv + 3, iv + 6);
		p.insert(p.end(), fv, fv + 9);
	}
	std::fclose(f);
	return set_triangles((int)(fr.size() / 3), p.data(), fr.data(), br.data(), err);
}

// ---------------------------------------------------------------------------
// Neighbour search.  The reference (rad.cpp:295-334) scans all pairs in the
// order (i, j, v, n) and, where vertex v of i equals vertex n of j, tests the
// next vertex in the same and the opposite direction; a coplanar enough pair
// (CosTheta > 0.85) writes i.neighbor[v] = j and j.neighbor[n or n+2] = i.
// Later writes overwrite earlier ones.  Here the equal-vertex pairs come
// from one sort of all vertices, and each (slot, edge) keeps the write with
// the greatest (i, j, v, n).
namespace {

struct VertexRef {
	uint32_t x, y, z;  // coordinate bits, -0 folded onto +0
	int code;          // 3 * slot + corner
	bool operator<(const VertexRef &o) const
	{
		if (x != o.x) return x < o.x;
		if (y != o.y) return y < o.y;
		if (z != o.z) return z < o.z;
		return code < o.code;
	}
	bool same_point(const VertexRef &o) const { return x == o.x && y == o.y && z == o.z; }
};

bool point_key(const float *p, VertexRef &r)
{
	if (std::isnan(p[0]) || std::isnan(p[1]) || std::isnan(p[2])) return false;  // never ==
	const float q[3] = {p[0] + 0.0f, p[1] + 0.0f, p[2] + 0.0f};
	std::memcpy(&r.x, &q[0], 4);
	std::memcpy(&r.y, &q[1], 4);
	std::memcpy(&r.z, &q[2], 4);
	return true;
}

inline bool same_vertex(const float *a, const float *b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }

// Keep the greatest order seen for one (slot, edge): 0 means none (an order
// is never 0: it holds two different slots).
inline void keep_max(uint64_t *cell, uint64_t order)
{
	uint64_t cur = __atomic_load_n(cell, __ATOMIC_RELAXED);
	while (cur < order &&
	       !__atomic_compare_exchange_n(cell, &cur, order, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
	}
}

template <bool F> void link_neighbors(PatchPool &T, int n)
{
	using A = Arith<F>;
	std::vector<VertexRef> refs((size_t)n * 3);
	std::vector<uint8_t> ok((size_t)n * 3);
#pragma omp parallel for schedule(static)
	for (int j = 0; j < n; j++)
		for (int c = 0; c < 3; c++) {
			VertexRef &r = refs[3 * (size_t)j + c];
			ok[3 * (size_t)j + c] = point_key(T[j].p[c], r);
			r.code = 3 * j + c;
		}
	size_t m = 0;  // drop NaN corners (never equal to anything)
	for (size_t k = 0; k < refs.size(); k++)
		if (ok[k]) refs[m++] = refs[k];
	refs.resize(m);
	__gnu_parallel::sort(refs.begin(), refs.end());
	std::vector<size_t> group;  // starts of the runs of one point
	for (size_t a = 0; a < refs.size(); a++)
		if (a == 0 || !refs[a].same_point(refs[a - 1])) group.push_back(a);
	group.push_back(refs.size());
	// order = (i, j, v, n) packed so that a later write of the reference's
	// loop is a greater number: i in bits 34.., j in bits 4..33, 3v + n below
	std::vector<uint64_t> best((size_t)n * 3, 0);
	const long ng = (long)group.size() - 1;
#pragma omp parallel for schedule(dynamic, 512)
	for (long g = 0; g < ng; g++) {
		const size_t a = group[g], b = group[g + 1];
		// every ordered pair (i, v) x (j, n) with i != j among refs[a, b)
		for (size_t u = a; u < b; u++) {
			const int i = refs[u].code / 3, v = refs[u].code % 3;
			const Patch &tp = T[i];
			for (size_t w = a; w < b; w++) {
				const int j = refs[w].code / 3, c = refs[w].code % 3;
				if (j == i) continue;
				const Patch &tn = T[j];
				const uint64_t order = ((uint64_t)i << 34) | ((uint64_t)j << 4) | (uint64_t)(3 * v + c);
				int other_edge;
				if (same_vertex(tp.p[(v + 1) % 3], tn.p[(c + 1) % 3])) other_edge = c;
				else if (same_vertex(tp.p[(v + 1) % 3], tn.p[(c + 2) % 3])) other_edge = (c + 2) % 3;
				else continue;
				const float cosine = A::dot(tp.n, tn.n) / A::len(tp.n) / A::len(tn.n);
				if (!((double)cosine > 0.85)) continue;
				keep_max(&best[(size_t)i * 3 + v], order);             // i.neighbor[v] = j
				keep_max(&best[(size_t)j * 3 + other_edge], order);    // j.neighbor[..] = i
			}
		}
	}
#pragma omp parallel for schedule(static)
	for (int s = 0; s < n; s++)
		for (int e = 0; e < 3; e++) {
			const uint64_t order = best[(size_t)s * 3 + e];
			if (!order) continue;
			const int i = (int)(order >> 34), j = (int)((order >> 4) & 0x3fffffffu);
			T[s].neighbor[e] = s == i ? j : i;
		}
}

}  // namespace

void HostScene::init_radiosity(bool link_on_host)
{
	float amin = 100000000.0f, amax = 0.0f;
	for (int i = 0; i < trinum; i++) {
		Patch &t = pool[i];
		if (t.area < amin) amin = t.area;
		if (amax < t.area) amax = t.area;
		t.deltaB[0] = t.deltaB[1] = t.deltaB[2] = 0.0f;
		if (emits(t)) {
			// emission = Brgb * -light; unshot energy 100x the emission
			for (int v = 0; v < 3; v++) {
				for (int c = 0; c < 3; c++) t.accB[v][c] = (float)t.Brgb[c] * o.light;
				for (int c = 0; c < 3; c++) t.deltaB[c] += 100 * t.accB[v][c] / 3;
			}
		} else {
			std::memset(t.accB, 0, sizeof t.accB);
		}
	}
	min_area = amin;
	max_area = amax;
	if (link_on_host) {
		if (o.fma) link_neighbors<true>(pool, trinum);
		else link_neighbors<false>(pool, trinum);
	}
	if (o.sample_area < 0.0f) o.sample_area = (float)((double)amax / 100.0);
	if (!o.quiet) {
		std::printf("[INFO] Max Patch Area = %f\n", amax);
		std::printf("[INFO] Min Patch Area = %f\n", amin);
		std::printf("[INFO] Config Table\n");
		std::printf("       -debug          %i\n", o.debug);
		std::printf("       -adapt_area     %f\n", o.adapt_area);
		std::printf("       -sample_area    %f\n", o.sample_area);
		std::printf("       -converge       %f\n", o.converge);
		std::printf("       -delta_ff       %f\n", o.delta_ff);
		std::printf("       -write_cycle    %d\n", o.write_cycle);
		std::printf("       -triangle       %d\n", o.triangle);
		std::printf("       -light          %f\n", o.light);
		std::printf("       -interactive    %d\n", o.interactive);
		std::printf("       -jobs           %d\n", o.jobs);
		std::printf("       -parallel_src   %d\n", o.parallel_src);
		std::printf("[INFO] Complete initialization\n");
		std::fflush(stdout);
	}
}

// ---------------------------------------------------------------------------
// Pre-partition.  Repeatedly bisect the largest non-emitting triangle at its
// longest edge together with the neighbour across that edge, until the
// largest remaining area is below 1, a split is refused, or the pool fills.
// The queue is an ordered set of (area, slot): entries are not refreshed
// when a slot changes, so an entry can outlive its triangle; popping such an
// entry re-splits whatever the slot now holds, exactly as the reference does
// (shade.cpp:561-595).
namespace {

struct Splitter {
	HostScene &S;
	explicit Splitter(HostScene &s) : S(s) {}
	PatchPool &T() { return S.pool; }

	bool pool_full(int need) const { return S.trinum + need >= std::min(S.o.max_tri, S.o.triangle); }
	int take_slot()
	{
		if (S.trinum == S.o.max_tri || S.trinum == S.o.triangle) {
			if (!S.warned_full) {
				S.log("\n**** Max Triangle exceed : %i ***\n", S.trinum);
				S.warned_full = true;
			}
			return -1;
		}
		return S.trinum++;
	}
	void give_back_slot() { S.trinum--; }
	void relink(int who, int from, int to)
	{
		if (who < 0) return;
		for (int e = 0; e < 3; e++)
			if (T()[who].neighbor[e] == from) T()[who].neighbor[e] = to;
	}
	// PartitionDestination, 2-way (shade.cpp:171-211): halves of slot s at
	// `edge` written into slots a and b.  Fields it does not set (se, Brgb,
	// deltaaccB, neighbor) keep whatever the slots held.
	void bisect(int s, int a, int b, int edge)
	{
		const Patch src = T()[s];
		Patch &A = T()[a], &B = T()[b];
		const int i0 = edge, i1 = (edge + 1) % 3, i2 = (edge + 2) % 3;
		float m[3], ma[3];
		midpoint(src.p[i0], src.p[i1], m);
		midpoint(src.accB[i0], src.accB[i1], ma);
		const float *pa[3] = {m, src.p[i2], src.p[i0]}, *pb[3] = {m, src.p[i1], src.p[i2]};
		const float *aa[3] = {ma, src.accB[i2], src.accB[i0]}, *ab[3] = {ma, src.accB[i1], src.accB[i2]};
		for (int v = 0; v < 3; v++) {
			std::memcpy(A.p[v], pa[v], 12);
			std::memcpy(B.p[v], pb[v], 12);
			std::memcpy(A.accB[v], aa[v], 12);
			std::memcpy(B.accB[v], ab[v], 12);
		}
		for (Patch *h : {&A, &B}) {
			std::memcpy(h->n, src.n, 12);
			std::memcpy(h->Frgb, src.Frgb, 12);
			std::memcpy(h->deltaB, src.deltaB, 12);
			centroid(h->p, h->c);
			h->area = (float)((double)src.area / 2.0);
			h->parent = src.parent;
		}
	}
	template <bool F> int longest_edge(const Patch &t, float &len) const
	{
		int edge = 0;
		len = 0.0f;
		for (int v = 0; v < 3; v++) {
			float d[3];
			for (int k = 0; k < 3; k++) d[k] = t.p[(v + 1) % 3][k] - t.p[v][k];
			const float l = Arith<F>::len2(d);
			if (len < l) { len = l; edge = v; }
		}
		return edge;
	}
	// PartitionTriangleAndStore (shade.cpp:311-431); out: the new/reused
	// slots t1, t2 (the split triangle's halves), t3, t4 (the neighbour's).
	template <bool F> int split(int s, int out[4])
	{
		out[0] = out[1] = out[2] = out[3] = -1;
		if (pool_full(4) || (double)T()[s].area < 1e-8) return -1;
		float maxlen;
		const int edge = longest_edge<F>(T()[s], maxlen);
		if (maxlen < 1) return -1;
		const int nb = T()[s].neighbor[edge];
		if (nb >= 0 && (double)T()[nb].area < 1e-1) return -1;
		int t1 = take_slot(), t2 = take_slot();
		bisect(s, t1, t2, edge);
		int n1 = T()[s].neighbor[(edge + 2) % 3], n2 = T()[s].neighbor[(edge + 1) % 3];
		T()[t1].neighbor[1] = n1;
		T()[t2].neighbor[1] = n2;
		T()[t1].neighbor[2] = T()[t2].neighbor[0] = -1;
		int replaced = 0;
		if (T()[s].parent == s) {
			T()[s].parent = -1;  // an original becomes logical
		} else {
			T()[s] = T()[t2];  // a child is replaced in place by its second half
			t2 = s;
			give_back_slot();
			replaced = 1;
		}
		relink(n1, s, t1);
		relink(n2, s, t2);
		T()[t1].neighbor[0] = t2;
		T()[t2].neighbor[2] = t1;
		out[0] = t1;
		out[1] = t2;
		if (nb < 0) return replaced;
		int t3 = take_slot(), t4 = take_slot();
		int ne = 0;
		for (int e = 0; e < 3; e++)
			if (T()[nb].neighbor[e] == s) ne = e;
		if (nb < s)  // undo its last received increment (shade.cpp:391-396)
			for (int v = 0; v < 3; v++)
				for (int c = 0; c < 3; c++) T()[nb].accB[v][c] -= T()[nb].deltaaccB[v][c];
		bisect(nb, t3, t4, ne);
		n1 = T()[nb].neighbor[(ne + 2) % 3];
		n2 = T()[nb].neighbor[(ne + 1) % 3];
		T()[t3].neighbor[1] = n1;
		T()[t4].neighbor[1] = n2;
		if (T()[nb].parent == nb) {
			T()[nb].parent = -1;
		} else {
			T()[nb] = T()[t4];
			t4 = nb;
			give_back_slot();
		}
		relink(n1, nb, t3);
		relink(n2, nb, t4);
		T()[t3].neighbor[0] = t4;
		T()[t4].neighbor[2] = t3;
		// shade.cpp:416 compares vertex-array addresses: equal only when the
		// neighbour is the triangle itself, split at the same edge
		if (nb == s && edge == ne) {
			T()[t1].neighbor[2] = t3;
			T()[t3].neighbor[2] = t1;
			T()[t2].neighbor[0] = t4;
			T()[t4].neighbor[0] = t2;
		} else {
			T()[t1].neighbor[2] = t4;
			T()[t4].neighbor[0] = t1;
			T()[t2].neighbor[0] = t3;
			T()[t3].neighbor[2] = t2;
		}
		out[2] = t3;
		out[3] = t4;
		return replaced;
	}

	// The children of a split light carry its emission (shade.cpp:580-590).
	void inherit_light(int s, const int *made)
	{
		if (!emits(T()[s])) return;
		const Patch &src = T()[s];
		for (int k = 0; k < 4; k++) {
			if (made[k] < 0) continue;
			Patch &h = T()[made[k]];
			std::memcpy(h.Brgb, src.Brgb, 12);
			for (int v = 0; v < 3; v++) {
				std::memcpy(h.accB[v], src.accB[v], 12);
				h.deltaB[v] = src.deltaB[v];
			}
		}
	}

	// The queue as the reference keeps it: a std::set of (area, slot),
	// largest first, entries left stale when a neighbour split changes a slot
	// (shade.cpp:561-595).
	template <bool F> void run_set()
	{
		std::set<std::pair<float, int>> queue;
		for (int i = 0; i < S.trinum; i++)
			if (!(T()[i].area < 1.f) && !emits(T()[i])) queue.insert({T()[i].area, i});
		if (queue.empty()) {
			// the reference dereferences the empty set here and crashes
			S.heap_was_empty = true;
			return;
		}
		for (;;) {
			if (queue.empty()) break;
			const std::pair<float, int> top = *queue.rbegin();
			queue.erase(top);
			const int s = top.second;
			if (T()[s].area < 1.0f) break;
			int made[4];
			if (split<F>(s, made) < 0) break;
			inherit_light(s, made);
			for (int k = 0; k < 4; k++)
				if (made[k] >= 0) queue.insert({T()[made[k]].area, made[k]});
		}
	}

	// The same queue as a binary max-heap of 64-bit keys (order-preserving
	// area bits, then the slot), several times faster at a million slots.
	// A set ignores the insertion of a pair it holds; the heap keeps both
	// copies, which are equal keys and so leave it together when they reach
	// the top: popping every copy of the top key is the set's single erase.
	// Areas only halve, so without a NaN among the originals' (whose set
	// order is the comparator's, not a total order) the orders agree.
	static uint64_t queue_key(float area, int slot)
	{
		uint32_t b;
		const float a = area + 0.0f;  // -0 and +0 compare equal
		std::memcpy(&b, &a, 4);
		b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
		return ((uint64_t)b << 32) | (uint32_t)slot;
	}
	template <bool F> void run_heap()
	{
		std::vector<uint64_t> heap;
		for (int i = 0; i < S.trinum; i++)
			if (!(T()[i].area < 1.f) && !emits(T()[i])) heap.push_back(queue_key(T()[i].area, i));
		if (heap.empty()) {
			S.heap_was_empty = true;
			return;
		}
		std::make_heap(heap.begin(), heap.end());
		while (!heap.empty()) {
			const uint64_t top = heap.front();
			do {
				std::pop_heap(heap.begin(), heap.end());
				heap.pop_back();
			} while (!heap.empty() && heap.front() == top);
			const int s = (int)(uint32_t)top;
			if (T()[s].area < 1.0f) break;
			int made[4];
			if (split<F>(s, made) < 0) break;
			inherit_light(s, made);
			for (int k = 0; k < 4; k++)
				if (made[k] >= 0) {
					heap.push_back(queue_key(T()[made[k]].area, made[k]));
					std::push_heap(heap.begin(), heap.end());
				}
		}
	}

	template <bool F> void run()
	{
		for (int i = 0; i < S.trinum; i++)
			if (std::isnan(T()[i].area)) {
				run_set<F>();
				return;
			}
		run_heap<F>();
	}
};

}  // namespace

void HostScene::prepartition()
{
	log("[DEBUG] Before PrePartitionTriangle %d\n", trinum);
	Splitter sp(*this);
	if (o.fma) sp.run<true>();
	else sp.run<false>();
	if (heap_was_empty) log("[WARN] no triangle qualifies for pre-partition (the reference crashes here)\n");
	log("[DEBUG] After PrePartitionTriangle %d\n", trinum);
}

}  // namespace rgpu
