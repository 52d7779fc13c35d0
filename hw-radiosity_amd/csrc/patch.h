// Host-side data model of the radiosity engine.
//
// A Patch is one pool slot.  Its field set and byte layout are the
// reference's Triangle record (include/rad.h:25-39), because the pool
// semantics of pre-partition copy, reuse and leave stale whole records
// (shade.cpp:357-367, 401-409) and because state is exchanged in that layout
// with callers and checkers.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>

namespace rgpu {

constexpr double kPi = 3.1415926535897932;  // rad.h:5
constexpr int kMaxLevel = 6;                 // rad.h:7
constexpr int kMaxTriPerNode = 10;           // rad.h:12

struct Patch {
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
	int parent;       // -1: logical (split) slot, else the original's index
	int neighbor[3];  // across edges p0p1, p1p2, p2p0
};
static_assert(sizeof(Patch) == 236, "Patch must keep the reference record layout");

inline bool emits(const Patch &t)  // isLightSource, rad.h:47-49
{
	return t.Brgb[0] == 255 || t.Brgb[1] == 255 || t.Brgb[2] == 255;
}

// The reference CLI options (rad.cpp:17-31, 527-528) and engine switches.
struct Options {
	int debug = 0;
	float adapt_area = 5.0f;
	float sample_area = -1.0f;
	float converge = 1200.0f;
	float delta_ff = 0.000005f;
	int write_cycle = 10;
	int zip = 0;
	int triangle = 1000000;
	float light = 1.0f;
	std::string out = "./test";
	int interactive = 0;
	int jobs = 2;
	int parallel_src = 1;
	// engine switches (not in the reference CLI)
	int fma = 1;            // reproduce GCC's FMA contraction sites (REF-FMA)
	int lowest = 0;         // tie-break on equal unshot energy: lowest index (V1)
	int max_tri = 1000000;  // the reference's pool size MaxTri (rad.h:10)
	int old_format = 0;     // per-vertex OLD_FORMAT output (rad.cpp:150-159)
	int quiet = 0;
	int batch = 32;         // shooting iterations queued per host check
	int kernel_timing = 0;  // time every shoot kernel with HIP events
	int graph = 0;          // replay shooting iterations from a captured HIP graph
	int neighbors_host = 0; // InitRadiosity's neighbour search on the host instead of the device
	int octree_host = 0;    // BuildTree on the host instead of the device
	int v3 = 0;             // the reference's V3 build: Delta-ff splitting live (config.h without DONT_PARTITION)
};

// Arithmetic at the reference compiler's contraction sites.  With F the
// products fuse exactly where GCC 11.4 -O2 -march=native fuses them
// (dot products vector.h:36,66,69; CrossProd vector.h:71-75; the ray
// points raycast.cpp:249,300,349); without F, source order.  Host code is
// compiled with -ffp-contract=off so nothing else fuses.
template <bool F> struct Arith {
	static float dot(const float *x, const float *y)
	{
		return F ? std::fma(x[2], y[2], std::fma(x[0], y[0], x[1] * y[1]))
		         : x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
	}
	static float det2(float a, float b, float c, float d)  // a*b - c*d
	{
		return F ? std::fma(a, b, -(c * d)) : a * b - c * d;
	}
	static void cross(const float *a, const float *b, float *out)
	{
		out[0] = det2(a[1], b[2], b[1], a[2]);
		out[1] = det2(a[2], b[0], b[2], a[0]);
		out[2] = det2(a[0], b[1], b[0], a[1]);
	}
	static float len(const float *v) { return std::sqrt(dot(v, v)); }
	static float len2(const float *v) { return dot(v, v); }
	static float lerp2(float a, float x, float b, float y)  // a*x + b*y (raycast.cpp:97)
	{
		return F ? std::fma(a, x, b * y) : a * x + b * y;
	}
};

// CalCenter (vector.h:53-56): float sum, /3.0 in double.
inline void centroid(const float p[3][3], float *c)
{
	for (int k = 0; k < 3; k++) c[k] = (float)((double)(p[0][k] + p[1][k] + p[2][k]) / 3.0);
}
// AverageVector (vector.h:27-30): float sum, /2.0 in double.
inline void midpoint(const float *a, const float *b, float *m)
{
	for (int k = 0; k < 3; k++) m[k] = (float)((double)(a[k] + b[k]) / 2.0);
}

bool parse_option(Options &o, int argc, const char *const *argv, int &i, std::string &err);

}  // namespace rgpu
