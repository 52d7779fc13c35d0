// Device side of the engine (HIP, gfx950).  Host code sees plain C++ only.
#pragma once

#include <string>

#include "octree.h"
#include "scene.h"

namespace rgpu {

// Device counters, indices into DeviceEngine::counters().
enum {
	CNT_RAYS = 0,      // rays traced
	CNT_REF_RAYS = 1,  // rays the reference casts (adds AdaptCalFF's discarded full-source ray)
	CNT_CELLS = 2,     // leaf cells visited
	CNT_TESTS = 3,     // ray/triangle tests
	CNT_LEVELS = 4,    // octree levels the reference descends for those cells
	CNT_PRECISION = 5, // "Precision Exceeded" steps (raycast.cpp:344-350)
	CNT_STALLED = 6,   // rays the reference never finishes (hit = -2)
	CNT_RECEIVERS = 7, // receiver updates (Shade calls)
	CNT_N = 8
};

// One ray for the cast entry point: origin, direction, banned and target
// original (RayHitted's arguments, raycast.cpp:335).
enum { VIS_SEGMENT = 0, VIS_REFERENCE = 1 };

struct RayIn {
	float p[3], v[3];
	int ban, target;
};

class DeviceEngine {
public:
	~DeviceEngine();
	// Selects the current HIP device and creates the engine's stream.
	static DeviceEngine *create(std::string &err);
	// The calling thread's current HIP device, or -1 when there is none.
	static int current_device();
	// create() on `device`, made current on the calling thread first (for a
	// thread other than the caller's).
	static DeviceEngine *create_on(int device, std::string &err);
	// Loads the engine's device code (one empty launch on the stream, waited
	// for), so that the first upload does not; may run on another thread
	// while the host prepares.
	int warm(std::string &err);
	// Copies the scene (after pre-partition) and the octree to the device and
	// selects the first source.
	int upload(const HostScene &s, const FlatOctree &t, std::string &err);
	// Runs up to max_shots shooting iterations on the device.
	int shoot(int max_shots, int *shots, int *converged, std::string &err);
	// Copies accB, deltaaccB and deltaB back into the host pool.
	int download(HostScene &s, std::string &err);
	// BuildTree (raycast.cpp:168-233) over the originals of s on the device.
	int build_octree(const HostScene &s, FlatOctree &t, std::string &err);
	// InitRadiosity's neighbour search (rad.cpp:295-334) for slots [0,
	// s.trinum) on the device: writes s.pool[i].neighbor.
	int link_neighbors(HostScene &s, std::string &err);
	int cast(int n, const RayIn *rays, int *hit, int *cells, int *tests, std::string &err);
	// cast() through the device's static scratch instead of allocations
	// (at most 1 << 18 rays; calls that overlap on one GPU race).
	int cast_static(int n, const RayIn *rays, int *hit, int *cells, int *tests, std::string &err);
	int ff_pairs(int n, const int *pairs, float *ff3, std::string &err);
	void counters(long long *out);
	// Explicit form factors of slots [0, n) into the device matrix dF
	// (row-major, leading dimension ld): CalFF's value from each centroid to
	// each whole patch, times the visibility vis (VIS_SEGMENT: the segment
	// between the centroids is clear, k_ff_seg; VIS_REFERENCE: RayHitted from
	// the centroid returns the patch, k_ff_rows); rows scaled to sum <= 1.
	// VIS_SEGMENT builds t's segment index on first use.
	int form_factors(float *dF, int n, int ld, int vis, FlatOctree &t, const HostScene &s, std::string &err);
	double shoot_ms() const { return shoot_ms_; }  // device time of the shooting kernels so far
	double kernel_ms() const { return kernel_ms_; }  // device time of the fused shoot kernel alone
	long long kernel_launches() const { return kernel_launches_; }
	long long kernel_timed() const { return kernel_timed_; }
	double cast_ms() const { return cast_ms_; }  // device time of the last cast kernel
	bool timing = false;  // time each shoot kernel with events (for the roofline)
	bool graphs = false;  // untimed shooting replays captured graphs of kGraphIters iterations (-graph 1)

private:
	DeviceEngine() = default;
	struct Impl;
	Impl *d_ = nullptr;
	double shoot_ms_ = 0, kernel_ms_ = 0, cast_ms_ = 0;
	long long kernel_launches_ = 0, kernel_timed_ = 0;
};

}  // namespace rgpu
