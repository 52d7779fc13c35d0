// Host stages of the engine: input, InitRadiosity and pre-partition.
#pragma once

#include <cstdlib>
#include <string>
#include <vector>

#include "patch.h"

namespace rgpu {

// The slot pool: zeroed memory from calloc, so a capacity of a million
// slots (236 MB) costs only the pages that pre-partition actually touches.
class PatchPool {
public:
	PatchPool() = default;
	PatchPool(const PatchPool &) = delete;
	PatchPool &operator=(const PatchPool &) = delete;
	~PatchPool() { std::free(p_); }
	// count zeroed slots; false when the allocation fails
	bool reset(size_t count)
	{
		std::free(p_);
		p_ = static_cast<Patch *>(std::calloc(count, sizeof(Patch)));
		n_ = p_ ? count : 0;
		return p_ != nullptr;
	}
	Patch &operator[](size_t i) { return p_[i]; }
	const Patch &operator[](size_t i) const { return p_[i]; }
	Patch *data() { return p_; }
	const Patch *data() const { return p_; }
	size_t size() const { return n_; }

private:
	Patch *p_ = nullptr;
	size_t n_ = 0;
};

struct HostScene {
	Options o;
	PatchPool pool;  // slots; [0, n_orig) are the originals
	int trinum = 0;
	int n_orig = 0;
	float G0[3] = {0, 0, 0}, G1[3] = {0, 0, 0};
	float min_area = 0, max_area = 0;
	bool warned_full = false;
	bool heap_was_empty = false;  // pre-partition found no splittable triangle

	// ReadTriangle (rad.cpp:47-111): derived fields and the grid corners.
	int set_triangles(int n, const float *p9, const int *frgb3, const int *brgb3, std::string &err);
	int load_tri(const char *path, std::string &err);
	// InitRadiosity (rad.cpp:259-356); the neighbour search (rad.cpp:295-334)
	// here when link_on_host, else the caller runs it on the device.
	void init_radiosity(bool link_on_host = true);
	// PrePartitionTriangles (shade.cpp:556-598).
	void prepartition();

	void log(const char *fmt, ...) const;
};

// Radius of the sphere about t's centroid that holds every point RayHitted
// can accept on t, for the ray-tracing kernel's exact early rejection: the
// largest vertex distance from the centroid when t's centroid, normal and
// side planes are exactly those its own vertices give (ReadTriangle's
// arithmetic) and no angle is so small that rounding in the side-plane tests
// moves its corner by more than max_amp times their fuzz
// (1 / sin(angle / 2) <= max_amp); +infinity otherwise (never rejected).
float cull_radius(const Patch &t, bool fma, float max_amp);
// cull_radius (max_amp 16) of every original of s into radius[0, n_orig);
// returns a bound on |c|_1 over the originals with a finite radius.
float cull_radii(const HostScene &s, float *radius);
// Slots [0, s.trinum) as 12 floats each (p[3][3], n[3]), and the inverse
// for the neighbour links (3 ints each), for the device neighbour search.
void pack_points(const HostScene &s, float *out12);
void unpack_neighbors(HostScene &s, const int *nb3);

}  // namespace rgpu
