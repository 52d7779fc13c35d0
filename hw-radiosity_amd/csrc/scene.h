// Host stages of the engine: input, InitRadiosity and pre-partition.
#pragma once

#include <string>
#include <vector>

#include "patch.h"

namespace rgpu {

struct HostScene {
	Options o;
	std::vector<Patch> pool;  // slots; [0, n_orig) are the originals
	int trinum = 0;
	int n_orig = 0;
	float G0[3] = {0, 0, 0}, G1[3] = {0, 0, 0};
	float min_area = 0, max_area = 0;
	bool warned_full = false;
	bool heap_was_empty = false;  // pre-partition found no splittable triangle

	// ReadTriangle (rad.cpp:47-111): derived fields and the grid corners.
	int set_triangles(int n, const float *p9, const int *frgb3, const int *brgb3, std::string &err);
	int load_tri(const char *path, std::string &err);
	// InitRadiosity (rad.cpp:259-356).
	void init_radiosity();
	// PrePartitionTriangles (shade.cpp:556-598).
	void prepartition();

	void log(const char *fmt, ...) const;
};

}  // namespace rgpu
