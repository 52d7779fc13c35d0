// host_check: the engine's host stages on one input, for sanitizer builds
// (make -C hw-radiosity_amd -f sanitize.mk: AddressSanitizer + UndefinedBehavior-
// Sanitizer, host code only, no device code linked).  Reads a .tri the way
// rad_load_tri does, runs InitRadiosity with the host neighbour search,
// pre-partition and the host octree, writes the .fin, then converts that
// .fin -- or, with -fin, a given .fin -- to the viewer's JSON.
//   host_check [-triangle N] [-fma 0|1] input.tri out_prefix
//   host_check -fin input.fin out.json
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "fin.h"
#include "octree.h"
#include "scene.h"
#include "viewer_json.h"

using namespace rgpu;

int main(int argc, char **argv)
{
	if (argc == 4 && !std::strcmp(argv[1], "-fin")) {
		int n = 0, tiny = 0;
		std::string err;
		const int rc = fin_to_json(argv[2], argv[3], n, tiny, err);
		std::printf("fin_to_json rc %d triangles %d %s\n", rc, n, err.c_str());
		return 0;
	}
	HostScene s;
	s.o.quiet = 1;
	int i = 1;
	for (; i + 1 < argc && argv[i][0] == '-'; i += 2) {
		if (!std::strcmp(argv[i], "-triangle")) s.o.triangle = std::atoi(argv[i + 1]);
		else if (!std::strcmp(argv[i], "-fma")) s.o.fma = std::atoi(argv[i + 1]);
	}
	if (i + 2 != argc) {
		std::fprintf(stderr, "usage: %s [-triangle N] [-fma 0|1] input.tri out_prefix | -fin in.fin out.json\n", argv[0]);
		return 2;
	}
	std::string err;
	if (s.load_tri(argv[i], err)) {
		std::printf("load_tri: %s\n", err.c_str());
		return 0;
	}
	FlatOctree t = build_octree(s.pool.data(), s.n_orig, s.G0, s.G1, s.o.fma != 0, 0);
	if (t.gt <= 0) {
		std::printf("octree: grid not tabulated (bounds not finite)\n");
		return 0;
	}
	s.init_radiosity(true);
	s.prepartition();
	FinWriter w;
	const std::string fin = std::string(argv[i + 1]) + ".fin";
	if (w.write(fin, s.pool.data(), s.trinum, false, err)) std::printf("write: %s\n", err.c_str());
	int n = 0, tiny = 0;
	const int rc = fin_to_json(fin, std::string(argv[i + 1]) + ".json", n, tiny, err);
	std::printf("triangles %d slots %d nodes %d lists %d json rc %d records %d\n", s.n_orig, s.trinum, t.n_nodes,
	            t.n_list, rc, n);
	return 0;
}
