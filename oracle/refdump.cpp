// Dump harness for the reference's own functions.  Test infrastructure only:
// linked against the reference's unmodified translation units (see ref.mk) so
// that golden vectors come from the reference itself, at float-bit level.
//
//   refdump <mode> <out> [rad options] input.tri
//
// modes
//   read       after InitModelFromFile (ReadTriangle + BuildTree, rad.cpp:244-252)
//   init       + InitRadiosity (rad.cpp:259-356)
//   prepart    + PrePartitionTriangles (shade.cpp:556-598)
//   final      DoRadiosity (rad.cpp:491-523), i.e. the full run; writes <-o>.fin
//   rays:F     after prepart, RayHitted (raycast.cpp:335-360) on the rays of F
//              (records: float p[3], float v[3], int ban, int target) -> int32 hits
//   shade:F    after prepart, Shade (shade.cpp:435-554) for the (src, dst) slot
//              pairs of F on a zeroed copy of dst; writes its deltaaccB (9 floats)
//
// State dumps: int32 magic 'RDMP', int32 trinum, float G0[3], float G1[3],
// int32 sizeof(Triangle), then trinum raw Triangle records (rad.h:25-39).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rad.h"
#include "raycast.h"
#include "shade.h"

void InitModelFromFile(FILE *fp);
void InitRadiosity(void);
void DoRadiosity(const char fileName[]);
char *ProcessOption(int argc, char *argv[], FILE *&fin);

static void dump_state(const char *path)
{
	FILE *f = fopen(path, "wb");
	if (!f) { perror(path); exit(2); }
	int hdr[2] = {0x504d4452, trinum};
	fwrite(hdr, 4, 2, f);
	fwrite(G0, 4, 3, f);
	fwrite(G1, 4, 3, f);
	int sz = (int)sizeof(Triangle);
	fwrite(&sz, 4, 1, f);
	fwrite(TriStore, sizeof(Triangle), trinum, f);
	fclose(f);
}

static std::vector<char> slurp(const char *path)
{
	FILE *f = fopen(path, "rb");
	if (!f) { perror(path); exit(2); }
	std::vector<char> buf;
	char tmp[1 << 16];
	size_t n;
	while ((n = fread(tmp, 1, sizeof tmp, f)) > 0)
		buf.insert(buf.end(), tmp, tmp + n);
	fclose(f);
	return buf;
}

int main(int argc, char **argv)
{
	if (argc < 5) {
		fprintf(stderr, "usage: %s <mode> <out> [rad options] input.tri\n", argv[0]);
		return 2;
	}
	std::string mode = argv[1];
	const char *out = argv[2];
	std::vector<char *> rargv;
	rargv.push_back(argv[0]);
	for (int i = 3; i < argc; i++)
		rargv.push_back(argv[i]);
	FILE *fin = NULL;
	const char *prefix = ProcessOption((int)rargv.size(), rargv.data(), fin);
	if (!fin) { fprintf(stderr, "cannot open input\n"); return 2; }
	InitModelFromFile(fin);
	fclose(fin);
	if (mode == "read") { dump_state(out); return 0; }
	if (mode == "final") { DoRadiosity(prefix); dump_state(out); return 0; }
	InitRadiosity();
	if (mode == "init") { dump_state(out); return 0; }
	PrePartitionTriangles();
	if (mode == "prepart") { dump_state(out); return 0; }
	if (mode.compare(0, 5, "rays:") == 0) {
		std::vector<char> raw = slurp(mode.c_str() + 5);
		struct Ray { float p[3], v[3]; int ban, target; };
		size_t n = raw.size() / sizeof(Ray);
		const Ray *r = (const Ray *)raw.data();
		std::vector<int> hits(n);
		for (size_t i = 0; i < n; i++) {
			Vector p = {r[i].p[0], r[i].p[1], r[i].p[2]};
			Vector v = {r[i].v[0], r[i].v[1], r[i].v[2]};
			hits[i] = RayHitted(p, v, r[i].ban, r[i].target);
		}
		FILE *f = fopen(out, "wb");
		fwrite(hits.data(), 4, n, f);
		fclose(f);
		return 0;
	}
	if (mode.compare(0, 6, "shade:") == 0) {
		std::vector<char> raw = slurp(mode.c_str() + 6);
		size_t n = raw.size() / 8;
		const int *pr = (const int *)raw.data();
		std::vector<float> res(n * 9);
		for (size_t i = 0; i < n; i++) {
			TrianglePtr src = &TriStore[pr[2 * i]];
			Triangle dst = TriStore[pr[2 * i + 1]];
			float keep[3] = {src->deltaB[0], src->deltaB[1], src->deltaB[2]};
			src->deltaB[0] = 1000.f; src->deltaB[1] = 2000.f; src->deltaB[2] = 3000.f;
			memset(dst.accB, 0, sizeof dst.accB);
			memset(dst.deltaaccB, 0, sizeof dst.deltaaccB);
			memset(dst.deltaB, 0, sizeof dst.deltaB);
			dst.Frgb[0] = dst.Frgb[1] = dst.Frgb[2] = 255;
			Shade(src, src->parent, &dst, dst.parent, pr[2 * i + 1]);
			memcpy(&res[9 * i], dst.deltaaccB, 36);
			memcpy(src->deltaB, keep, sizeof keep);
		}
		FILE *f = fopen(out, "wb");
		fwrite(res.data(), 4, res.size(), f);
		fclose(f);
		return 0;
	}
	fprintf(stderr, "unknown mode %s\n", mode.c_str());
	return 2;
}
