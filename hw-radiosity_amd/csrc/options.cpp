// Option parsing (the reference's ProcessOption, rad.cpp:525-631, without
// the input-file rule) plus the engine switches.
#include <cstdio>
#include <cstring>

#include "patch.h"

namespace rgpu {

bool parse_option(Options &o, int argc, const char *const *argv, int &i, std::string &err)
{
	const std::string a = argv[i];
	auto value = [&](const char *&out) -> bool {
		if (i + 1 >= argc) {
			err = "missing value for " + a;
			return false;
		}
		out = argv[++i];
		return true;
	};
	auto as_float = [&](float &f) -> bool {
		const char *v;
		if (!value(v)) return false;
		if (std::sscanf(v, "%f", &f) != 1) { err = "bad value for " + a + ": " + v; return false; }
		return true;
	};
	auto as_int = [&](int &x) -> bool {
		const char *v;
		if (!value(v)) return false;
		if (std::sscanf(v, "%d", &x) != 1) { err = "bad value for " + a + ": " + v; return false; }
		return true;
	};
	const char *v = nullptr;
	if (a == "-debug") return as_int(o.debug);
	if (a == "-adapt_area") return as_float(o.adapt_area);
	if (a == "-sample_area") return as_float(o.sample_area);
	if (a == "-converge") return as_float(o.converge);
	if (a == "-delta_ff") return as_float(o.delta_ff);
	if (a == "-write_cycle") return as_int(o.write_cycle);
	if (a == "-zip") { o.zip = 1; return true; }
	if (a == "-triangle") return as_int(o.triangle);
	if (a == "-light") return as_float(o.light);
	if (a == "-o") { if (!value(v)) return false; o.out = v; return true; }
	if (a == "-interactive") { o.interactive = 1; return true; }
	if (a == "-jobs") return as_int(o.jobs);
	if (a == "-parallel_src") {
		// listCanditate (rad.cpp:400) takes min(ceil(|C|/5), value) sources:
		// 0 shoots nothing; a negative value is undefined behaviour there
		if (!as_int(o.parallel_src)) return false;
		if (o.parallel_src < 0) { err = "bad value for -parallel_src: must be >= 0"; return false; }
		return true;
	}
	if (a == "-fma") return as_int(o.fma);
	if (a == "-max_tri") return as_int(o.max_tri);
	if (a == "-old_format") return as_int(o.old_format);
	if (a == "-quiet") return as_int(o.quiet);
	if (a == "-batch") return as_int(o.batch);
	if (a == "-kernel_timing") return as_int(o.kernel_timing);
	if (a == "-graph") return as_int(o.graph);
	if (a == "-variant") {
		if (!value(v)) return false;
		if (!std::strcmp(v, "v3")) o.v3 = 1;
		else if (!std::strcmp(v, "v0")) o.v3 = 0;
		else { err = std::string("bad value for -variant: ") + v; return false; }
		return true;
	}
	if (a == "-octree") {
		if (!value(v)) return false;
		if (!std::strcmp(v, "host")) o.octree_host = 1;
		else if (!std::strcmp(v, "device")) o.octree_host = 0;
		else { err = std::string("bad value for -octree: ") + v; return false; }
		return true;
	}
	if (a == "-neighbors") {
		if (!value(v)) return false;
		if (!std::strcmp(v, "host")) o.neighbors_host = 1;
		else if (!std::strcmp(v, "device")) o.neighbors_host = 0;
		else { err = std::string("bad value for -neighbors: ") + v; return false; }
		return true;
	}
	if (a == "-select") {
		if (!value(v)) return false;
		if (!std::strcmp(v, "low")) o.lowest = 1;
		else if (!std::strcmp(v, "high")) o.lowest = 0;
		else { err = std::string("bad value for -select: ") + v; return false; }
		return true;
	}
	err = "unknown option " + a;
	return false;
}

}  // namespace rgpu
