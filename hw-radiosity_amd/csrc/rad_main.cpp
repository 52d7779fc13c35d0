// `rad`: the drop-in command line (rad.cpp:525-661) over libradgpu.
//   rad [options] input.tri -o prefix
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/radiosity_gpu.h"

namespace {

const char *kUsage =
	"Usage: %s [options] input_file -o output_file\n"
	"OPTIONS\n"
	"   Debug Options\n"
	"       -debug <integer>        Output intermediate result according debug level.\n"
	"   Radiosity Options\n"
	"       -adapt_area <float>     The threahold of adaptive splitting algorithm.\n"
	"       -sample_area <float>    Sampling area for the form factor difference test.\n"
	"       -converge <float>       Stop when no patch has more unshot energy than this.\n"
	"       -delta_ff <float>       Form factor differences below this count as equal.\n"
	"       -write_cycle <integer>  Write the state every write_cycle iterations.\n"
	"       -triangle <integer>     The maximum number of triangles in the model.\n"
	"       -light <float>          Scale of the light emission.\n"
	"       -jobs <integer>         Host threads for set-up.\n"
	"       -parallel_src <integer> Sources shot per iteration.\n"
	"   Output Options\n"
	"       -zip                    Compress output files with gzip.\n"
	"       -o </<path>/file>       Output prefix (default ./test).\n"
	"       -interactive            After each output file run ./WebGL_demo/run.sh on it.\n"
	"   Engine Options\n"
	"       -fma 0|1  -select high|low  -max_tri N  -old_format 0|1  -batch N  -quiet 0|1  -graph 0|1\n";

bool takes_value(const char *a)
{
	static const char *with[] = {"-debug", "-adapt_area", "-sample_area", "-converge", "-delta_ff", "-write_cycle",
		"-triangle", "-light", "-o", "-jobs", "-parallel_src", "-fma", "-select", "-max_tri", "-old_format",
		"-quiet", "-batch", "-kernel_timing", "-graph", "-neighbors", "-octree", "-variant", nullptr};
	for (int k = 0; with[k]; k++)
		if (!std::strcmp(a, with[k])) return true;
	return false;
}
bool is_flag(const char *a) { return !std::strcmp(a, "-zip") || !std::strcmp(a, "-interactive"); }

}  // namespace

int main(int argc, char **argv)
{
	if (argc < 3) {
		std::fprintf(stderr, kUsage, argv[0]);
		return 1;
	}
	std::vector<const char *> opts;
	const char *input = nullptr, *prefix = "./test";
	for (int i = 1; i < argc; i++) {
		if (takes_value(argv[i]) && i + 1 < argc) {
			if (!std::strcmp(argv[i], "-o")) prefix = argv[i + 1];
			opts.push_back(argv[i]);
			opts.push_back(argv[++i]);
		} else if (is_flag(argv[i])) {
			opts.push_back(argv[i]);
		} else {
			input = argv[i];  // any other word is the input; the last wins (rad.cpp:625-626)
		}
	}
	std::fprintf(stderr, "[INFO] Input Model %s and Output Prefix %s\n", input ? input : "(null)", prefix);
	rad_ctx *ctx = nullptr;
	if (rad_create(&ctx, (int)opts.size(), opts.data()) != RAD_OK) return 1;
	auto t0 = std::chrono::steady_clock::now();
	if (!input || rad_load_tri(ctx, input) != RAD_OK) {
		std::fprintf(stderr, "File open error::%s\n", argv[2]);
		rad_destroy(ctx);
		return 1;
	}
	auto t1 = std::chrono::steady_clock::now();
	std::fprintf(stderr, "[INFO] Init took %f sec. time.\n", std::chrono::duration<double>(t1 - t0).count());
	int iters = 0;
	const rad_status st = rad_run(ctx, prefix, &iters);
	auto t2 = std::chrono::steady_clock::now();
	if (st != RAD_OK) {
		std::fprintf(stderr, "[ERR] %s\n", rad_last_error(ctx));
		rad_destroy(ctx);
		return 1;
	}
	std::fprintf(stderr, "[INFO] Radiosity Method took %f sec. time.\n", std::chrono::duration<double>(t2 - t1).count());
	long long cnt[RAD_CNT_N];
	double tm[7];
	rad_counters(ctx, cnt);
	rad_timings(ctx, tm);
	std::fprintf(stderr, "[STAT] rays %lld (reference-equivalent %lld) cells %lld tests %lld precision %lld stalled %lld; "
	             "setup %.1f ms upload %.1f ms device shooting %.1f ms\n",
	             cnt[RAD_CNT_RAYS], cnt[RAD_CNT_REF_RAYS], cnt[RAD_CNT_CELLS], cnt[RAD_CNT_TESTS],
	             cnt[RAD_CNT_PRECISION], cnt[RAD_CNT_STALLED], tm[0], tm[1], tm[2]);
	rad_destroy(ctx);
	return 0;
}
