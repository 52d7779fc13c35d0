// refsem CLI: the oracle driven like the reference's `rad` (rad.cpp:525-661).
// TEST INFRASTRUCTURE ONLY (used by tests/ and the benchmark's CPU baseline).
//   refsem [rad options] [restatement switches] input.tri -o prefix
#include <omp.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "refsem.h"

int main(int argc, char **argv)
{
	static const char *takes_value[] = {"-debug", "-adapt_area", "-sample_area", "-converge", "-delta_ff",
		"-write_cycle", "-triangle", "-light", "-o", "-jobs", "-parallel_src", "-fma", "-select",
		"-neighbors", "-skip_discarded", "-max_tri", "-old_format", "-quiet", "-variant", nullptr};
	static const char *flags[] = {"-zip", "-interactive", nullptr};
	std::vector<const char *> opts;
	const char *input = nullptr, *prefix = "./test";
	for (int i = 1; i < argc; i++) {
		bool used = false;
		for (int k = 0; takes_value[k]; k++)
			if (!strcmp(argv[i], takes_value[k]) && i + 1 < argc) {
				if (!strcmp(argv[i], "-o")) prefix = argv[i + 1];
				opts.push_back(argv[i]);
				opts.push_back(argv[++i]);
				used = true;
				break;
			}
		for (int k = 0; !used && flags[k]; k++)
			if (!strcmp(argv[i], flags[k])) { opts.push_back(argv[i]); used = true; }
		if (!used) input = argv[i];
	}
	if (!input) { fprintf(stderr, "Usage: %s [options] input_file -o output_file\n", argv[0]); return 1; }
	rs_ctx *c = rs_new((int)opts.size(), opts.data());
	if (!c) return 1;
	double t0 = omp_get_wtime();
	if (rs_load_file(c, input) != 0) { fprintf(stderr, "%s\n", rs_error(c)); return 1; }
	rs_build_tree(c);
	double t1 = omp_get_wtime();
	fprintf(stderr, "[INFO] Init took %f sec. time.\n", t1 - t0);
	int loops = rs_run(c, prefix);
	double t2 = omp_get_wtime();
	fprintf(stderr, "[INFO] Radiosity Method took %f sec. time.\n", t2 - t1);
	long long cnt[RS_CNT_N];
	rs_counters(c, cnt);
	fprintf(stderr, "[STAT] iterations %d rays %lld cells %lld tests %lld precision %lld\n", loops, cnt[RS_CNT_RAYS],
		cnt[RS_CNT_CELLS], cnt[RS_CNT_TESTS], cnt[RS_CNT_PRECISION]);
	rs_free(c);
	return loops > 0 ? 0 : 1;
}
