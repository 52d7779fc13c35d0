// tri2json: the WebGL viewer's model converter, a drop-in for the
// reference's WebGL_demo/model-converter (main.cpp:248-290): same options,
// messages and output file.  A record the converter cannot read ends the run
// with status 1 and a message (the reference aborts on an assertion).
#include <cstdio>
#include <cstring>
#include <string>

#include "viewer_json.h"

int main(int argc, char *argv[])
{
	if (argc == 2) {
		if (!std::strcmp("--help", argv[1])) {
			std::fprintf(stderr, "\t\t--format\t\tModel File Format\n\t\t\t\t"
			                     "--format SIMPLE / --format COLOR\n");
			return 0;
		}
		return 1;
	}
	if (argc < 5) {
		std::fprintf(stderr, "./tri2json -i <filename.tri> -o <jsonname.json> [--format]\n");
		std::fprintf(stderr, "./tri2json --help\n");
		return 1;
	}
	std::string in, out, format = "COLOR";
	for (int i = 1; i < argc; i++) {
		if (!std::strcmp(argv[i], "-i") && i + 1 < argc) {
			in = argv[++i];
		} else if (!std::strcmp(argv[i], "-o") && i + 1 < argc) {
			out = argv[++i];
		} else if (!std::strcmp(argv[i], "--format") && i + 1 < argc) {
			format = argv[++i];
			if (format != "COLOR" && format != "SIMPLE") {
				std::fprintf(stderr, "Please check file format\n");
				return 1;
			}
		}
	}
	if (in.empty() || out.empty()) {
		std::fprintf(stderr, in.empty() ? "Please give input file name\n" : "Please give output file name\n");
		return 1;
	}
	if (format != "COLOR") {
		std::fprintf(stderr, "Unsupport format\n");
		return 1;
	}
	int n = 0, tiny = 0;
	std::string err;
	if (rgpu::fin_to_json(in, out, n, tiny, err)) {
		std::fprintf(stderr, "%s\n", err.c_str());
		return 1;
	}
	for (int k = 0; k < tiny; k++) std::fprintf(stderr, "[Warning] Triangle area is too small to show\n");
	std::fprintf(stderr, "#Triangle %d\n", n);
	std::fprintf(stderr, "Model-Converter Success\n");
	return 0;
}
