#include "viewer_json.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

namespace rgpu {

namespace {

// One triangle as the converter holds it (main.cpp:6-24): colours and
// vertex positions / normals in double.
struct ViewTri {
	double fc[3], bc[3];
	double v[3][3], n[3][3];
};

// Six numbers from one line as `stringstream >> double` reads them
// (main.cpp:72-80): each after optional white space, the longest decimal
// prefix (strtod's decimal grammar; the stream reads no hexadecimal,
// infinity or NaN, so those stop at their first letter); what follows the
// sixth number is ignored.
bool six_numbers(const char *a, const char *b, double *f)
{
	const std::string line(a, b);
	const char *s = line.c_str();
	for (int i = 0; i < 6; i++) {
		while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\v' || *s == '\f' || *s == '\n') s++;
		const char *d = s + ((*s == '+' || *s == '-') ? 1 : 0);
		if (!((*d >= '0' && *d <= '9') || *d == '.')) return false;
		if (d[0] == '0' && (d[1] == 'x' || d[1] == 'X')) {  // "0x..." reads as 0
			f[i] = (*s == '-') ? -0.0 : 0.0;
			s = d + 1;
			continue;
		}
		char *stop = nullptr;
		f[i] = std::strtod(s, &stop);
		if (stop == s) return false;
		s = stop;
	}
	return true;
}

double tri_area(const ViewTri &t)  // Tri::area (main.cpp:12-23)
{
	double a[3], b[3], c[3];
	for (int i = 0; i < 3; i++) {
		a[i] = t.v[1][i] - t.v[0][i];
		b[i] = t.v[2][i] - t.v[0][i];
	}
	c[0] = a[1] * b[2] - b[1] * a[2];
	c[1] = a[2] * b[0] - b[2] * a[0];
	c[2] = a[0] * b[1] - b[0] * a[1];
	return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) / 2;
}

// "%g,%g,%g" for each vertex of each triangle, the groups joined by ','
// (main.cpp:90-150), formatted in parallel chunks and concatenated in order.
template <typename Fn> std::string section(size_t ntri, Fn put)
{
	const size_t kChunk = 8192;
	const size_t nch = (ntri + kChunk - 1) / kChunk;
	std::vector<std::string> part(nch);
#pragma omp parallel for schedule(dynamic, 1)
	for (long c = 0; c < (long)nch; c++) {
		std::string &o = part[c];
		char buf[160];
		for (size_t i = c * kChunk; i < std::min(ntri, (c + 1) * kChunk); i++)
			for (int j = 0; j < 3; j++) {
				const int len = put(i, j, buf);
				if (i || j) o += ',';
				o.append(buf, len);
			}
	}
	std::string all;
	size_t total = 0;
	for (const std::string &p : part) total += p.size();
	all.reserve(total);
	for (const std::string &p : part) all += p;
	return all;
}

}  // namespace

int fin_to_json(const std::string &fin_path, const std::string &json_path, int &triangles, int &tiny, std::string &err)
{
	FILE *fi = std::fopen(fin_path.c_str(), "rb");
	if (!fi) {
		err = "Input file not exited: " + fin_path;
		return -1;
	}
	std::string text;
	{
		char buf[1 << 16];
		size_t got;
		while ((got = std::fread(buf, 1, sizeof buf, fi)) > 0) text.append(buf, got);
		std::fclose(fi);
	}
	// readColorfulTriangle (main.cpp:165-221): records of "Triangle ", a
	// colour line and three vertex lines; lines are indexed first, records
	// read in parallel, and the first bad line (in file order) reported
	std::vector<size_t> start;  // line starts, then one past the end
	for (size_t q = 0; q < text.size();) {
		start.push_back(q);
		const void *nl = std::memchr(text.data() + q, '\n', text.size() - q);
		q = nl ? (size_t)(static_cast<const char *>(nl) - text.data()) + 1 : text.size();
	}
	const size_t nlines = start.size();
	start.push_back(text.size());
	auto line_end = [&](size_t l) {  // the line without its '\n'
		size_t e = start[l + 1];
		if (e > start[l] && text[e - 1] == '\n') e--;
		return e;
	};
	const size_t n = (nlines + 4) / 5;  // five lines per record
	std::vector<ViewTri> T(n);
	long bad = LONG_MAX;  // first bad line, LONG_MAX for none
#pragma omp parallel for schedule(static) reduction(min : bad)
	for (long r = 0; r < (long)n; r++) {
		ViewTri &t = T[r];
		const size_t l0 = 5 * (size_t)r;
		long here = -1;
		if (std::string(text.data() + start[l0], line_end(l0) - start[l0]) != "Triangle ") here = (long)l0;
		double f[6];
		for (int q = 0; q < 4 && here < 0; q++) {
			const size_t l = l0 + 1 + q;
			if (l >= nlines || !six_numbers(text.data() + start[l], text.data() + line_end(l), f)) {
				here = (long)l;
				break;
			}
			if (q == 0) {
				for (int i = 0; i < 3; i++) {
					t.fc[i] = f[i];
					t.bc[i] = f[i + 3];
				}
			} else {
				for (int j = 0; j < 3; j++) {
					t.v[q - 1][j] = f[j];
					t.n[q - 1][j] = f[j + 3];
				}
			}
		}
		if (here >= 0) bad = std::min(bad, here);
	}
	if (bad != LONG_MAX) {
		err = (size_t)bad < nlines ? "unreadable line " + std::to_string(bad + 1) + ": " +
		                                 std::string(text.data() + start[bad], line_end(bad) - start[bad])
		                           : std::string("the last record is cut short");
		return -1;
	}
	double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
	for (size_t r = 0; r < n; r++)  // extendMin / extendMax in file order (main.cpp:50-55)
		for (int i = 0; i < 3; i++)
			for (int k = 0; k < 3; k++) {
				const double x = T[r].v[i][k];
				if (r == 0 && i == 0) {
					lo[k] = hi[k] = x;
				} else {
					lo[k] = x < lo[k] ? x : lo[k];
					hi[k] = hi[k] < x ? x : hi[k];
				}
			}
	// into a cube of side 10 centred on the origin (main.cpp:224-240)
	const double d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
	const double m01 = d0 < d1 ? d1 : d0;
	const double max_side = m01 < d2 ? d2 : m01;
	const double view = 10.0, scale = view / max_side;
	tiny = 0;
	for (ViewTri &t : T) {
		for (int j = 0; j < 3; j++)
			for (int k = 0; k < 3; k++) {
				t.v[j][k] -= lo[k];
				t.v[j][k] *= scale;
				t.v[j][k] -= 0.5 * view;
			}
		if (tri_area(t) < 1e-6) tiny++;
	}
	triangles = (int)n;
	// writeColorfulJson (main.cpp:81-164)
	std::string out = "{\n\t\"colorful\": true,\n\t\"vertexPositions\" : [";
	out += section(n, [&](size_t i, int j, char *s) {
		return std::snprintf(s, 160, "%g,%g,%g", T[i].v[j][0], T[i].v[j][1], T[i].v[j][2]);
	});
	out += "],\n\t\"vertexNormals\" : [";
	out += section(n, [&](size_t i, int j, char *s) {
		return std::snprintf(s, 160, "%g,%g,%g", T[i].n[j][0], T[i].n[j][1], T[i].n[j][2]);
	});
	out += "],\n\t\"vertexTextureCoords\" : [";
	out += section(n, [&](size_t, int, char *s) {
		std::memcpy(s, "1,1,1", 5);
		return 5;
	});
	out += "],\n\t\"vertexFrontColors\" : [";
	out += section(n, [&](size_t i, int, char *s) {
		return std::snprintf(s, 160, "%g,%g,%g", T[i].fc[0] / 255.0, T[i].fc[1] / 255.0, T[i].fc[2] / 255.0);
	});
	out += "],\n\t\"vertexBackColors\" : [";
	out += section(n, [&](size_t i, int, char *s) {
		return std::snprintf(s, 160, "%g,%g,%g", T[i].bc[0] / 255.0, T[i].bc[1] / 255.0, T[i].bc[2] / 255.0);
	});
	out += "],\n\t\"indices\" : [";
	out += section(n, [&](size_t i, int j, char *s) { return std::snprintf(s, 160, "%zu", 3 * i + j); });
	out += "]\n}\n";
	FILE *fo = std::fopen(json_path.c_str(), "wb");
	if (!fo) {
		err = "Output file path failed";
		return -1;
	}
	const bool ok = std::fwrite(out.data(), 1, out.size(), fo) == out.size();
	if (std::fclose(fo) != 0 || !ok) {
		err = "Output file write failed: " + json_path;
		return -1;
	}
	return 0;
}

}  // namespace rgpu
