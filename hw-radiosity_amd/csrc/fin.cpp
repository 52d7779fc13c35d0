#include "fin.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace rgpu {

int format_truncated(float x, char *out)
{
	uint32_t bits;
	std::memcpy(&bits, &x, 4);
	const uint32_t biased = (bits >> 23) & 0xffu, frac = bits & 0x7fffffu;
	const bool negative = (bits >> 31) != 0;
	if (biased == 0xffu || (biased == 0 && frac != 0)) return 0;
	char *o = out;
	if (negative) *o++ = '-';
	if (biased == 0) {  // signed zero
		std::memcpy(o, "0.0", 4);
		return (int)(o - out) + 3;
	}
	// x = sig * 2^e exactly
	const uint64_t sig = frac | (1u << 23);
	const int e = (int)biased - 150;
	if (e >= 41) return 0;
	uint64_t whole;
	if (e >= 0) whole = sig << e;
	else whole = (e < -23) ? 0 : (sig >> -e);
	char rev[24];
	int nd = 0;
	do {
		rev[nd++] = (char)('0' + whole % 10);
		whole /= 10;
	} while (whole);
	while (nd) *o++ = rev[--nd];
	*o++ = '.';
	// fraction digits: only for 1 <= -e <= 60, as a (-e)-bit fixed point
	int kept = 0;
	char digit[8];
	if (e < 0 && e >= -60) {
		const int fb = -e;
		const uint64_t mask = (fb == 64) ? ~0ull : ((1ull << fb) - 1);
		uint64_t f = (fb >= 24) ? sig : (sig & mask);  // sig < 2^24 <= 2^fb when fb >= 24
		for (int k = 0; k < 8 && f; k++) {
			// f < 2^fb <= 2^60, so 10 * f fits
			f *= 10;
			digit[k] = (char)('0' + (f >> fb));
			f &= mask;
			if (digit[k] != '0') kept = k + 1;
		}
	}
	if (kept == 0) *o++ = '0';
	for (int k = 0; k < kept; k++) *o++ = digit[k];
	*o = 0;
	return (int)(o - out);
}

const char *FinWriter::text(float x)
{
	char tmp[48];
	if (format_truncated(x, tmp)) std::strcpy(last_, tmp);
	return last_;
}

int FinWriter::write(const std::string &path, const Patch *pool, int trinum, bool old_format, std::string &err)
{
	FILE *f = std::fopen(path.c_str(), "w");
	if (!f) {
		err = "Error: output file " + path + " open failed";
		return -1;
	}
	std::vector<char> buf(1 << 20);
	std::setvbuf(f, buf.data(), _IOFBF, buf.size());
	char line[256];
	for (int i = 0; i < trinum; i++) {
		const Patch &t = pool[i];
		if (t.parent < 0) continue;  // logical slots are not written
		std::fputs("Triangle \n", f);
		if (old_format) {
			for (int v = 0; v < 3; v++) {
				std::fprintf(f, "%f %f %f ", t.p[v][0], t.p[v][1], t.p[v][2]);
				std::fprintf(f, "%i %i %i 0 0 0 \n", (int)t.accB[v][0], (int)t.accB[v][1], (int)t.accB[v][2]);
			}
			continue;
		}
		int rgb[3];
		for (int c = 0; c < 3; c++) {
			float m = 0;
			for (int v = 0; v < 3; v++) m += t.accB[v][c];
			m /= 3;
			if (255.0f < m) m = 255.0f;  // std::min(m, 255.0f)
			if (m < 0.0f) m = 0.0f;      // std::max(m, 0.0f)
			if (std::isnan(m)) m = 0;
			rgb[c] = (int)m;
		}
		std::fprintf(f, "%d %d %d 0 0 0\n", rgb[0], rgb[1], rgb[2]);
		// the normal text is formatted before the vertices and reused
		std::string nrm;
		for (int k = 0; k < 3; k++) {
			nrm += text(t.n[k]);
			nrm += k < 2 ? " " : "\n";
		}
		for (int v = 0; v < 3; v++) {
			int len = 0;
			for (int k = 0; k < 3; k++) {
				const char *s = text(t.p[v][k]);
				const size_t l = std::strlen(s);
				std::memcpy(line + len, s, l);
				len += (int)l;
				line[len++] = ' ';
			}
			std::fwrite(line, 1, len, f);
			std::fputs(nrm.c_str(), f);
		}
	}
	std::fclose(f);
	return 0;
}

}  // namespace rgpu
