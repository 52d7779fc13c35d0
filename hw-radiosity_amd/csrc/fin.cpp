#include "fin.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include <omp.h>

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

namespace {

// One record in the reference's order: "Triangle", the mean colour, then per
// vertex its coordinates and the normal, written at o (at most kRecordMax
// bytes); returns the end.  last (ll chars) = the text of the last value
// that converted (the reference prints its static buffer again for a value
// that does not convert); updated in place.
constexpr int kRecordMax = 1024;

inline char *put_literal(char *o, const char *s, size_t l)
{
	std::memcpy(o, s, l);
	return o + l;
}

inline char *put_uint(char *o, unsigned v)
{
	char rev[10];
	int nd = 0;
	do {
		rev[nd++] = (char)('0' + v % 10);
		v /= 10;
	} while (v);
	while (nd) *o++ = rev[--nd];
	return o;
}

// x's text at o, or the text of the last value that converted (last, ll
// chars, NUL-terminated; updated when x converts); returns the end
inline char *put_value(char *o, float x, char *last, int &ll)
{
	const int l = format_truncated(x, o);
	if (l) {
		std::memcpy(last, o, (size_t)l);
		last[l] = 0;
		ll = l;
		return o + l;
	}
	return put_literal(o, last, (size_t)ll);
}

char *format_record(const Patch &t, bool old_format, char *last, int &ll, char *o)
{
	o = put_literal(o, "Triangle \n", 10);
	if (old_format) {
		for (int v = 0; v < 3; v++) {
			o += std::snprintf(o, 400, "%f %f %f ", t.p[v][0], t.p[v][1], t.p[v][2]);
			o += std::snprintf(o, 64, "%i %i %i 0 0 0 \n", (int)t.accB[v][0], (int)t.accB[v][1], (int)t.accB[v][2]);
		}
		return o;
	}
	for (int c = 0; c < 3; c++) {
		float m = 0;
		for (int v = 0; v < 3; v++) m += t.accB[v][c];
		m /= 3;
		if (255.0f < m) m = 255.0f;  // std::min(m, 255.0f)
		if (m < 0.0f) m = 0.0f;      // std::max(m, 0.0f)
		if (std::isnan(m)) m = 0;
		o = put_uint(o, (unsigned)(int)m);  // 0 .. 255, as %d prints it
		*o++ = ' ';
	}
	o = put_literal(o, "0 0 0\n", 6);
	// the normal text is formatted before the vertices and reused
	char nrm[160];
	char *e = nrm;
	for (int k = 0; k < 3; k++) {
		e = put_value(e, t.n[k], last, ll);
		*e++ = k < 2 ? ' ' : '\n';
	}
	const size_t nl = (size_t)(e - nrm);
	for (int v = 0; v < 3; v++) {
		for (int k = 0; k < 3; k++) {
			o = put_value(o, t.p[v][k], last, ll);
			*o++ = ' ';
		}
		o = put_literal(o, nrm, nl);
	}
	return o;
}

// The state `last` after formatting record t, given it before: the last of
// its values (formatting order n[0..2], p[0][0..2], .., p[2][0..2]) that
// converts.  Returns false, leaving `last` as it was, when none converts.
bool last_after(const Patch &t, char *last)
{
	char tmp[48];
	for (int q = 8; q >= 0; q--)
		if (format_truncated(t.p[q / 3][q % 3], tmp)) {
			std::strcpy(last, tmp);
			return true;
		}
	for (int k = 2; k >= 0; k--)
		if (format_truncated(t.n[k], tmp)) {
			std::strcpy(last, tmp);
			return true;
		}
	return false;
}

}  // namespace

int FinWriter::write(const std::string &path, const Patch *pool, int trinum, bool old_format, std::string &err)
{
	FILE *f = std::fopen(path.c_str(), "w");
	if (!f) {
		err = "Error: output file " + path + " open failed";
		return -1;
	}
	// records are formatted in chunks in parallel; each chunk starts from the
	// `last` state that the records before it leave (so any chunking gives the
	// same bytes): about four chunks per thread, of 256 to 4096 records
	const int nt = std::max(1, omp_get_max_threads());
	const int chunk = std::min(4096, std::max(256, (trinum + 4 * nt - 1) / (4 * nt)));
	const int nch = (trinum + chunk - 1) / chunk;
	std::vector<std::array<char, 48>> start(nch);
	for (int c = 0; c < nch; c++) {
		std::memcpy(start[c].data(), last_, sizeof last_);
		if (old_format) continue;
		// the latest written record of the chunk with a value that converts
		for (int i = std::min(trinum, (c + 1) * chunk) - 1; i >= c * chunk; i--)
			if (pool[i].parent >= 0 && last_after(pool[i], last_)) break;
	}
	bool ok = true;
	// chunks are formatted in parallel and written in order as they finish;
	// each thread formats into one buffer of its own
#pragma omp parallel
	{
		std::unique_ptr<char[]> buf(new char[(size_t)chunk * kRecordMax]);  // not zeroed
#pragma omp for ordered schedule(dynamic, 1)
		for (int c = 0; c < nch; c++) {
			char last[48];
			std::memcpy(last, start[c].data(), sizeof last);
			int ll = (int)std::strlen(last);
			char *o = buf.get();
			for (int i = c * chunk; i < std::min(trinum, (c + 1) * chunk); i++)
				if (pool[i].parent >= 0) o = format_record(pool[i], old_format, last, ll, o);  // logical slots are not written
			const size_t len = (size_t)(o - buf.get());
#pragma omp ordered
			if (ok) ok = std::fwrite(buf.get(), 1, len, f) == len;
		}
	}
	ok = (std::fclose(f) == 0) && ok;
	if (!ok) {
		err = "Error: output file " + path + " write failed";
		return -1;
	}
	return 0;
}

}  // namespace rgpu
