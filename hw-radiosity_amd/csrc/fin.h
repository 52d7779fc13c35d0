// Output files: <prefix>.fin and the -write_cycle snapshots <prefix>.<loop>
// (rad.cpp:117-242).
#pragma once

#include <string>

#include "patch.h"

namespace rgpu {

// Decimal text of a float as the reference prints it (float2string.h): the
// exact binary value truncated to 8 fraction digits, trailing zeros dropped,
// at least one digit on each side of the point.  Returns the length, or 0
// where the reference prints nothing (NaN, inf, subnormal, |x| >= 2^64).
int format_truncated(float x, char *out);

class FinWriter {
public:
	// Writes the non-logical slots of pool[0, trinum).  Returns 0, or -1
	// with err set when the file cannot be opened.
	int write(const std::string &path, const Patch *pool, int trinum, bool old_format, std::string &err);

private:
	// The reference formats into one static buffer and prints it even when
	// a value is not converted, so such a value repeats the previous text.
	char last_[48] = {0};
};

}  // namespace rgpu
