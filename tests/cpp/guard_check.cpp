// Host-only check of the non-finite guard's logic (brass-pan_amd/csrc/bp_guard.h),
// compiled with g++ without HIP: which positions are flagged, which slot is
// named, and the message.  Prints "guard ok" on success.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "bp_guard.h"

static int failures = 0;
static void expect(bool ok, const char *what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

int main() {
    const float nan = std::numeric_limits<float>::quiet_NaN(), inf = std::numeric_limits<float>::infinity();
    // finite extremes are not flagged
    expect(!bp::guard_flags(FLT_MAX, -FLT_MAX, 0.0f), "FLT_MAX is finite");
    expect(!bp::guard_flags(FLT_TRUE_MIN, -FLT_MIN, -0.0f), "denormals and -0 are finite");
    // every non-finite value in every coordinate is flagged
    const float bad[3] = {nan, inf, -inf};
    for (float b : bad) {
        expect(bp::guard_flags(b, 0.f, 0.f), "x non-finite");
        expect(bp::guard_flags(0.f, b, 0.f), "y non-finite");
        expect(bp::guard_flags(0.f, 0.f, b), "z non-finite");
    }
    // the lowest flagged slot is named; w is not a coordinate
    const int n = 10;
    std::vector<float> x(4 * n, 1.0f);
    expect(bp::guard_first_flagged(x.data(), n) == n, "no flagged slot gives n");
    x[4 * 2 + 3] = nan;  // w of slot 2
    expect(bp::guard_first_flagged(x.data(), n) == n, "w is not checked");
    x[4 * 7 + 1] = inf;
    x[4 * 4 + 2] = nan;
    expect(bp::guard_first_flagged(x.data(), n) == 4, "lowest flagged slot");
    x[4 * 0 + 0] = -inf;
    expect(bp::guard_first_flagged(x.data(), n) == 0, "slot 0");
    expect(bp::guard_first_flagged(x.data(), 0) == 0, "empty input");
    expect(bp::guard_message(4, 17) == "non-finite position at slot 4 (original id 17)", "message");
    if (failures == 0) std::printf("guard ok\n");
    return failures == 0 ? 0 : 1;
}
