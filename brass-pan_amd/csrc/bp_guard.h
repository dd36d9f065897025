// bp_guard.h -- the optional non-finite guard (the reference's CHECK_NAN
// printIfNan, kernel/cubrasstower.cuh:15), shared by the device scan and host
// code: which positions it flags, which one it names, and its message.  Host
// code includes it without HIP (the guard's CPU test compiles it with g++).
#pragma once

#include <cmath>
#include <string>

#if defined(__HIPCC__)
#define BP_GUARD_HD __host__ __device__
#else
#define BP_GUARD_HD
#endif

namespace bp {

// A position is flagged when x, y or z is NaN or infinite.  (The fourth word
// of a record is the inverse scaled mass, not a coordinate.)
BP_GUARD_HD inline bool guard_flags(float x, float y, float z) {
#if defined(__HIP_DEVICE_COMPILE__)
    return !(isfinite(x) && isfinite(y) && isfinite(z));
#else
    return !(std::isfinite(x) && std::isfinite(y) && std::isfinite(z));
#endif
}

// The slot the guard names: the lowest flagged one among n float4 records
// (x, y, z, w), or n when none is flagged -- what the device scan returns.
inline int guard_first_flagged(const float *x4, int n) {
    for (int i = 0; i < n; ++i)
        if (guard_flags(x4[4 * i], x4[4 * i + 1], x4[4 * i + 2])) return i;
    return n;
}

inline std::string guard_message(int slot, int original_id) {
    return "non-finite position at slot " + std::to_string(slot) + " (original id " + std::to_string(original_id) +
           ")";
}

}  // namespace bp
