// mathcheck.cpp -- test-only: the device library's restatements of glibc's
// expf and logf (exp_ref, log_ref in aaags-claude_amd/csrc/gs_device.h),
// compiled from the same source for the host, compared bit for bit with the
// system's expf and logf over ranges of float bit patterns.  Built into
// tests/cpp/_build/libgs_mathcheck.so.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "gs_device.h"

static float as_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

static bool same(float a, float b) {
    if (std::isnan(a) && std::isnan(b)) return true;
    uint32_t x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    return x == y;
}

// Mismatches over the bit patterns lo, lo + stride, ... < hi.
extern "C" long long check_log_ref(uint32_t lo, uint64_t hi, uint32_t stride) {
    long long bad = 0;
    for (uint64_t u = lo; u < hi; u += stride) {
        const float x = as_float((uint32_t)u);
        if (!same(gsd::log_ref(x), logf(x))) bad++;
    }
    return bad;
}

extern "C" long long check_exp_ref(uint32_t lo, uint64_t hi, uint32_t stride) {
    long long bad = 0;
    for (uint64_t u = lo; u < hi; u += stride) {
        const float x = as_float((uint32_t)u);
        if (!same(gsd::exp_ref(x), expf(x))) bad++;
    }
    return bad;
}
