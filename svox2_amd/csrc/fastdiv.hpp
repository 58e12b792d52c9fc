// Division of non-negative 31-bit integers by a divisor fixed at launch time,
// as one multiply-high, one add and one shift (Granlund and Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, section 4).
//
// For a divisor 1 <= d <= 2^31 let l = ceil(log2 d) and m = floor(2^(32+l) / d) - 2^32 + 1
// (it fits 32 bits).  Then for every 0 <= n < 2^31,
//     n / d = (umulhi(m, n) + n) >> l,
// and the sum cannot overflow 32 bits because umulhi(m, n) < n < 2^31.
//
// The header needs no HIP runtime so that host code (and the CPU test that
// checks it against plain division) can include it.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SVOX2_HD __host__ __device__
#else
#define SVOX2_HD
#endif

namespace svox2amd {

struct FastDiv {
    uint32_t d;      // the divisor
    uint32_t m;      // multiplier
    uint32_t shift;  // l

    FastDiv() : d(1), m(1), shift(0) {}
    // Divisors outside [1, 2^31] are clamped into it (callers check them).
    explicit FastDiv(uint32_t divisor)
        : d(divisor == 0 ? 1u : (divisor > 0x80000000u ? 0x80000000u : divisor)) {
        uint32_t l = 0;
        while ((uint64_t{1} << l) < d) ++l;
        shift = l;
        m = static_cast<uint32_t>(((uint64_t{1} << (32 + l)) / d) - (uint64_t{1} << 32) + 1);
    }

    SVOX2_HD uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t hi = __umulhi(m, n);
#else
        const uint32_t hi = static_cast<uint32_t>((static_cast<uint64_t>(m) * n) >> 32);
#endif
        return (hi + n) >> shift;
    }
    // q = n / d and r = n - q d.
    SVOX2_HD uint32_t divmod(uint32_t n, uint32_t& r) const {
        const uint32_t q = div(n);
        r = n - q * d;
        return q;
    }
};

}  // namespace svox2amd
