// Philox4x32-10 counter-based generator (J. K. Salmon, M. A. Moraes,
// R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11), written from the published description.  Usable on host and device.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define TFG_HD __host__ __device__
#else
#define TFG_HD
#endif

namespace tfg {

TFG_HD inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = static_cast<uint64_t>(M0) * c[0];
        const uint64_t p1 = static_cast<uint64_t>(M1) * c[2];
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

// Zero-mean, unit-variance uniform variate for the DPD pair (a, b) at step t.
// The counter is symmetric in the pair, so both sides draw the same number.
// The value before the sqrt(3) scale is exactly representable in FP32.
template <typename real>
TFG_HD inline real dpd_xi(uint32_t a, uint32_t b, uint64_t step, uint32_t seed_lo, uint32_t seed_hi) {
    uint32_t c[4];
    c[0] = a < b ? a : b;
    c[1] = a < b ? b : a;
    c[2] = static_cast<uint32_t>(step);
    c[3] = static_cast<uint32_t>(step >> 32);
    philox4x32_10(c, seed_lo, seed_hi);
    const int m = static_cast<int>(c[0] >> 9) - 4194304;   // [-2^22, 2^22)
    const real t = (static_cast<real>(m) + static_cast<real>(0.5)) * static_cast<real>(2.384185791015625e-07);
    return t * static_cast<real>(1.7320508075688772);
}

}  // namespace tfg
