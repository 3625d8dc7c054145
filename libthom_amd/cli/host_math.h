// Host-side number formats and the seeded input generator of the CLI.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace cli {

// IEEE binary32 -> binary16, round to nearest even (subnormals, inf, nan kept).
inline uint16_t f32_to_f16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t exp = (x >> 23) & 0xffu;
    const uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) return (uint16_t)(sign | (man ? 0x7e00u | (man >> 13) : 0x7c00u));
    const int e = (int)exp - 112;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        const uint32_t m = man | 0x800000u;
        const int shift = 14 - e;
        uint32_t keep = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (keep & 1u))) keep++;
        return (uint16_t)(sign | keep);
    }
    uint32_t out = sign | ((uint32_t)e << 10) | (man >> 13);
    const uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (out & 1u))) out++;
    return (uint16_t)out;
}

inline float f16_to_f32(uint16_t h) {
    const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu, man = h & 0x3ffu, x;
    if (exp == 0) {
        if (man == 0) {
            x = sign;
        } else {
            int e = -1;
            do { man <<= 1; e++; } while (!(man & 0x400u));
            x = sign | ((uint32_t)(112 - e) << 23) | ((man & 0x3ffu) << 13);
        }
    } else if (exp == 0x1fu) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp + 112) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// SplitMix64, counter form: element i of stream s under seed is output i of a
// run started at seed + s * 0xD1B54A32D192ED03 (the same generator as
// lthip_fill_uniform_* and libthom_amd/inputs.py).
inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t i) {
    uint64_t z = seed + stream * 0xD1B54A32D192ED03ull + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double rng_unit(uint64_t seed, uint64_t stream, uint64_t i) {
    return (double)(rng_u64(seed, stream, i) >> 40) * 0x1p-24;
}
inline double rng_normal(uint64_t seed, uint64_t stream, uint64_t i) {
    const uint64_t pair = i & ~1ull;
    const uint64_t a = rng_u64(seed, stream, pair), b = rng_u64(seed, stream, pair + 1);
    const double u1 = (double)((a >> 11) + 1) * 0x1p-53, u2 = (double)(b >> 11) * 0x1p-53;
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    return (i & 1) ? r * std::sin(th) : r * std::cos(th);
}
inline int rng_int(uint64_t seed, uint64_t stream, uint64_t i, int lo, int hi) {
    const uint64_t span = (uint64_t)(hi - lo + 1);
    return lo + (int)(((rng_u64(seed, stream, i) >> 32) * span) >> 32);
}

constexpr uint64_t kSeed = 1234;   // the reference seeds every generator with 1234ULL

}  // namespace cli
