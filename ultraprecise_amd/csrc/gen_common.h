// gen_common.h -- seeded synthetic compact columns, identical on host and
// device.  Counter-based: every value is a pure function of (seed, column
// id, row, draw index), so row ranges can be produced independently, on
// either side, in any order.
#ifndef UPDEC_GEN_COMMON_H
#define UPDEC_GEN_COMMON_H
#include <stdint.h>

#if defined(__HIPCC__)
#define UPD_HD __host__ __device__ inline
#else
#define UPD_HD inline
#endif

namespace updec_gen {

// splitmix64 finalizer (Steele, Lea, Flood: "Fast splittable pseudorandom
// number generators", OOPSLA 2014)
UPD_HD uint64_t mix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

UPD_HD uint64_t column_key(uint64_t seed, uint32_t col) { return mix(seed * 0x100000001B3ull + col + 1); }

UPD_HD uint64_t draw(uint64_t key, uint64_t row, uint32_t k) { return mix(key ^ mix((row << 8) | k)); }

enum { GEN_RANGE = 0, GEN_DIGITS = 1, GEN_BELOW = 2 };
enum { GEN_MAX_BOUND_WORDS = 16 };

struct Desc {
    int32_t kind, precision, digits, neg_permille;
    uint64_t lo, hi, step, seed;
    uint32_t column_id;
    int32_t bound_words;
    uint32_t bound[GEN_MAX_BOUND_WORDS];
};

// Write one compact row of `nbytes` bytes (the reference's layout: little
// endian magnitude, sign in bit 7 of the last byte).
UPD_HD void gen_row(const Desc& d, uint64_t key, uint64_t row, uint8_t* out, int nbytes)
{
    uint32_t m[48];
    for (int i = 0; i < 48; i++) m[i] = 0;
    uint32_t neg = (draw(key, row, 0) % 1000u) < (uint64_t)d.neg_permille ? 1u : 0u;
    if (d.kind == GEN_RANGE) {
        uint64_t span = d.hi - d.lo + 1u;
        uint64_t r = draw(key, row, 1);
        uint64_t v = d.lo + (span ? r % span : r);
        v *= d.step;
        m[0] = (uint32_t)v;
        m[1] = (uint32_t)(v >> 32);
    } else if (d.kind == GEN_BELOW) {
        // floor(R * bound / 2^(32 (n + 2))) for a random R of n + 2 words:
        // uniform on [0, bound) up to a relative bias below 2^-64
        const int n = d.bound_words;
        uint32_t r[GEN_MAX_BOUND_WORDS + 2];
        for (int i = 0; i < n + 2; i += 2) {
            uint64_t x = draw(key, row, 1 + (uint32_t)(i / 2));
            r[i] = (uint32_t)x;
            if (i + 1 < n + 2) r[i + 1] = (uint32_t)(x >> 32);
        }
        uint32_t prod[2 * GEN_MAX_BOUND_WORDS + 2];
        for (int i = 0; i < 2 * n + 2; i++) prod[i] = 0;
        for (int i = 0; i < n + 2; i++) {
            uint64_t c = 0;
            for (int j = 0; j < n; j++) {
                uint64_t t = (uint64_t)r[i] * d.bound[j] + prod[i + j] + c;
                prod[i + j] = (uint32_t)t;
                c = t >> 32;
            }
            prod[i + n] = (uint32_t)c;
        }
        for (int i = 0; i < n; i++) m[i] = prod[n + 2 + i];
    } else {
        // exactly `digits` decimal digits, leading digit nonzero: a top
        // chunk of t digits in [10^(t-1), 10^t), then 9-digit chunks
        int nchunk = (d.digits + 8) / 9;
        int t = d.digits - 9 * (nchunk - 1);
        uint64_t p = 1;
        for (int i = 1; i < t; i++) p *= 10u;
        uint64_t top = p + draw(key, row, 1) % (9u * p);
        m[0] = (uint32_t)top;
        m[1] = (uint32_t)(top >> 32);
        for (int c = 1; c < nchunk; c++) {
            uint64_t chunk = draw(key, row, 1 + (uint32_t)c) % 1000000000u;
            uint64_t carry = chunk;
            for (int i = 0; i < 48; i++) {
                uint64_t s = (uint64_t)m[i] * 1000000000u + carry;
                m[i] = (uint32_t)s;
                carry = s >> 32;
            }
        }
    }
    for (int b = 0; b < nbytes; b++) out[b] = (uint8_t)(m[b >> 2] >> (8 * (b & 3)));
    if (nbytes > 0) out[nbytes - 1] |= (uint8_t)(neg << 7);
}

UPD_HD int32_t gen_int32(uint64_t key, uint64_t row, int32_t lo, int32_t hi)
{
    uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
    return (int32_t)((int64_t)lo + (int64_t)(draw(key, row, 1) % span));
}

UPD_HD uint8_t gen_byte(uint64_t key, uint64_t row, const uint8_t* choices, int n)
{
    return choices[draw(key, row, 1) % (uint64_t)n];
}

}  // namespace updec_gen
#endif
