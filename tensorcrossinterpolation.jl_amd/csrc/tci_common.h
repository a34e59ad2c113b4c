// tci_common.h — definitions shared by the solve engine and the register-tile
// rrLU: status codes, the argmax key, Julia's NaN-propagating max, the
// team-shared scalars and the project PRNG.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define ENG __device__
#define ENG_INLINE __device__ __forceinline__
#define ENG_NOINLINE __device__ __attribute__((noinline))
#else
#define ENG
#define ENG_INLINE inline
#define ENG_NOINLINE __attribute__((noinline))
#endif

namespace tci {

enum {
    ST_OK = 0, ST_NAN_L = 1, ST_NAN_U = 2, ST_ARG = 3, ST_ZERO_MAXSAMPLE = 4,
    ST_NAN_TENSOR = 5, ST_NONSQUARE = 6, ST_DIM = 7, ST_GENERIC = 8, ST_SINGULAR = 9,
    ST_CAPACITY = 10
};

struct ArgBest {
    double v;      // abs2 value, -1 = no candidate (NaN entries never count)
    int64_t lin;   // column-major linear position
    double a;      // the entry itself
    int32_t orig;  // original row + 65536·original column (register-tile rrLU)
    int32_t pad;
};

ENG inline bool arg_better(const ArgBest &a, const ArgBest &b)
{
    return a.v > b.v || (a.v == b.v && a.lin < b.lin);
}

// bit casts usable on host and device
ENG_INLINE int64_t eng_as_i64(double x)
{
    int64_t i;
    __builtin_memcpy(&i, &x, sizeof i);
    return i;
}
ENG_INLINE double eng_as_f64(int64_t i)
{
    double x;
    __builtin_memcpy(&x, &i, sizeof x);
    return x;
}

ENG_INLINE uint32_t eng_umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// Base.max on Float64: NaN-propagating
ENG inline double jl_max(double a, double b)
{
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

// Scalars shared by the team (written by thread 0 between barriers).
constexpr int kPivErrCap = 256;
struct EngShared {
    double maxsample;
    double abstol;
    double ws[6];
    uint64_t rng[4];
    uint64_t tph[8];   // clock ticks per phase (thread 0): see PH_*
    int32_t status;
    int32_t cnt;       // scratch count (union compaction)
    int32_t nf;        // global pivots found
    int32_t npe;       // length of pivoterrors
    int32_t rk[4];     // :rook search: |I0|, |J0|, a flag, a count
    double pe[kPivErrCap];   // pivoterrors while the solve runs, when they fit
    uint64_t cache_lookups, cache_hits;   // memo table of f (CachedFunction)
};

// ---- 256-bit keys of the memo table (four 64-bit limbs, least significant
// first): k += c·s for a key coefficient c and a small factor s
ENG inline void key_muladd(uint64_t *k, const uint64_t *c, uint64_t s)
{
    uint64_t carry = 0;
    for (int q = 0; q < 4; ++q) {
        // the 128-bit product c[q]·s + carry, split in 32-bit halves
        const uint64_t lo = (c[q] & 0xffffffffULL) * s, hi = (c[q] >> 32) * s;
        const uint64_t mid = (lo >> 32) + (hi & 0xffffffffULL);
        const uint64_t plo = (lo & 0xffffffffULL) | (mid << 32);
        const uint64_t phi = (hi >> 32) + (mid >> 32);
        const uint64_t t = plo + carry;
        const uint64_t c1 = t < plo;
        const uint64_t u = k[q] + t;
        const uint64_t c2 = u < t;
        k[q] = u;
        carry = phi + c1 + c2;
    }
}

// a hash of a 256-bit key (SplitMix64's finalizer over the folded limbs)
ENG inline uint64_t key_hash(const uint64_t *k)
{
    uint64_t x = k[0] ^ (k[1] * 0x9e3779b97f4a7c15ULL) ^ (k[2] * 0xbf58476d1ce4e5b9ULL) ^
                 (k[3] * 0x94d049bb133111ebULL);
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// ---- project PRNG: SplitMix64-seeded xoshiro256** --------------------------
ENG inline uint64_t eng_splitmix(uint64_t *x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
ENG inline void eng_rng_seed(uint64_t *s, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = eng_splitmix(&x);
}
ENG inline uint64_t eng_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
ENG inline uint64_t eng_rng_next(uint64_t *s)
{
    uint64_t r = eng_rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = eng_rotl(s[3], 45);
    return r;
}


}  // namespace tci
