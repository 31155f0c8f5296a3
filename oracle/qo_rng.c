/*
 * qo_rng.c -- random-number plumbing of the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * (1) std::mt19937_64 and the three libstdc++ distributions the reference draws
 *     from: uniform_real_distribution<double>(0,1) (src/heatbath/heatbath.cpp:9),
 *     uniform_int_distribution<int>(0,L-1) (src/mpi/Shift.cpp:166) and
 *     normal_distribution<double>(0,1) (src/su3/utils.cpp:14).  Pinned in
 *     tests/test_oracle_rng.py against the C++ standard library of this image
 *     (a compiled probe) and the published mt19937_64 10000th-output value.
 * (2) Philox4x32-10 (Salmon et al., SC'11), the counter-based generator of the
 *     GPU path.  Pinned by the Random123 known-answer vectors.
 *
 * Philox counter scheme (identical in ecmc-qncb_amd/csrc/qcb_rng.hpp):
 *   key  = { seed & 0xffffffff, seed >> 32 }
 *   heatbath: ctr = { cb_site, sweep_idx, (hs<<24)|(hit<<8)|(subgroup<<4)|0, call }
 *             call a -> attempt a: (delta, r1, r2, r3) = u32(w0..w3);
 *             sphere of subgroup m = 3*hit + subgroup: call 0x80000000 at the
 *             counter of subgroup m - (m&1), (phi, cos theta) = u32(w[2(m&1)]),
 *             u32(w[2(m&1)+1]) -- two subgroups share one call;
 *             u32(w) = (w + 1/2) 2^-32, exact in double and strictly inside (0,1);
 *             the float path takes the float nearest to the same value, so both
 *             precisions see the same uniforms.
 *   hot start: ctr = { lex_site, mu, 1, call 0..8 }  (Box-Muller pairs)
 *              with u53(w0,w1), u53(w2,w3): top 52 bits of (hi:lo) plus half
 *              an ulp, strictly inside (0,1).
 */
#include "qo.h"

#include <math.h>
#include <time.h>

#define MT_N 312
#define MT_M 156

void qo_mt64_seed(qo_mt64* g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < MT_N; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = MT_N;
}

static void mt64_twist(qo_mt64* g) {
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    for (int i = 0; i < MT_N; ++i) {
        uint64_t y = (g->mt[i] & UM) | (g->mt[(i + 1) % MT_N] & LM);
        uint64_t v = g->mt[(i + MT_M) % MT_N] ^ (y >> 1);
        if (y & 1ULL) v ^= 0xB5026F5AA96619E9ULL;
        g->mt[i] = v;
    }
    g->idx = 0;
}

uint64_t qo_mt64_next(qo_mt64* g) {
    if (g->idx >= MT_N) mt64_twist(g);
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

/* generate_canonical<double, 53> on a 64-bit engine: a single draw divided by
 * 2^64 (the draw is first rounded to double); a result of 1 is pulled down to
 * the largest double below 1. */
double qo_mt64_canonical(qo_mt64* g) {
    double r = (double)qo_mt64_next(g) / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* uniform_int_distribution over [a,b] with a 64-bit engine: Lemire's
 * multiply-and-reject with a 128-bit product (range+1 < 2^64). */
uint64_t qo_mt64_uniform_int(qo_mt64* g, uint64_t a, uint64_t b) {
    uint64_t range = b - a + 1;
    unsigned __int128 prod = (unsigned __int128)qo_mt64_next(g) * range;
    uint64_t low = (uint64_t)prod;
    if (low < range) {
        uint64_t thresh = (0 - range) % range;
        while (low < thresh) {
            prod = (unsigned __int128)qo_mt64_next(g) * range;
            low = (uint64_t)prod;
        }
    }
    return a + (uint64_t)(prod >> 64);
}

/* normal_distribution: Marsaglia polar method; the first call of a fresh
 * distribution returns y*mult and keeps x*mult for the second call. */
void qo_mt64_normal_pair(qo_mt64* g, double* first, double* second) {
    double x, y, r2;
    do {
        x = 2.0 * qo_mt64_canonical(g) - 1.0;
        y = 2.0 * qo_mt64_canonical(g) - 1.0;
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    double mult = sqrt(-2.0 * log(r2) / r2);
    *first = y * mult;
    *second = x * mult;
}

void qo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double qo_u53(uint32_t hi, uint32_t lo) {
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 12) * 0x1.0p-52 + 0x1.0p-53;
}

double qo_u32(uint32_t w) { return ((double)w + 0.5) * 0x1.0p-32; }

double qo_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
