/*
 * orc_rng.c -- std::mt19937 and libstdc++'s uniform_real_distribution, as
 * used by the reference tests to make their inputs (tests/grid2d_frnn.cu:211-
 * 218, tests/test_frnn.cu:21-28, tests/test_closest_neighbors.cu:21-28)
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * mt19937: Matsumoto & Nishimura (1998), parameters of the C++ standard
 * [rand.predef]; the standard's check value (10000th output of the default
 * seed 5489 is 4123659995) pins it.
 * uniform_real_distribution<T>(a, b): libstdc++ computes
 *   generate_canonical<T, digits(T)>: k = ceil(digits / 32) draws,
 *   sum = sum_i T(x_i) * 2^(32 i), u = sum / 2^(32 k), u = nextafter(1, 0) if u >= 1,
 *   then returns u * (b - a) + a in T arithmetic.
 */
#include "oracle.h"
#include <math.h>

void orc_mt_seed(orc_mt19937* g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

uint32_t orc_mt_next(orc_mt19937* g) {
    if (g->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

float orc_uniform_float(orc_mt19937* g, float a, float b) {
    float sum = (float)orc_mt_next(g);
    float u = sum / 4294967296.0f;
    if (u >= 1.0f) u = nextafterf(1.0f, 0.0f);
    return u * (b - a) + a;
}

double orc_uniform_double(orc_mt19937* g, double a, double b) {
    double x0 = (double)orc_mt_next(g);
    double x1 = (double)orc_mt_next(g);
    double sum = x0 + x1 * 4294967296.0;
    double u = sum / 18446744073709551616.0;
    if (u >= 1.0) u = nextafter(1.0, 0.0);
    return u * (b - a) + a;
}

void orc_mt_uniform_floats(uint32_t seed, float a, float b, int n, float* out) {
    orc_mt19937 g;
    orc_mt_seed(&g, seed);
    for (int i = 0; i < n; i++) out[i] = orc_uniform_float(&g, a, b);
}
void orc_mt_uniform_doubles(uint32_t seed, double a, double b, int n, double* out) {
    orc_mt19937 g;
    orc_mt_seed(&g, seed);
    for (int i = 0; i < n; i++) out[i] = orc_uniform_double(&g, a, b);
}
void orc_mt_raw(uint32_t seed, int n, uint32_t* out) {
    orc_mt19937 g;
    orc_mt_seed(&g, seed);
    for (int i = 0; i < n; i++) out[i] = orc_mt_next(&g);
}
