// Accuracy of galhip_math.h's gh_log_ge1 and gh_atan2, built for the host
// with the same restriction as the device build (no contraction): reports
// the maximum error in double ulps against 64-bit-mantissa references
// (logl, atan2l), and how often the FP32 rounding of the value the cell
// index uses differs from the one formed with glibc's double log/atan2.
// Usage: math_accuracy <samples> <seed>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../galaxy_amd/csrc/galhip_math.h"

static uint64_t rng_state;
static uint64_t next_u64() {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return rng_state;
}
static double uni() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

static double ulp_err(double got, long double ref) {
    if (ref == 0.0L) return got == 0.0 ? 0.0 : INFINITY;
    int e;
    frexpl(ref, &e);
    long double ulp = ldexpl(1.0L, e - 53);
    return (double)(fabsl((long double)got - ref) / ulp);
}

int main(int argc, char **argv) {
    long n = argc > 1 ? atol(argv[1]) : 1000000;
    rng_state = argc > 2 ? strtoull(argv[2], 0, 10) : 88172645463325252ULL;
    double max_log = 0, max_atan = 0;
    long mis_log = 0, mis_t = 0, mis_u = 0;
    const int na = 512;
    const float alpha = (float)(2.0 * 3.141592653589793238462643 / na);
    for (long i = 0; i < n; i++) {
        // radii spread over [0, 1e4] on a log scale, as float
        float r = (float)(expm1(uni() * 9.3));
        if (i % 7 == 0) r = (float)(uni() * 1e-3);
        float v = 1.0f + r;
        double lg = gh_log_ge1((double)v);
        double e1 = ulp_err(lg, logl((long double)v));
        if (e1 > max_log) max_log = e1;
        if ((float)lg != (float)log((double)v)) mis_log++;
        float u1 = (float)lg / alpha, u2 = (float)log((double)v) / alpha;
        if (u1 != u2) mis_u++;
        float x = (float)((uni() * 2 - 1) * 200), y = (float)((uni() * 2 - 1) * 200);
        if (i % 11 == 0) y = (float)((uni() * 2 - 1) * 1e-6);
        if (i % 13 == 0) x = -x * 1e-7f;
        double at = gh_atan2((double)y, (double)x);
        double e2 = ulp_err(at, atan2l((long double)y, (long double)x));
        if (e2 > max_atan) max_atan = e2;
        float t1 = (float)(at * (double)na / (2.0 * 3.141592653589793238462643));
        float t2 = (float)(atan2((double)y, (double)x) * (double)na / (2.0 * 3.141592653589793238462643));
        if (t1 != t2) mis_t++;
    }
    // signed zeros
    int zeros_ok = gh_atan2(0.0, 1.0) == 0.0 && !signbit(gh_atan2(0.0, 1.0)) &&
                   signbit(gh_atan2(-0.0, 1.0)) && gh_atan2(0.0, -1.0) == atan2(0.0, -1.0) &&
                   gh_atan2(-0.0, -1.0) == atan2(-0.0, -1.0) &&
                   gh_atan2(0.0, -0.0) == atan2(0.0, -0.0) &&
                   gh_atan2(1.0, 0.0) == atan2(1.0, 0.0) && gh_atan2(-1.0, 0.0) == atan2(-1.0, 0.0);
    printf("%ld %.4f %.4f %ld %ld %ld %d\n", n, max_log, max_atan, mis_log, mis_u, mis_t, zeros_ok);
    return 0;
}
