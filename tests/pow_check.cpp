// Compares se_pow::pow_int (the device's power function, compiled here as host
// code) with the C library's pow(x, k) bit for bit.  Prints the mismatch count.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "glibc_pow.h"

static uint64_t state = 0x9E3779B97F4A7C15ULL;
static uint64_t next() {  // SplitMix64
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static long bad = 0, total = 0;
static void compare(double x, int k, double a, double b) {
    total++;
    if (std::memcmp(&a, &b, 8) != 0) {
        if (bad < 5) std::printf("x=%a k=%d libm=%a here=%a\n", x, k, a, b);
        bad++;
    }
}
static void check(double x, int k) { compare(x, k, pow(x, (double)k), se_pow::pow_int(x, k)); }
// every power 0..12 of x from one shared base (one log), as the Legendre
// derivative evaluates them
static void check_shared(double x) {
    const se_pow::PowBase b = se_pow::pow_base(x);
    for (int k = 0; k <= 12; k++) compare(x, k, pow(x, (double)k), se_pow::pow_int_at(b, k));
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
    for (long t = 0; t < n; t++) {
        double x = (double)(next() >> 11) * 0x1p-53 * 2.0 - 1.0;  // uniform [-1, 1)
        if (t % 4 == 1) x = std::ldexp(x, -(int)(next() % 40));
        check(x, 1 + (int)(next() % 12));
        if (t % 8 == 0) check_shared(x);
    }
    // around +-1, around GLL-like interior points, tiny and subnormal powers, specials
    const double centres[] = {1.0, -1.0, 0.5, 0.2852315164806451, -0.7650553239294647};
    for (int k = 0; k <= 12; k++) {
        for (double c : centres)
            for (long j = -2000; j <= 2000; j++) {
                uint64_t u;
                std::memcpy(&u, &c, 8);
                u += j;
                double x;
                std::memcpy(&x, &u, 8);
                if (std::fabs(x) <= 1.0) check(x, k);
            }
        for (int e = 0; e < 1080; e += 3) check(std::ldexp(1.37, -e), k), check(-std::ldexp(1.91, -e), k);
        const double specials[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 1.0, -1.0, 0x1p-1074, -0x1p-1060};
        for (double s : specials) check(s, k);
    }
    const double specials[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 1.0, -1.0, 0x1p-1074, -0x1p-1060, -0.5};
    for (double s : specials) check_shared(s);
    std::printf("mismatches %ld of %ld\n", bad, total);
    return bad != 0;
}
