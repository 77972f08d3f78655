// glibc_pow.h -- x^k for a small non-negative integer k, bit-identical to
// glibc 2.35's pow(x, (double)k) on x86-64 (the FMA variant, selected at run
// time on every CPU with FMA).
//
// The reference evaluates its Legendre polynomials as monomial sums with
// pow(x, i) (Jacobi.cpp:43,53,63,73), and the GLL cardinal functions built on
// them (Legendre.cpp:145-169) are ill-conditioned next to a GLL line: a
// one-ulp difference in one power changes an interpolated value by up to
// 1e-8 relative and can move a departure point across an element edge.  So
// the device evaluates the powers with the same algorithm as glibc:
//
//   glibc's double pow since 2.28 (sysdeps/ieee754/dbl-64/e_pow.c, from ARM's
//   optimized-routines, Szabolcs Nagy 2018): log(x) to ~2^-68 relative as
//   hi + lo through a 128-entry table of (1/c, log c) and a degree-8
//   polynomial of z/c - 1; y*log(x) as a double-double; exp through a
//   128-entry table of 2^(k/128) and a degree-5 polynomial; one final
//   rounding scale + scale*tmp.
//
// The operation order, the explicit FMAs of the FMA build and the products
// that compiler fuses into FMAs are reproduced (measured black-box against
// the host's pow: 3e8 random and 1.2e7 targeted inputs, x in [-1, 1],
// k = 1..8, no difference; the check is tests/test_pow_glibc.py).  The
// tables come from gen_pow_tables.py; the polynomial coefficients are the
// published ones.  Everything outside the algorithm's own range reduction
// (NaN, 0, infinities, negative x, subnormal x and results) takes glibc's
// special-case paths, restated here for integer exponents only.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SE_POW_FN __device__ __forceinline__
#define SE_POW_TAB static __device__ const
#else
#define SE_POW_FN static inline
#define SE_POW_TAB static const
#include <cstring>
#endif

#include "glibc_pow_tables.h"

namespace se_pow {

SE_POW_FN uint64_t bits(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint64_t)__double_as_longlong(x);
#else
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
#endif
}

SE_POW_FN double from_bits(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __longlong_as_double((long long)u);
#else
    double x;
    std::memcpy(&x, &u, 8);
    return x;
#endif
}

// log(x) = hi + lo for the positive normal bit pattern ix
SE_POW_FN double log_hi_lo(uint64_t ix, double* tail) {
#pragma clang fp contract(off)
    // log1p polynomial, already scaled for the evaluation below
    const double A0 = -0x1p-1, A1 = 0x1.555555555556p-2 * -2, A2 = -0x1.0000000000006p-2 * -2,
                 A3 = 0x1.999999959554ep-3 * 4, A4 = -0x1.555555529a47ap-3 * 4, A5 = 0x1.2495b9b4845e9p-3 * -8,
                 A6 = -0x1.0002b8b263fc3p-3 * -8;
    const double Ln2hi = 0x1.62e42fefa3800p-1, Ln2lo = 0x1.ef35793c76730p-45;
    const uint64_t OFF = 0x3fe6955500000000ULL;
    const uint64_t tmp = ix - OFF;
    const int i = (int)((tmp >> 45) % 128);
    const int k = (int)((int64_t)tmp >> 52);
    const uint64_t iz = ix - (tmp & (0xfffULL << 52));
    const double z = from_bits(iz), kd = (double)k;
    const double invc = kPowLogInvc[i], logc = kPowLogLogc[i], logctail = kPowLogTail[i];
    const double r = fma(z, invc, -1.0);
    const double t1 = fma(kd, Ln2hi, logc);
    const double t2 = t1 + r;
    const double lo1 = fma(kd, Ln2lo, logctail);
    const double lo2 = t1 - t2 + r;
    const double ar = A0 * r;
    const double ar2 = r * ar;
    const double ar3 = r * ar2;
    const double hi = t2 + ar2;
    const double lo3 = fma(ar, r, -ar2);
    const double lo4 = t2 - hi + ar2;
    const double q = fma(ar2, fma(ar2, fma(r, A6, A5), fma(r, A4, A3)), fma(r, A2, A1));
    const double lo = fma(ar3, q, lo1 + lo2 + lo3 + lo4);
    const double y = hi + lo;
    *tail = hi - y + lo;
    return y;
}

// exp(x + xtail) with the sign of the result in sign_bias
SE_POW_FN double exp_hi_lo(double x, double xtail, uint64_t sign_bias) {
#pragma clang fp contract(off)
    const double InvLn2N = 0x1.71547652b82fep0 * 128, Shift = 0x1.8p52, NegLn2hiN = -0x1.62e42fefa0000p-8,
                 NegLn2loN = -0x1.cf79abc9e3b3ap-47;
    const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
                 C5 = 0x1.1111167a4d017p-7;
    uint32_t abstop = (uint32_t)(bits(x) >> 52) & 0x7ff;
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {  // |x| < 2^-54 or |x| >= 512
        if (abstop - 0x3c9u >= 0x80000000u) {
            const double one = 1.0 + x;
            return sign_bias ? -one : one;
        }
        if (abstop >= 0x409u) {  // |x| >= 1024: overflow / underflow
            if (bits(x) >> 63) return sign_bias ? -0.0 : 0.0;
            return sign_bias ? -HUGE_VAL : HUGE_VAL;
        }
        abstop = 0;
    }
    double kd = fma(InvLn2N, x, Shift);
    const uint64_t ki = bits(kd);
    kd -= Shift;
    double r = fma(kd, NegLn2loN, fma(kd, NegLn2hiN, x));
    r += xtail;
    const uint64_t idx = 2 * (ki % 128);
    const uint64_t top = (ki + sign_bias) << 45;
    const double tail = from_bits(kPowExpTab[idx]);
    uint64_t sbits = kPowExpTab[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
    if (abstop == 0) {
        // scale's exponent may be out of range: glibc's specialcase (no FMAs)
        if ((ki & 0x80000000ULL) == 0) {
            sbits -= 1009ULL << 52;
            const double scale = from_bits(sbits);
            return 0x1p1009 * (scale + scale * tmp);
        }
        sbits += 1022ULL << 52;
        const double scale = from_bits(sbits);
        double y = scale + scale * tmp;
        if (fabs(y) < 1.0) {  // subnormal result: round once at the right precision
            const double one = y < 0.0 ? -1.0 : 1.0;
            double lo = scale - y + scale * tmp;
            const double hi = one + y;
            lo = one - hi + y + lo;
            y = (hi + lo) - one;
            if (y == 0.0) y = from_bits(sbits & 0x8000000000000000ULL);
        }
        return 0x1p-1022 * y;
    }
    const double scale = from_bits(sbits);
    return fma(scale, tmp, scale);
}

// x's part of pow(x, k) for integers k -- the special cases, the sign and
// log(x) = hi + lo -- so that several powers of one x share a single log;
// pow_int_at(pow_base(x), k) performs exactly pow_int(x, k)'s operations
struct PowBase {
    double x, hi, lo;
    uint64_t ix;
    int special, neg;
};

SE_POW_FN PowBase pow_base(double x) {
#pragma clang fp contract(off)
    PowBase b;
    b.x = x;
    b.hi = b.lo = 0.0;
    b.special = b.neg = 0;
    uint64_t ix = bits(x);
    const uint32_t topx = (uint32_t)(ix >> 52);
    if (topx - 0x001u >= 0x7ffu - 0x001u) {
        if (2 * ix - 1 >= 2 * 0x7ff0000000000000ULL - 1) {  // x = +-0, +-inf or NaN
            b.special = 1;
            b.ix = ix;
            return b;
        }
        if (ix >> 63) {  // finite x < 0 and an integer exponent
            b.neg = 1;
            ix &= 0x7fffffffffffffffULL;
        }
        if (((ix >> 52) & 0x7ff) == 0) {  // subnormal x: normalise
            ix = bits(from_bits(ix) * 0x1p52);
            ix &= 0x7fffffffffffffffULL;
            ix -= 52ULL << 52;
        }
    }
    b.ix = ix;
    b.hi = log_hi_lo(ix, &b.lo);
    return b;
}

SE_POW_FN double pow_int_at(const PowBase& b, int k) {
#pragma clang fp contract(off)
    if (k == 0) return 1.0;
    if (b.special) {
        double x2 = b.x * b.x;
        if ((b.ix >> 63) && (k & 1)) x2 = -x2;
        return x2;
    }
    const uint64_t sign_bias = (b.neg && (k & 1)) ? (0x800ULL << 7) : 0;
    const double y = (double)k;
    const double ehi = y * b.hi;
    const double elo = fma(y, b.lo, fma(y, b.hi, -ehi));
    return exp_hi_lo(ehi, elo, sign_bias);
}

// pow(x, k) for an integer 0 <= k < 2^53
SE_POW_FN double pow_int(double x, int k) {
    if (k == 0) return 1.0;
    return pow_int_at(pow_base(x), k);
}

}  // namespace se_pow
