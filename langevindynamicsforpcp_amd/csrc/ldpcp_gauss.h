/*
 * ldpcp_gauss.h -- FP64 Box-Muller transform built from IEEE operations only.
 *
 * The Gaussian noise of the FP64 paths is a Box-Muller pair
 *     r = sqrt(-2 ln u1),  g0 = r cos(2 pi u2),  g1 = r sin(2 pi u2)
 * of two Philox words mapped to u = (w + 1/2) 2^-32.  The logarithm and the
 * sine/cosine here use only +, -, *, fma, division, sqrt, rint and exponent
 * bit manipulation, each of which IEEE 754 rounds exactly one way, so the
 * device code and a host build of this same header produce the same bits.
 * (Library transcendentals differ between the GPU's math library and the host
 * C library by an ulp on some inputs, which is enough to move a barrier
 * crossing by a step.)  Accuracy: a few ulp against a correctly rounded
 * result over the whole input range.
 *
 * Plain C99/C++ (no HIP types): hipcc compiles it for host and device, a C
 * compiler for the host alone.  Build every includer with
 * -ffp-contract=off: the fused operations below are explicit fma() calls.
 */
#ifndef LDPCP_GAUSS_H
#define LDPCP_GAUSS_H
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define LDPCP_HD __host__ __device__ inline
#else
#define LDPCP_HD static inline
#endif

LDPCP_HD uint64_t ldpcp_f64_bits(double x)
{
    uint64_t b;
    memcpy(&b, &x, sizeof(b));
    return b;
}

LDPCP_HD double ldpcp_bits_f64(uint64_t b)
{
    double x;
    memcpy(&x, &b, sizeof(x));
    return x;
}

/* (w + 1/2) / 2^32: exact, in (0, 1). */
LDPCP_HD double ldpcp_unit_f64(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

/*
 * Natural logarithm of a positive normal x.
 * x = 2^k m with m in [sqrt(1/2), sqrt(2)); with f = (m - 1)/(m + 1),
 * ln m = 2 atanh f = 2 (f + f^3/3 + f^5/5 + ...), |f| <= 0.1716, f^2 <= 0.0295;
 * the series through f^25 leaves a relative truncation error below 1e-19.
 * ln x = k ln2_hi + (k ln2_lo + ln m), ln2_hi with 32 significant bits so that
 * k ln2_hi is exact for |k| < 2^21.
 */
LDPCP_HD double ldpcp_log(double x)
{
    const uint64_t b = ldpcp_f64_bits(x);
    int k = (int)((b >> 52) & 0x7ff) - 1023;
    double m = ldpcp_bits_f64((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull); /* [1, 2) */
    if (m >= 1.4142135623730951) {
        m = m * 0.5;
        k += 1;
    }
    const double f = (m - 1.0) / (m + 1.0);
    const double z = f * f;
    /* 1/(2j+1), j = 12 .. 1, Horner in z */
    double p = 0.04; /* 1/25 */
    p = fma(p, z, 0.043478260869565216);  /* 1/23 */
    p = fma(p, z, 0.047619047619047616);  /* 1/21 */
    p = fma(p, z, 0.05263157894736842);  /* 1/19 */
    p = fma(p, z, 0.058823529411764705);  /* 1/17 */
    p = fma(p, z, 0.06666666666666667);  /* 1/15 */
    p = fma(p, z, 0.07692307692307693);  /* 1/13 */
    p = fma(p, z, 0.09090909090909091);  /* 1/11 */
    p = fma(p, z, 0.1111111111111111);    /* 1/9 */
    p = fma(p, z, 0.14285714285714285);   /* 1/7 */
    p = fma(p, z, 0.2);                   /* 1/5 */
    p = fma(p, z, 0.3333333333333333);   /* 1/3 */
    /* ln m = 2f + 2f z p */
    const double f2 = f + f;
    const double lnm = fma(f2 * z, p, f2);
    const double ln2_hi = 0.6931471803691238;     /* 32 significant bits */
    const double ln2_lo = 1.9082149292705877e-10;  /* ln 2 - ln2_hi */
    const double dk = (double)k;
    return fma(dk, ln2_hi, fma(dk, ln2_lo, lnm));
}

/*
 * sin(2 pi u) and cos(2 pi u) for u in [0, 1] a multiple of 2^-52 (any
 * u = (w + 1/2) 2^-32 is).  The quadrant comes out exactly in turns:
 * n = rint(4u), r = u - n/4 in [-1/8, 1/8] (exact), so the
 * polynomials see |2 pi r| <= pi/4 and the Taylor series through degree 19
 * (sine) and 18 (cosine) leave truncation errors below 1e-18.
 */
LDPCP_HD void ldpcp_sincos_turn(double u, double* s_out, double* c_out)
{
    const double n = rint(4.0 * u);
    const double r = u - 0.25 * n;
    const double t = 6.283185307179586 * r; /* 2 pi r */
    const double z = t * t;
    /* sin t = t (1 - z/3! + z^2/5! - ...), through t^19 */
    double ps = -8.22063524662433e-18;     /* -1/19! */
    ps = fma(ps, z, 2.8114572543455206e-15);  /* 1/17! */
    ps = fma(ps, z, -7.647163731819816e-13); /* -1/15! */
    ps = fma(ps, z, 1.6059043836821613e-10);  /* 1/13! */
    ps = fma(ps, z, -2.505210838544172e-08);  /* -1/11! */
    ps = fma(ps, z, 2.7557319223985893e-06);  /* 1/9! */
    ps = fma(ps, z, -0.0001984126984126984); /* -1/7! */
    ps = fma(ps, z, 0.008333333333333333);   /* 1/5! */
    ps = fma(ps, z, -0.16666666666666666);    /* -1/3! */
    const double sn = fma(t * z, ps, t);
    /* cos t = 1 - z/2! + z^2/4! - ..., through t^18 */
    double pc = 1.5619206968586225e-16;      /* 1/18! */
    pc = fma(pc, z, -4.779477332387385e-14); /* -1/16! */
    pc = fma(pc, z, 1.1470745597729725e-11);  /* 1/14! */
    pc = fma(pc, z, -2.08767569878681e-09);   /* -1/12! */
    pc = fma(pc, z, 2.755731922398589e-07);  /* 1/10! */
    pc = fma(pc, z, -2.48015873015873e-05); /* -1/8! */
    pc = fma(pc, z, 0.001388888888888889);   /* 1/6! */
    pc = fma(pc, z, -0.041666666666666664);   /* -1/4! */
    pc = fma(pc, z, 0.5);                     /* 1/2! (subtracted below) */
    const double cs = fma(-z, pc, 1.0);
    /* rotate by the quadrant n (mod 4) */
    const int q = ((int)n) & 3;
    double s, c;
    if (q == 0) {
        s = sn;
        c = cs;
    } else if (q == 1) {
        s = cs;
        c = -sn;
    } else if (q == 2) {
        s = -sn;
        c = -cs;
    } else {
        s = -cs;
        c = sn;
    }
    *s_out = s;
    *c_out = c;
}

/* Box-Muller pair of two Philox words: g0 = r cos(2 pi u2), g1 = r sin(2 pi u2). */
LDPCP_HD void ldpcp_box_muller(uint32_t a, uint32_t b, double* g0, double* g1)
{
    const double u1 = ldpcp_unit_f64(a), u2 = ldpcp_unit_f64(b);
    const double r = sqrt(-2.0 * ldpcp_log(u1));
    double s, c;
    ldpcp_sincos_turn(u2, &s, &c);
    *g0 = r * c;
    *g1 = r * s;
}

#undef LDPCP_HD
#endif
