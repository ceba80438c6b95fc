// glibc_trig.h -- single-precision sine and cosine that return, bit for bit,
// what the reference's host build returns for sin/cos of a float.
//
// The reference computes DH frames with unqualified sin/cos on float
// arguments (prm/src/lib/collision4/cuda_head.h:25-28, geo4.h:184-187, 195-198);
// with g++ 11 / glibc 2.35 these bind to glibc's sinf/cosf (merged to sincosf).
// glibc 2.35's sinf/cosf (sysdeps/ieee754/flt-32/s_sinf.c, s_cosf.c, sincosf.h,
// sincosf_data.c, from ARM optimized-routines) evaluate a short polynomial in
// double precision after a Cody-Waite (|x| < 120) or table (|x| >= 120)
// reduction.  On x86-64 CPUs with FMA, glibc selects its FMA build, in which
// every `a*b + c` of that algorithm is one fused multiply-add.  This file
// restates that published algorithm with explicit fma(), so a GPU (or any
// host) reproduces the reference's floats exactly.  The polynomial and
// reduction constants are those of glibc's sincosf_data.c; they have to be
// identical for identical results.
//
// Pinned exhaustively against the glibc of the build container over every
// float (tests/test_trig.py); not correctly rounded, on purpose.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define GPRM_HD __host__ __device__ __forceinline__
#define GPRM_FMA(a, b, c) __builtin_fma((a), (b), (c))
#else
#include <math.h>
#define GPRM_HD static inline
#define GPRM_FMA(a, b, c) fma((a), (b), (c))
#endif

namespace gprm_trig {

struct Poly {
  double c0, c1, c2, c3, c4, s1, s2, s3;
};

// Cosine polynomial negated in the second entry (quadrants 2 and 3).
#define GPRM_TRIG_POLY(sgn)                                                                     \
  {                                                                                              \
    sgn 0x1p0, sgn - 0x1.ffffffd0c621cp-2, sgn 0x1.55553e1068f19p-5, sgn - 0x1.6c087e89a359dp-10, \
        sgn 0x1.99343027bf8c3p-16, -0x1.555545995a603p-3, 0x1.1107605230bc4p-7,                  \
        -0x1.994eb3774cf24p-13                                                                   \
  }

GPRM_HD uint32_t asu(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
GPRM_HD int abstop12(float x) { return (int)((asu(x) >> 20) & 0x7ff); }

GPRM_HD Poly poly_table(int second) {
  Poly p0 = GPRM_TRIG_POLY(+);
  Poly p1 = GPRM_TRIG_POLY(-);
  return second ? p1 : p0;
}

// sinf_poly: n even -> sine polynomial, odd -> cosine polynomial.
GPRM_HD float sinf_poly(double x, double x2, const Poly& p, int n) {
  if ((n & 1) == 0) {
    double x3 = x * x2;
    double s1 = GPRM_FMA(x2, p.s3, p.s2);
    double x7 = x3 * x2;
    double s = GPRM_FMA(x3, p.s1, x);
    return (float)GPRM_FMA(x7, s1, s);
  } else {
    double x4 = x2 * x2;
    double c2 = GPRM_FMA(x2, p.c4, p.c3);
    double c1 = GPRM_FMA(x2, p.c1, p.c0);
    double x6 = x4 * x2;
    double c = GPRM_FMA(x4, p.c2, c1);
    return (float)GPRM_FMA(x6, c2, c);
  }
}

// |x| < 120: x - n*pi/2 with n from a 2^24-scaled 2/pi product.
GPRM_HD double reduce_fast(double x, int* np) {
  const double hpi_inv = 0x1.45F306DC9C883p+23;
  const double hpi = 0x1.921FB54442D18p0;
  double r = x * hpi_inv;
  int n = ((int32_t)r + 0x800000) >> 24;
  *np = n;
  return GPRM_FMA(-(double)n, hpi, x);
}

// |x| >= 120: 2.62 fixed-point product with 2/pi bits (8-bit stepped windows).
GPRM_HD double reduce_large(uint32_t xi, int* np) {
  const uint32_t inv[24] = {0x000000a2, 0x0000a2f9, 0x00a2f983, 0xa2f9836e, 0xf9836e4e, 0x836e4e44,
                            0x6e4e4415, 0x4e441529, 0x441529fc, 0x1529fc27, 0x29fc2757, 0xfc2757d1,
                            0x2757d1f5, 0x57d1f534, 0xd1f534dd, 0xf534ddc0, 0x34ddc0db, 0xddc0db62,
                            0xc0db6295, 0xdb629599, 0x6295993c, 0x95993c43, 0x993c4390, 0x3c439041};
  const uint32_t* arr = &inv[(xi >> 26) & 15];
  int shift = (xi >> 23) & 7;
  xi = (xi & 0xffffff) | 0x800000;
  xi <<= shift;
  uint64_t res0 = (uint32_t)(xi * arr[0]);
  uint64_t res1 = (uint64_t)xi * arr[4];
  uint64_t res2 = (uint64_t)xi * arr[8];
  res0 = (res2 >> 32) | (res0 << 32);
  res0 += res1;
  uint64_t n = (res0 + (1ULL << 61)) >> 62;
  res0 -= n << 62;
  double x = (double)(int64_t)res0;
  *np = (int)n;
  return x * 0x1.921FB54442D18p-62;
}

GPRM_HD float sinf_ref(float y) {
  double x = y;
  int n;
  const float pio4 = 0x1.921FB6p-1f;
  if (abstop12(y) < abstop12(pio4)) {
    double s = x * x;
    if (abstop12(y) < abstop12(0x1p-12f)) return y;
    return sinf_poly(x, s, poly_table(0), 0);
  } else if (abstop12(y) < abstop12(120.0f)) {
    x = reduce_fast(x, &n);
    double s = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0;
    return sinf_poly(x * s, x * x, poly_table(n & 2), n);
  } else if (abstop12(y) < abstop12(__builtin_inff())) {
    uint32_t xi = asu(y);
    int sign = (int)(xi >> 31);
    x = reduce_large(xi, &n);
    int t = (n + sign) & 3;
    double s = (t == 1 || t == 2) ? -1.0 : 1.0;
    return sinf_poly(x * s, x * x, poly_table(t & 2), n);
  }
  return (y - y) / (y - y);
}

GPRM_HD float cosf_ref(float y) {
  double x = y;
  int n;
  const float pio4 = 0x1.921FB6p-1f;
  if (abstop12(y) < abstop12(pio4)) {
    double x2 = x * x;
    if (abstop12(y) < abstop12(0x1p-12f)) return 1.0f;
    return sinf_poly(x, x2, poly_table(0), 1);
  } else if (abstop12(y) < abstop12(120.0f)) {
    x = reduce_fast(x, &n);
    double s = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0;
    return sinf_poly(x * s, x * x, poly_table(n & 2), n ^ 1);
  } else if (abstop12(y) < abstop12(__builtin_inff())) {
    uint32_t xi = asu(y);
    int sign = (int)(xi >> 31);
    x = reduce_large(xi, &n);
    int t = (n + sign) & 3;
    double s = (t == 1 || t == 2) ? -1.0 : 1.0;
    return sinf_poly(x * s, x * x, poly_table(t & 2), n ^ 1);
  }
  return (y - y) / (y - y);
}

#undef GPRM_TRIG_POLY

}  // namespace gprm_trig
