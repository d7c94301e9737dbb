// Signed 64.64 fixed-point representation used for order-independent exact
// sums (the MLS force spreading and the mean of the spread forces).
//
// A double v is held as the 128-bit two's-complement integer
// trunc(v * 2^64) split into (hi, lo) 64-bit words.  The split works on the
// integer mantissa, so there is no floating-point rounding in it and no value
// (however close to an integer or to zero) can saturate a word.  Integer
// addition with carry is associative, so a sum of such values does not depend
// on the order of the additions.  |v| < 2^62 is assumed.
//
// The header has no HIP dependency so that the host test can compile it with
// a plain C++ compiler (MELT_HD defined empty).
#pragma once

#include <cmath>

#ifndef MELT_HD
#define MELT_HD __host__ __device__
#endif

MELT_HD inline void acc_split(double v, unsigned long long& lo, unsigned long long& hi) {
  lo = 0;
  hi = 0;
  if (v == 0.0) return;
  int e;
  double m = frexp(fabs(v), &e);                            // |v| = m 2^e, m in [0.5, 1)
  unsigned long long mi = (unsigned long long)ldexp(m, 53);  // exact 53-bit integer
  int sh = e + 11;                                           // |v| 2^64 = mi 2^sh
  if (sh >= 64) {
    hi = mi << (sh - 64);
  } else if (sh > 0) {
    lo = mi << sh;
    hi = mi >> (64 - sh);
  } else if (sh == 0) {
    lo = mi;
  } else if (sh > -64) {
    lo = mi >> (-sh);
  }
  if (v < 0.0) {
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0 ? 1ull : 0ull);
  }
}

// (lo, hi) += (l2, h2) modulo 2^128.
MELT_HD inline void acc_add_words(unsigned long long& lo, unsigned long long& hi, unsigned long long l2, unsigned long long h2) {
  unsigned long long nl = lo + l2;
  hi = hi + h2 + (nl < lo ? 1ull : 0ull);
  lo = nl;
}

// The double nearest (within one ulp) to the signed value (hi:lo) / 2^64.
MELT_HD inline double acc_value(unsigned long long lo, unsigned long long hi) {
  bool neg = (long long)hi < 0;
  if (neg) {
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0 ? 1ull : 0ull);
  }
  double f = (double)lo * 5.421010862427522170037264e-20;  // 2^-64
  double m = hi ? (double)hi + f : f;
  return neg ? -m : m;
}
