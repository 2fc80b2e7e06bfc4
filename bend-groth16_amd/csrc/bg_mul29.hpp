// Montgomery product a * b * 2^-256 mod p for 8 x 32-bit word operands, computed with
// 29-bit digits (reduced-radix product scanning).
//
// Each operand is split into 9 digits of 29 bits (a is taken as a * 2^5, so that a
// radix-2^29 Montgomery reduction, which divides by 2^261, yields a * b * 2^-256).  A
// digit product is < 2^58, so a column of up to 18 of them plus the carried-in column
// fits one 64-bit accumulator: every partial product is one v_mad_u64_u32 with no carry
// fold, and no carry passes through scalar registers.  The reduction digit of column k is
// m_k = (acc * -p^-1) mod 2^29, formed in the same column.  The result is packed back to
// 8 words.
//
// Range: for a, b < 2p (p < 2^254): a * 2^5 < 2^261 (9 digits), and the result
// t < (32a * b) / 2^261 + p < 128 p^2 / 2^261 + p < 2p, so values stay in [0, 2p) with no
// final subtraction, as the 32-bit-digit product it replaces.
//
// Plain C++ on purpose: the same code runs on the host in the unit test of the digit
// arithmetic.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define BG_HD __host__ __device__ __forceinline__
#else
#define BG_HD inline
#endif

namespace bg {
namespace m29 {

static constexpr uint32_t MASK = (1u << 29) - 1;

struct Digits {
  uint32_t d[9];
};

// digits of w * 2^S (S = 0 or 5) for a 256-bit w (w * 2^S < 2^261)
template <int S>
BG_HD Digits split(const uint32_t* w) {
  Digits r;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int b = 29 * i - S;  // first bit of digit i in w
    uint32_t v;
    if (b < 0) {
      v = w[0] << (-b);
    } else {
      const int wi = b >> 5, sh = b & 31;
      v = w[wi] >> sh;
      if (sh > 3 && wi + 1 < 8) v |= w[wi + 1] << (32 - sh);
    }
    r.d[i] = v & MASK;
  }
  return r;
}

// 8 words of a 9-digit value below 2^256 (digits below 2^29)
BG_HD void pack(const Digits& a, uint32_t* w) {
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int b = 32 * j, li = b / 29, sh = b % 29;
    uint32_t v = a.d[li] >> sh;
    if (li + 1 < 9) v |= a.d[li + 1] << (29 - sh);
    if (sh > 26 && li + 2 < 9) v |= a.d[li + 2] << (58 - sh);
    w[j] = v;
  }
}

// The column accumulator is a template parameter only so that the host unit test can run
// the products with an accumulator that reports overflow; the device uses uint64_t.

// x * y * 2^-261 mod p (not reduced below p); p29: p in 29-bit digits, n0 = -p^-1 mod 2^29
template <class Acc = uint64_t>
BG_HD Digits mont(const Digits& x, const Digits& y, const uint32_t* p29, uint32_t n0) {
  uint32_t m[9];
  Digits r;
  Acc acc = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = 0; i < k; i++) {
      acc += (uint64_t)x.d[i] * y.d[k - i];
      acc += (uint64_t)m[i] * p29[k - i];
    }
    acc += (uint64_t)x.d[k] * y.d[0];
    m[k] = ((uint32_t)acc * n0) & MASK;
    acc += (uint64_t)m[k] * p29[0];
    acc >>= 29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int i = k - 8; i < 9; i++) {
      acc += (uint64_t)x.d[i] * y.d[k - i];
      acc += (uint64_t)m[i] * p29[k - i];
    }
    r.d[k - 9] = (uint32_t)acc & MASK;
    acc >>= 29;
  }
  r.d[8] = (uint32_t)acc;
  return r;
}

// (x1 * y1 + x2 * y2) * 2^-261 mod p, one reduction for the sum of two products (Fq2
// products).  A column holds up to 18 digit products and 9 reduction products, each
// below 2^58: below 2^63 with the carry.  For x_k = 32 a_k, a_k, y_k < 2p the result is
// below 256 p^2 / 2^261 + p < 2.6 p (not reduced below 2p).
template <class Acc = uint64_t>
BG_HD Digits mont2(const Digits& x1, const Digits& y1, const Digits& x2, const Digits& y2, const uint32_t* p29,
                   uint32_t n0) {
  uint32_t m[9];
  Digits r;
  Acc acc = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = 0; i < k; i++) {
      acc += (uint64_t)x1.d[i] * y1.d[k - i];
      acc += (uint64_t)x2.d[i] * y2.d[k - i];
      acc += (uint64_t)m[i] * p29[k - i];
    }
    acc += (uint64_t)x1.d[k] * y1.d[0];
    acc += (uint64_t)x2.d[k] * y2.d[0];
    m[k] = ((uint32_t)acc * n0) & MASK;
    acc += (uint64_t)m[k] * p29[0];
    acc >>= 29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int i = k - 8; i < 9; i++) {
      acc += (uint64_t)x1.d[i] * y1.d[k - i];
      acc += (uint64_t)x2.d[i] * y2.d[k - i];
      acc += (uint64_t)m[i] * p29[k - i];
    }
    r.d[k - 9] = (uint32_t)acc & MASK;
    acc >>= 29;
  }
  r.d[8] = (uint32_t)acc;
  return r;
}

// (x1 y1 + x2 y2 + x3 y3 + x4 y4) * 2^-261 mod p, one reduction for four products (the
// G2 y coordinate).  A column holds up to 36 digit products and 9 reduction products,
// each below 2^58: below 2^63.5 with the carry.  For x_k = 32 a_k, a_k, y_k < 2p the
// result is below 512 p^2 / 2^261 + p < 4.1 p.
template <class Acc = uint64_t>
BG_HD Digits mont4(const Digits* x, const Digits* y, const uint32_t* p29, uint32_t n0) {
  uint32_t m[9];
  Digits r;
  Acc acc = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = 0; i < k; i++) {
#pragma unroll
      for (int q = 0; q < 4; q++) acc += (uint64_t)x[q].d[i] * y[q].d[k - i];
      acc += (uint64_t)m[i] * p29[k - i];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) acc += (uint64_t)x[q].d[k] * y[q].d[0];
    m[k] = ((uint32_t)acc * n0) & MASK;
    acc += (uint64_t)m[k] * p29[0];
    acc >>= 29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int i = k - 8; i < 9; i++) {
#pragma unroll
      for (int q = 0; q < 4; q++) acc += (uint64_t)x[q].d[i] * y[q].d[k - i];
      acc += (uint64_t)m[i] * p29[k - i];
    }
    r.d[k - 9] = (uint32_t)acc & MASK;
    acc >>= 29;
  }
  r.d[8] = (uint32_t)acc;
  return r;
}

// x^2 * 2^-261 mod p for x with digits below 2^29: each cross product x_i x_j (i < j)
// once, against the doubled digit (below 2^30), plus the squares on the diagonal -- 45
// digit products instead of 81 (the 81 reduction products stay).
template <class Acc = uint64_t>
BG_HD Digits msqr(const Digits& x, const uint32_t* p29, uint32_t n0) {
  uint32_t x2[9], m[9];
#pragma unroll
  for (int i = 0; i < 9; i++) x2[i] = x.d[i] << 1;
  Digits r;
  Acc acc = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) {
#pragma unroll
    for (int i = 0; i < k; i++) acc += (uint64_t)m[i] * p29[k - i];
#pragma unroll
    for (int i = 0; 2 * i < k; i++) acc += (uint64_t)x.d[i] * x2[k - i];
    if ((k & 1) == 0) acc += (uint64_t)x.d[k / 2] * x.d[k / 2];
    m[k] = ((uint32_t)acc * n0) & MASK;
    acc += (uint64_t)m[k] * p29[0];
    acc >>= 29;
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
#pragma unroll
    for (int i = k - 8; i < 9; i++) acc += (uint64_t)m[i] * p29[k - i];
#pragma unroll
    for (int i = k - 8; 2 * i < k; i++) acc += (uint64_t)x.d[i] * x2[k - i];
    if ((k & 1) == 0) acc += (uint64_t)x.d[k / 2] * x.d[k / 2];
    r.d[k - 9] = (uint32_t)acc & MASK;
    acc >>= 29;
  }
  r.d[8] = (uint32_t)acc;
  return r;
}

// a * b * 2^-256 mod p on 8-word operands
BG_HD void mul_words(const uint32_t* a, const uint32_t* b, uint32_t* out, const uint32_t* p29, uint32_t n0) {
  pack(mont(split<5>(a), split<0>(b), p29, n0), out);
}

// (a1 * b1 + a2 * b2) * 2^-256 mod p on 8-word operands below 2p, result below 2.6 p
BG_HD void mul2_words(const uint32_t* a1, const uint32_t* b1, const uint32_t* a2, const uint32_t* b2, uint32_t* out,
                      const uint32_t* p29, uint32_t n0) {
  pack(mont2(split<5>(a1), split<0>(b1), split<5>(a2), split<0>(b2), p29, n0), out);
}

// (sum of a_k * b_k, k < 4) * 2^-256 mod p on 8-word operands below 2p, result below 4.1 p
BG_HD void mul4_words(const uint32_t* const* a, const uint32_t* const* b, uint32_t* out, const uint32_t* p29,
                      uint32_t n0) {
  Digits x[4], y[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    x[q] = split<5>(a[q]);
    y[q] = split<0>(b[q]);
  }
  pack(mont4(x, y, p29, n0), out);
}

}  // namespace m29
}  // namespace bg
