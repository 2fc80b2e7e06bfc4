// Digit-domain field elements for the bucket accumulation: 9 digits of 29 bits held in
// registers from one Montgomery product to the next, with no 8-word packing in between.
//
// An element x is held as an integer D = x * 2^261 (mod p), D < 2^261, in digits d_0..d_8
// (bits 29i ..); the product of bg_mul29.hpp then maps digit-domain operands to a
// digit-domain result (D1 D2 / 2^261 = x1 x2 2^261), so the 8-word split and pack around
// every product (about 40 % of a product's non-multiply instructions) is paid only where an
// element enters or leaves the accumulation.  Sums and differences are formed lazily,
// digit by digit: a difference a - b is a + B - b with B a multiple of p whose digits are
// raised (borrowing from the digit above) so that no digit goes negative; carries are
// propagated only where a product needs normalised digits.
//
// Bounds.  "Normalised" = d_0..d_7 < 2^29 (d_8 holds the rest).  A product's column adds
// at most 9 digit products and 9 reduction products in a 64-bit accumulator: with one
// operand normalised and the other's digits below 2^31 it stays below 2^63.5 (mont2: with
// one operand of each pair normalised and the other's digits below 1.5 * 2^30, below
// 2^64).  The value of a product's result is below p + a b / 2^261 = p (1 + alpha beta
// eps) for operands below alpha p and beta p, eps = p / 2^261 = 0.00591 for both BN254
// fields; each formula below states the value bound of its terms.
//
// Word-domain <-> digit-domain: a word-domain value W = x 2^256 (8 words, below 2p) gives
// the digit-domain integer 32 W (the 29-bit split of W shifted by 5), which is a valid
// product operand (below 64 p) but not yet reduced; a product with 2^261 mod p reduces it
// below 2p.  Back: the product with 2^256 mod p gives W = D / 32 (mod p) below 2p, packed
// to 8 words.
//
// Plain C++ on purpose: the host unit test runs the same code.
#pragma once
#include <stdint.h>
#include "bg_mul29.hpp"

namespace bg {
namespace fd {

using Fd = m29::Digits;
static constexpr uint32_t MASK = m29::MASK;

struct Consts9 {
  uint32_t d[9];
};

// k p in normalised digits, then each of the low 8 digits raised by t 2^29 (t from the
// digit above): digits 0..7 in [t 2^29 - t, (t + 1) 2^29), the top digit lowered by t.
// As a subtrahend's complement it absorbs any digits below t 2^29 - t + 1 (i >= 1), and
// a top digit up to its own.
constexpr Consts9 bias_digits(const uint32_t* p29, uint32_t k, uint32_t t) {
  Consts9 s{};
  uint64_t carry = 0;
  for (int i = 0; i < 8; i++) {
    const uint64_t v = (uint64_t)p29[i] * k + carry;
    s.d[i] = (uint32_t)(v & MASK);
    carry = v >> 29;
  }
  s.d[8] = (uint32_t)((uint64_t)p29[8] * k + carry);
  Consts9 b{};
  b.d[0] = s.d[0] + (t << 29);
  for (int i = 1; i < 8; i++) b.d[i] = s.d[i] + (t << 29) - t;
  b.d[8] = s.d[8] - t;
  return b;
}

// Per-field digit-domain constants: p in 29-bit digits, n0 = -p^-1 mod 2^29, one =
// 2^261 mod p (1 in the digit domain), c256 = 2^256 mod p (the factor back to the word
// domain), and the subtrahend complements Bias<K, T> = K p raised by T (bias_digits).
struct FqDig {
  static constexpr uint32_t p29[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                                      0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr uint32_t n0_29 = 0x04866389u;
  static constexpr uint32_t top2p = 0x60c89ce5u;  // top word of 2p
  static constexpr Consts9 one = {{0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u, 0x1d4240ceu,
                                   0x11d54c07u, 0x052ac7a8u, 0x000dc836u}};
  static constexpr Consts9 c256 = {{0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u, 0x11b7bc3cu,
                                    0x1cbd99bau, 0x183340fbu, 0x000e0a77u}};
};
struct FrDig {
  static constexpr uint32_t p29[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u,
                                      0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr uint32_t n0_29 = 0x0fffffffu;
  static constexpr uint32_t top2p = 0x60c89ce5u;
  static constexpr Consts9 one = {{0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu, 0x1d4240ceu,
                                   0x11d54c07u, 0x052ac7a8u, 0x000dc836u}};
  static constexpr Consts9 c256 = {{0x0ffffffbu, 0x04b1a0e2u, 0x18334a6bu, 0x18ed2b3eu, 0x1462e36fu, 0x11b7bc3cu,
                                    0x1cbd99bau, 0x183340fbu, 0x000e0a77u}};
};

template <class P, uint32_t K, uint32_t T>
struct Bias {
  static constexpr Consts9 v = bias_digits(P::p29, K, T);
};

BG_HD Fd from_consts(const Consts9& c) {
  Fd r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.d[i] = c.d[i];
  return r;
}

// carries out of digits 0..7 (all digits non-negative)
BG_HD Fd normalize(Fd a) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    a.d[i + 1] += a.d[i] >> 29;
    a.d[i] &= MASK;
  }
  return a;
}

// a + B - b, digit by digit (B = bias digits covering b's digits)
BG_HD Fd sub_lazy(const Fd& a, const Fd& b, const Consts9& B) {
  Fd r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.d[i] = a.d[i] + B.d[i] - b.d[i];
  return r;
}
// B - b
BG_HD Fd neg_lazy(const Fd& b, const Consts9& B) {
  Fd r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.d[i] = B.d[i] - b.d[i];
  return r;
}
BG_HD Fd add_lazy(const Fd& a, const Fd& b) {
  Fd r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
BG_HD Fd dbl_lazy(const Fd& a) {
  Fd r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.d[i] = a.d[i] << 1;
  return r;
}
BG_HD Fd select(bool c, const Fd& x, const Fd& y) {
  Fd r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.d[i] = c ? x.d[i] : y.d[i];
  return r;
}

// a = 0 mod p for a normalised a below 2p (0 or p)
template <class P>
BG_HD bool is_zero(const Fd& a) {
  uint32_t x = 0, y = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    x |= a.d[i];
    y |= a.d[i] ^ P::p29[i];
  }
  return x == 0 || y == 0;
}

// The products' column accumulator (the host unit test substitutes one that reports
// overflow).
#ifndef BG_FD_ACC
#define BG_FD_ACC uint64_t
#endif

template <class P>
BG_HD Fd mul(const Fd& a, const Fd& b) {
  return m29::mont<BG_FD_ACC>(a, b, P::p29, P::n0_29);
}
template <class P>
BG_HD Fd sqr(const Fd& a) {
  return m29::msqr<BG_FD_ACC>(a, P::p29, P::n0_29);
}
template <class P>
BG_HD Fd mul2(const Fd& a, const Fd& b, const Fd& c, const Fd& d) {
  return m29::mont2<BG_FD_ACC>(a, b, c, d, P::p29, P::n0_29);
}
template <class P>
BG_HD Fd mul4(const Fd* x, const Fd* y) {
  return m29::mont4<BG_FD_ACC>(x, y, P::p29, P::n0_29);
}

// unreduced (below 2^261, digits normalised or below 2^31) -> below 2.1 p
template <class P>
BG_HD Fd reduce_raw(const Fd& raw) {
  return mul<P>(raw, from_consts(P::one));
}
// word-domain value (8 words) -> digit-domain product operand below 64 p: 32 W itself when
// W's top word is below 2p's (then W < 2p), else 32 W reduced (any 8-word input)
template <class P>
BG_HD Fd load_raw(const uint32_t* w) {
  Fd x = m29::split<5>(w);
  if (w[7] >= P::top2p) x = reduce_raw<P>(x);
  return x;
}
// Value below 2^262 with digits below 2^31 -> the same residue below 1.01 p, by one
// subtraction of q p with q = floor(d_8 / (p_8 + 1)) (q p never exceeds the value; the
// remainder is below p + 2^240), signed carries.  Used where a run starts: the point's
// 32 W (below 64 p) becomes the first accumulator state.
template <class P>
BG_HD Fd reduce_fast(const Fd& a) {
  const uint32_t q = a.d[8] / (P::p29[8] + 1u);
  Fd r;
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int64_t v = (int64_t)a.d[i] - (int64_t)q * P::p29[i] + c;
    r.d[i] = (uint32_t)v & MASK;
    c = v >> 29;  // arithmetic
  }
  r.d[8] = (uint32_t)((int64_t)a.d[8] - (int64_t)q * P::p29[8] + c);
  return r;
}

// digit domain (normalised, below 33 p) -> word domain below 2p, 8 words: W = D / 32 mod p
// computed exactly as (D + k p) / 32 with k = -D p^-1 mod 32 (one 5-bit Montgomery step:
// D + k p is a multiple of 32), then packed from bit 5.  (D + 31 p) / 32 < 2p.
template <class P>
BG_HD void to_words(const Fd& a, uint32_t* w) {
  const uint32_t k = (a.d[0] * P::n0_29) & 31u;
  uint32_t t[9];
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint64_t v = (uint64_t)k * P::p29[i] + a.d[i] + c;
    t[i] = i < 8 ? (uint32_t)v & MASK : (uint32_t)v;
    c = v >> 29;
  }
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int b = 32 * j + 5, li = b / 29, sh = b % 29;
    uint32_t v = t[li] >> sh;
    if (li + 1 < 9) v |= t[li + 1] << (29 - sh);
    if (sh > 26 && li + 2 < 9) v |= t[li + 2] << (58 - sh);
    w[j] = v;
  }
}

// ------------------------------------------------------------------ Fq2 = Fq[u]/(u^2 + 1)

struct Fd2 {
  Fd c0, c1;
};

// a b for normalised a and b, b's components below (KB - 1) p:
//   c0 = a0 b0 + a1 (KB p - b1),  c1 = a0 b1 + a1 b0  (one reduction each)
template <class P, uint32_t KB>
BG_HD Fd2 mul2x(const Fd2& a, const Fd2& b) {
  const Fd nb1 = neg_lazy(b.c1, Bias<P, KB, 1>::v);
  return {mul2<P>(a.c0, b.c0, a.c1, nb1), mul2<P>(a.c0, b.c1, a.c1, b.c0)};
}
// a^2 for normalised a with a1 below (KA - 1) p:
//   c0 = (a0 + a1)(a0 + KA p - a1),  c1 = 2 a0 a1
template <class P, uint32_t KA>
BG_HD Fd2 sqr2x(const Fd2& a) {
  return {mul<P>(normalize(add_lazy(a.c0, a.c1)), sub_lazy(a.c0, a.c1, Bias<P, KA, 1>::v)),
          mul<P>(dbl_lazy(a.c0), a.c1)};
}
template <class P>
BG_HD bool is_zero2(const Fd2& a) {
  return is_zero<P>(a.c0) && is_zero<P>(a.c1);
}

}  // namespace fd
}  // namespace bg
