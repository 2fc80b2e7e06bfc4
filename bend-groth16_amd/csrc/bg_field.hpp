// BN254 prime-field arithmetic for gfx950: 8 x 32-bit little-endian limbs held in VGPRs,
// Montgomery form with R = 2^256 (the same integers as the reference's 16 x 16-bit limbs,
// bend2/FORMAT.md:7-8, so repacking is a pure bit pack).
//
// The Montgomery product (bg_mul29.hpp) splits its operands into 29-bit digits and
// accumulates every column of a*b + m*p in one 64-bit v_mad_u64_u32 accumulator, the
// reduction digit formed in the same column: 171 32x32->64 multiplies (81 a*b, 81 m*p,
// 9 m digits) and no carry folds.
//
// Values are kept lazily reduced, in [0, 2p): p < 2^254, so 4p < 2^256 = R and a
// Montgomery product of inputs below 2p is already below 2p (no final subtraction); add and
// subtract reduce modulo 2p.  is_zero / eq test congruence mod p; canon() gives the
// canonical representative, which every value leaving the device path goes through.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bg_mul29.hpp"

namespace bg {

struct Fp {
  uint32_t v[8];
};

// Field constants.  p < 2^254 for both BN254 fields: a Montgomery product of inputs < 2p is
// < 2p and never needs a 9th word.
struct FrP {
  static constexpr uint32_t p[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                                    0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr uint32_t p2[8] = {0xe0000002u, 0x87c3eb27u, 0xf372e122u, 0x5067d090u,
                                     0x0302b0bau, 0x70a08b6du, 0xc2634053u, 0x60c89ce5u};  // 2p
  static constexpr uint32_t n0 = 0xefffffffu;
  static constexpr uint32_t one[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                      0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr uint32_t r2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                     0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
  // p in 29-bit digits and -p^-1 mod 2^29 (the product's digit form)
  static constexpr uint32_t p29[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u,
                                      0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr uint32_t n0_29 = 0x0fffffffu;
};
struct FqP {
  static constexpr uint32_t p[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                                    0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr uint32_t p2[8] = {0xb0f9fa8eu, 0x7841182du, 0xd0e3951au, 0x2f02d522u,
                                     0x0302b0bbu, 0x70a08b6du, 0xc2634053u, 0x60c89ce5u};  // 2p
  static constexpr uint32_t n0 = 0xe4866389u;
  static constexpr uint32_t one[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                      0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr uint32_t r2[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                     0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
  static constexpr uint32_t p29[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                                      0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr uint32_t n0_29 = 0x04866389u;
};

// ------------------------------------------------------------------ carry helpers

// 32-bit add / subtract steps with the carry or borrow in an SGPR pair (VOP3 forms), so a
// 256-bit chain is one instruction per limb; the compiler otherwise builds each limb from
// 64-bit adds and register moves (about 3x the instructions).  ALU-only assembly.
__device__ __forceinline__ uint32_t add_co(uint32_t a, uint32_t b, uint64_t& co) {
  uint32_t r;
  asm("v_add_co_u32_e64 %0, %1, %2, %3" : "=v"(r), "=s"(co) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ uint32_t addc_co(uint32_t a, uint32_t b, uint64_t ci, uint64_t& co) {
  uint32_t r;
  asm("v_addc_co_u32_e64 %0, %1, %2, %3, %4" : "=v"(r), "=s"(co) : "v"(a), "v"(b), "s"(ci));
  return r;
}
__device__ __forceinline__ uint32_t sub_co(uint32_t a, uint32_t b, uint64_t& bo) {
  uint32_t r;
  asm("v_sub_co_u32_e64 %0, %1, %2, %3" : "=v"(r), "=s"(bo) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ uint32_t subb_co(uint32_t a, uint32_t b, uint64_t bi, uint64_t& bo) {
  uint32_t r;
  asm("v_subb_co_u32_e64 %0, %1, %2, %3, %4" : "=v"(r), "=s"(bo) : "v"(a), "v"(b), "s"(bi));
  return r;
}
// same with a wave-uniform constant as the second operand: an SGPR where the step has no
// carry in, a VGPR where it has one (one scalar operand per VOP3 instruction)
__device__ __forceinline__ uint32_t add_co_k(uint32_t a, uint32_t k, uint64_t& co) {
  uint32_t r;
  asm("v_add_co_u32_e64 %0, %1, %2, %3" : "=v"(r), "=s"(co) : "v"(a), "s"(k));
  return r;
}
__device__ __forceinline__ uint32_t addc_co_k(uint32_t a, uint32_t k, uint64_t ci, uint64_t& co) {
  uint32_t r;
  asm("v_addc_co_u32_e64 %0, %1, %2, %3, %4" : "=v"(r), "=s"(co) : "v"(a), "v"(k), "s"(ci));
  return r;
}
__device__ __forceinline__ uint32_t sub_co_k(uint32_t a, uint32_t k, uint64_t& bo) {
  uint32_t r;
  asm("v_sub_co_u32_e64 %0, %1, %2, %3" : "=v"(r), "=s"(bo) : "v"(a), "s"(k));
  return r;
}
__device__ __forceinline__ uint32_t subb_co_k(uint32_t a, uint32_t k, uint64_t bi, uint64_t& bo) {
  uint32_t r;
  asm("v_subb_co_u32_e64 %0, %1, %2, %3, %4" : "=v"(r), "=s"(bo) : "v"(a), "v"(k), "s"(bi));
  return r;
}
// m ? y : x per lane
__device__ __forceinline__ uint32_t select(uint64_t m, uint32_t x, uint32_t y) {
  uint32_t r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "s"(m));
  return r;
}

// r = a - p if a >= p else a (a < 2p)
template <class P>
__device__ __forceinline__ void reduce_once(uint32_t* r, const uint32_t* a) {
  uint32_t d[8];
  uint64_t b;
  d[0] = sub_co_k(a[0], P::p[0], b);
#pragma unroll
  for (int i = 1; i < 8; i++) d[i] = subb_co_k(a[i], P::p[i], b, b);
  // borrow out -> a < p: keep a
#pragma unroll
  for (int i = 0; i < 8; i++) r[i] = select(b, d[i], a[i]);
}

// r = a - 2p if a >= 2p else a (a < 4p)
template <class P>
__device__ __forceinline__ void reduce_2p(uint32_t* r, const uint32_t* a) {
  uint32_t d[8];
  uint64_t b;
  d[0] = sub_co_k(a[0], P::p2[0], b);
#pragma unroll
  for (int i = 1; i < 8; i++) d[i] = subb_co_k(a[i], P::p2[i], b, b);
#pragma unroll
  for (int i = 0; i < 8; i++) r[i] = select(b, d[i], a[i]);
}

// ------------------------------------------------------------------ field ops

template <class P>
struct Field {
  using E = Fp;

  __device__ __forceinline__ static Fp zero() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    return r;
  }
  __device__ __forceinline__ static Fp one() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = P::one[i];
    return r;
  }
  // a = 0 mod p (a in [0, 2p): 0 or p)
  __device__ __forceinline__ static bool is_zero(const Fp& a) {
    uint32_t x = 0, y = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      x |= a.v[i];
      y |= a.v[i] ^ P::p[i];
    }
    return x == 0 || y == 0;
  }
  __device__ __forceinline__ static bool eq(const Fp& a, const Fp& b) { return is_zero(sub(a, b)); }
  // canonical representative in [0, p)
  __device__ __forceinline__ static Fp canon(const Fp& a) {
    Fp r;
    reduce_once<P>(r.v, a.v);
    return r;
  }

  __device__ __forceinline__ static Fp add(const Fp& a, const Fp& b) {
    uint32_t s[8];
    uint64_t c;
    s[0] = add_co(a.v[0], b.v[0], c);
#pragma unroll
    for (int i = 1; i < 8; i++) s[i] = addc_co(a.v[i], b.v[i], c, c);
    Fp r;
    reduce_2p<P>(r.v, s);  // a + b < 4p < 2^256: no carry out of word 7
    return r;
  }

  __device__ __forceinline__ static Fp sub(const Fp& a, const Fp& b) {
    uint32_t d[8], t[8];
    uint64_t br, c;
    d[0] = sub_co(a.v[0], b.v[0], br);
#pragma unroll
    for (int i = 1; i < 8; i++) d[i] = subb_co(a.v[i], b.v[i], br, br);
    // borrow -> add 2p back
    t[0] = add_co_k(d[0], P::p2[0], c);
#pragma unroll
    for (int i = 1; i < 8; i++) t[i] = addc_co_k(d[i], P::p2[i], c, c);
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = select(br, d[i], t[i]);
    return r;
  }

  __device__ __forceinline__ static Fp neg(const Fp& a) { return sub(zero(), a); }

  __device__ __forceinline__ static Fp dbl(const Fp& a) { return add(a, a); }

  // Montgomery product a * b * 2^-256 mod p with 29-bit digits (bg_mul29.hpp): no carry
  // passes through scalar registers, so the product issues ~250 VALU instructions and no
  // hazard wait states (a 32-bit-digit product scanning with the carries folded through
  // SGPR pairs: ~295 VALU instructions plus ~115 s_nop).
  __device__ __forceinline__ static Fp mul(const Fp& a, const Fp& b) {
    Fp o;
    m29::mul_words(a.v, b.v, o.v, P::p29, P::n0_29);
    return o;
  }

  __device__ __forceinline__ static Fp sqr(const Fp& a) { return mul(a, a); }

  // a b - c d as a b + c (2p - d) with one reduction for both products (below 2.6 p, then
  // one conditional subtraction): the y coordinate of every XYZZ addition and doubling
  __device__ __forceinline__ static Fp mul_sub(const Fp& a, const Fp& b, const Fp& c, const Fp& d) {
    const Fp nd = neg(d);
    Fp t, r;
    m29::mul2_words(a.v, b.v, c.v, nd.v, t.v, P::p29, P::n0_29);
    reduce_2p<P>(r.v, t.v);
    return r;
  }

  // x^e for an exponent given as 8 little-endian words (left to right)
  __device__ __noinline__ static Fp pow(const Fp& x, const uint32_t* e) {
    Fp acc = one();
    for (int i = 255; i >= 0; i--) {
      acc = sqr(acc);
      if ((e[i >> 5] >> (i & 31)) & 1) acc = mul(acc, x);
    }
    return acc;
  }

  // Fermat inverse a^(p-2); inv(0) = 0
  __device__ static Fp inv(const Fp& a) {
    uint32_t e[8];
    uint32_t br = 2;
    for (int i = 0; i < 8; i++) {
      uint64_t t = (uint64_t)P::p[i] - br;
      e[i] = (uint32_t)t;
      br = (uint32_t)(t >> 63);
    }
    return pow(a, e);
  }

  // Fermat inverse a^(p-2) fully inlined (uniform loop over the exponent's bits, no
  // device calls), for kernels that must stay call-free; inv(0) = 0
  __device__ __forceinline__ static Fp inv_inline(const Fp& a) {
    Fp acc = one();
#pragma unroll
    for (int wi = 7; wi >= 0; wi--) {
      const uint32_t w = wi == 0 ? P::p[0] - 2u : P::p[wi];  // p - 2 (p[0] > 2: no borrow)
#pragma unroll 1
      for (int b = 31; b >= 0; b--) {
        acc = sqr(acc);
        if ((w >> b) & 1u) acc = mul(acc, a);
      }
    }
    return acc;
  }

  // canonical -> Montgomery (x * R^2 * R^-1)
  __device__ __forceinline__ static Fp to_mont(const Fp& a) {
    Fp r2;
#pragma unroll
    for (int i = 0; i < 8; i++) r2.v[i] = P::r2[i];
    return mul(a, r2);
  }
  // Montgomery -> canonical (x * 1 * R^-1)
  __device__ __forceinline__ static Fp from_mont(const Fp& a) {
    Fp o = zero();
    o.v[0] = 1;
    return mul(a, o);
  }
};

using Fr = Field<FrP>;
using Fq = Field<FqP>;

// ------------------------------------------------------------------ Fq2 = Fq[u]/(u^2 + 1)

struct Fp2 {
  Fp c0, c1;
};

struct Fq2 {
  using E = Fp2;
  __device__ __forceinline__ static Fp2 zero() { return {Fq::zero(), Fq::zero()}; }
  __device__ __forceinline__ static Fp2 one() { return {Fq::one(), Fq::zero()}; }
  __device__ __forceinline__ static bool is_zero(const Fp2& a) { return Fq::is_zero(a.c0) && Fq::is_zero(a.c1); }
  __device__ __forceinline__ static bool eq(const Fp2& a, const Fp2& b) { return Fq::eq(a.c0, b.c0) && Fq::eq(a.c1, b.c1); }
  __device__ __forceinline__ static Fp2 canon(const Fp2& a) { return {Fq::canon(a.c0), Fq::canon(a.c1)}; }
  __device__ __forceinline__ static Fp2 add(const Fp2& a, const Fp2& b) { return {Fq::add(a.c0, b.c0), Fq::add(a.c1, b.c1)}; }
  __device__ __forceinline__ static Fp2 sub(const Fp2& a, const Fp2& b) { return {Fq::sub(a.c0, b.c0), Fq::sub(a.c1, b.c1)}; }
  __device__ __forceinline__ static Fp2 neg(const Fp2& a) { return {Fq::neg(a.c0), Fq::neg(a.c1)}; }
  __device__ __forceinline__ static Fp2 dbl(const Fp2& a) { return {Fq::dbl(a.c0), Fq::dbl(a.c1)}; }
  // Two sums of two products, one Montgomery reduction each (bg_mul29.hpp mont2):
  //   c0 = a0 b0 + a1 (2p - b1),  c1 = a0 b1 + a1 b0,
  // each below 2.6 p and brought below 2p by one conditional subtraction.  4 x 81 digit
  // products and 2 x 81 reduction products: as many multiplies as Karatsuba's 3 x (81 + 81),
  // but two reductions instead of three and no Karatsuba sums and differences.  The
  // scheduling barrier keeps the two halves from being interleaved (two operand sets live
  // at once).
  __device__ __forceinline__ static Fp2 mul(const Fp2& a, const Fp2& b) {
    const Fp nb1 = Fq::neg(b.c1);
    Fp t, c0, c1;
    m29::mul2_words(a.c0.v, b.c0.v, a.c1.v, nb1.v, t.v, FqP::p29, FqP::n0_29);
    reduce_2p<FqP>(c0.v, t.v);
    __builtin_amdgcn_sched_barrier(0);
    m29::mul2_words(a.c0.v, b.c1.v, a.c1.v, b.c0.v, t.v, FqP::p29, FqP::n0_29);
    reduce_2p<FqP>(c1.v, t.v);
    __builtin_amdgcn_sched_barrier(0);
    return {c0, c1};
  }
  // a b - c d with one reduction per component (bg_mul29.hpp mont4):
  //   r0 = a0 b0 + a1 (2p - b1) + c0 (2p - d0) + c1 d1
  //   r1 = a0 b1 + a1 b0 + c0 (2p - d1) + c1 (2p - d0)
  // each below 4.1 p, brought below 2p by two conditional subtractions
  __device__ __forceinline__ static Fp2 mul_sub(const Fp2& a, const Fp2& b, const Fp2& c, const Fp2& d) {
    const Fp nb1 = Fq::neg(b.c1), nd0 = Fq::neg(d.c0), nd1 = Fq::neg(d.c1);
    Fp t, r0, r1;
    {
      const uint32_t* x[4] = {a.c0.v, a.c1.v, c.c0.v, c.c1.v};
      const uint32_t* y[4] = {b.c0.v, nb1.v, nd0.v, d.c1.v};
      m29::mul4_words(x, y, t.v, FqP::p29, FqP::n0_29);
      reduce_2p<FqP>(t.v, t.v);
      reduce_2p<FqP>(r0.v, t.v);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
      const uint32_t* x[4] = {a.c0.v, a.c1.v, c.c0.v, c.c1.v};
      const uint32_t* y[4] = {b.c1.v, b.c0.v, nd1.v, nd0.v};
      m29::mul4_words(x, y, t.v, FqP::p29, FqP::n0_29);
      reduce_2p<FqP>(t.v, t.v);
      reduce_2p<FqP>(r1.v, t.v);
    }
    __builtin_amdgcn_sched_barrier(0);
    return {r0, r1};
  }
  // complex squaring: 2 Fq products
  __device__ __forceinline__ static Fp2 sqr(const Fp2& a) {
    Fp c0 = Fq::mul(Fq::add(a.c0, a.c1), Fq::sub(a.c0, a.c1));
    __builtin_amdgcn_sched_barrier(0);
    Fp c1 = Fq::mul(Fq::dbl(a.c0), a.c1);
    __builtin_amdgcn_sched_barrier(0);
    return {c0, c1};
  }
  __device__ __forceinline__ static Fp2 inv_inline(const Fp2& a) {
    Fp n = Fq::add(Fq::sqr(a.c0), Fq::sqr(a.c1));
    Fp ni = Fq::inv_inline(n);
    return {Fq::mul(a.c0, ni), Fq::neg(Fq::mul(a.c1, ni))};
  }
  __device__ __noinline__ static Fp2 inv(const Fp2& a) {
    Fp n = Fq::add(Fq::sqr(a.c0), Fq::sqr(a.c1));
    Fp ni = Fq::inv(n);
    return {Fq::mul(a.c0, ni), Fq::neg(Fq::mul(a.c1, ni))};
  }
};

// ------------------------------------------------------------------ loads / stores

__device__ __forceinline__ Fp load_fp(const uint32_t* __restrict__ p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 x = q[0], y = q[1];
  return Fp{{x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w}};
}
__device__ __forceinline__ void store_fp(uint32_t* __restrict__ p, const Fp& a) {
  uint4* q = reinterpret_cast<uint4*>(p);
  q[0] = make_uint4(a.v[0], a.v[1], a.v[2], a.v[3]);
  q[1] = make_uint4(a.v[4], a.v[5], a.v[6], a.v[7]);
}
__device__ __forceinline__ Fp2 load_fp2(const uint32_t* __restrict__ p) { return {load_fp(p), load_fp(p + 8)}; }
__device__ __forceinline__ void store_fp2(uint32_t* __restrict__ p, const Fp2& a) {
  store_fp(p, a.c0);
  store_fp(p + 8, a.c1);
}

template <class E>
struct Ld;
template <>
struct Ld<Fp> {
  static constexpr int words = 8;
  __device__ __forceinline__ static Fp load(const uint32_t* p) { return load_fp(p); }
  __device__ __forceinline__ static void store(uint32_t* p, const Fp& a) { store_fp(p, a); }
};
template <>
struct Ld<Fp2> {
  static constexpr int words = 16;
  __device__ __forceinline__ static Fp2 load(const uint32_t* p) { return load_fp2(p); }
  __device__ __forceinline__ static void store(uint32_t* p, const Fp2& a) { store_fp2(p, a); }
};

}  // namespace bg
