// Short-Weierstrass (a = 0) point arithmetic over Fq (G1) and Fq2 (G2) for gfx950.
//
// The MSM works in extended Jacobian "XYZZ" coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2;
// infinity = ZZ = 0), which make the mixed add 8M + 2S and the full add 12M + 2S with no
// Z^3 products (EFD "add-2008-s", "madd-2008-s", "dbl-2008-s-1", "mdbl-2008-s-1").
// Affine inputs use the reference encoding: infinity = (0, 0) (bend2/FORMAT.md:12-13).
// Every special case (infinity operand, P = Q, P = -Q) is resolved exactly, so any
// accumulation order yields the same group element and hence the same affine bits.
//
// JacRef restates the reference's Jacobian formulas (bend2/gen/gen_curve.py:86-184:
// dbl-2009-l, madd-2007-bl) for the on-device known-answer chain only.
#pragma once
#include "bg_field.hpp"

namespace bg {

template <class F>
struct Aff {
  typename F::E x, y;
};

template <class F>
struct Xyzz {
  typename F::E X, Y, ZZ, ZZZ;
};

template <class F>
struct Curve {
  using E = typename F::E;
  using A = Aff<F>;
  using P = Xyzz<F>;

  __device__ __forceinline__ static P infinity() { return {F::zero(), F::zero(), F::zero(), F::zero()}; }
  __device__ __forceinline__ static bool is_inf(const P& p) { return F::is_zero(p.ZZ); }
  __device__ __forceinline__ static bool aff_is_inf(const A& a) { return F::is_zero(a.x) && F::is_zero(a.y); }
  __device__ __forceinline__ static P from_affine(const A& a) {
    if (aff_is_inf(a)) return infinity();
    return {a.x, a.y, F::one(), F::one()};
  }
  __device__ __forceinline__ static A aff_neg(const A& a) { return {a.x, F::neg(a.y)}; }
  // canonical coordinates (values leaving the device path)
  __device__ __forceinline__ static P canon(const P& p) {
    return {F::canon(p.X), F::canon(p.Y), F::canon(p.ZZ), F::canon(p.ZZZ)};
  }
  __device__ __forceinline__ static A aff_canon(const A& a) { return {F::canon(a.x), F::canon(a.y)}; }
  __device__ __forceinline__ static P neg(const P& p) { return {p.X, F::neg(p.Y), p.ZZ, p.ZZZ}; }

  // 2 * affine (a, not infinity): mdbl-2008-s-1
  __device__ __forceinline__ static P dbl_aff(const A& a) {
    E u = F::dbl(a.y);
    E v = F::sqr(u);
    E w = F::mul(u, v);
    E s = F::mul(a.x, v);
    E x2 = F::sqr(a.x);
    E m = F::add(F::dbl(x2), x2);
    E x3 = F::sub(F::sqr(m), F::dbl(s));
    E y3 = F::mul_sub(m, F::sub(s, x3), w, a.y);
    return {x3, y3, v, w};
  }

  // 2 * p: dbl-2008-s-1 (infinity stays infinity since V*ZZ1 = 0)
  __device__ __forceinline__ static P dbl(const P& p) {
    E u = F::dbl(p.Y);
    E v = F::sqr(u);
    E w = F::mul(u, v);
    E s = F::mul(p.X, v);
    E x2 = F::sqr(p.X);
    E m = F::add(F::dbl(x2), x2);
    E x3 = F::sub(F::sqr(m), F::dbl(s));
    E y3 = F::mul_sub(m, F::sub(s, x3), w, p.Y);
    return {x3, y3, F::mul(v, p.ZZ), F::mul(w, p.ZZZ)};
  }

  // p + a (mixed): madd-2008-s with exact special cases
  __device__ __forceinline__ static P madd(const P& p, const A& a) {
    if (aff_is_inf(a)) return p;
    if (is_inf(p)) return {a.x, a.y, F::one(), F::one()};
    E u2 = F::mul(a.x, p.ZZ);
    E s2 = F::mul(a.y, p.ZZZ);
    E pp_ = F::sub(u2, p.X);
    E r = F::sub(s2, p.Y);
    if (F::is_zero(pp_)) {
      if (F::is_zero(r)) return dbl_aff(a);
      return infinity();
    }
    E pp = F::sqr(pp_);
    E ppp = F::mul(pp_, pp);
    E q = F::mul(p.X, pp);
    E x3 = F::sub(F::sub(F::sqr(r), ppp), F::dbl(q));
    E y3 = F::mul_sub(r, F::sub(q, x3), p.Y, ppp);
    return {x3, y3, F::mul(p.ZZ, pp), F::mul(p.ZZZ, ppp)};
  }

  // p + q by the generic add-2008-s formula, for p and q not infinity; sets bad when
  // x(p) = x(q), where the result is not the sum.
  __device__ __forceinline__ static P add_generic(const P& p, const P& q, uint32_t& bad) {
    E u1 = F::mul(p.X, q.ZZ);
    E u2 = F::mul(q.X, p.ZZ);
    E pp_ = F::sub(u2, u1);
    bad |= F::is_zero(pp_) ? 1u : 0u;
    E s1 = F::mul(p.Y, q.ZZZ);
    E s2 = F::mul(q.Y, p.ZZZ);
    E r = F::sub(s2, s1);
    E pp = F::sqr(pp_);
    E ppp = F::mul(pp_, pp);
    E q_ = F::mul(u1, pp);
    E zz3 = F::mul(F::mul(p.ZZ, q.ZZ), pp);
    E zzz3 = F::mul(F::mul(p.ZZZ, q.ZZZ), ppp);
    E x3 = F::sub(F::sub(F::sqr(r), ppp), F::dbl(q_));
    E y3 = F::mul_sub(r, F::sub(q_, x3), s1, ppp);
    return {x3, y3, zz3, zzz3};
  }

  // p + q where either may be infinity (ZZ = 0): the generic formula, with the infinite
  // operand cases selected around it; bad as in add_generic.
  __device__ __forceinline__ static P add_lite(const P& p, const P& q, uint32_t& bad) {
    bool pi = is_inf(p), qi = is_inf(q);
    if (pi) return q;
    if (qi) return p;
    return add_generic(p, q, bad);
  }

  // p + q: add-2008-s with exact special cases
  __device__ __forceinline__ static P add(const P& p, const P& q) {
    if (is_inf(p)) return q;
    if (is_inf(q)) return p;
    E u1 = F::mul(p.X, q.ZZ);
    E u2 = F::mul(q.X, p.ZZ);
    E s1 = F::mul(p.Y, q.ZZZ);
    E s2 = F::mul(q.Y, p.ZZZ);
    E pp_ = F::sub(u2, u1);
    E r = F::sub(s2, s1);
    if (F::is_zero(pp_)) {
      if (F::is_zero(r)) return dbl(p);
      return infinity();
    }
    E pp = F::sqr(pp_);
    E ppp = F::mul(pp_, pp);
    E q_ = F::mul(u1, pp);
    E x3 = F::sub(F::sub(F::sqr(r), ppp), F::dbl(q_));
    E y3 = F::mul_sub(r, F::sub(q_, x3), s1, ppp);
    return {x3, y3, F::mul(F::mul(p.ZZ, q.ZZ), pp), F::mul(F::mul(p.ZZZ, q.ZZZ), ppp)};
  }

  // k doublings
  __device__ static P dbl_n(P p, int k) {
    for (int i = 0; i < k; i++) p = dbl(p);
    return p;
  }

  // XYZZ -> affine with one inversion: t = 1/(ZZ*ZZZ); x = X*ZZZ*t; y = Y*ZZ*t
  __device__ __noinline__ static A to_affine(const P& p) {
    if (is_inf(p)) return {F::zero(), F::zero()};
    E t = F::inv(F::mul(p.ZZ, p.ZZZ));
    return {F::mul(F::mul(p.X, p.ZZZ), t), F::mul(F::mul(p.Y, p.ZZ), t)};
  }

  // the same, fully inlined (no device calls)
  __device__ __forceinline__ static A to_affine_inline(const P& p) {
    if (is_inf(p)) return {F::zero(), F::zero()};
    E t = F::inv_inline(F::mul(p.ZZ, p.ZZZ));
    return {F::mul(F::mul(p.X, p.ZZZ), t), F::mul(F::mul(p.Y, p.ZZ), t)};
  }

  // memory layout of an XYZZ point: X, Y, ZZ, ZZZ (4 elements)
  static constexpr int elem_words = Ld<E>::words;
  static constexpr int xyzz_words = 4 * elem_words;
  static constexpr int aff_words = 2 * elem_words;

  __device__ __forceinline__ static P load(const uint32_t* q) {
    return {Ld<E>::load(q), Ld<E>::load(q + elem_words), Ld<E>::load(q + 2 * elem_words),
            Ld<E>::load(q + 3 * elem_words)};
  }
  __device__ __forceinline__ static void store(uint32_t* q, const P& p) {
    Ld<E>::store(q, p.X);
    Ld<E>::store(q + elem_words, p.Y);
    Ld<E>::store(q + 2 * elem_words, p.ZZ);
    Ld<E>::store(q + 3 * elem_words, p.ZZZ);
  }
  __device__ __forceinline__ static A load_aff(const uint32_t* q) {
    return {Ld<E>::load(q), Ld<E>::load(q + elem_words)};
  }
  __device__ __forceinline__ static void store_aff(uint32_t* q, const A& a) {
    Ld<E>::store(q, a.x);
    Ld<E>::store(q + elem_words, a.y);
  }
};

// ---------------------------------------------------------------- quad-cooperative sums
// For the latency-bound reductions (few waves in flight, each waiting on a chain of
// dependent additions) a sum gets the four lanes of a quad (lanes 4i .. 4i+3 of a wave):
// every lane holds both operands, computes one of the four independent products of a step
// on operands selected by its lane index -- one product's instruction stream yields four
// products -- and the four products are shared by DPP quad broadcasts.  The add-2008-s
// sum then takes four product steps instead of 14 products, the dbl-2008-s-1 doubling
// four instead of eight; every lane ends with the whole result.  The products are the same
// field operations as Curve's (the same group element; representatives may differ within
// [0, 2p)).  Callers keep control flow uniform within each quad.
// (by masks: a select between whole operands lets the compiler select their addresses and
// keep them in private memory)
__device__ __forceinline__ Fp qsel(uint32_t r, const Fp& a0, const Fp& a1, const Fp& a2, const Fp& a3) {
  const uint32_t m0 = 0u - (uint32_t)(r == 0), m1 = 0u - (uint32_t)(r == 1), m2 = 0u - (uint32_t)(r == 2),
                 m3 = 0u - (uint32_t)(r == 3);
  Fp o;
#pragma unroll
  for (int i = 0; i < 8; i++) o.v[i] = (a0.v[i] & m0) | (a1.v[i] & m1) | (a2.v[i] & m2) | (a3.v[i] & m3);
  return o;
}
__device__ __forceinline__ Fp2 qsel(uint32_t r, const Fp2& a0, const Fp2& a1, const Fp2& a2, const Fp2& a3) {
  return {qsel(r, a0.c0, a1.c0, a2.c0, a3.c0), qsel(r, a0.c1, a1.c1, a2.c1, a3.c1)};
}
// lane S of each quad to all four (DPP quad_perm [S, S, S, S])
template <int S>
__device__ __forceinline__ Fp qbc(const Fp& x) {
  Fp o;
#pragma unroll
  for (int i = 0; i < 8; i++) o.v[i] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x.v[i], S * 0x55, 0xF, 0xF, false);
  return o;
}
template <int S>
__device__ __forceinline__ Fp2 qbc(const Fp2& x) {
  return {qbc<S>(x.c0), qbc<S>(x.c1)};
}

template <class F>
struct Quad {
  using C = Curve<F>;
  using E = typename F::E;
  using P = typename C::P;
  __device__ __forceinline__ static uint32_t role() { return threadIdx.x & 3u; }

  // Curve::add_lite on a quad (p, q the same in its four lanes)
  __device__ __forceinline__ static P add_lite(const P& p, const P& q, uint32_t& bad) {
    if (C::is_inf(p)) return q;
    if (C::is_inf(q)) return p;
    const uint32_t r = role();
    // u1 = X1 ZZ2, u2 = X2 ZZ1, s1 = Y1 ZZZ2, s2 = Y2 ZZZ1
    E m = F::mul(qsel(r, p.X, q.X, p.Y, q.Y), qsel(r, q.ZZ, p.ZZ, q.ZZZ, p.ZZZ));
    const E u1 = qbc<0>(m), u2 = qbc<1>(m), s1 = qbc<2>(m), s2 = qbc<3>(m);
    const E pp_ = F::sub(u2, u1), rr = F::sub(s2, s1);
    bad |= F::is_zero(pp_) ? 1u : 0u;
    // pp = P^2, r^2, ZZ1 ZZ2, ZZZ1 ZZZ2
    m = F::mul(qsel(r, pp_, rr, p.ZZ, p.ZZZ), qsel(r, pp_, rr, q.ZZ, q.ZZZ));
    const E pp = qbc<0>(m), r2 = qbc<1>(m), zz12 = qbc<2>(m), zzz12 = qbc<3>(m);
    // ppp = P pp, q = u1 pp, ZZ3 = ZZ1 ZZ2 pp
    m = F::mul(qsel(r, pp_, u1, zz12, zz12), pp);
    const E ppp = qbc<0>(m), q_ = qbc<1>(m), zz3 = qbc<2>(m);
    const E x3 = F::sub(F::sub(r2, ppp), F::dbl(q_));
    // Y3 = r (q - X3) - s1 ppp, ZZZ3 = ZZZ1 ZZZ2 ppp - 0
    const E z = F::zero(), t = F::sub(q_, x3);
    m = F::mul_sub(qsel(r, rr, zzz12, zzz12, zzz12), qsel(r, t, ppp, ppp, ppp), qsel(r, s1, z, z, z), ppp);
    return {x3, qbc<0>(m), zz3, qbc<1>(m)};
  }

  // Curve::add (exact special cases) on a quad
  __device__ __forceinline__ static P add(const P& p, const P& q) {
    if (C::is_inf(p)) return q;
    if (C::is_inf(q)) return p;
    const uint32_t r = role();
    E m = F::mul(qsel(r, p.X, q.X, p.Y, q.Y), qsel(r, q.ZZ, p.ZZ, q.ZZZ, p.ZZZ));
    const E u1 = qbc<0>(m), u2 = qbc<1>(m), s1 = qbc<2>(m), s2 = qbc<3>(m);
    const E pp_ = F::sub(u2, u1), rr = F::sub(s2, s1);
    if (F::is_zero(pp_)) {
      if (F::is_zero(rr)) return dbl(p);
      return C::infinity();
    }
    m = F::mul(qsel(r, pp_, rr, p.ZZ, p.ZZZ), qsel(r, pp_, rr, q.ZZ, q.ZZZ));
    const E pp = qbc<0>(m), r2 = qbc<1>(m), zz12 = qbc<2>(m), zzz12 = qbc<3>(m);
    m = F::mul(qsel(r, pp_, u1, zz12, zz12), pp);
    const E ppp = qbc<0>(m), q_ = qbc<1>(m), zz3 = qbc<2>(m);
    const E x3 = F::sub(F::sub(r2, ppp), F::dbl(q_));
    const E z = F::zero(), t = F::sub(q_, x3);
    m = F::mul_sub(qsel(r, rr, zzz12, zzz12, zzz12), qsel(r, t, ppp, ppp, ppp), qsel(r, s1, z, z, z), ppp);
    return {x3, qbc<0>(m), zz3, qbc<1>(m)};
  }

  // Curve::dbl on a quad
  __device__ __forceinline__ static P dbl(const P& p) {
    const uint32_t r = role();
    const E u = F::dbl(p.Y);
    // v = u^2, x2 = X^2
    E m = F::mul(qsel(r, u, p.X, u, u), qsel(r, u, p.X, u, u));
    const E v = qbc<0>(m), x2 = qbc<1>(m);
    // w = u v, s = X v, ZZ3 = v ZZ
    m = F::mul(qsel(r, u, p.X, p.ZZ, p.ZZ), v);
    const E w = qbc<0>(m), s = qbc<1>(m), zz3 = qbc<2>(m);
    const E mm = F::add(F::dbl(x2), x2);
    // mm^2, ZZZ3 = w ZZZ
    m = F::mul(qsel(r, mm, w, w, w), qsel(r, mm, p.ZZZ, p.ZZZ, p.ZZZ));
    const E m2 = qbc<0>(m), zzz3 = qbc<1>(m);
    const E x3 = F::sub(m2, F::dbl(s));
    const E y3 = F::mul_sub(mm, F::sub(s, x3), w, p.Y);
    return {x3, y3, zz3, zzz3};
  }
};

using G1 = Curve<Fq>;
using G2 = Curve<Fq2>;

// ---------------------------------------------------------------- reference Jacobian
// gen_curve.py:86-102 (dbl-2009-l) and :147-184 (madd-2007-bl + special cases);
// infinity = (0, 1, 0).  Used only by the known-answer chain kernel.
template <class F>
struct JacRef {
  using E = typename F::E;
  struct J {
    E x, y, z;
  };
  __device__ static J infinity() { return {F::zero(), F::one(), F::zero()}; }
  __device__ static J dbl(const J& p) {
    E a = F::sqr(p.x);
    E b = F::sqr(p.y);
    E c = F::sqr(b);
    E d0 = F::sub(F::sub(F::sqr(F::add(p.x, b)), a), c);
    E d = F::add(d0, d0);
    E e = F::add(F::add(a, a), a);
    E f = F::sqr(e);
    E x3 = F::sub(f, F::add(d, d));
    E c2 = F::add(c, c);
    E c4 = F::add(c2, c2);
    E y3 = F::sub(F::mul(e, F::sub(d, x3)), F::add(c4, c4));
    E z3 = F::mul(F::add(p.y, p.y), p.z);
    return {x3, y3, z3};
  }
  __device__ static J madd(const J& p, const Aff<F>& q) {
    if (F::is_zero(p.z)) {
      if (F::is_zero(q.x) && F::is_zero(q.y)) return infinity();
      return {q.x, q.y, F::one()};
    }
    if (F::is_zero(q.x) && F::is_zero(q.y)) return p;
    E z1z1 = F::sqr(p.z);
    E u2 = F::mul(q.x, z1z1);
    E s2 = F::mul(F::mul(q.y, p.z), z1z1);
    E h = F::sub(u2, p.x);
    E rr = F::sub(s2, p.y);
    if (F::is_zero(h)) {
      if (F::is_zero(rr)) return dbl(p);
      return infinity();
    }
    E hh = F::sqr(h);
    E hh2 = F::add(hh, hh);
    E i = F::add(hh2, hh2);
    E j = F::mul(h, i);
    E r = F::add(rr, rr);
    E v = F::mul(p.x, i);
    E x3 = F::sub(F::sub(F::sqr(r), j), F::add(v, v));
    E y1j = F::mul(p.y, j);
    E y3 = F::sub(F::mul(r, F::sub(v, x3)), F::add(y1j, y1j));
    E z3 = F::sub(F::sub(F::sqr(F::add(p.z, h)), z1z1), hh);
    return {x3, y3, z3};
  }
};

}  // namespace bg
