// Bucket accumulation in the digit domain (bg_fd.hpp): the running XYZZ sum of a bucket
// run stays in 29-bit digits from one mixed addition to the next, and is packed to the
// word domain (8 words per coordinate, below 2p, the layout every other kernel reads)
// only when the run ends.
//
// The addition is madd-2008-s (U2 = x ZZ, S2 = y ZZZ, P = U2 - X, R = S2 - Y, PP = P^2,
// PPP = P PP, Q = X PP, X3 = R^2 - PPP - 2Q, Y3 = R (Q - X3) - Y PPP, ZZ3 = ZZ PP,
// ZZZ3 = ZZZ PPP) on a run that never meets infinity; x(acc) = x(point) (doubling or
// cancellation) is detected as ZZ3 = 0 and reported through `bad` for an exact
// recomputation, as with the word-domain formula.  A negated point (signed digit) enters
// through R = -(S2 + Y).
//
// Value bounds, in units of p (interval analysis over the fixed point of the accumulator
// state, restated with these constants in the CPU test of the bounds):
//   G1: X < 6.2, Y < 1.4, ZZ, ZZZ < 1.01; point operands below 64 (load_raw);
//       U2, S2 < 1.4; P < 9.4; R < 4.4; PP < 1.6; PPP < 1.1; Q < 1.06; R^2 < 1.2
//   G2: X < 8.0, Y < 1.9, ZZ < 1.07, ZZZ < 1.03; U2, S2 < 2.6; P < 11.6; R < 6;
//       PP < 4.4; PPP < 1.8; Q < 1.5; R^2 < 2.0
// and every subtraction's bias multiple K below covers its subtrahend (K - 1 above its
// bound; T above its digit width).
#pragma once
#include "bg_fd.hpp"

namespace bg {
namespace fd {

template <class P>
struct G1Acc {
  static constexpr uint32_t KX = 8, KY = 3, KYN = 4, KX3 = 5, KT = 8, KU = 3, KRAW = 65;
  static constexpr int aff_words = 16, xyzz_words = 32;
  struct Pt {
    Fd X, Y, ZZ, ZZZ;
  };
  struct Raw {
    Fd x, y;
  };
  BG_HD static Raw load(const uint32_t* w) { return {load_raw<P>(w), load_raw<P>(w + 8)}; }
  // the run's first point (y negated for a negative digit), ZZ = ZZZ = 1
  BG_HD static Pt start(const Raw& a, bool neg) {
    const Fd y = select(neg, neg_lazy(a.y, Bias<P, KRAW, 1>::v), a.y);
    const Fd one = from_consts(P::one);
    return {reduce_fast<P>(a.x), reduce_fast<P>(y), one, one};
  }
  BG_HD static Pt madd(const Pt& p, const Raw& a, bool neg, uint32_t& bad) {
    const Fd u2 = mul<P>(a.x, p.ZZ);
    const Fd s2 = mul<P>(a.y, p.ZZZ);
    const Fd pp_ = normalize(sub_lazy(u2, p.X, Bias<P, KX, 1>::v));
    const Fd r = normalize(
        select(neg, neg_lazy(add_lazy(s2, p.Y), Bias<P, KYN, 2>::v), sub_lazy(s2, p.Y, Bias<P, KY, 1>::v)));
    const Fd pp = sqr<P>(pp_);
    const Fd zz3 = mul<P>(p.ZZ, pp);
    bad |= is_zero<P>(zz3) ? 1u : 0u;
    const Fd ppp = mul<P>(pp_, pp);
    const Fd zzz3 = mul<P>(p.ZZZ, ppp);
    const Fd q = mul<P>(p.X, pp);
    const Fd r2 = sqr<P>(r);
    const Fd x3 = normalize(sub_lazy(r2, add_lazy(ppp, dbl_lazy(q)), Bias<P, KX3, 3>::v));
    const Fd y3 = mul2<P>(r, sub_lazy(q, x3, Bias<P, KT, 1>::v), p.Y, neg_lazy(ppp, Bias<P, KU, 1>::v));
    return {x3, y3, zz3, zzz3};
  }
  // word-domain XYZZ (X, Y, ZZ, ZZZ, 8 words each)
  BG_HD static void store(uint32_t* w, const Pt& p) {
    to_words<P>(p.X, w);
    to_words<P>(p.Y, w + 8);
    to_words<P>(p.ZZ, w + 16);
    to_words<P>(p.ZZZ, w + 24);
  }
  // one digit-form coordinate (the k-th Fd of Pt) to its 8 words (words 8k .. 8k + 7 of store)
  BG_HD static void store_fd(uint32_t* w, const Fd& a) { to_words<P>(a, w); }
};

template <class P>
struct G2Acc {
  static constexpr uint32_t KX = 9, KY = 3, KYN = 6, KX3 = 6, KT = 9, KNT = 12, KU = 3, KRAW = 65;
  // bias multiples of the Fq2 products' second operands and of the squarings
  static constexpr uint32_t KZ = 3, KPP = 6, KSP = 13, KSR = 8;
  static constexpr int aff_words = 32, xyzz_words = 64;
  struct Pt {
    Fd2 X, Y, ZZ, ZZZ;
  };
  struct Raw {
    Fd2 x, y;
  };
  BG_HD static Raw load(const uint32_t* w) {
    return {{load_raw<P>(w), load_raw<P>(w + 8)}, {load_raw<P>(w + 16), load_raw<P>(w + 24)}};
  }
  BG_HD static Pt start(const Raw& a, bool neg) {
    const Fd y0 = select(neg, neg_lazy(a.y.c0, Bias<P, KRAW, 1>::v), a.y.c0);
    const Fd y1 = select(neg, neg_lazy(a.y.c1, Bias<P, KRAW, 1>::v), a.y.c1);
    const Fd one = from_consts(P::one);
    Fd zero;
#pragma unroll
    for (int i = 0; i < 9; i++) zero.d[i] = 0;
    return {{reduce_fast<P>(a.x.c0), reduce_fast<P>(a.x.c1)}, {reduce_fast<P>(y0), reduce_fast<P>(y1)}, {one, zero},
            {one, zero}};
  }
  BG_HD static Fd r_of(const Fd& s2, const Fd& y, bool neg) {
    return normalize(select(neg, neg_lazy(add_lazy(s2, y), Bias<P, KYN, 2>::v), sub_lazy(s2, y, Bias<P, KY, 1>::v)));
  }
  BG_HD static Pt madd(const Pt& p, const Raw& a, bool neg, uint32_t& bad) {
    const Fd2 u2 = mul2x<P, KZ>(a.x, p.ZZ);
    const Fd2 s2 = mul2x<P, KZ>(a.y, p.ZZZ);
    const Fd2 pp_ = {normalize(sub_lazy(u2.c0, p.X.c0, Bias<P, KX, 1>::v)),
                     normalize(sub_lazy(u2.c1, p.X.c1, Bias<P, KX, 1>::v))};
    const Fd2 r = {r_of(s2.c0, p.Y.c0, neg), r_of(s2.c1, p.Y.c1, neg)};
    const Fd2 pp = sqr2x<P, KSP>(pp_);
    const Fd2 zz3 = mul2x<P, KPP>(p.ZZ, pp);
    bad |= is_zero2<P>(zz3) ? 1u : 0u;
    const Fd2 ppp = mul2x<P, KPP>(pp_, pp);
    const Fd2 zzz3 = mul2x<P, KZ>(p.ZZZ, ppp);
    const Fd2 q = mul2x<P, KPP>(p.X, pp);
    const Fd2 r2 = sqr2x<P, KSR>(r);
    const Fd2 x3 = {normalize(sub_lazy(r2.c0, add_lazy(ppp.c0, dbl_lazy(q.c0)), Bias<P, KX3, 3>::v)),
                    normalize(sub_lazy(r2.c1, add_lazy(ppp.c1, dbl_lazy(q.c1)), Bias<P, KX3, 3>::v))};
    // Y3 = R T - Y PPP with T = Q - X3, one reduction per component (all operands normalised)
    const Fd t0 = normalize(sub_lazy(q.c0, x3.c0, Bias<P, KT, 1>::v));
    const Fd t1 = normalize(sub_lazy(q.c1, x3.c1, Bias<P, KT, 1>::v));
    const Fd nt1 = normalize(neg_lazy(t1, Bias<P, KNT, 1>::v));
    const Fd nppp0 = normalize(neg_lazy(ppp.c0, Bias<P, KU, 1>::v));
    const Fd nppp1 = normalize(neg_lazy(ppp.c1, Bias<P, KU, 1>::v));
    const Fd xs[4] = {r.c0, r.c1, p.Y.c0, p.Y.c1};
    const Fd y0s[4] = {t0, nt1, nppp0, ppp.c1};
    const Fd y1s[4] = {t1, t0, nppp1, nppp0};
    return {x3, {mul4<P>(xs, y0s), mul4<P>(xs, y1s)}, zz3, zzz3};
  }
  // one digit-form coordinate (the k-th Fd of Pt) to its 8 words (words 8k .. 8k + 7 of store)
  BG_HD static void store_fd(uint32_t* w, const Fd& a) { to_words<P>(a, w); }
  BG_HD static void store(uint32_t* w, const Pt& p) {
    const Fd2* c[4] = {&p.X, &p.Y, &p.ZZ, &p.ZZZ};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      to_words<P>(c[k]->c0, w + 16 * k);
      to_words<P>(c[k]->c1, w + 16 * k + 8);
    }
  }
};

}  // namespace fd
}  // namespace bg
