// Groth16 verification with an own BN254 optimal-ate pairing (host; C ABI: bg_verify,
// bg_pairing_check).
//
// The reference verifies through arkworks (ref/src/verify.rs:8-27: prepare_verifying_key +
// Groth16::verify_proof, i.e. e(A, B) = e(alpha, beta) e(IC(x), gamma) e(C, delta)); this is
// the same check written from the textbook construction:
//   Fq12 = Fq6[w]/(w^2 - v), Fq6 = Fq2[v]/(v^3 - xi), xi = 9 + u (D-type sextic twist
//   E': y^2 = x^3 + 3/xi), Miller loop over 6x + 2 (x = 4965661367192848881) with affine
//   lines, the two Frobenius-twisted corrections, and the final exponentiation
//   (p^6 - 1)(p^2 + 1) followed by a plain exponentiation by (p^4 - p^2 + 1)/r.
// Points are checked on their curves and B in the order-r subgroup ([r]B = O); a point
// that fails is an error (BG_ERR_ARG), not a rejection, as arkworks treats it.
#include "../../include/bg16.h"
#include "bg_host_ec.hpp"
#include "bg_runtime.hpp"
#include <string.h>
#include <vector>

namespace {

using bgh::Fq;
using bgh::Fq2;
using bgh::Fr;

Fq fq_small(uint64_t v) { return Fq::to_mont(Fq::raw(v)); }

// ---------------------------------------------------------------- Fq2 helpers
Fq2 conj(const Fq2& a) { return {a.c0, -a.c1}; }
Fq2 mul_fq(const Fq2& a, const Fq& s) { return {a.c0 * s, a.c1 * s}; }
// a * xi, xi = 9 + u: (9 a0 - a1) + (a0 + 9 a1) u
Fq2 mul_xi(const Fq2& a) {
  Fq a0_9 = a.c0.dbl().dbl().dbl() + a.c0, a1_9 = a.c1.dbl().dbl().dbl() + a.c1;
  return {a0_9 - a.c1, a.c0 + a1_9};
}
Fq2 pow2(const Fq2& a, const uint64_t* e, int words) {
  Fq2 r = Fq2::one();
  for (int i = 64 * words - 1; i >= 0; i--) {
    r = r.sqr();
    if ((e[i / 64] >> (i % 64)) & 1) r = r * a;
  }
  return r;
}

// ---------------------------------------------------------------- Fq6, Fq12
struct Fq6 {
  Fq2 c0, c1, c2;
  static Fq6 zero() { return {Fq2::zero(), Fq2::zero(), Fq2::zero()}; }
  static Fq6 one() { return {Fq2::one(), Fq2::zero(), Fq2::zero()}; }
  Fq6 operator+(const Fq6& b) const { return {c0 + b.c0, c1 + b.c1, c2 + b.c2}; }
  Fq6 operator-(const Fq6& b) const { return {c0 - b.c0, c1 - b.c1, c2 - b.c2}; }
  Fq6 operator-() const { return {-c0, -c1, -c2}; }
  Fq6 operator*(const Fq6& b) const {
    Fq2 t0 = c0 * b.c0, t1 = c1 * b.c1, t2 = c2 * b.c2;
    return {t0 + mul_xi(c1 * b.c2 + c2 * b.c1), c0 * b.c1 + c1 * b.c0 + mul_xi(t2), c0 * b.c2 + t1 + c2 * b.c0};
  }
  Fq6 mul_v() const { return {mul_xi(c2), c0, c1}; }  // * v, v^3 = xi
  Fq6 inv() const {
    Fq2 A = c0.sqr() - mul_xi(c1 * c2);
    Fq2 B = mul_xi(c2.sqr()) - c0 * c1;
    Fq2 C = c1.sqr() - c0 * c2;
    Fq2 F = c0 * A + mul_xi(c2 * B + c1 * C);
    Fq2 fi = F.inv();
    return {A * fi, B * fi, C * fi};
  }
  bool is_one() const { return c0 == Fq2::one() && c1.is_zero() && c2.is_zero(); }
  bool is_zero() const { return c0.is_zero() && c1.is_zero() && c2.is_zero(); }
};

struct Fq12 {
  Fq6 c0, c1;  // c0 + c1 w, w^2 = v
  static Fq12 one() { return {Fq6::one(), Fq6::zero()}; }
  Fq12 operator*(const Fq12& b) const {
    Fq6 t0 = c0 * b.c0, t1 = c1 * b.c1;
    return {t0 + t1.mul_v(), (c0 + c1) * (b.c0 + b.c1) - t0 - t1};
  }
  Fq12 sqr() const { return *this * *this; }
  Fq12 conj() const { return {c0, -c1}; }
  Fq12 inv() const {
    Fq6 t = c0 * c0 - (c1 * c1).mul_v();
    Fq6 ti = t.inv();
    return {c0 * ti, -(c1 * ti)};
  }
  bool is_one() const { return c0.is_one() && c1.is_zero(); }
};

// Frobenius constants gamma_{k,i} = xi^(i (p^k - 1) / 6) for the w^i coefficients
struct Frob {
  Fq2 g[4][6];
  Frob() {
    // (p - 1) / 6
    static const uint64_t P16[4] = {0x34b017592414d4e1ULL, 0xee9591c2e6bda1c2ULL, 0xf40d60f3c0403964ULL,
                                    0x0810b7bdd032f006ULL};
    Fq2 xi = {fq_small(9), fq_small(1)};
    Fq2 g1 = pow2(xi, P16, 4);
    Fq2 acc = Fq2::one();
    for (int i = 0; i < 6; i++) {
      g[1][i] = acc;
      acc = acc * g1;
    }
    for (int i = 0; i < 6; i++) {
      g[0][i] = Fq2::one();
      g[2][i] = g[1][i] * conj(g[1][i]);  // (w^i)^(p^2) = conj(g1_i) g1_i w^i
      g[3][i] = g[2][i] * g[1][i];        // (w^i)^(p^3) = conj(g2_i) g1_i w^i, g2_i in Fq
    }
  }
};
const Frob& frob_consts() {
  static const Frob F;
  return F;
}

// w^i coefficients (i = 0..5) of c0 + c1 w with c_j = c_j0 + c_j1 v + c_j2 v^2, v = w^2
void to_coeffs(const Fq12& a, Fq2* k) {
  k[0] = a.c0.c0; k[1] = a.c1.c0; k[2] = a.c0.c1; k[3] = a.c1.c1; k[4] = a.c0.c2; k[5] = a.c1.c2;
}
Fq12 from_coeffs(const Fq2* k) { return {{k[0], k[2], k[4]}, {k[1], k[3], k[5]}}; }
Fq12 frobenius(const Fq12& a, int k) {
  Fq2 c[6];
  to_coeffs(a, c);
  const Frob& F = frob_consts();
  for (int i = 0; i < 6; i++) c[i] = ((k & 1) ? conj(c[i]) : c[i]) * F.g[k][i];
  return from_coeffs(c);
}

// ---------------------------------------------------------------- curves
struct G1a {
  Fq x, y;
  bool inf;
};
struct G2a {
  Fq2 x, y;
  bool inf;
};

Fq2 twist_b() {  // 3 / xi
  Fq2 xi = {fq_small(9), fq_small(1)};
  return mul_fq(xi.inv(), fq_small(3));
}
bool on_g1(const G1a& p) { return p.inf || p.y.sqr() == p.x.sqr() * p.x + fq_small(3); }
bool on_g2(const G2a& q) { return q.inf || q.y.sqr() == q.x.sqr() * q.x + twist_b(); }

// affine G2 arithmetic on the twist (exceptional cases handled)
G2a g2_add(const G2a& a, const G2a& b) {
  if (a.inf) return b;
  if (b.inf) return a;
  Fq2 lam;
  if (a.x == b.x) {
    if (!(a.y == b.y) || a.y.is_zero()) return {Fq2::zero(), Fq2::zero(), true};
    Fq2 x2 = a.x.sqr();
    lam = (x2.dbl() + x2) * a.y.dbl().inv();
  } else {
    lam = (b.y - a.y) * (b.x - a.x).inv();
  }
  Fq2 x3 = lam.sqr() - a.x - b.x;
  return {x3, lam * (a.x - x3) - a.y, false};
}
bool in_g2_subgroup(const G2a& q) {
  if (q.inf) return true;
  // [r] q = O, double-and-add over r's bits
  static const uint64_t R[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                                0x30644e72e131a029ULL};
  G2a acc = {Fq2::zero(), Fq2::zero(), true};
  for (int i = 255; i >= 0; i--) {
    acc = g2_add(acc, acc);
    if ((R[i / 64] >> (i % 64)) & 1) acc = g2_add(acc, q);
  }
  return acc.inf;
}

// line through T and S (tangent if S = T) on the twist, evaluated at P; advances T
Fq12 line_step(G2a& T, const G2a& S, const G1a& P, bool tangent) {
  Fq2 lam;
  if (tangent) {
    Fq2 x2 = T.x.sqr();
    lam = (x2.dbl() + x2) * T.y.dbl().inv();
  } else {
    lam = (S.y - T.y) * (S.x - T.x).inv();
  }
  // y_P - lam x_P w + (lam x_T - y_T) w^3
  Fq2 c[6] = {{P.y, Fq::zero()}, -mul_fq(lam, P.x), Fq2::zero(), lam * T.x - T.y, Fq2::zero(), Fq2::zero()};
  Fq2 x3 = lam.sqr() - T.x - S.x;
  Fq2 y3 = lam * (T.x - x3) - T.y;
  T = {x3, y3, false};
  return from_coeffs(c);
}

Fq12 miller_loop(const G1a& P, const G2a& Q) {
  if (P.inf || Q.inf) return Fq12::one();
  static const uint64_t S[2] = {0x9d797039be763ba8ULL, 0x1ULL};  // 6x + 2, 65 bits
  Fq12 f = Fq12::one();
  G2a T = Q;
  for (int i = 63; i >= 0; i--) {
    f = f.sqr() * line_step(T, T, P, true);
    if ((S[i / 64] >> (i % 64)) & 1) f = f * line_step(T, Q, P, false);
  }
  const Frob& F = frob_consts();
  G2a Q1 = {conj(Q.x) * F.g[1][2], conj(Q.y) * F.g[1][3], false};  // pi(Q)
  G2a Q2 = {Q.x * F.g[2][2], -(Q.y * F.g[2][3]), false};           // -pi^2(Q)
  f = f * line_step(T, Q1, P, false);
  f = f * line_step(T, Q2, P, false);
  return f;
}

Fq12 final_exp(const Fq12& f) {
  // easy part: f^((p^6 - 1)(p^2 + 1))
  Fq12 t = f.conj() * f.inv();
  t = frobenius(t, 2) * t;
  // hard part: t^((p^4 - p^2 + 1) / r)
  static const uint64_t E[12] = {0xe81bb482ccdf42b1ULL, 0x5abf5cc4f49c36d4ULL, 0xf1154e7e1da014fdULL,
                                 0xdcc7b44c87cdbacfULL, 0xaaa441e3954bcf8aULL, 0x6b887d56d5095f23ULL,
                                 0x79581e16f3fd90c6ULL, 0x3b1b1355d189227dULL, 0x4e529a5861876f6bULL,
                                 0x6c0eb522d5b12278ULL, 0x331ec15183177fafULL, 0x01baaa710b0759adULL};
  Fq12 r = Fq12::one();
  for (int i = 64 * 12 - 1; i >= 0; i--) {
    r = r.sqr();
    if ((E[i / 64] >> (i % 64)) & 1) r = r * t;
  }
  return r;
}

// decoding (words16 or packed; Montgomery)
struct Reader {
  const uint32_t* p;
  bool w16;
  Fq fq() {
    uint32_t w[8];
    if (w16) {
      for (int i = 0; i < 8; i++) {
        BG_REQUIRE(((p[2 * i] | p[2 * i + 1]) >> 16) == 0, BG_ERR_ARG, "coordinate word above 16 bits");
        w[i] = p[2 * i] | (p[2 * i + 1] << 16);
      }
      p += 16;
    } else {
      memcpy(w, p, 32);
      p += 8;
    }
    Fq x = Fq::from_u32(w);
    BG_REQUIRE(!Fq::geq_p(x.v), BG_ERR_ARG, "coordinate not below q");
    return x;
  }
  Fr fr() {
    uint32_t w[8];
    if (w16) {
      for (int i = 0; i < 8; i++) {
        BG_REQUIRE(((p[2 * i] | p[2 * i + 1]) >> 16) == 0, BG_ERR_ARG, "input word above 16 bits");
        w[i] = p[2 * i] | (p[2 * i + 1] << 16);
      }
      p += 16;
    } else {
      memcpy(w, p, 32);
      p += 8;
    }
    Fr x = Fr::from_u32(w);
    BG_REQUIRE(!Fr::geq_p(x.v), BG_ERR_ARG, "public input not below r");
    return x;
  }
  G1a g1() {
    Fq x = fq(), y = fq();
    return {x, y, x.is_zero() && y.is_zero()};
  }
  G2a g2() {
    Fq2 x = {fq(), fq()}, y = {fq(), fq()};
    return {x, y, x.is_zero() && y.is_zero()};
  }
  uint32_t u32() { return *p++; }
};

bgh::P1 to_pt(const G1a& a) { return a.inf ? bgh::P1::inf() : bgh::P1::from_affine(a.x, a.y); }
G1a from_pt(const bgh::P1& p) {
  if (p.is_inf()) return {Fq::zero(), Fq::zero(), true};
  Fq x, y;
  p.to_affine(x, y);
  return {x, y, false};
}

bool pairing_product_is_one(const std::vector<G1a>& ps, const std::vector<G2a>& qs) {
  Fq12 f = Fq12::one();
  for (size_t i = 0; i < ps.size(); i++) f = f * miller_loop(ps[i], qs[i]);
  return final_exp(f).is_one();
}

}  // namespace

using namespace bg;

extern "C" int bg_pairing_check(const uint32_t* g1s, const uint32_t* g2s, uint32_t n, int flags) {
  int result = 0;
  int rc = guard([&] {
    BG_REQUIRE(n == 0 || (g1s && g2s), BG_ERR_ARG, "bg_pairing_check: null argument");
    const bool w16 = (flags & 1) == BG_FMT_WORDS16;
    Reader r1{g1s, w16}, r2{g2s, w16};
    std::vector<G1a> ps;
    std::vector<G2a> qs;
    for (uint32_t i = 0; i < n; i++) {
      G1a p = r1.g1();
      G2a q = r2.g2();
      BG_REQUIRE(on_g1(p), BG_ERR_ARG, "bg_pairing_check: G1 point %u is not on the curve", i);
      BG_REQUIRE(on_g2(q) && in_g2_subgroup(q), BG_ERR_ARG, "bg_pairing_check: G2 point %u is not in G2", i);
      ps.push_back(p);
      qs.push_back(q);
    }
    result = pairing_product_is_one(ps, qs) ? 1 : 0;
  });
  return rc < 0 ? rc : result;
}

extern "C" int bg_verify(const uint32_t* vk, uint64_t vk_words, const uint32_t* proof, const uint32_t* inputs,
                         uint32_t n_inputs, int flags) {
  int result = 0;
  int rc = guard([&] {
    BG_REQUIRE(vk && proof && (n_inputs == 0 || inputs), BG_ERR_ARG, "bg_verify: null argument");
    const bool w16 = (flags & 1) == BG_FMT_WORDS16;
    const uint64_t ew = w16 ? 16 : 8;
    BG_REQUIRE(vk_words >= 14 * ew + 1, BG_ERR_ARG, "bg_verify: verifying key too short");
    Reader rv{vk, w16};
    G1a alpha = rv.g1();
    G2a beta = rv.g2(), gamma = rv.g2(), delta = rv.g2();
    uint32_t count = rv.u32();
    BG_REQUIRE(count == n_inputs + 1, BG_ERR_ARG, "bg_verify: key has %u IC points for %u public inputs", count,
               n_inputs);
    BG_REQUIRE(vk_words == 14 * ew + 1 + 2ull * ew * count, BG_ERR_ARG, "bg_verify: verifying key size mismatch");
    std::vector<G1a> ic(count);
    for (auto& p : ic) p = rv.g1();
    Reader rp{proof, w16};
    G1a A = rp.g1();
    G2a B = rp.g2();
    G1a C = rp.g1();
    BG_REQUIRE(on_g1(alpha) && on_g1(A) && on_g1(C), BG_ERR_ARG, "bg_verify: G1 point not on the curve");
    for (auto& p : ic) BG_REQUIRE(on_g1(p), BG_ERR_ARG, "bg_verify: IC point not on the curve");
    for (const G2a* q : {&beta, &gamma, &delta, &B})
      BG_REQUIRE(on_g2(*q) && in_g2_subgroup(*q), BG_ERR_ARG, "bg_verify: G2 point not in G2");
    // IC(x) = IC_0 + sum_i x_i IC_i (inputs Montgomery Fr, canonical scalars for the sum)
    Reader ri{inputs, w16};
    bgh::P1 acc = to_pt(ic[0]);
    for (uint32_t i = 0; i < n_inputs; i++) {
      Fr x = ri.fr().from_mont();
      acc = acc.add(to_pt(ic[i + 1]).mul(x.v));
    }
    G1a icx = from_pt(acc);
    // e(-A, B) e(alpha, beta) e(IC(x), gamma) e(C, delta) = 1
    G1a negA = A;
    if (!negA.inf) negA.y = -negA.y;
    result = pairing_product_is_one({negA, alpha, icx, C}, {B, beta, gamma, delta}) ? 1 : 0;
  });
  return rc < 0 ? rc : result;
}
