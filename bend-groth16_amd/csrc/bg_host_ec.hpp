// Host-side BN254 arithmetic for the small serial tails of the device pipeline: the
// final weighted bucket sums and window combination of each MSM, affine normalisation,
// the proof combination (bend2/src/groth16.bend:298-308) and the key generator.
// 4 x 64-bit limbs, Montgomery R = 2^256 -- the same bytes as the device's 8 x u32.
#pragma once
#include <stdint.h>
#include <string.h>
#include <vector>

namespace bgh {

typedef unsigned __int128 u128;

struct Mod {
  uint64_t p[4];
  uint64_t n0;
  uint64_t one[4];
  uint64_t r2[4];
};

// BN254 base field q and scalar field r
static constexpr Mod MQ = {{0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
                           0x87d20782e4866389ULL,
                           {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL},
                           {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL}};
static constexpr Mod MR = {{0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
                           0xc2e1f593efffffffULL,
                           {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL},
                           {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL}};

template <const Mod& M>
struct F {
  uint64_t v[4];

  static F zero() { return F{{0, 0, 0, 0}}; }
  static F one() { return F{{M.one[0], M.one[1], M.one[2], M.one[3]}}; }
  static F raw(uint64_t a0, uint64_t a1 = 0, uint64_t a2 = 0, uint64_t a3 = 0) { return F{{a0, a1, a2, a3}}; }
  static F from_u32(const uint32_t* w) {
    F r;
    memcpy(r.v, w, 32);
    return r;
  }
  void to_u32(uint32_t* w) const { memcpy(w, v, 32); }
  bool is_zero() const { return (v[0] | v[1] | v[2] | v[3]) == 0; }
  bool operator==(const F& o) const { return memcmp(v, o.v, 32) == 0; }
  bool operator!=(const F& o) const { return !(*this == o); }

  static bool geq_p(const uint64_t* a) {
    for (int i = 3; i >= 0; i--) {
      if (a[i] > M.p[i]) return true;
      if (a[i] < M.p[i]) return false;
    }
    return true;
  }
  static void sub_p(uint64_t* a) {
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) {
      u128 t = (u128)a[i] - M.p[i] - br;
      a[i] = (uint64_t)t;
      br = (uint64_t)(t >> 64) & 1;
    }
  }
  F operator+(const F& b) const {
    F r;
    uint64_t c = 0;
    for (int i = 0; i < 4; i++) {
      u128 t = (u128)v[i] + b.v[i] + c;
      r.v[i] = (uint64_t)t;
      c = (uint64_t)(t >> 64);
    }
    if (c || geq_p(r.v)) sub_p(r.v);
    return r;
  }
  F operator-(const F& b) const {
    F r;
    uint64_t br = 0;
    for (int i = 0; i < 4; i++) {
      u128 t = (u128)v[i] - b.v[i] - br;
      r.v[i] = (uint64_t)t;
      br = (uint64_t)(t >> 64) & 1;
    }
    if (br) {
      uint64_t c = 0;
      for (int i = 0; i < 4; i++) {
        u128 t = (u128)r.v[i] + M.p[i] + c;
        r.v[i] = (uint64_t)t;
        c = (uint64_t)(t >> 64);
      }
    }
    return r;
  }
  F operator-() const { return zero() - *this; }
  F operator*(const F& b) const {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
      u128 c = 0;
      for (int j = 0; j < 4; j++) {
        c = (u128)v[j] * b.v[i] + t[j] + (uint64_t)(c >> 64);
        t[j] = (uint64_t)c;
      }
      c = (u128)t[4] + (uint64_t)(c >> 64);
      t[4] = (uint64_t)c;
      t[5] = (uint64_t)(c >> 64);
      uint64_t m = t[0] * M.n0;
      c = (u128)m * M.p[0] + t[0];
      for (int j = 1; j < 4; j++) {
        c = (u128)m * M.p[j] + t[j] + (uint64_t)(c >> 64);
        t[j - 1] = (uint64_t)c;
      }
      c = (u128)t[4] + (uint64_t)(c >> 64);
      t[3] = (uint64_t)c;
      t[4] = t[5] + (uint64_t)(c >> 64);
    }
    F r;
    memcpy(r.v, t, 32);
    if (t[4] || geq_p(r.v)) sub_p(r.v);
    return r;
  }
  F sqr() const { return *this * *this; }
  F dbl() const { return *this + *this; }
  // x^e, e given as 4 little-endian words
  F pow(const uint64_t* e) const {
    F acc = one();
    for (int i = 255; i >= 0; i--) {
      acc = acc.sqr();
      if ((e[i >> 6] >> (i & 63)) & 1) acc = acc * *this;
    }
    return acc;
  }
  F inv() const {
    uint64_t e[4];
    memcpy(e, M.p, 32);
    uint64_t br = 2;
    for (int i = 0; i < 4; i++) {
      u128 t = (u128)e[i] - br;
      e[i] = (uint64_t)t;
      br = (uint64_t)(t >> 64) & 1;
    }
    return pow(e);
  }
  static F to_mont(const F& canon) { return canon * F{{M.r2[0], M.r2[1], M.r2[2], M.r2[3]}}; }
  F from_mont() const { return *this * raw(1); }
};

using Fq = F<MQ>;
using Fr = F<MR>;

struct Fq2 {
  Fq c0, c1;
  static Fq2 zero() { return {Fq::zero(), Fq::zero()}; }
  static Fq2 one() { return {Fq::one(), Fq::zero()}; }
  bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
  bool operator==(const Fq2& o) const { return c0 == o.c0 && c1 == o.c1; }
  Fq2 operator+(const Fq2& b) const { return {c0 + b.c0, c1 + b.c1}; }
  Fq2 operator-(const Fq2& b) const { return {c0 - b.c0, c1 - b.c1}; }
  Fq2 operator-() const { return {-c0, -c1}; }
  Fq2 operator*(const Fq2& b) const {
    Fq t0 = c0 * b.c0, t1 = c1 * b.c1;
    Fq t2 = (c0 + c1) * (b.c0 + b.c1);
    return {t0 - t1, t2 - t0 - t1};
  }
  Fq2 sqr() const { return {(c0 + c1) * (c0 - c1), (c0 + c0) * c1}; }
  Fq2 dbl() const { return *this + *this; }
  Fq2 inv() const {
    Fq ni = (c0.sqr() + c1.sqr()).inv();
    return {c0 * ni, -(c1 * ni)};
  }
  static Fq2 from_u32(const uint32_t* w) { return {Fq::from_u32(w), Fq::from_u32(w + 8)}; }
  void to_u32(uint32_t* w) const {
    c0.to_u32(w);
    c1.to_u32(w + 8);
  }
};

// XYZZ points (x = X/ZZ, y = Y/ZZZ; infinity: ZZ = 0) with the same formulas as the
// device (EFD add-2008-s, madd-2008-s, dbl-2008-s-1) and exact special cases.
template <class E>
struct Pt {
  E X, Y, ZZ, ZZZ;
  static Pt inf() { return {E::zero(), E::zero(), E::zero(), E::zero()}; }
  bool is_inf() const { return ZZ.is_zero(); }
  static Pt from_affine(const E& x, const E& y) {
    if (x.is_zero() && y.is_zero()) return inf();
    return {x, y, E::one(), E::one()};
  }
  Pt dbl() const {
    if (is_inf()) return *this;
    E u = Y.dbl(), v = u.sqr(), w = u * v, s = X * v, x2 = X.sqr();
    E m = x2.dbl() + x2;
    E x3 = m.sqr() - s.dbl();
    E y3 = m * (s - x3) - w * Y;
    return {x3, y3, v * ZZ, w * ZZZ};
  }
  Pt add(const Pt& q) const {
    if (is_inf()) return q;
    if (q.is_inf()) return *this;
    E u1 = X * q.ZZ, u2 = q.X * ZZ, s1 = Y * q.ZZZ, s2 = q.Y * ZZZ;
    E p = u2 - u1, r = s2 - s1;
    if (p.is_zero()) return r.is_zero() ? dbl() : inf();
    E pp = p.sqr(), ppp = p * pp, qq = u1 * pp;
    E x3 = r.sqr() - ppp - qq.dbl();
    E y3 = r * (qq - x3) - s1 * ppp;
    return {x3, y3, ZZ * q.ZZ * pp, ZZZ * q.ZZZ * ppp};
  }
  // mixed add with an affine point (x, y) that is not infinity: madd-2008-s
  Pt madd(const E& x, const E& y) const {
    if (is_inf()) return {x, y, E::one(), E::one()};
    E u2 = x * ZZ, s2 = y * ZZZ;
    E p = u2 - X, r = s2 - Y;
    if (p.is_zero()) return r.is_zero() ? from_affine(x, y).dbl() : inf();
    E pp = p.sqr(), ppp = p * pp, qq = X * pp;
    E x3 = r.sqr() - ppp - qq.dbl();
    E y3 = r * (qq - x3) - Y * ppp;
    return {x3, y3, ZZ * pp, ZZZ * ppp};
  }
  Pt neg() const { return {X, -Y, ZZ, ZZZ}; }
  Pt dbl_n(int k) const {
    Pt r = *this;
    for (int i = 0; i < k; i++) r = r.dbl();
    return r;
  }
  // affine (x, y); infinity -> (0, 0)
  void to_affine(E& x, E& y) const {
    if (is_inf()) {
      x = E::zero();
      y = E::zero();
      return;
    }
    E t = (ZZ * ZZZ).inv();
    x = X * ZZZ * t;
    y = Y * ZZ * t;
  }
  // double-and-add by a canonical 256-bit scalar (4 little-endian words)
  Pt mul(const uint64_t* k) const {
    Pt acc = inf();
    for (int i = 255; i >= 0; i--) {
      acc = acc.dbl();
      if ((k[i >> 6] >> (i & 63)) & 1) acc = acc.add(*this);
    }
    return acc;
  }
};

using P1 = Pt<Fq>;
using P2 = Pt<Fq2>;

// element I/O in packed u32 words
inline Fq ld(const uint32_t* w, Fq*) { return Fq::from_u32(w); }
inline Fq2 ld(const uint32_t* w, Fq2*) { return Fq2::from_u32(w); }
template <class E>
constexpr int words() { return sizeof(E) / 4; }

template <class E>
Pt<E> load_xyzz(const uint32_t* w) {
  const int k = words<E>();
  return {ld(w, (E*)nullptr), ld(w + k, (E*)nullptr), ld(w + 2 * k, (E*)nullptr), ld(w + 3 * k, (E*)nullptr)};
}
template <class E>
Pt<E> load_affine(const uint32_t* w) {
  const int k = words<E>();
  return Pt<E>::from_affine(ld(w, (E*)nullptr), ld(w + k, (E*)nullptr));
}
template <class E>
void store_affine(const Pt<E>& p, uint32_t* w) {
  const int k = words<E>();
  E x, y;
  p.to_affine(x, y);
  x.to_u32(w);
  y.to_u32(w + k);
}

// Affine normalisation of a batch with one inversion (Montgomery's trick).
template <class E>
void batch_to_affine(const Pt<E>* pts, size_t n, E* xs, E* ys) {
  std::vector<E> pre(n);
  E run = E::one();
  for (size_t i = 0; i < n; i++) {
    pre[i] = run;
    if (!pts[i].is_inf()) run = run * (pts[i].ZZ * pts[i].ZZZ);
  }
  E inv = run.inv();
  for (size_t i = n; i-- > 0;) {
    if (pts[i].is_inf()) {
      xs[i] = E::zero();
      ys[i] = E::zero();
      continue;
    }
    E d = pts[i].ZZ * pts[i].ZZZ;
    E t = inv * pre[i];
    inv = inv * d;
    xs[i] = pts[i].X * pts[i].ZZZ * t;
    ys[i] = pts[i].Y * pts[i].ZZ * t;
  }
}

}  // namespace bgh
