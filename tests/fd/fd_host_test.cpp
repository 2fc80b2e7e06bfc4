// Host test of the digit-domain bucket accumulation (bg_fd.hpp, bg_curve_fd.hpp): the
// same code the device runs, compiled for the CPU with a product accumulator that flags
// any column above 2^64.  Checked against the word-domain host arithmetic
// (bg_host_ec.hpp, exact special cases):
//   - field products, squarings and the Fq2 products on random operands, including
//     operands re-represented near the bounds the accumulation allows;
//   - G1 and G2 runs of 1..40 mixed additions with random signs, against the exact sum,
//     with the accumulator's coordinates re-represented at random multiples of p up to
//     their stated bounds before every addition, and points given as W, W + p or W + kp
//     with a top word at or above 2p's (the reducing load);
//   - the doubling / cancellation flag (a repeated or opposite point in a run).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

static bool g_overflow = false;
struct CAcc {
  unsigned __int128 v;
  CAcc(uint64_t x = 0) : v(x) {}
  CAcc& operator+=(uint64_t x) {
    v += x;
    if (v >> 64) g_overflow = true;
    return *this;
  }
  CAcc& operator>>=(int s) {
    v >>= s;
    return *this;
  }
  explicit operator uint32_t() const { return (uint32_t)v; }
};
#define BG_FD_ACC CAcc
#include "bg_curve_fd.hpp"
#include "bg_host_ec.hpp"

using namespace bg::fd;
using bgh::Fq;
using bgh::Fq2;
using bgh::P1;
using bgh::P2;

static int g_fail = 0;
#define CHECK(c)                                                \
  do {                                                          \
    if (!(c)) {                                                 \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      if (++g_fail > 20) return;                                \
    }                                                           \
  } while (0)

static std::mt19937_64 rng(20261204);

// ---------------------------------------------------------------- small big integers
// 5 x 64-bit limbs (values below 2^261)
struct Big {
  uint64_t w[5];
};
static Big from_digits(const Fd& a) {
  Big b{};
  for (int i = 0; i < 9; i++) {
    // add a.d[i] << (29 i)
    const int bit = 29 * i, li = bit / 64, sh = bit % 64;
    unsigned __int128 add = (unsigned __int128)a.d[i] << sh;
    for (int k = li; k < 5 && add; k++) {
      unsigned __int128 t = (unsigned __int128)b.w[k] + (uint64_t)add;
      b.w[k] = (uint64_t)t;
      add = (add >> 64) + (t >> 64);
    }
  }
  return b;
}
static Fd to_digits(const Big& b) {
  Fd r;
  for (int i = 0; i < 9; i++) {
    const int bit = 29 * i, li = bit / 64, sh = bit % 64;
    unsigned __int128 v = b.w[li] >> sh;
    if (sh > 35 && li + 1 < 5) v |= (unsigned __int128)b.w[li + 1] << (64 - sh);
    r.d[i] = i < 8 ? (uint32_t)(v & MASK) : (uint32_t)v;
  }
  return r;
}
static const Big& pbig() {
  static Big p = {{bgh::MQ.p[0], bgh::MQ.p[1], bgh::MQ.p[2], bgh::MQ.p[3], 0}};
  return p;
}
static bool geq(const Big& a, const Big& b) {
  for (int i = 4; i >= 0; i--)
    if (a.w[i] != b.w[i]) return a.w[i] > b.w[i];
  return true;
}
static Big add(const Big& a, const Big& b) {
  Big r;
  unsigned __int128 c = 0;
  for (int i = 0; i < 5; i++) {
    c += (unsigned __int128)a.w[i] + b.w[i];
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
static Big sub(const Big& a, const Big& b) {
  Big r;
  uint64_t br = 0;
  for (int i = 0; i < 5; i++) {
    unsigned __int128 t = (unsigned __int128)a.w[i] - b.w[i] - br;
    r.w[i] = (uint64_t)t;
    br = (uint64_t)(t >> 64) & 1;
  }
  return r;
}
// canonical residue and the multiple of p removed
static Big canon(Big a, int* k = nullptr) {
  int n = 0;
  while (geq(a, pbig())) {
    a = sub(a, pbig());
    n++;
  }
  if (k) *k = n;
  return a;
}
// value in units of p (for bound checks)
static double units(const Fd& a) {
  const Big b = from_digits(a);
  double v = 0, pv = 0;
  for (int i = 4; i >= 0; i--) {
    v = v * 18446744073709551616.0 + (double)b.w[i];
    pv = pv * 18446744073709551616.0 + (double)pbig().w[i];
  }
  return v / pv;
}
// the same residue plus a random multiple of p, staying below bound p
static Fd rerep(const Fd& a, double bound) {
  const Big c = canon(from_digits(a));
  const double cu = units(to_digits(c));
  int jmax = (int)(bound - cu - 1e-9);
  if (jmax < 0) jmax = 0;
  const int j = (int)(rng() % (uint64_t)(jmax + 1));
  Big v = c;
  for (int i = 0; i < j; i++) v = add(v, pbig());
  return to_digits(v);
}

// digit domain -> word-domain host element (canonical)
static Fq host_of(const Fd& a) {
  uint32_t w[8];
  to_words<FqDig>(a, w);
  Fq x = Fq::from_u32(w);
  uint64_t v[4];
  memcpy(v, x.v, 32);
  while (Fq::geq_p(v)) Fq::sub_p(v);
  Fq r;
  memcpy(r.v, v, 32);
  return r;
}
static Fd dig_of(const Fq& x) {
  uint32_t w[8];
  x.to_u32(w);
  return reduce_raw<FqDig>(bg::m29::split<5>(w));
}
static Fq rand_fq() {
  return Fq::to_mont(Fq::raw(rng(), rng(), rng(), rng() % 0x30644e72e131a029ULL));
}

// ---------------------------------------------------------------- field checks
static void field_checks() {
  for (int it = 0; it < 3000; it++) {
    const Fq a = rand_fq(), b = rand_fq(), c = rand_fq(), d = rand_fq();
    // operands near the bounds the formulas use (10 p normalised, 2p)
    const Fd A = rerep(dig_of(a), 11.5), B = rerep(dig_of(b), 2.0), C = rerep(dig_of(c), 6.0),
             D = rerep(dig_of(d), 2.0);
    CHECK(host_of(mul<FqDig>(A, B)) == a * b);
    CHECK(host_of(sqr<FqDig>(A)) == a.sqr());
    CHECK(units(sqr<FqDig>(A)) < 1.8);
    CHECK(host_of(mul2<FqDig>(A, B, C, D)) == a * b + c * d);
    // lazy differences feed products
    const Fd t = sub_lazy(B, C, Bias<FqDig, 8, 1>::v);  // b - c (c < 7p)
    CHECK(host_of(mul<FqDig>(A, t)) == a * (b - c));
    const Fd u = neg_lazy(D, Bias<FqDig, 3, 1>::v);
    CHECK(host_of(mul2<FqDig>(A, t, C, u)) == a * (b - c) - c * d);
    Fd xs[4] = {A, B, C, D}, ys[4] = {D, C, B, A};
    CHECK(host_of(mul4<FqDig>(xs, ys)) == a * d + b * c + c * b + d * a);
    const Fd2 x2 = {A, C}, y2 = {B, D};
    const Fd2 p2 = mul2x<FqDig, 3>(x2, y2);
    const Fq2 hx{a, c}, hy{b, d}, hp = hx * hy;
    CHECK(host_of(p2.c0) == hp.c0 && host_of(p2.c1) == hp.c1);
    const Fd2 s2 = sqr2x<FqDig, 13>(Fd2{rerep(dig_of(a), 11.5), rerep(dig_of(c), 11.5)});
    const Fq2 hs = hx.sqr();
    CHECK(host_of(s2.c0) == hs.c0 && host_of(s2.c1) == hs.c1);
    // word-domain loads: W, W + p, and W + kp at or above 2p (the reducing path)
    uint32_t w[8];
    a.to_u32(w);
    Big wb{};
    memcpy(wb.w, a.v, 32);
    const int k = (int)(rng() % 5);  // W + 4p < 2^256
    for (int i = 0; i < k; i++) wb = add(wb, pbig());
    uint32_t w2[8];
    memcpy(w2, wb.w, 32);
    const Fd r = load_raw<FqDig>(w2);
    CHECK(units(r) < 64.0);
    CHECK(host_of(mul<FqDig>(r, from_consts(FqDig::one))) == a);
    CHECK(!g_overflow);
  }
}

// ---------------------------------------------------------------- accumulation runs
template <class E>
struct HostSide;
template <>
struct HostSide<Fq> {
  using Acc = G1Acc<FqDig>;
  using HP = P1;
  static P1 gen() { return P1::from_affine(Fq::one(), Fq::one() + Fq::one()); }
  static void words(const P1& p, uint32_t* w) {
    Fq x, y;
    p.to_affine(x, y);
    x.to_u32(w);
    y.to_u32(w + 8);
  }
  // re-represent the accumulator state at random multiples of p inside its bounds
  static void stress(Acc::Pt& a) {
    a.X = rerep(a.X, 6.2);
    a.Y = rerep(a.Y, 1.39);
    a.ZZ = rerep(a.ZZ, 1.01);
    a.ZZZ = rerep(a.ZZZ, 1.01);
  }
  static P1 host(const Acc::Pt& a) {
    uint32_t w[32];
    Acc::store(w, a);
    return bgh::load_xyzz<Fq>(w);
  }
  // what the biases need: X < 7p (KX, KT), Y < 2p (KY), ZZ, ZZZ < 2p
  static void bounds(const Acc::Pt& a, int& fail) {
    if (!(units(a.X) < 7 && units(a.Y) < 2 && units(a.ZZ) < 2 && units(a.ZZZ) < 2)) fail++;
  }
};
template <>
struct HostSide<Fq2> {
  using Acc = G2Acc<FqDig>;
  using HP = P2;
  static P2 gen() {
    Fq2 x{Fq::to_mont(Fq::raw(0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL,
                              0x1800deef121f1e76ULL)),
          Fq::to_mont(Fq::raw(0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL,
                              0x198e9393920d483aULL))};
    Fq2 y{Fq::to_mont(Fq::raw(0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL,
                              0x12c85ea5db8c6debULL)),
          Fq::to_mont(Fq::raw(0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL,
                              0x090689d0585ff075ULL))};
    return P2::from_affine(x, y);
  }
  static void words(const P2& p, uint32_t* w) {
    Fq2 x, y;
    p.to_affine(x, y);
    x.to_u32(w);
    y.to_u32(w + 16);
  }
  static void stress(Acc::Pt& a) {
    Fd2* c[4] = {&a.X, &a.Y, &a.ZZ, &a.ZZZ};
    const double b[4] = {7.99, 1.85, 1.065, 1.03};
    for (int k = 0; k < 4; k++) {
      c[k]->c0 = rerep(c[k]->c0, b[k]);
      c[k]->c1 = rerep(c[k]->c1, b[k]);
    }
  }
  static P2 host(const Acc::Pt& a) {
    uint32_t w[64];
    Acc::store(w, a);
    return bgh::load_xyzz<Fq2>(w);
  }
  static void bounds(const Acc::Pt& a, int& fail) {
    const Fd2* c[4] = {&a.X, &a.Y, &a.ZZ, &a.ZZZ};
    const double b[4] = {8, 2, 2, 2};  // X < 8p (KX, KT), Y < 2p (KY), ZZ, ZZZ < 2p (KZ)
    for (int k = 0; k < 4; k++)
      if (!(units(c[k]->c0) < b[k] && units(c[k]->c1) < b[k])) fail++;
  }
};

template <class E>
static bool same_point(const bgh::Pt<E>& a, const bgh::Pt<E>& b) {
  E ax, ay, bx, by;
  a.to_affine(ax, ay);
  b.to_affine(bx, by);
  return ax == bx && ay == by;
}

// W + kp representation of a word-domain point (k random; a top word at or above 2p's
// takes the reducing load)
static void perturb(uint32_t* w, int elems) {
  for (int e = 0; e < elems; e++) {
    Big b{};
    memcpy(b.w, w + 8 * e, 32);
    const int k = (int)(rng() % 5);  // W + 4p < 2^256
    for (int i = 0; i < k; i++) b = add(b, pbig());
    memcpy(w + 8 * e, b.w, 32);
  }
}

template <class E>
static void runs(int nruns) {
  using H = HostSide<E>;
  using Acc = typename H::Acc;
  using HP = typename H::HP;
  constexpr int AW = Acc::aff_words;
  const HP g = H::gen();
  // a pool of points k G
  std::vector<HP> pool;
  std::vector<uint32_t> pw;
  for (int i = 0; i < 64; i++) {
    uint64_t k[4] = {rng(), rng(), rng(), rng() >> 4};
    const HP p = g.mul(k);
    pool.push_back(p);
    uint32_t w[AW];
    H::words(p, w);
    pw.insert(pw.end(), w, w + AW);
  }
  int bound_fail = 0;
  for (int run = 0; run < nruns; run++) {
    const int len = 1 + (int)(rng() % 40);
    const bool stress = (run & 1) != 0;
    HP want = HP::inf();
    typename Acc::Pt acc;
    uint32_t bad = 0;
    // distinct points of the pool (a repeated one could make acc = +-point: the flagged case)
    std::vector<int> idx(pool.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
    std::shuffle(idx.begin(), idx.end(), rng);
    for (int i = 0; i < len; i++) {
      const int j = idx[(size_t)i];
      const bool neg = (rng() & 1) != 0;
      uint32_t w[AW];
      memcpy(w, &pw[(size_t)j * AW], sizeof w);
      if (stress) perturb(w, AW / 8);
      const typename Acc::Raw raw = Acc::load(w);
      const HP pt = neg ? pool[j].neg() : pool[j];
      if (i == 0) {
        acc = Acc::start(raw, neg);
      } else {
        if (stress) H::stress(acc);
        acc = Acc::madd(acc, raw, neg, bad);
      }
      H::bounds(acc, bound_fail);
      want = want.add(pt);
    }
    // a random run of distinct points never meets x(acc) = x(point)
    CHECK(bad == 0);
    CHECK(same_point<E>(H::host(acc), want));
    CHECK(!g_overflow);
  }
  CHECK(bound_fail == 0);
  // the flag: P + P and P + (-P), also after other points
  for (int neg = 0; neg < 2; neg++) {
    uint32_t w[AW];
    memcpy(w, &pw[0], sizeof w);
    const typename Acc::Raw raw = Acc::load(w);
    memcpy(w, &pw[AW], sizeof w);
    const typename Acc::Raw other = Acc::load(w);
    uint32_t bad = 0;
    typename Acc::Pt acc = Acc::start(raw, false);
    acc = Acc::madd(acc, raw, neg != 0, bad);
    CHECK(bad == 1);
    bad = 0;
    acc = Acc::start(other, false);
    acc = Acc::madd(acc, raw, false, bad);
    CHECK(bad == 0);
    // acc = other + P; adding -(other + P)'s x-twin: add -other then -P -> cancellation at the end
    acc = Acc::madd(acc, other, true, bad);  // = P
    CHECK(bad == 0 && same_point<E>(H::host(acc), pool[0]));
    acc = Acc::madd(acc, raw, true, bad);  // P - P
    CHECK(bad == 1);
  }
}

int main() {
  // the overflow detector itself: digits of 2^31 - 1 on both sides overflow a column
  {
    Fd big;
    for (int i = 0; i < 9; i++) big.d[i] = 0x7FFFFFFFu;
    (void)mul<FqDig>(big, big);
    if (!g_overflow) {
      std::fprintf(stderr, "the column overflow detector did not fire\n");
      return 1;
    }
    g_overflow = false;
  }
  field_checks();
  runs<Fq>(3000);
  runs<Fq2>(600);
  if (g_overflow) std::fprintf(stderr, "a product column exceeded 2^64\n");
  if (g_fail) {
    std::fprintf(stderr, "%d checks failed\n", g_fail);
    return 1;
  }
  std::printf("digit-domain accumulation checks OK\n");
  return 0;
}
