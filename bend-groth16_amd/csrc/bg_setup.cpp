// Trapdoor Groth16 setup for SquareChain(2^log2) writing the FORMAT.md key directory
// (C ABI: bg_setup_square_chain).  Host-only; no device is touched.
//
// Restates what ref/src/gen.rs:95-164 produces with arkworks (the reference's own key
// generator), with this library's own seeded generator in place of arkworks' StdRng:
//   circuit   ref/src/circuit.rs:37-73: x_{i+1} = x_i^2, z = [1, x_N, x_0 .. x_{N-1}],
//             row i: A = B = [(2+i, 1)], C = [(3+i, 1)] or [(1, 1)] for the last row
//   domain    d = smallest power of two >= N + 2 (ref/src/gen.rs:91-93)
//   QAP       libsnark reduction: u_i(tau) = sum_rows A[k][i] L_k(tau) plus L_{N+i}(tau) for
//             the instance variables i < 2 (input-consistency rows), v, w from B, C
//   key       a_query = u_i G1, b_g1_query = v_i G1, b_g2_query = v_i G2,
//             h_query[k] = tau^k (tau^d - 1) / delta G1 (k < d - 1),
//             l_query[j] = (beta u + alpha v + w)_{2+j} / delta G1,
//             alpha G1, beta G1, beta G2, delta G1, delta G2
// plus trapdoor.bin (tau, alpha, beta, gamma, delta as Montgomery Fr, FORMAT.md words)
// so tests can recompute the proof in closed form.
#include "bg_host_ec.hpp"
#include "../../include/bg16.h"
#include <stdarg.h>
#include <stdio.h>
#include <algorithm>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

namespace bg {
void set_error(const char* fmt, ...);
}

namespace {

using bgh::Fq;
using bgh::Fq2;
using bgh::Fr;
using P1 = bgh::P1;
using P2 = bgh::P2;

struct SetupError {
  int code;
};

// xoshiro256** seeded by splitmix64
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (auto& x : s) {
      seed += 0x9e3779b97f4a7c15ULL;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      x = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0, r) by rejection on 254 bits; returned in Montgomery form
  Fr fr() {
    for (;;) {
      Fr x = Fr::raw(next(), next(), next(), next() & ((1ULL << 62) - 1));
      if (!Fr::geq_p(x.v)) return Fr::to_mont(x);
    }
  }
};

template <class Fn>
void parallel_for(size_t n, int threads, Fn&& fn) {
  if (threads <= 1 || n < 2) {
    fn(0, n);
    return;
  }
  size_t chunk = (n + threads - 1) / threads;
  std::vector<std::thread> th;
  for (int t = 0; t < threads; t++) {
    size_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
  }
  for (auto& x : th) x.join();
}

// canonical 4-word scalar of a Montgomery Fr
void canon(const Fr& x, uint64_t* k) {
  Fr c = x.from_mont();
  memcpy(k, c.v, 32);
}

// Fixed-base batch: out affine k_i * base for Montgomery scalars, 8-bit windows over
// tables T[w][d] = d 2^(8w) base (normalised once), XYZZ accumulation, batched affine.
template <class E>
void fixed_base(const bgh::Pt<E>& base, const std::vector<Fr>& ks, std::vector<E>& xs, std::vector<E>& ys,
                int threads) {
  const int NW = 32;
  std::vector<bgh::Pt<E>> tab(NW * 256);
  bgh::Pt<E> b = base;
  for (int w = 0; w < NW; w++) {
    tab[w * 256] = bgh::Pt<E>::inf();
    for (int d = 1; d < 256; d++) tab[w * 256 + d] = tab[w * 256 + d - 1].add(b);
    b = b.dbl_n(8);
  }
  std::vector<E> tx(tab.size()), ty(tab.size());
  bgh::batch_to_affine(tab.data(), tab.size(), tx.data(), ty.data());
  size_t n = ks.size();
  xs.resize(n);
  ys.resize(n);
  parallel_for(n, threads, [&](size_t lo, size_t hi) {
    const size_t CH = 1024;
    std::vector<bgh::Pt<E>> acc(CH);
    for (size_t c0 = lo; c0 < hi; c0 += CH) {
      size_t m = std::min(CH, hi - c0);
      for (size_t i = 0; i < m; i++) {
        uint64_t k[4];
        canon(ks[c0 + i], k);
        bgh::Pt<E> a = bgh::Pt<E>::inf();
        for (int w = 0; w < NW; w++) {
          unsigned dg = (unsigned)((k[w / 8] >> (8 * (w % 8))) & 0xFF);
          if (dg) a = a.madd(tx[w * 256 + dg], ty[w * 256 + dg]);
        }
        acc[i] = a;
      }
      bgh::batch_to_affine(acc.data(), m, &xs[c0], &ys[c0]);
    }
  });
}

// ---------------------------------------------------------------- FORMAT.md writer
struct Writer {
  std::string dir;
  void file(const std::string& name, const std::vector<uint32_t>& w) {
    std::string p = dir + "/" + name;
    std::ofstream f(p, std::ios::binary | std::ios::trunc);
    if (!f.good()) {
      bg::set_error("cannot write %s", p.c_str());
      throw SetupError{BG_ERR_IO};
    }
    f.write(reinterpret_cast<const char*>(w.data()), (std::streamsize)(w.size() * 4));
    if (!f.good()) {
      bg::set_error("write failed: %s", p.c_str());
      throw SetupError{BG_ERR_IO};
    }
  }
};

// one field element (Montgomery, 4 x u64) -> 16 words of 16-bit limbs
void put16(std::vector<uint32_t>& out, const uint64_t* v) {
  for (int j = 0; j < 16; j++) out.push_back((uint32_t)((v[j / 4] >> (16 * (j % 4))) & 0xFFFF));
}
void put_fq(std::vector<uint32_t>& o, const Fq& x) { put16(o, x.v); }
void put_fr(std::vector<uint32_t>& o, const Fr& x) { put16(o, x.v); }
void put_fq2(std::vector<uint32_t>& o, const Fq2& x) {
  put16(o, x.c0.v);
  put16(o, x.c1.v);
}

}  // namespace

extern "C" int bg_setup_square_chain(const char* dir, int log2, uint64_t seed, int threads) {
  try {
    if (!dir || log2 < 1 || log2 > 24) {
      bg::set_error("bg_setup_square_chain: bad argument (log2 in 1..24)");
      return BG_ERR_ARG;
    }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    const uint64_t N = 1ull << log2;
    uint64_t d = 1;
    int lg = 0;
    while (d < N + 2) {
      d <<= 1;
      lg++;
    }
    const uint64_t ni = 2, nw = N, nz = ni + nw;
    Writer W{dir};

    // ---- witness (ref/src/circuit.rs:37-41: its own seed, independent of the setup)
    Rng wr(seed ^ 0x51574e455353ULL);
    std::vector<Fr> xs(N + 1);
    xs[0] = wr.fr();
    for (uint64_t i = 0; i < N; i++) xs[i + 1] = xs[i].sqr();
    std::vector<Fr> z(nz);
    z[0] = Fr::one();
    z[1] = xs[N];
    for (uint64_t i = 0; i < N; i++) z[2 + i] = xs[i];

    // ---- R1CS as CSR, one unit coefficient per row
    auto csr = [&](auto col_of) {
      std::vector<uint32_t> w;
      w.reserve(2 + N + 17 * N);
      w.push_back((uint32_t)N);
      for (uint64_t i = 0; i <= N; i++) w.push_back((uint32_t)i);
      for (uint64_t i = 0; i < N; i++) {
        w.push_back((uint32_t)col_of(i));
        put_fr(w, Fr::one());
      }
      return w;
    };
    W.file("r1cs_a.bin", csr([](uint64_t i) { return 2 + i; }));
    W.file("r1cs_b.bin", csr([](uint64_t i) { return 2 + i; }));
    W.file("r1cs_c.bin", csr([N](uint64_t i) { return i + 1 == N ? (uint64_t)1 : 3 + i; }));
    {
      std::vector<uint32_t> w;
      for (auto& x : z) put_fr(w, x);
      W.file("witness.bin", w);
    }

    // ---- trapdoor and r, s
    Rng sr(seed);
    Fr tau = sr.fr(), alpha = sr.fr(), beta = sr.fr(), gamma = sr.fr(), delta = sr.fr();
    Fr r = sr.fr(), s = sr.fr();
    {
      std::vector<uint32_t> w;
      put_fr(w, r);
      put_fr(w, s);
      W.file("rs.bin", w);
      std::vector<uint32_t> t;
      for (const Fr* x : {&tau, &alpha, &beta, &gamma, &delta}) put_fr(t, *x);
      W.file("trapdoor.bin", t);
    }

    // ---- Lagrange basis at tau: L_k = (tau^d - 1)/d * w^k / (tau - w^k)
    Fr om = Fr::to_mont(Fr::raw(0x9bd61b6e725b19f0ULL, 0x402d111e41112ed4ULL, 0x00e0a7eb8ef62abcULL,
                                0x2a3c09f0a58a7e85ULL));
    for (int i = 0; i < 28 - lg; i++) om = om.sqr();
    uint64_t de[4] = {d, 0, 0, 0};
    Fr zt = tau.pow(de) - Fr::one();
    Fr dinv = Fr::to_mont(Fr::raw(d)).inv();
    std::vector<Fr> wk(d), den(d), L(d);
    Fr wpow = Fr::one();
    for (uint64_t k = 0; k < d; k++) {
      wk[k] = wpow;
      den[k] = tau - wpow;
      wpow = wpow * om;
    }
    {
      std::vector<Fr> pre(d);
      Fr run = Fr::one();
      for (uint64_t k = 0; k < d; k++) {
        pre[k] = run;
        run = run * den[k];
      }
      Fr inv = run.inv();
      Fr c = zt * dinv;
      for (uint64_t k = d; k-- > 0;) {
        Fr di = inv * pre[k];
        inv = inv * den[k];
        L[k] = c * wk[k] * di;
      }
    }
    // ---- QAP polynomials at tau
    std::vector<Fr> u(nz, Fr::zero()), v(nz, Fr::zero()), w(nz, Fr::zero());
    for (uint64_t i = 0; i < N; i++) {
      u[2 + i] = u[2 + i] + L[i];
      v[2 + i] = v[2 + i] + L[i];
      uint64_t cc = i + 1 == N ? 1 : 3 + i;
      w[cc] = w[cc] + L[i];
    }
    for (uint64_t i = 0; i < ni; i++) u[i] = u[i] + L[N + i];

    Fr dlt_inv = delta.inv();
    std::vector<Fr> hs(d - 1), ls(nw);
    {
      Fr t = zt * dlt_inv;
      for (uint64_t k = 0; k + 1 < d; k++) {
        hs[k] = t;
        t = t * tau;
      }
      for (uint64_t j = 0; j < nw; j++) ls[j] = (beta * u[ni + j] + alpha * v[ni + j] + w[ni + j]) * dlt_inv;
    }

    // ---- group elements
    P1 g1 = P1::from_affine(Fq::one(), Fq::one() + Fq::one());
    Fq2 g2x{Fq::to_mont(Fq::raw(0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL, 0x1800deef121f1e76ULL)),
            Fq::to_mont(Fq::raw(0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL, 0x198e9393920d483aULL))};
    Fq2 g2y{Fq::to_mont(Fq::raw(0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL, 0x12c85ea5db8c6debULL)),
            Fq::to_mont(Fq::raw(0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL, 0x090689d0585ff075ULL))};
    P2 g2 = P2::from_affine(g2x, g2y);

    auto g1_file = [&](const char* name, const std::vector<Fr>& ks) {
      std::vector<Fq> px, py;
      fixed_base(g1, ks, px, py, threads);
      std::vector<uint32_t> out;
      out.reserve(ks.size() * 32);
      for (size_t i = 0; i < ks.size(); i++) {
        put_fq(out, px[i]);
        put_fq(out, py[i]);
      }
      W.file(name, out);
    };
    auto g2_file = [&](const char* name, const std::vector<Fr>& ks) {
      std::vector<Fq2> px, py;
      fixed_base(g2, ks, px, py, threads);
      std::vector<uint32_t> out;
      out.reserve(ks.size() * 64);
      for (size_t i = 0; i < ks.size(); i++) {
        put_fq2(out, px[i]);
        put_fq2(out, py[i]);
      }
      W.file(name, out);
    };
    g1_file("pk_g1.bin", {alpha, beta, delta});
    g2_file("pk_g2.bin", {beta, delta});
    g1_file("pk_a.bin", u);
    g1_file("pk_b1.bin", v);
    g2_file("pk_b2.bin", v);
    g1_file("pk_h.bin", hs);
    g1_file("pk_l.bin", ls);
    {
      // verifying key (ref/src/gen.rs:159-164 keeps arkworks' vk; this library's verifier reads
      // vk.bin): alpha G1, beta G2, gamma G2, delta G2, count, gamma_abc[count] with
      // gamma_abc_i = (beta u_i + alpha v_i + w_i) / gamma G1 for the instance variables
      Fr gam_inv = gamma.inv();
      std::vector<Fr> k1 = {alpha};
      for (uint64_t i = 0; i < ni; i++) k1.push_back((beta * u[i] + alpha * v[i] + w[i]) * gam_inv);
      std::vector<Fq> ax, ay;
      std::vector<Fq2> bx, by;
      fixed_base(g1, k1, ax, ay, 1);
      fixed_base(g2, std::vector<Fr>{beta, gamma, delta}, bx, by, 1);
      std::vector<uint32_t> out;
      put_fq(out, ax[0]);
      put_fq(out, ay[0]);
      for (int i = 0; i < 3; i++) {
        put_fq2(out, bx[i]);
        put_fq2(out, by[i]);
      }
      out.push_back((uint32_t)ni);
      for (uint64_t i = 0; i < ni; i++) {
        put_fq(out, ax[1 + i]);
        put_fq(out, ay[1 + i]);
      }
      W.file("vk.bin", out);
    }
    std::vector<uint32_t> hdr = {(uint32_t)N,      (uint32_t)ni,     (uint32_t)nw,      (uint32_t)lg,
                                 (uint32_t)nz,     (uint32_t)nz,     (uint32_t)(d - 1), (uint32_t)nw};
    W.file("header.bin", hdr);
    return BG_OK;
  } catch (const SetupError& e) {
    return e.code;
  } catch (const std::bad_alloc&) {
    bg::set_error("bg_setup_square_chain: host allocation failed");
    return BG_ERR_NOMEM;
  }
}

// ---------------------------------------------------------------- synthetic bases
namespace {

P1 g1_generator() { return P1::from_affine(Fq::one(), Fq::one() + Fq::one()); }

P2 g2_generator() {
  Fq2 x{Fq::to_mont(Fq::raw(0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL, 0x1800deef121f1e76ULL)),
        Fq::to_mont(Fq::raw(0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL, 0x198e9393920d483aULL))};
  Fq2 y{Fq::to_mont(Fq::raw(0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL, 0x12c85ea5db8c6debULL)),
        Fq::to_mont(Fq::raw(0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL, 0x090689d0585ff075ULL))};
  return P2::from_affine(x, y);
}

// affine Montgomery coordinates -> packed words (8 x u32 per Fq)
void pack_fq(const Fq& x, uint32_t* o) { x.to_u32(o); }
void pack_fq2(const Fq2& x, uint32_t* o) {
  x.c0.to_u32(o);
  x.c1.to_u32(o + 8);
}

int fixed_base_packed(int group, const std::vector<Fr>& ks, uint32_t* out, int threads) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (group == 1) {
    std::vector<Fq> px, py;
    fixed_base(g1_generator(), ks, px, py, threads);
    for (size_t i = 0; i < ks.size(); i++) {
      pack_fq(px[i], out + 16 * i);
      pack_fq(py[i], out + 16 * i + 8);
    }
  } else {
    std::vector<Fq2> px, py;
    fixed_base(g2_generator(), ks, px, py, threads);
    for (size_t i = 0; i < ks.size(); i++) {
      pack_fq2(px[i], out + 32 * i);
      pack_fq2(py[i], out + 32 * i + 16);
    }
  }
  return BG_OK;
}

}  // namespace

// k_i * G (G1 or G2 generator) for canonical packed scalars: the closed form of an MSM
// over bases k_i G, and the generator of synthetic bases.
extern "C" int bg_fixed_base_mul(int group, const uint32_t* scalars, uint64_t n, uint32_t* out_affine, int threads) {
  try {
    if ((group != 1 && group != 2) || (n && (!scalars || !out_affine))) {
      bg::set_error("bg_fixed_base_mul: bad argument");
      return BG_ERR_ARG;
    }
    std::vector<Fr> ks(n);
    for (uint64_t i = 0; i < n; i++) {
      const uint32_t* s = scalars + 8 * i;
      ks[i] = Fr::to_mont(Fr::raw((uint64_t)s[0] | (uint64_t)s[1] << 32, (uint64_t)s[2] | (uint64_t)s[3] << 32,
                                  (uint64_t)s[4] | (uint64_t)s[5] << 32, (uint64_t)s[6] | (uint64_t)s[7] << 32));
    }
    return fixed_base_packed(group, ks, out_affine, threads);
  } catch (const std::bad_alloc&) {
    bg::set_error("bg_fixed_base_mul: host allocation failed");
    return BG_ERR_NOMEM;
  }
}

// n bases k_i G with k_i uniform in [0, r) from the seeded generator; k_out (optional)
// receives the k_i canonical packed.
extern "C" int bg_gen_bases(int group, uint64_t seed, uint64_t n, uint32_t* out_affine, uint32_t* k_out, int threads) {
  try {
    if ((group != 1 && group != 2) || (n && !out_affine)) {
      bg::set_error("bg_gen_bases: bad argument");
      return BG_ERR_ARG;
    }
    Rng rng(seed);
    std::vector<Fr> ks(n);
    for (auto& k : ks) k = rng.fr();
    if (k_out)
      for (uint64_t i = 0; i < n; i++) {
        Fr c = ks[i].from_mont();
        for (int j = 0; j < 4; j++) {
          k_out[8 * i + 2 * j] = (uint32_t)c.v[j];
          k_out[8 * i + 2 * j + 1] = (uint32_t)(c.v[j] >> 32);
        }
      }
    return fixed_base_packed(group, ks, out_affine, threads);
  } catch (const std::bad_alloc&) {
    bg::set_error("bg_gen_bases: host allocation failed");
    return BG_ERR_NOMEM;
  }
}
