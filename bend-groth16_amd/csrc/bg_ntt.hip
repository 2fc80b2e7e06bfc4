// Radix-2 NTT over BN254 Fr on gfx950 (C ABI: bg_ntt, bg_qap_h).  See bg_ntt.hpp.
//
// A transform is a short list of passes; each pass runs up to tile (2^9..2^11 elements) radix-2
// stages of one workgroup's elements in LDS (one global read and write per pass).
// Decimation in frequency takes natural order to bit-reversed order and decimation in
// time takes bit-reversed order back to natural order, so qap_h chains inverse (DIF) and
// coset (DIT) transforms with no reordering pass between them; a final gather puts h
// into natural order.  A 2^19 transform is two passes.  Results equal the DFT that
// bend2/gen/ref_qap.py:27-54 computes, with the same omega = root^(2^(28 - log2)).
#include "bg_ntt.hpp"
#include "bg_msm.hpp"
#include "bg_field.hpp"
#include "bg_host_ec.hpp"
#include "bg_tune.hpp"
#include <map>
#include <memory>

namespace bg {

// elements per workgroup: 2^TILE (default 1024, 32 KB of LDS; the tuning table may pick 512
// or 2048 per size), stages two at a time in registers (radix 4) or one at a time
static constexpr int NTT_TILE_DEFAULT = 10;
static constexpr int NTT_THREADS = 256;

__global__ void k_geom(uint32_t* __restrict__ out, uint64_t n, const uint32_t* __restrict__ lo,
                       const uint32_t* __restrict__ hi, int lo_bits, int canon) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fp x = Fr::mul(load_fp(hi + 8 * (i >> lo_bits)), load_fp(lo + 8 * (i & ((1ull << lo_bits) - 1))));
  if (canon) x = Fr::from_mont(x);
  store_fp(out + 8 * i, x);
}

// per-stage twiddles: out[n - n/2^k + j] = tw[j << k] for j < n/2^(k+1), k < log2
__global__ void k_stage_twiddles(const uint32_t* __restrict__ tw, uint32_t* __restrict__ out, int log2) {
  const uint64_t n = 1ull << log2;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  uint64_t j = i;
  int k = 0;
  while (j >= (n >> (k + 1))) {
    j -= n >> (k + 1);
    k++;
  }
  store_fp(out + 8 * i, load_fp(tw + 8 * (j << k)));
}

__device__ __forceinline__ uint32_t bitrev(uint32_t i, int log2) { return log2 ? (__brev(i) >> (32 - log2)) : 0; }

// out[k] = in[rev(k)] * (scale ? scale[k] : c)   (bit-reversed order -> natural order)
// blockIdx.y: transform of a batch of contiguous 2^log2-element arrays
__global__ void k_gather(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int log2,
                         const uint32_t* __restrict__ scale, Fp c, int use_c) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (1u << log2)) return;
  in += 8 * ((uint64_t)blockIdx.y << log2);
  out += 8 * ((uint64_t)blockIdx.y << log2);
  Fp x = load_fp(in + 8ull * bitrev(k, log2));
  if (scale)
    x = Fr::mul(x, load_fp(scale + 8ull * k));
  else if (use_c)
    x = Fr::mul(x, c);
  store_fp(out + 8ull * k, Fr::canon(x));  // transform outputs leave the device path canonical
}

// One pass: radix-2 stages k = k0 .. k0+t-1 (half-size m = 2^(log2-1-k)).  The stages
// only couple elements whose indices differ by multiples of S = 2^(log2-k0-t) inside an
// aligned block of 2^(log2-k0) elements, so a workgroup takes the 2^t elements
// base + j*S + q (j < 2^t) for each of C = 2^logC consecutive offsets q (C consecutive
// elements per row keep the global accesses coalesced).  blockIdx.y selects the transform
// of a batch of contiguous 2^log2-element arrays (the tables are shared).
//   DIF: k ascending,  (u, v) -> (u + v, (u - v) w)
//   DIT: k descending, (u, v) -> (u + v w, u - v w)
// with w = omega^((i mod m) * 2^k), i the upper element's index, read from stage k's part
// of the per-stage table tw.  scale (optional) multiplies each element by scale[index] as
// it is loaded.
// src and dst may alias (in-place pass).
// The stages of one pass over the tile in LDS (cnt elements: rows of C = 2^logC columns),
// each round ending with a barrier.
template <bool DIT, bool R4>
__device__ __forceinline__ void ntt_stages(Fp* sh, int log2, int k0, int t, int logC, int S_log, uint64_t q0,
                                           uint32_t cnt, const uint32_t* __restrict__ tw) {
  const uint32_t C = 1u << logC;
  // twiddle of stage k for local row residue jl: stage k's table holds omega^(j 2^k)
  // contiguously from n - n/2^k, so the lanes of a row (consecutive c) read consecutive
  // twiddles
  auto twid = [&](int k, uint32_t jl, uint32_t c) {
    return load_fp(tw + 8 * (((1ull << log2) - (1ull << (log2 - k))) + (((uint64_t)jl) << S_log) + q0 + c));
  };
  // Stages two at a time (radix 4): a thread takes the four rows j0, j0 + 2^ls, j0 + 2^(ls+1),
  // j0 + 2^(ls+1) + 2^ls of one column, where 2^ls is the smaller of the two stages' local
  // row distances, and applies both stages in registers: half the LDS round trips and
  // barriers of one stage at a time, the same four twiddle products.
  int s = 0;
  for (; R4 && s + 1 < t; s += 2) {
    const int ka = DIT ? k0 + t - 1 - s : k0 + s;  // applied first
    const int kb = DIT ? ka - 1 : ka + 1;          // applied second
    const int ls = log2 - 1 - (DIT ? ka : kb) - S_log;
    const uint32_t ng = cnt >> 2;
    for (uint32_t g = threadIdx.x; g < ng; g += NTT_THREADS) {
      const uint32_t c = g & (C - 1), gj = g >> logC;
      const uint32_t jl = gj & ((1u << ls) - 1);
      const uint32_t j0 = ((gj >> ls) << (ls + 2)) + jl;
      const uint32_t e0 = (j0 << logC) + c, ds = 1u << (ls + logC), db = ds << 1;
      Fp x0 = sh[e0], x1 = sh[e0 + ds], x2 = sh[e0 + db], x3 = sh[e0 + db + ds];
      if (DIT) {
        // stage ka (distance ds): (x0, x1), (x2, x3); stage kb (distance db): (y0, y2), (y1, y3)
        const Fp wa = twid(ka, jl, c);
        Fp u = Fr::mul(x1, wa), v = Fr::mul(x3, wa);
        const Fp y0 = Fr::add(x0, u), y1 = Fr::sub(x0, u), y2 = Fr::add(x2, v), y3 = Fr::sub(x2, v);
        u = Fr::mul(y2, twid(kb, jl, c));
        v = Fr::mul(y3, twid(kb, jl + (1u << ls), c));
        sh[e0] = Fr::add(y0, u);
        sh[e0 + db] = Fr::sub(y0, u);
        sh[e0 + ds] = Fr::add(y1, v);
        sh[e0 + db + ds] = Fr::sub(y1, v);
      } else {
        // stage ka (distance db): (x0, x2), (x1, x3); stage kb (distance ds): (y0, y1), (y2, y3)
        const Fp y0 = Fr::add(x0, x2), y1 = Fr::add(x1, x3);
        const Fp y2 = Fr::mul(Fr::sub(x0, x2), twid(ka, jl, c));
        const Fp y3 = Fr::mul(Fr::sub(x1, x3), twid(ka, jl + (1u << ls), c));
        const Fp wb = twid(kb, jl, c);
        sh[e0] = Fr::add(y0, y1);
        sh[e0 + ds] = Fr::mul(Fr::sub(y0, y1), wb);
        sh[e0 + db] = Fr::add(y2, y3);
        sh[e0 + db + ds] = Fr::mul(Fr::sub(y2, y3), wb);
      }
    }
    __syncthreads();
  }
  for (; s < t; s++) {
    // radix-2 stages: the one an odd stage count leaves, or all of them without radix 4
    const int k = DIT ? k0 + t - 1 - s : k0 + s;
    const int lmr = log2 - 1 - k - S_log;  // local row distance m / S = 2^lmr
    const uint32_t nbf = cnt >> 1;
    for (uint32_t b = threadIdx.x; b < nbf; b += NTT_THREADS) {
      uint32_t c = b & (C - 1), bj = b >> logC;
      uint32_t jl = bj & ((1u << lmr) - 1);
      uint32_t j0 = ((bj >> lmr) << (lmr + 1)) + jl;
      uint32_t e0 = (j0 << logC) + c, e1 = e0 + (1u << (lmr + logC));
      Fp w = twid(k, jl, c);
      Fp u = sh[e0], v = sh[e1];
      if (DIT) {
        v = Fr::mul(v, w);
        sh[e0] = Fr::add(u, v);
        sh[e1] = Fr::sub(u, v);
      } else {
        sh[e0] = Fr::add(u, v);
        sh[e1] = Fr::mul(Fr::sub(u, v), w);
      }
    }
    __syncthreads();
  }
}

template <bool DIT, int TILE, bool R4>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_pass(const uint32_t* src, uint32_t* dst, int log2, int k0, int t,
                                                          int logC, const uint32_t* __restrict__ tw,
                                                          const uint32_t* __restrict__ scale) {
  __shared__ Fp sh[1 << TILE];
  src += 8 * ((uint64_t)blockIdx.y << log2);
  dst += 8 * ((uint64_t)blockIdx.y << log2);
  const int S_log = log2 - k0 - t;
  const uint32_t C = 1u << logC;
  const uint32_t qgroups = 1u << (S_log - logC);
  const uint64_t blk = blockIdx.x >> (S_log - logC);
  const uint64_t q0 = (uint64_t)(blockIdx.x & (qgroups - 1)) << logC;
  const uint64_t base = (blk << (log2 - k0)) + q0;
  const uint32_t cnt = 1u << (t + logC);
  for (uint32_t e = threadIdx.x; e < cnt; e += NTT_THREADS) {
    uint64_t idx = base + ((uint64_t)(e >> logC) << S_log) + (e & (C - 1));
    Fp x = load_fp(src + 8 * idx);
    if (scale) x = Fr::mul(x, load_fp(scale + 8 * idx));
    sh[e] = x;
  }
  __syncthreads();
  ntt_stages<DIT, R4>(sh, log2, k0, t, logC, S_log, q0, cnt, tw);
  for (uint32_t e = threadIdx.x; e < cnt; e += NTT_THREADS) {
    uint64_t idx = base + ((uint64_t)(e >> logC) << S_log) + (e & (C - 1));
    store_fp(dst + 8 * idx, sh[e]);
  }
}

// The last pass of a DIF transform and the first of the DIT transform that follows it
// (the same stages over the same tiles) in one: DIF stages with tw_dif, each element times
// scale[index] (the DIT pass's scale on load), DIT stages with tw_dit; one load and one
// store of the tile instead of two each.
template <int TILE, bool R4>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_mid(const uint32_t* src, uint32_t* dst, int log2, int k0, int t,
                                                         int logC, const uint32_t* __restrict__ tw_dif,
                                                         const uint32_t* __restrict__ scale,
                                                         const uint32_t* __restrict__ tw_dit) {
  __shared__ Fp sh[1 << TILE];
  src += 8 * ((uint64_t)blockIdx.y << log2);
  dst += 8 * ((uint64_t)blockIdx.y << log2);
  const int S_log = log2 - k0 - t;
  const uint32_t C = 1u << logC;
  const uint32_t qgroups = 1u << (S_log - logC);
  const uint64_t blk = blockIdx.x >> (S_log - logC);
  const uint64_t q0 = (uint64_t)(blockIdx.x & (qgroups - 1)) << logC;
  const uint64_t base = (blk << (log2 - k0)) + q0;
  const uint32_t cnt = 1u << (t + logC);
  for (uint32_t e = threadIdx.x; e < cnt; e += NTT_THREADS) {
    uint64_t idx = base + ((uint64_t)(e >> logC) << S_log) + (e & (C - 1));
    sh[e] = load_fp(src + 8 * idx);
  }
  __syncthreads();
  ntt_stages<false, R4>(sh, log2, k0, t, logC, S_log, q0, cnt, tw_dif);
  for (uint32_t e = threadIdx.x; e < cnt; e += NTT_THREADS) {
    uint64_t idx = base + ((uint64_t)(e >> logC) << S_log) + (e & (C - 1));
    sh[e] = Fr::mul(sh[e], load_fp(scale + 8 * idx));
  }
  __syncthreads();
  ntt_stages<true, R4>(sh, log2, k0, t, logC, S_log, q0, cnt, tw_dit);
  for (uint32_t e = threadIdx.x; e < cnt; e += NTT_THREADS) {
    uint64_t idx = base + ((uint64_t)(e >> logC) << S_log) + (e & (C - 1));
    store_fp(dst + 8 * idx, sh[e]);
  }
}

// a[i] = (a[i] * b[i] - c[i]) * zinv  (ntt.bend:325-379, Ntt.point)
__global__ void k_qap_point(uint32_t* __restrict__ a, const uint32_t* __restrict__ b, const uint32_t* __restrict__ c,
                            uint64_t n, Fp zinv) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fp x = Fr::mul(load_fp(a + 8 * i), load_fp(b + 8 * i));
  store_fp(a + 8 * i, Fr::mul(Fr::sub(x, load_fp(c + 8 * i)), zinv));
}

static unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------------------------ domain tables

static bgh::Fr fr_from_u64(uint64_t v) { return bgh::Fr::to_mont(bgh::Fr::raw(v)); }

static bgh::Fr omega_of(int log2, bool inverse) {
  // 2-adic root of unity of order 2^28 (bend2/gen/gen_ntt.py:9)
  bgh::Fr w = bgh::Fr::to_mont(
      bgh::Fr::raw(0x9bd61b6e725b19f0ULL, 0x402d111e41112ed4ULL, 0x00e0a7eb8ef62abcULL, 0x2a3c09f0a58a7e85ULL));
  if (inverse) w = w.inv();
  for (int i = 0; i < 28 - log2; i++) w = w.sqr();
  return w;
}

// Host halves of the geometric table c * x^i: lo[i] = c x^i (i < nlo), hi[j] = x^(2^10 j)
// (j < nhi), as packed words.
static constexpr int GEOM_LO_BITS = 10;
static void geom_host(uint32_t* lo, uint64_t nlo, uint32_t* hi, uint64_t nhi, bgh::Fr x, bgh::Fr c) {
  bgh::Fr v = c;
  for (uint64_t i = 0; i < nlo; i++) {
    v.to_u32(&lo[8 * i]);
    v = v * x;
  }
  bgh::Fr step = bgh::Fr::one();
  for (int i = 0; i < (1 << GEOM_LO_BITS); i++) step = step * x;  // x^1024
  bgh::Fr h = bgh::Fr::one();
  for (uint64_t i = 0; i < nhi; i++) {
    h.to_u32(&hi[8 * i]);
    h = h * step;
  }
}

// out[i] = c * x^i for i < n on the device, from the host halves staged through the
// device buffers dlo / dhi (2^10 and n / 2^10 elements); blocks until done, so the host
// halves may be reused at once.
static void geom_fill(uint32_t* out, uint64_t n, bgh::Fr x, bgh::Fr c, bool canon, uint32_t* hlo, uint32_t* hhi,
                      uint32_t* dlo, uint32_t* dhi, hipStream_t s) {
  const uint64_t nlo = std::min<uint64_t>(n, 1ull << GEOM_LO_BITS);
  const uint64_t nhi = (n + (1ull << GEOM_LO_BITS) - 1) >> GEOM_LO_BITS;
  geom_host(hlo, nlo, hhi, nhi, x, c);
  BG_HIP(hipMemcpyAsync(dlo, hlo, nlo * 32, hipMemcpyHostToDevice, s));
  BG_HIP(hipMemcpyAsync(dhi, hhi, nhi * 32, hipMemcpyHostToDevice, s));
  k_geom<<<nblk(n, 256), 256, 0, s>>>(out, n, dlo, dhi, GEOM_LO_BITS, canon ? 1 : 0);
  BG_HIP(hipGetLastError());
  BG_HIP(hipStreamSynchronize(s));
}

static void build_geom(DevArray& out, uint64_t n, bgh::Fr x, bgh::Fr c, bool canon, hipStream_t s) {
  const uint64_t nlo = std::min<uint64_t>(n, 1ull << GEOM_LO_BITS);
  const uint64_t nhi = (n + (1ull << GEOM_LO_BITS) - 1) >> GEOM_LO_BITS;
  std::vector<uint32_t> lo(8 * nlo), hi(8 * nhi);
  DevArray dlo, dhi;
  dlo.alloc(lo.size() * 4);
  dhi.alloc(hi.size() * 4);
  out.alloc(n * 32);
  geom_fill(out.as<uint32_t>(), n, x, c, canon, lo.data(), hi.data(), dlo.as<uint32_t>(), dhi.as<uint32_t>(), s);
}

NttDomain& ntt_domain(int log2) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<NttDomain>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(log2);
  if (it != cache.end()) return *it->second;
  BG_REQUIRE(log2 >= 1 && log2 <= 28, BG_ERR_ARG, "ntt: log2 %d out of range 1..28", log2);
  auto dom = std::make_unique<NttDomain>();
  dom->log2 = log2;
  dom->n = 1ull << log2;
  hipStream_t s = ctx().stream;
  uint64_t n = dom->n;
  bgh::Fr w = omega_of(log2, false), wi = omega_of(log2, true);
  bgh::Fr g = fr_from_u64(5), gi = g.inv();
  bgh::Fr ninv = fr_from_u64(n).inv();
  for (int dir = 0; dir < 2; dir++) {
    DevArray flat;  // omega^j, j < n/2
    build_geom(flat, n / 2, dir ? wi : w, bgh::Fr::one(), false, s);
    DevArray& st = dir ? dom->tw_inv : dom->tw_fwd;
    st.alloc(n * 32);
    k_stage_twiddles<<<nblk(n, 256), 256, 0, s>>>(flat.as<uint32_t>(), st.as<uint32_t>(), log2);
    BG_HIP(hipGetLastError());
    BG_HIP(hipStreamSynchronize(s));  // flat is freed at the end of this scope
  }
  build_geom(dom->sc_g, n, g, bgh::Fr::one(), false, s);
  build_geom(dom->sc_ninv_g, n, g, ninv, false, s);
  build_geom(dom->sc_ninv_gi, n, gi, ninv, false, s);
  build_geom(dom->sc_ninv_gi_c, n, gi, ninv, true, s);
  // n^-1 g^rev(k): the coset scale for coefficients held in bit-reversed order
  dom->sc_ninv_g_rev.alloc(n * 32);
  k_gather<<<nblk(n, 256), 256, 0, s>>>(dom->sc_ninv_g.as<uint32_t>(), dom->sc_ninv_g_rev.as<uint32_t>(), log2,
                                        nullptr, Fp{}, 0);
  BG_HIP(hipGetLastError());
  BG_HIP(hipStreamSynchronize(s));
  ninv.to_u32(dom->ninv);
  // zinv = (g^n - 1)^-1: the vanishing polynomial x^n - 1 on the coset g H
  uint64_t e4[4] = {n, 0, 0, 0};
  (g.pow(e4) - bgh::Fr::one()).inv().to_u32(dom->zinv);
  NttDomain& ref = *dom;
  cache[log2] = std::move(dom);
  return ref;
}

static Fp to_fp(const uint32_t* w) {
  Fp r;
  for (int i = 0; i < 8; i++) r.v[i] = w[i];
  return r;
}

// Pass list of a 2^log2 transform in DIF order (DIT runs it backwards): strided passes
// of at most tile - 1 stages, then one contiguous pass of tile stages (at most 4 passes up
// to 2^28 for tiles of 2^9 and more).
struct NttPass {
  int k0, t, logC;
};
struct NttPlan {
  NttPass p[4];
  int count = 0;
  int tile = NTT_TILE_DEFAULT;
  bool r4 = true;
};
static NttPlan ntt_plan(int log2) {
  NttPlan out;
  NttTune tt;
  if (tune_ntt(log2, tt)) {
    if (tt.tile_log2) out.tile = tt.tile_log2;
    if (tt.radix) out.r4 = tt.radix == 4;
  }
  const int T = out.tile;
  if (log2 <= T) {
    out.p[out.count++] = {0, log2, 0};
    return out;
  }
  const int rem = log2 - T;
  const int np = (rem + T - 2) / (T - 1);
  int k0 = 0;
  for (int q = 0; q < np; q++) {
    int t = rem / np + (q < rem % np ? 1 : 0);
    int S_log = log2 - k0 - t;
    out.p[out.count++] = {k0, t, std::min(T - t, S_log)};
    k0 += t;
  }
  out.p[out.count++] = {k0, T, 0};
  return out;
}

template <bool DIT, int TILE>
static void launch_pass_t(bool r4, dim3 wgs, const uint32_t* src, uint32_t* dst, int log2, const NttPass& p,
                          const uint32_t* tw, const uint32_t* scale, hipStream_t s) {
  if (r4)
    k_ntt_pass<DIT, TILE, true><<<wgs, NTT_THREADS, 0, s>>>(src, dst, log2, p.k0, p.t, p.logC, tw, scale);
  else
    k_ntt_pass<DIT, TILE, false><<<wgs, NTT_THREADS, 0, s>>>(src, dst, log2, p.k0, p.t, p.logC, tw, scale);
}

// batch: number of contiguous 2^log2-element transforms (grid y)
static void launch_pass(bool dit, const uint32_t* src, uint32_t* dst, int log2, const NttPlan& plan,
                        const NttPass& p, const uint32_t* tw, const uint32_t* scale, hipStream_t s, unsigned batch) {
  const dim3 wgs(1u << (log2 - p.t - p.logC), batch);
  switch (plan.tile * 2 + (dit ? 1 : 0)) {
    case 18: launch_pass_t<false, 9>(plan.r4, wgs, src, dst, log2, p, tw, scale, s); break;
    case 19: launch_pass_t<true, 9>(plan.r4, wgs, src, dst, log2, p, tw, scale, s); break;
    case 22: launch_pass_t<false, 11>(plan.r4, wgs, src, dst, log2, p, tw, scale, s); break;
    case 23: launch_pass_t<true, 11>(plan.r4, wgs, src, dst, log2, p, tw, scale, s); break;
    case 21: launch_pass_t<true, 10>(plan.r4, wgs, src, dst, log2, p, tw, scale, s); break;
    default: launch_pass_t<false, 10>(plan.r4, wgs, src, dst, log2, p, tw, scale, s); break;
  }
  BG_HIP(hipGetLastError());
}

template <int TILE>
static void launch_mid_t(bool r4, dim3 wgs, uint32_t* v, int log2, const NttPass& p, const uint32_t* tw_dif,
                         const uint32_t* scale, const uint32_t* tw_dit, hipStream_t s) {
  if (r4)
    k_ntt_mid<TILE, true><<<wgs, NTT_THREADS, 0, s>>>(v, v, log2, p.k0, p.t, p.logC, tw_dif, scale, tw_dit);
  else
    k_ntt_mid<TILE, false><<<wgs, NTT_THREADS, 0, s>>>(v, v, log2, p.k0, p.t, p.logC, tw_dif, scale, tw_dit);
}

// In place: a DIF transform with tw_dif (natural -> bit-reversed order), each element times
// scale[index], a DIT transform with tw_dit (-> natural order); the DIF's last pass and the
// DIT's first run as one kernel.
static void dif_scale_dit(uint32_t* v, int log2, const uint32_t* tw_dif, const uint32_t* scale,
                          const uint32_t* tw_dit, hipStream_t s, unsigned batch) {
  const NttPlan plan = ntt_plan(log2);
  for (int i = 0; i + 1 < plan.count; i++)
    launch_pass(false, v, v, log2, plan, plan.p[i], tw_dif, nullptr, s, batch);
  const NttPass& m = plan.p[plan.count - 1];
  const dim3 wgs(1u << (log2 - m.t - m.logC), batch);
  switch (plan.tile) {
    case 9: launch_mid_t<9>(plan.r4, wgs, v, log2, m, tw_dif, scale, tw_dit, s); break;
    case 11: launch_mid_t<11>(plan.r4, wgs, v, log2, m, tw_dif, scale, tw_dit, s); break;
    default: launch_mid_t<10>(plan.r4, wgs, v, log2, m, tw_dif, scale, tw_dit, s); break;
  }
  BG_HIP(hipGetLastError());
  for (int i = plan.count - 2; i >= 0; i--)
    launch_pass(true, v, v, log2, plan, plan.p[i], tw_dit, nullptr, s, batch);
}

// natural order in -> bit-reversed order out; scale multiplies the (natural) input
static void dif(const uint32_t* src, uint32_t* dst, int log2, const uint32_t* tw, const uint32_t* scale,
                hipStream_t s, unsigned batch = 1) {
  const NttPlan plan = ntt_plan(log2);
  for (int i = 0; i < plan.count; i++)
    launch_pass(false, i == 0 ? src : dst, dst, log2, plan, plan.p[i], tw, i == 0 ? scale : nullptr, s, batch);
}

// bit-reversed order in -> natural order out; scale multiplies the (bit-reversed) input
static void dit(const uint32_t* src, uint32_t* dst, int log2, const uint32_t* tw, const uint32_t* scale,
                hipStream_t s, unsigned batch = 1) {
  const NttPlan plan = ntt_plan(log2);
  for (int i = plan.count; i-- > 0;)
    launch_pass(true, i + 1 == plan.count ? src : dst, dst, log2, plan, plan.p[i], tw,
                i + 1 == plan.count ? scale : nullptr, s, batch);
}

static void gather(const uint32_t* in, uint32_t* out, int log2, const uint32_t* scale, const uint32_t* c,
                   hipStream_t s, unsigned batch = 1) {
  Fp cf{};
  if (c) cf = to_fp(c);
  k_gather<<<dim3(nblk(1ull << log2, 256), batch), 256, 0, s>>>(in, out, log2, scale, cf, c ? 1 : 0);
  BG_HIP(hipGetLastError());
}

void ntt_run(uint32_t* d, uint32_t* tmp, int log2, int direction, hipStream_t s) {
  NttDomain& D = ntt_domain(log2);
  bool inv = direction == BG_NTT_IFFT || direction == BG_NTT_COSET_IFFT;
  const uint32_t* pre = direction == BG_NTT_COSET_FFT ? D.sc_g.as<uint32_t>() : nullptr;
  dif(d, tmp, log2, inv ? D.tw_inv.as<uint32_t>() : D.tw_fwd.as<uint32_t>(), pre, s);
  if (direction == BG_NTT_IFFT)
    gather(tmp, d, log2, nullptr, D.ninv, s);
  else if (direction == BG_NTT_COSET_IFFT)
    gather(tmp, d, log2, D.sc_ninv_gi.as<uint32_t>(), nullptr, s);
  else
    gather(tmp, d, log2, nullptr, nullptr, s);
}

// The tables qap_h reads, wherever they live.
struct QapTables {
  const uint32_t* tw_fwd;
  const uint32_t* tw_inv;
  const uint32_t* ninv_g_rev;  // n^-1 g^rev(j)
  const uint32_t* ninv_gi;     // n^-1 g^-j (Montgomery out) or its canonical-out form
  uint32_t zinv[8];
};

static void qap_h_core(uint32_t* a, uint32_t* b, uint32_t* c, uint32_t* h, int log2, const QapTables& T,
                       hipStream_t s, unsigned batch) {
  const uint64_t n = 1ull << log2;
  uint32_t* v3[3] = {a, b, c};
  // evaluations -> coefficients (bit-reversed, times n) -> n^-1 g^k -> coset evaluations.
  // Domains up to 2^17 fuse the middle passes (K = 10 .. 16 proofs 1-2 % faster); the 2^19
  // domain keeps them apart: fused, the 2^18 proof's kernels measured 1 % less serialised
  // but the proof 0.5 % slower overlapped (its longer blocks wait for SIMDs beside the MSMs)
  for (uint32_t* v : v3) {
    if (log2 <= 17) {
      dif_scale_dit(v, log2, T.tw_inv, T.ninv_g_rev, T.tw_fwd, s, batch);
    } else {
      dif(v, v, log2, T.tw_inv, nullptr, s, batch);
      dit(v, v, log2, T.tw_fwd, T.ninv_g_rev, s, batch);
    }
  }
  k_qap_point<<<nblk(n * batch, 256), 256, 0, s>>>(a, b, c, n * batch, to_fp(T.zinv));
  BG_HIP(hipGetLastError());
  dif(a, a, log2, T.tw_inv, nullptr, s, batch);
  gather(a, h, log2, T.ninv_gi, nullptr, s, batch);
}

void qap_h_run(uint32_t* a, uint32_t* b, uint32_t* c, uint32_t* h, uint32_t* tmp, int log2, bool canonical_out,
               hipStream_t s, unsigned batch) {
  (void)tmp;
  NttDomain& D = ntt_domain(log2);
  QapTables T{D.tw_fwd.as<uint32_t>(), D.tw_inv.as<uint32_t>(), D.sc_ninv_g_rev.as<uint32_t>(),
              canonical_out ? D.sc_ninv_gi_c.as<uint32_t>() : D.sc_ninv_gi.as<uint32_t>(), {}};
  memcpy(T.zinv, D.zinv, 32);
  qap_h_core(a, b, c, h, log2, T, s, batch);
}

// ------------------------------------------------------------------ allocation-free qap_h
// bg_qap_h_noalloc keeps everything beyond its caller's buffers in one file-scope device
// array: copies of a, b, c (transformed in place), the domain's tables and the staging
// of the tables' host halves.  Sized for domains up to 2^NOALLOC_QAP_MAX_LOG2; the tables
// of the last domain stay in the array and are rebuilt when a call asks for another size.
static constexpr int NOALLOC_QAP_MAX_LOG2 = 17;
static constexpr uint64_t NA_N = 1ull << NOALLOC_QAP_MAX_LOG2;
static constexpr uint64_t NA_LO = 8ull << GEOM_LO_BITS;            // words of the low half
static constexpr uint64_t NA_HI = 8 * (NA_N >> GEOM_LO_BITS);      // words of the high half
// a, b, c, tw_fwd, tw_inv, n^-1 g^rev(j), n^-1 g^-j (8 words each per element), the
// n / 2 flat powers of omega, the staged halves
static constexpr uint64_t NA_WORDS = 7 * 8 * NA_N + 4 * NA_N + NA_LO + NA_HI;
static __device__ __align__(256) uint32_t scratch[NA_WORDS];

static constexpr int NA_MAX_DEVICES = 64;
static uint32_t* g_na_ptr[NA_MAX_DEVICES];   // scratch's device address, per device
static int g_na_log2[NA_MAX_DEVICES];        // domain whose tables scratch holds (0: none)
static uint32_t g_na_hlo[NA_LO], g_na_hhi[NA_HI];  // host halves (reused per table)

}  // namespace bg

using namespace bg;

extern "C" int bg_qap_h_noalloc(const uint32_t* a, const uint32_t* b, const uint32_t* cc, int log2n, uint32_t* h) {
  return guard([&] {
    BG_REQUIRE(a && b && cc && h, BG_ERR_ARG, "bg_qap_h_noalloc: null argument");
    BG_REQUIRE(log2n >= 1 && log2n <= NOALLOC_QAP_MAX_LOG2, BG_ERR_ARG,
               "bg_qap_h_noalloc: log2n %d out of range 1..%d (the static scratch holds 2^%d-element domains)",
               log2n, NOALLOC_QAP_MAX_LOG2, NOALLOC_QAP_MAX_LOG2);
    Ctx& c = ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    BG_REQUIRE(c.device >= 0 && c.device < NA_MAX_DEVICES, BG_ERR_ARG, "bg_qap_h_noalloc: device %d", c.device);
    uint32_t*& base = g_na_ptr[c.device];
    if (!base) {
      // the array's address on this device (each device holds its own copy)
      void* p = nullptr;
      BG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      base = static_cast<uint32_t*>(p);
    }
    const uint64_t n = 1ull << log2n;
    uint32_t* da = base;
    uint32_t* db = da + 8 * NA_N;
    uint32_t* dc = db + 8 * NA_N;
    uint32_t* tw_fwd = dc + 8 * NA_N;
    uint32_t* tw_inv = tw_fwd + 8 * NA_N;
    uint32_t* g_rev = tw_inv + 8 * NA_N;
    uint32_t* ninv_gi = g_rev + 8 * NA_N;
    uint32_t* flat = ninv_gi + 8 * NA_N;  // 4 NA_N words
    uint32_t* dlo = flat + 4 * NA_N;
    uint32_t* dhi = dlo + NA_LO;
    hipStream_t s = c.stream;
    const bgh::Fr g = bgh::Fr::to_mont(bgh::Fr::raw(5)), gi = g.inv();
    const bgh::Fr ninv = bgh::Fr::to_mont(bgh::Fr::raw(n)).inv();
    if (g_na_log2[c.device] != log2n) {
      g_na_log2[c.device] = 0;  // tables incomplete until the end of this block
      for (int dir = 0; dir < 2; dir++) {
        geom_fill(flat, n / 2, omega_of(log2n, dir != 0), bgh::Fr::one(), false, g_na_hlo, g_na_hhi, dlo, dhi, s);
        k_stage_twiddles<<<nblk(n, 256), 256, 0, s>>>(flat, dir ? tw_inv : tw_fwd, log2n);
        BG_HIP(hipGetLastError());
      }
      // n^-1 g^j staged in the a copy, then put in bit-reversed order
      geom_fill(da, n, g, ninv, false, g_na_hlo, g_na_hhi, dlo, dhi, s);
      k_gather<<<nblk(n, 256), 256, 0, s>>>(da, g_rev, log2n, nullptr, Fp{}, 0);
      BG_HIP(hipGetLastError());
      geom_fill(ninv_gi, n, gi, ninv, false, g_na_hlo, g_na_hhi, dlo, dhi, s);
      g_na_log2[c.device] = log2n;
    }
    QapTables T{tw_fwd, tw_inv, g_rev, ninv_gi, {}};
    uint64_t e4[4] = {n, 0, 0, 0};
    (g.pow(e4) - bgh::Fr::one()).inv().to_u32(T.zinv);
    BG_HIP(hipMemcpyAsync(da, a, 32 * n, hipMemcpyDeviceToDevice, s));
    BG_HIP(hipMemcpyAsync(db, b, 32 * n, hipMemcpyDeviceToDevice, s));
    BG_HIP(hipMemcpyAsync(dc, cc, 32 * n, hipMemcpyDeviceToDevice, s));
    qap_h_core(da, db, dc, h, log2n, T, s, 1);
    BG_HIP(hipStreamSynchronize(s));
  });
}

extern "C" int bg_ntt(uint32_t* data, int log2n, int direction, int flags) {
  return guard([&] {
    BG_REQUIRE(data && log2n >= 1 && log2n <= 28 && direction >= 0 && direction <= 3, BG_ERR_ARG,
               "bg_ntt: bad argument");
    bool words16 = (flags & 1) == BG_FMT_WORDS16;
    bool dev = (flags & BG_MEM_DEVICE) != 0;
    BG_REQUIRE(!(dev && words16), BG_ERR_ARG, "bg_ntt: device buffers must be packed");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    uint64_t n = 1ull << log2n;
    uint32_t* tmp = c.scratch[3].as<uint32_t>(8 * n);
    uint32_t* d = data;
    if (!dev) {
      d = c.scratch[4].as<uint32_t>(8 * n);
      if (words16) {
        uint32_t* raw = c.scratch[5].as<uint32_t>(16 * n);
        BG_HIP(hipMemcpyAsync(raw, data, 64 * n, hipMemcpyHostToDevice, c.stream));
        launch_pack16(raw, d, 8 * n, c.stream);
      } else {
        BG_HIP(hipMemcpyAsync(d, data, 32 * n, hipMemcpyHostToDevice, c.stream));
      }
    }
    {
      ScopedKTime kt("ntt", c.stream);
      ntt_run(d, tmp, log2n, direction, c.stream);
    }
    if (!dev) {
      if (words16) {
        uint32_t* raw = c.scratch[5].as<uint32_t>(16 * n);
        launch_unpack16(d, raw, 8 * n, c.stream);
        BG_HIP(hipMemcpyAsync(data, raw, 64 * n, hipMemcpyDeviceToHost, c.stream));
      } else {
        BG_HIP(hipMemcpyAsync(data, d, 32 * n, hipMemcpyDeviceToHost, c.stream));
      }
    }
    BG_HIP(hipStreamSynchronize(c.stream));
  });
}

extern "C" int bg_qap_h(const uint32_t* a, const uint32_t* b, const uint32_t* cc, int log2n, uint32_t* h,
                        int flags) {
  return guard([&] {
    BG_REQUIRE(a && b && cc && h && log2n >= 1 && log2n <= 28, BG_ERR_ARG, "bg_qap_h: bad argument");
    bool words16 = (flags & 1) == BG_FMT_WORDS16;
    bool dev = (flags & BG_MEM_DEVICE) != 0;
    BG_REQUIRE(!(dev && words16), BG_ERR_ARG, "bg_qap_h: device buffers must be packed");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    uint64_t n = 1ull << log2n;
    uint32_t* tmp = c.scratch[3].as<uint32_t>(8 * n);
    uint32_t* buf = c.scratch[4].as<uint32_t>(4 * 8 * n);  // a | b | c | h (device copies)
    uint32_t* da = buf;
    uint32_t* db = buf + 8 * n;
    uint32_t* dc = buf + 16 * n;
    uint32_t* dh = dev ? h : buf + 24 * n;
    const uint32_t* src[3] = {a, b, cc};
    uint32_t* dst[3] = {da, db, dc};
    for (int k = 0; k < 3; k++) {
      if (dev) {
        BG_HIP(hipMemcpyAsync(dst[k], src[k], 32 * n, hipMemcpyDeviceToDevice, c.stream));
      } else if (words16) {
        uint32_t* raw = c.scratch[5].as<uint32_t>(16 * n);
        BG_HIP(hipMemcpyAsync(raw, src[k], 64 * n, hipMemcpyHostToDevice, c.stream));
        launch_pack16(raw, dst[k], 8 * n, c.stream);
      } else {
        BG_HIP(hipMemcpyAsync(dst[k], src[k], 32 * n, hipMemcpyHostToDevice, c.stream));
      }
    }
    {
      ScopedKTime kt("qap_h", c.stream);
      qap_h_run(da, db, dc, dh, tmp, log2n, false, c.stream);
    }
    if (!dev) {
      if (words16) {
        uint32_t* raw = c.scratch[5].as<uint32_t>(16 * n);
        launch_unpack16(dh, raw, 8 * n, c.stream);
        BG_HIP(hipMemcpyAsync(h, raw, 64 * n, hipMemcpyDeviceToHost, c.stream));
      } else {
        BG_HIP(hipMemcpyAsync(h, dh, 32 * n, hipMemcpyDeviceToHost, c.stream));
      }
    }
    BG_HIP(hipStreamSynchronize(c.stream));
  });
}
