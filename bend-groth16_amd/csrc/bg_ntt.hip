// Radix-2 NTT over BN254 Fr on gfx950 (C ABI: bg_ntt, bg_qap_h).  See bg_ntt.hpp.
//
// A transform is a bit-reversal gather (optionally fused with a coset / n^-1 scale)
// followed by log2 n radix-2 DIT stages: the first TILE_LOG2 stages of every 2^TILE_LOG2
// contiguous block run in LDS (one kernel), the remaining stages as global passes.
// Output is the DFT in natural order (bend2/gen/ref_qap.py:27-54 computes the same
// transform with the same omega = root^(2^(28 - log2))).
#include "bg_ntt.hpp"
#include "bg_msm.hpp"
#include "bg_field.hpp"
#include "bg_host_ec.hpp"
#include <map>
#include <memory>

namespace bg {

static constexpr int TILE_LOG2 = 10;
static constexpr int NTT_THREADS = 256;

__global__ void k_geom(uint32_t* __restrict__ out, uint64_t n, const uint32_t* __restrict__ lo,
                       const uint32_t* __restrict__ hi, int lo_bits, int canon) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fp x = Fr::mul(load_fp(hi + 8 * (i >> lo_bits)), load_fp(lo + 8 * (i & ((1ull << lo_bits) - 1))));
  if (canon) x = Fr::from_mont(x);
  store_fp(out + 8 * i, x);
}

__device__ __forceinline__ uint32_t bitrev(uint32_t i, int log2) { return log2 ? (__brev(i) >> (32 - log2)) : 0; }

// out[rev(i)] = in[i] * scale[i] (scale optional)
__global__ void k_bitrev(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int log2,
                         const uint32_t* __restrict__ scale) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << log2)) return;
  Fp x = load_fp(in + 8ull * i);
  if (scale) x = Fr::mul(x, load_fp(scale + 8ull * i));
  store_fp(out + 8ull * bitrev(i, log2), x);
}

// the first min(log2, TILE_LOG2) DIT stages of each contiguous tile, in LDS
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_tile(uint32_t* __restrict__ d, int log2,
                                                          const uint32_t* __restrict__ tw) {
  __shared__ Fp sh[1 << TILE_LOG2];
  const int tl = log2 < TILE_LOG2 ? log2 : TILE_LOG2;
  const uint32_t tile = 1u << tl;
  const uint64_t base = (uint64_t)blockIdx.x * tile;
  for (uint32_t i = threadIdx.x; i < tile; i += NTT_THREADS) sh[i] = load_fp(d + 8 * (base + i));
  __syncthreads();
  for (int st = 0; st < tl; st++) {
    const uint32_t m = 1u << st;
    const int tws = log2 - 1 - st;  // twiddle stride n / (2m) = 2^(log2 - 1 - st)
    for (uint32_t b = threadIdx.x; b < tile / 2; b += NTT_THREADS) {
      uint32_t j = b & (m - 1);
      uint32_t i0 = ((b >> st) << (st + 1)) + j;
      uint32_t i1 = i0 + m;
      Fp w = load_fp(tw + 8ull * ((uint64_t)j << tws));
      Fp u = sh[i0];
      Fp v = Fr::mul(sh[i1], w);
      sh[i0] = Fr::add(u, v);
      sh[i1] = Fr::sub(u, v);
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < tile; i += NTT_THREADS) store_fp(d + 8 * (base + i), sh[i]);
}

// one global DIT stage with half-size m = 2^st
__global__ void k_ntt_stage(uint32_t* __restrict__ d, int log2, int st, const uint32_t* __restrict__ tw) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (1u << (log2 - 1))) return;
  const uint32_t m = 1u << st;
  uint32_t j = b & (m - 1);
  uint64_t i0 = ((uint64_t)(b >> st) << (st + 1)) + j;
  uint64_t i1 = i0 + m;
  Fp w = load_fp(tw + 8ull * ((uint64_t)j << (log2 - 1 - st)));
  Fp u = load_fp(d + 8 * i0);
  Fp v = Fr::mul(load_fp(d + 8 * i1), w);
  store_fp(d + 8 * i0, Fr::add(u, v));
  store_fp(d + 8 * i1, Fr::sub(u, v));
}

// out[i] = in[i] * (table ? table[i] : c)
__global__ void k_scale(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n,
                        const uint32_t* __restrict__ table, Fp c) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fp x = load_fp(in + 8 * i);
  store_fp(out + 8 * i, Fr::mul(x, table ? load_fp(table + 8 * i) : c));
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

// out[i] = c * x^i for i < n, built from 2^10-entry low / high tables on the device
static void build_geom(DevArray& out, uint64_t n, bgh::Fr x, bgh::Fr c, bool canon, hipStream_t s) {
  const int lb = 10;
  uint64_t nlo = std::min<uint64_t>(n, 1ull << lb);
  uint64_t nhi = (n + (1ull << lb) - 1) >> lb;
  std::vector<uint32_t> lo(8 * nlo), hi(8 * nhi);
  bgh::Fr v = c;
  for (uint64_t i = 0; i < nlo; i++) {
    v.to_u32(&lo[8 * i]);
    v = v * x;
  }
  bgh::Fr step = bgh::Fr::one();
  for (int i = 0; i < (1 << lb); i++) step = step * x;  // x^1024
  bgh::Fr h = bgh::Fr::one();
  for (uint64_t i = 0; i < nhi; i++) {
    h.to_u32(&hi[8 * i]);
    h = h * step;
  }
  DevArray dlo, dhi;
  dlo.alloc(lo.size() * 4);
  dhi.alloc(hi.size() * 4);
  BG_HIP(hipMemcpyAsync(dlo.ptr, lo.data(), lo.size() * 4, hipMemcpyHostToDevice, s));
  BG_HIP(hipMemcpyAsync(dhi.ptr, hi.data(), hi.size() * 4, hipMemcpyHostToDevice, s));
  out.alloc(n * 32);
  k_geom<<<nblk(n, 256), 256, 0, s>>>(out.as<uint32_t>(), n, dlo.as<uint32_t>(), dhi.as<uint32_t>(), lb, canon ? 1 : 0);
  BG_HIP(hipGetLastError());
  BG_HIP(hipStreamSynchronize(s));  // host tables go out of scope
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
  build_geom(dom->tw_fwd, n / 2, w, bgh::Fr::one(), false, s);
  build_geom(dom->tw_inv, n / 2, wi, bgh::Fr::one(), false, s);
  build_geom(dom->sc_g, n, g, bgh::Fr::one(), false, s);
  build_geom(dom->sc_ninv_g, n, g, ninv, false, s);
  build_geom(dom->sc_ninv_gi, n, gi, ninv, false, s);
  build_geom(dom->sc_ninv_gi_c, n, gi, ninv, true, s);
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

static void core(uint32_t* d, int log2, const uint32_t* tw, hipStream_t s) {
  uint64_t n = 1ull << log2;
  int tl = log2 < TILE_LOG2 ? log2 : TILE_LOG2;
  k_ntt_tile<<<(unsigned)(n >> tl), NTT_THREADS, 0, s>>>(d, log2, tw);
  for (int st = tl; st < log2; st++) k_ntt_stage<<<nblk(n / 2, 256), 256, 0, s>>>(d, log2, st, tw);
  BG_HIP(hipGetLastError());
}

void ntt_run(uint32_t* d, uint32_t* tmp, int log2, int direction, hipStream_t s) {
  NttDomain& D = ntt_domain(log2);
  uint64_t n = D.n;
  bool inv = direction == BG_NTT_IFFT || direction == BG_NTT_COSET_IFFT;
  const uint32_t* pre = direction == BG_NTT_COSET_FFT ? D.sc_g.as<uint32_t>() : nullptr;
  k_bitrev<<<nblk(n, 256), 256, 0, s>>>(d, tmp, log2, pre);
  core(tmp, log2, inv ? D.tw_inv.as<uint32_t>() : D.tw_fwd.as<uint32_t>(), s);
  if (direction == BG_NTT_IFFT)
    k_scale<<<nblk(n, 256), 256, 0, s>>>(tmp, d, n, nullptr, to_fp(D.ninv));
  else if (direction == BG_NTT_COSET_IFFT)
    k_scale<<<nblk(n, 256), 256, 0, s>>>(tmp, d, n, D.sc_ninv_gi.as<uint32_t>(), Fp{});
  else
    BG_HIP(hipMemcpyAsync(d, tmp, n * 32, hipMemcpyDeviceToDevice, s));
  BG_HIP(hipGetLastError());
}

void qap_h_run(uint32_t* a, uint32_t* b, uint32_t* c, uint32_t* h, uint32_t* tmp, int log2, bool canonical_out,
               hipStream_t s) {
  NttDomain& D = ntt_domain(log2);
  uint64_t n = D.n;
  uint32_t* v3[3] = {a, b, c};
  for (uint32_t* v : v3) {
    // ifft (omega^-1 core) -> scale n^-1 g^j fused into the next bit reversal -> fft
    k_bitrev<<<nblk(n, 256), 256, 0, s>>>(v, tmp, log2, nullptr);
    core(tmp, log2, D.tw_inv.as<uint32_t>(), s);
    k_bitrev<<<nblk(n, 256), 256, 0, s>>>(tmp, v, log2, D.sc_ninv_g.as<uint32_t>());
    core(v, log2, D.tw_fwd.as<uint32_t>(), s);
  }
  k_qap_point<<<nblk(n, 256), 256, 0, s>>>(a, b, c, n, to_fp(D.zinv));
  k_bitrev<<<nblk(n, 256), 256, 0, s>>>(a, tmp, log2, nullptr);
  core(tmp, log2, D.tw_inv.as<uint32_t>(), s);
  k_scale<<<nblk(n, 256), 256, 0, s>>>(tmp, h, n,
                                       canonical_out ? D.sc_ninv_gi_c.as<uint32_t>() : D.sc_ninv_gi.as<uint32_t>(),
                                       Fp{});
  BG_HIP(hipGetLastError());
}

}  // namespace bg

using namespace bg;

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
