// Per-group MSM kernels (accumulation, partial merge, two-level bucket reduction) and
// the device pipeline driver, instantiated once per group in bg_msm_g1.hip /
// bg_msm_g2.hip so the two large instantiations compile in parallel.
#pragma once
#include "bg_msm.hpp"
#include "bg_curve.hpp"

namespace bg {

inline unsigned blocks_of(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

static constexpr uint32_t NOKEY_C = 0xFFFFFFFFu;

// Each thread sums L consecutive sorted entries.  Runs wholly inside the slice are
// complete buckets; a run touching a slice edge whose key continues into the neighbour
// is a partial: slot 2t+1 if it continues into the next slice, else slot 2t.
template <class F>
__global__ __launch_bounds__(256) void k_accumulate(const uint32_t* __restrict__ points,
                                                    const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ svals,
                                                    const uint32_t* __restrict__ total_ptr, uint32_t L,
                                                    uint32_t nthreads, uint32_t* __restrict__ buckets,
                                                    uint32_t* __restrict__ pkeys, uint32_t* __restrict__ pflags,
                                                    uint32_t* __restrict__ ppts) {
  using C = Curve<F>;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nthreads) return;
  pkeys[2 * t] = NOKEY_C;
  pkeys[2 * t + 1] = NOKEY_C;
  uint32_t T = *total_ptr;
  uint64_t start = (uint64_t)t * L;
  if (start >= T) return;
  uint64_t end = min(start + L, (uint64_t)T);
  uint32_t prevkey = start > 0 ? skeys[start - 1] : NOKEY_C;
  uint32_t nextkey = end < T ? skeys[end] : NOKEY_C;
  uint32_t cur = skeys[start];
  bool first = true;
  typename C::P acc = C::infinity();
  for (uint64_t e = start; e < end; e++) {
    uint32_t k = skeys[e];
    if (k != cur) {
      bool cp = first && prevkey == cur;
      if (!cp) {
        C::store(buckets + (uint64_t)C::xyzz_words * cur, acc);
      } else {
        pkeys[2 * t] = cur;
        pflags[2 * t] = 1;
        C::store(ppts + (uint64_t)C::xyzz_words * (2 * t), acc);
      }
      first = false;
      cur = k;
      acc = C::infinity();
    }
    uint32_t v = svals[e];
    typename C::A pt = C::load_aff(points + (uint64_t)C::aff_words * (v & 0x7FFFFFFFu));
    if (v >> 31) pt = C::aff_neg(pt);
    acc = C::madd(acc, pt);
  }
  bool cp = first && prevkey == cur;
  bool cn = nextkey == cur;
  if (!cp && !cn) {
    C::store(buckets + (uint64_t)C::xyzz_words * cur, acc);
  } else {
    uint32_t slot = cn ? 2 * t + 1 : 2 * t;
    pkeys[slot] = cur;
    pflags[slot] = cp ? 1u : 0u;
    C::store(ppts + (uint64_t)C::xyzz_words * slot, acc);
  }
}

// A group of partials for one bucket starts at a tail slot whose run does not continue
// from the previous slice; it spans following slices' 'both' tails and ends at a head.
template <class F>
__global__ __launch_bounds__(256) void k_merge(const uint32_t* __restrict__ pkeys, const uint32_t* __restrict__ pflags,
                                               const uint32_t* __restrict__ ppts, uint32_t nthreads,
                                               uint32_t* __restrict__ buckets) {
  using C = Curve<F>;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nthreads) return;
  uint32_t slot = 2 * t + 1;
  uint32_t k = pkeys[slot];
  if (k == NOKEY_C || pflags[slot]) return;
  typename C::P acc = C::load(ppts + (uint64_t)C::xyzz_words * slot);
  for (uint32_t u = t + 1; u < nthreads; u++) {
    uint32_t s2 = 2 * u + 1;
    if (pkeys[s2] == k && pflags[s2]) {
      acc = C::add(acc, C::load(ppts + (uint64_t)C::xyzz_words * s2));
      continue;
    }
    if (pkeys[2 * u] == k) acc = C::add(acc, C::load(ppts + (uint64_t)C::xyzz_words * (2 * u)));
    break;
  }
  C::store(buckets + (uint64_t)C::xyzz_words * k, acc);
}

// Level 1: thread g sums buckets i = j*m + l (l < m) of window w:
//   S = sum_l b_l,  T = sum_l l * b_l   (running sums from the top)
template <class F>
__global__ __launch_bounds__(256) void k_reduce1(const uint32_t* __restrict__ buckets,
                                                 const uint32_t* __restrict__ offsets, uint32_t B, uint32_t m,
                                                 uint32_t ngroups, uint32_t* __restrict__ lvlS,
                                                 uint32_t* __restrict__ lvlT) {
  using C = Curve<F>;
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  uint32_t G = B / m;
  uint32_t w = g / G, j = g % G;
  uint64_t base = (uint64_t)w * B + (uint64_t)j * m;
  typename C::P R = C::infinity(), T = C::infinity();
  for (int l = (int)m - 1; l >= 0; l--) {
    uint64_t key = base + l;
    if (offsets[key + 1] > offsets[key]) R = C::add(R, C::load(buckets + (uint64_t)C::xyzz_words * key));
    if (l > 0) T = C::add(T, R);
  }
  C::store(lvlS + (uint64_t)C::xyzz_words * g, R);
  C::store(lvlT + (uint64_t)C::xyzz_words * g, T);
}

// Level 2: a block of K = blockDim.x threads (power of two) takes K consecutive level-1
// groups of one window:
//   U = sum_k (T_k + S_k),  V = sum_k k*S_k (suffix scan + tree),  Z = sum_k S_k
//   out P = U + m*V, Q = Z
template <class F>
__global__ __launch_bounds__(256) void k_reduce2(const uint32_t* __restrict__ lvlS, const uint32_t* __restrict__ lvlT,
                                                 int log2m, uint32_t* __restrict__ outPQ) {
  using C = Curve<F>;
  using P = typename C::P;
  extern __shared__ __align__(16) unsigned char red2_smem[];
  const uint32_t K = blockDim.x;
  P* shA = reinterpret_cast<P*>(red2_smem);
  P* shU = shA + K;
  uint32_t k = threadIdx.x;
  uint64_t g = (uint64_t)blockIdx.x * K + k;
  P S = C::load(lvlS + (uint64_t)C::xyzz_words * g);
  P T = C::load(lvlT + (uint64_t)C::xyzz_words * g);
  P a = S;
  shA[k] = a;
  shU[k] = C::add(T, S);
  __syncthreads();
  for (uint32_t d = 1; d < K; d <<= 1) {
    P x = (k + d < K) ? shA[k + d] : C::infinity();
    __syncthreads();
    a = C::add(a, x);
    shA[k] = a;
    __syncthreads();
  }
  P Z = shA[0];
  __syncthreads();
  if (k == 0) shA[0] = C::infinity();
  __syncthreads();
  for (uint32_t s = K / 2; s > 0; s >>= 1) {
    if (k < s) {
      shA[k] = C::add(shA[k], shA[k + s]);
      shU[k] = C::add(shU[k], shU[k + s]);
    }
    __syncthreads();
  }
  if (k == 0) {
    P V = shA[0];
    for (int i = 0; i < log2m; i++) V = C::dbl(V);
    P out = C::add(shU[0], V);
    C::store(outPQ + (uint64_t)C::xyzz_words * (2 * blockIdx.x), out);
    C::store(outPQ + (uint64_t)C::xyzz_words * (2 * blockIdx.x + 1), Z);
  }
}

template <class F>
void enqueue_curve(const uint32_t* d_points, const uint32_t* d_scalars, uint64_t n, int window_bits,
                         hipStream_t s, MsmWork& w, int group) {
  using C = Curve<F>;
  BG_REQUIRE(n < (1ull << 31), BG_ERR_ARG, "msm: n = %llu exceeds 2^31 - 1", (unsigned long long)n);
  MsmPlan p = msm_plan(n, group, window_bits);
  BG_REQUIRE(p.c >= 2 && p.c <= 22, BG_ERR_ARG, "msm: window_bits %d out of range", p.c);
  w.plan = p;
  w.group = group;
  const uint64_t xw = C::xyzz_words;
  uint32_t* keys = w.keys.as<uint32_t>(p.entries);
  uint32_t* vals = w.vals.as<uint32_t>(p.entries);
  uint32_t* counts = w.counts.as<uint32_t>(2 * ((uint64_t)p.nb + 1));  // counts | offsets
  uint32_t* offsets = counts + p.nb + 1;
  uint32_t nblk = msm_scan_blocks(p.nb);
  uint32_t* tmp = w.scan_tmp.as<uint32_t>((uint64_t)nblk + 8 + p.nb);
  uint32_t* bsum = tmp;
  uint32_t* cursor = tmp + nblk + 8;
  uint32_t* skeys = w.skeys.as<uint32_t>(p.entries);
  uint32_t* svals = w.svals.as<uint32_t>(p.entries);
  uint32_t* buckets = w.buckets.as<uint32_t>((uint64_t)p.nb * xw);
  uint32_t* pkeys = w.pkeys.as<uint32_t>(2ull * p.acc_threads);
  uint32_t* pflags = w.pflags.as<uint32_t>(2ull * p.acc_threads);
  uint32_t* ppts = w.ppts.as<uint32_t>(2ull * p.acc_threads * xw);
  uint64_t ngroups = (uint64_t)p.W * p.G;
  uint32_t* lvlS = w.lvlS.as<uint32_t>(ngroups * xw);
  uint32_t* lvlT = w.lvlT.as<uint32_t>(ngroups * xw);
  uint64_t nout = (uint64_t)p.W * p.H;
  uint32_t* outPQ = w.outPQ.as<uint32_t>(2 * nout * xw);
  uint32_t* hout = static_cast<uint32_t*>(w.host_out.get(2 * nout * xw * 4));

  if (n == 0) {
    // empty MSM: every bucket empty -> zero outputs
    launch_zero_u32(outPQ, 2 * nout * xw, s);
  } else {
    msm_sort(d_scalars, n, p, s, group, keys, vals, counts, offsets, bsum, cursor, skeys, svals);
    {
      ScopedKTime kt(group == 1 ? "msm_g1_accumulate" : "msm_g2_accumulate", s);
      k_accumulate<F><<<blocks_of(p.acc_threads, 256), 256, 0, s>>>(d_points, skeys, svals, offsets + p.nb, p.L,
                                                                     p.acc_threads, buckets, pkeys, pflags, ppts);
      k_merge<F><<<blocks_of(p.acc_threads, 256), 256, 0, s>>>(pkeys, pflags, ppts, p.acc_threads, buckets);
      BG_HIP(hipGetLastError());
    }
    {
      ScopedKTime kt(group == 1 ? "msm_g1_reduce" : "msm_g2_reduce", s);
      k_reduce1<F><<<blocks_of(ngroups, 256), 256, 0, s>>>(buckets, offsets, p.B, p.m, (uint32_t)ngroups, lvlS, lvlT);
      int log2m = 0;
      while ((1u << log2m) < p.m) log2m++;
      unsigned nblocks = (unsigned)(ngroups / p.K);
      size_t smem = 2 * (size_t)p.K * sizeof(typename C::P);
      BG_REQUIRE(smem <= 65536, BG_ERR_INTERNAL, "msm: level-2 block %u needs %zu B of LDS", p.K, smem);
      k_reduce2<F><<<nblocks, p.K, smem, s>>>(lvlS, lvlT, log2m, outPQ);
      BG_HIP(hipGetLastError());
    }
  }
  BG_HIP(hipMemcpyAsync(hout, outPQ, 2 * nout * xw * 4, hipMemcpyDeviceToHost, s));
  if (!w.done) BG_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
  BG_HIP(hipEventRecord(w.done, s));
}


}  // namespace bg
