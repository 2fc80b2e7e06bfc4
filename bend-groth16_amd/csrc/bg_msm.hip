// Pippenger MSM on gfx950 (C ABI: bg_msm_g1, bg_msm_g2).  See bg_msm.hpp for the shape.
#include "bg_msm.hpp"
#include <chrono>
#include "bg_curve.hpp"
#include "bg_host_ec.hpp"
#include "bg_tune.hpp"
#include <algorithm>
#include <map>
#include <memory>
#include <thread>
#include <vector>

namespace bg {

static constexpr uint32_t NOKEY = 0xFFFFFFFFu;

// counting sort (k_count): LDS histogram words (128 KB of 16-bit counter pairs), buckets
// one block counts, threads per block, largest chunk of scalars per block
static constexpr uint32_t HIST_WORDS = 32768;
static constexpr uint32_t HIST_BUCKETS = 2 * HIST_WORDS;
static constexpr int COUNT_T = 1024;
static constexpr int MAX_CHUNK_LOG2 = 15;

// ------------------------------------------------------------------ planning

// Tuning overrides (BG16_MSM_C_G1 / BG16_MSM_C_G2 / BG16_MSM_L / BG16_MSM_M / BG16_MSM_K);
// 0 or unset keeps the planner's choice.
static int tune_env(const char* name) {
  const char* e = getenv(name);
  return e ? atoi(e) : 0;
}

uint64_t accumulate_threads_g1();
uint64_t accumulate_threads_g2();
uint32_t reduce2_quad_kmax_g1();
uint32_t reduce2_quad_kmax_g2();
uint64_t accumulate_threads_per_round(int group) {
  return group == 1 ? accumulate_threads_g1() : accumulate_threads_g2();
}

// reduction shape and accumulation slicing, given c, W, Wd, M, n
static void finish_plan(MsmPlan& p, int group) {
  p.B = 1u << (p.c - 1);
  p.kspan = (uint32_t)p.W * p.B;
  p.nb = p.M * p.kspan;
  // buckets per level-1 group: 2, 4 or 8 (the four-role level-1 sums need m <= 8)
  // G1's level 1 runs one wave per binary-weight role, so wide groups cost no depth; G2's
  // runs one lane per group (a running sum of depth 2m), so with one set of buckets
  // (fixed-base tables) its groups are narrower; with a bucket set per window the level-2
  // blocks of narrow groups would outnumber what the device holds at once.  A fixed-base
  // set of 2^16 or more buckets takes groups of 4: in pairs its level 2 would hold 512
  // waves of G2 registers (a SIMD's file each) for ~0.65 ms while the other MSMs of a
  // proof accumulate (2^18 proof 5.22 -> 5.12 ms; below 2^16 buckets pairs stay faster)
  // the device-tuned table's entry for this batch, if any (environment overrides win)
  MsmTune tt;
  const bool tuned = tune_msm(group, p.fixed, p.n, tt);
  uint32_t mm = (uint32_t)tune_env(group == 1 ? "BG16_MSM_M" : "BG16_MSM_M_G2");
  if (mm == 0 && tuned) mm = (uint32_t)tt.m;
  if (mm == 0) mm = group == 1 || p.W > 1 ? 8 : (p.B >= 65536 ? 4 : 2);
  mm = mm >= 8 ? 8 : (mm >= 4 ? 4 : 2);
  p.m = std::min<uint32_t>(mm, p.B);
  p.G = p.B / p.m;
  // level-2 block: as wide as the LDS holds, except for a G2 fixed-base set of at most
  // 1024 groups (the B2 of proofs up to 2^11), where the level 2 is a few blocks whose
  // chain of 2 log2 K G2 additions is the latency: blocks of 32 shorten it and leave the
  // host at most 32 outputs (2^10 proof 1.36 -> 1.21 ms; at 2048 groups the extra host
  // work cancels it)
  uint32_t kmax = group == 1 ? 256 : 128;
  if (group == 2 && p.W == 1 && p.G <= 1024) kmax = 32;
  // The bucket reductions with a quad of lanes per group (k_reduce1q / k_reduce1_serialq,
  // k_reduce2q; level-2 blocks as wide as that kernel takes) where they are latency-bound:
  // at most 2^15 level-1 groups (a windowed 2^19 G1 MSM's 69,632 groups fill the device and
  // its level 2 measured 0.235 -> 0.254 ms with quads).  Measured per proof, alternating
  // with the one-lane kernels: K=10 0.97-0.98 -> 0.78-0.79 ms, K=12 1.08-1.09 -> 0.92-0.94,
  // K=14 1.29-1.31 -> 1.14-1.17, K=16 1.85 -> 1.74-1.79, K=18 4.04-4.05 -> 3.97-4.00.
  // BG16_QUAD_MIN_LOG2 keeps batches below 2^n points on the one-lane kernels;
  // BG16_QUAD_GROUPS_LOG2 / BG16_QUAD_GROUPS_LOG2_G2 move the group-count limit.
  const uint32_t kq = group == 1 ? reduce2_quad_kmax_g1() : reduce2_quad_kmax_g2();
  const int qmin = tune_env("BG16_QUAD_MIN_LOG2");
  int qgroups = tune_env(group == 1 ? "BG16_QUAD_GROUPS_LOG2" : "BG16_QUAD_GROUPS_LOG2_G2");
  if (qgroups <= 0) qgroups = 15;
  p.quad2 = kq != 0 && (qmin <= 0 || p.n >= (1ull << std::min(qmin, 40))) &&
            (uint64_t)p.M * p.W * p.G <= (1ull << std::min(qgroups, 40));
  if (p.quad2) kmax = std::min(kmax, kq);
  // override: rounded down to a power of two (the level-2 scan and trees need one, and
  // K must divide the group count)
  uint32_t kk = (uint32_t)std::max(0, tune_env("BG16_MSM_K"));
  if (!kk && tuned) kk = (uint32_t)tt.K;
  if (kk) {
    uint32_t k2 = 1;
    while (k2 * 2 <= kk) k2 *= 2;
    kmax = std::min(kmax, k2);
  }
  p.K = std::min<uint32_t>(kmax, p.G);
  p.H = p.G / p.K;
  p.entries = (uint64_t)p.Wd * p.n;
  // entries per accumulation thread: about 40, rounded so that the launch fills a whole
  // number of device rounds (threads resident at once), which leaves no tail of a few
  // waves.  Longer slices leave fewer partial runs to merge: one round of ~40-60 entries
  // instead of two of ~20 costs the accumulation ~5 % and halves the first merge level
  // (2^18 proof 4.74 -> 4.67 ms, G1 MSM 2^20 3.08 -> 2.84 ms, G2 MSM 2^20 7.01 -> 6.54 ms).
  uint32_t lenv = (uint32_t)tune_env(group == 1 ? "BG16_MSM_L" : "BG16_MSM_L_G2");
  uint64_t L;
  if (lenv) {
    L = lenv;
  } else if (p.fixed && tuned && tt.L) {
    // a fixed-base table's tuned slice length is taken as is, a partial device round
    // included: the 2^18 proof's B2 MSM (3.9 M entries) in 60-entry slices uses half the
    // device beside the G1 batch on its own stream (3.874 -> 3.779 ms per proof) where the
    // rounding below would give a full round of 30
    L = tt.L;
  } else {
    const uint64_t cap = accumulate_threads_per_round(group);
    const uint64_t target = tuned && tt.L ? (uint64_t)tt.L : 40;
    const uint64_t k = std::max<uint64_t>(1, p.entries / (target * cap));
    L = (p.entries + k * cap - 1) / (k * cap);
    // G2 MSMs of less than a device round: the rule above gives 2-3 entries a slice, and
    // the merge tree then adds the runs of a bucket in several dependent rounds of G2
    // additions; 8-entry slices trade them for sequential mixed adds (K=12 / 14 proofs
    // 6-10 % faster; below 2^16 entries the rule is kept)
    if (group == 2 && p.entries >= (1u << 16)) L = std::max<uint64_t>(L, 8);
  }
  p.L = (uint32_t)std::max<uint64_t>(2, std::min<uint64_t>(64, L));
  p.acc_threads = (uint32_t)((p.entries + p.L - 1) / p.L);
}

// Window-width cost in tenths of a mixed add: accumulation (10 per entry), digit counting
// (about 1.1 per entry per bucket part: windows wider than 17 bits are counted in several
// passes), chunk histograms (about 0.65 per word: written, scanned, offset, read by the
// scatter) and the bucket reduction (28 per bucket).  Weights from measured kernel times.
static double window_cost(const uint64_t* ns, uint32_t M, int c, uint32_t Wd, uint32_t Wb) {
  const uint64_t B = 1ull << (c - 1);
  const double parts = B > HIST_BUCKETS ? (double)(B / HIST_BUCKETS) : 1.0;
  double cost = (double)M * Wb * (double)B * 28.0;
  for (uint32_t i = 0; i < M; i++) {
    if (!ns[i]) continue;
    const uint64_t want = std::max<uint64_t>((ns[i] + 31) / 32, B / 2);
    const uint64_t ch = want > (1ull << 14) ? (1ull << MAX_CHUNK_LOG2) : (1ull << 14);
    const double hist_words = (double)Wd * (double)((ns[i] + ch - 1) / ch) * (double)B;
    cost += (double)ns[i] * Wd * (10.0 + 1.1 * parts) + 0.65 * hist_words;
  }
  return cost;
}

MsmPlan msm_plan_batch(const uint64_t* ns, uint32_t M, int group, int window_bits) {
  MsmPlan p;
  if (window_bits <= 0) window_bits = tune_env(group == 1 ? "BG16_MSM_C_G1" : "BG16_MSM_C_G2");
  p.M = M;
  uint64_t n = 0;
  for (uint32_t i = 0; i < M; i++) n += ns[i];
  p.n = n;
  int c = window_bits;
  MsmTune tt;
  if (c <= 0 && tune_msm(group, false, n, tt) && tt.c) c = tt.c;
  if (c <= 0) {
    // every window has its own buckets
    double best = 1e300;
    for (int cc = 4; cc <= 20; cc++) {
      const uint32_t W = (255 + cc - 1) / cc;
      const double cost = window_cost(ns, M, cc, W, W);
      if (cost < best) {
        best = cost;
        c = cc;
      }
    }
  }
  p.c = c;
  p.W = p.Wd = (255 + c - 1) / c;
  finish_plan(p, group);
  return p;
}

int fixed_window_bits(const uint64_t* ns, uint32_t M, int group) {
  int env = tune_env(group == 1 ? "BG16_FIXED_C_G1" : "BG16_FIXED_C_G2");
  if (env >= 2 && env <= 22) return env;  // other values: the planner's width
  uint64_t total = 0;
  for (uint32_t i = 0; i < M; i++) total += ns[i];
  MsmTune tt;
  if (tune_msm(group, true, total, tt) && tt.c) return tt.c;
  // every window shares one set of buckets
  double best = 1e300;
  int c = 16;
  for (int cc = 8; cc <= 20; cc++) {
    const uint32_t Wd = (255 + cc - 1) / cc;
    const double cost = window_cost(ns, M, cc, Wd, 1);
    if (cost < best) {
      best = cost;
      c = cc;
    }
  }
  return c;
}

MsmPlan msm_plan_fixed(const uint64_t* ns, uint32_t M, int group, int c) {
  MsmPlan p;
  p.M = M;
  for (uint32_t i = 0; i < M; i++) p.n += ns[i];
  p.c = c;
  p.fixed = true;
  p.W = 1;
  p.Wd = (255 + c - 1) / c;
  finish_plan(p, group);
  return p;
}

MsmPlan msm_plan(uint64_t n, int group, int window_bits) { return msm_plan_batch(&n, 1, group, window_bits); }

uint64_t msm_work_bytes(const MsmPlan& p, int group) {
  const uint64_t xb = group == 1 ? 128 : 256;  // XYZZ point bytes
  const uint64_t ngroups = (uint64_t)p.M * p.W * p.G, nout = (uint64_t)p.M * p.W * p.H;
  const uint64_t mcap = (p.acc_threads + 127) / 128;
  uint64_t b = p.entries * 20                    // keys, vals, ranks, sorted pairs
               + 8 * ((uint64_t)p.nb + 1)        // counts | offsets
               + 4 * (1 + 2ull * p.nb)           // fix-up flags and list
               + 4 * (2 + ngroups + ngroups / p.K)  // reduction exception lists
               + (uint64_t)p.nb * xb             // buckets
               + 2ull * p.acc_threads * (8 + xb)  // partial slots
               + 4 * mcap * (8 + xb)             // merge levels
               + 2 * ngroups * xb + 2 * nout * xb + (1u << 16)
               + 32 * p.n                         // reduced scalars
               + 4 * (ngroups + 256) * xb          // level-1 role sums
               + 4ull * p.Wd * p.B * (33 * p.M + 2 * p.n / (p.B > 1 ? p.B : 1) + 1);  // chunk histograms (bound)
  return b + b / 8;  // grow-only buffers keep 1/8 slack
}

// ------------------------------------------------------------------ kernels

// canonical scalar (any value < 2^256) -> value mod r, then signed c-bit digits
__device__ __forceinline__ void reduce_mod_r(uint32_t* s) {
  for (int it = 0; it < 5; it++) {
    uint32_t d[8];
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint64_t t = (uint64_t)s[i] - FrP::p[i] - br;
      d[i] = (uint32_t)t;
      br = (uint32_t)(t >> 63);
    }
    if (br) break;  // s < r
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] = d[i];
  }
}

__device__ __forceinline__ uint32_t get_bits(const uint32_t* s, int lo, int c) {
  int k = lo >> 5, off = lo & 31;
  uint32_t v = s[k] >> off;
  if (off + c > 32 && k + 1 < 8) v |= s[k + 1] << (32 - off);
  return v & ((1u << c) - 1u);
}

// Reduced scalars: s mod r, or 0 when the base is the point at infinity (all-zero affine
// encoding), so such a base gets no bucket entries and the accumulation never meets an
// infinite addend.
__global__ void k_prep_scalars(const uint32_t* __restrict__ scalars, const uint32_t* __restrict__ points,
                               int aff_words, uint32_t n, uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t s[8];
  const uint4* q = reinterpret_cast<const uint4*>(scalars + 8ull * i);
  uint4 a = q[0], b = q[1];
  s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
  s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
  reduce_mod_r(s);
  const uint4* pp = reinterpret_cast<const uint4*>(points + (uint64_t)aff_words * i);
  uint32_t nz = 0;
  for (int k = 0; k < aff_words / 4; k++) {
    uint4 v = pp[k];
    nz |= v.x | v.y | v.z | v.w;
  }
  uint4* o = reinterpret_cast<uint4*>(out + 8ull * i);
  o[0] = nz ? make_uint4(s[0], s[1], s[2], s[3]) : make_uint4(0, 0, 0, 0);
  o[1] = nz ? make_uint4(s[4], s[5], s[6], s[7]) : make_uint4(0, 0, 0, 0);
}

// Signed c-bit digits, window by window: digit_w = bits_w + carry_w, minus 2^c when it
// exceeds 2^(c-1) (then carrying 1 into window w+1).  The carry into window w has a
// closed form, so each window is computed on its own: carry_w = 1 exactly when the low
// w*c bits of s exceed the constant with bit j*c + c-1 set for every j < w (the largest
// value the digits below window w reach without a carry).
__device__ __forceinline__ uint32_t window_carry(const uint32_t* s, const uint32_t* pat, int nbits) {
  uint32_t gt = 0, decided = 0;
#pragma unroll
  for (int k = 7; k >= 0; k--) {
    uint32_t m = nbits >= 32 * (k + 1) ? 0xFFFFFFFFu : (nbits <= 32 * k ? 0u : (1u << (nbits - 32 * k)) - 1u);
    uint32_t x = s[k] & m, y = pat[k];
    uint32_t diff = (x != y) & !decided;
    gt |= diff & (x > y);
    decided |= diff;
  }
  return gt;
}

// Counting sort of (bucket, point) entries without global atomics (radix-sort layout):
// block (chunk, window, part) ranks its chunk's digits for one window in an LDS
// histogram (see k_count), writes (key, index | sign, rank in block) for
// each entry it owns and its histogram to hist[(window * nchunks + chunk) * B + bucket];
// k_column_scan turns the histograms into per-chunk bucket bases and bucket totals.
// The histogram takes 128 KB of gfx950's 160 KB LDS: 65536 16-bit counters, two to a
// word.  Chunks hold at most 2^15 scalars, so no counter reaches 2^16 even when every
// scalar of a chunk lands in one bucket.  Windows wider than 17 bits split the bucket range
// over blockIdx.z (each part reads the chunk again; the window planners charge for it).
// The rank word also carries the window in bits 24..31, so the scatter finds the entry's
// chunk histogram without dividing.

__global__ __launch_bounds__(COUNT_T) void k_count(const uint32_t* __restrict__ sc, uint32_t n, int c,
                                                   uint32_t chunk, uint32_t nchunks, uint32_t key_base, int fixed,
                                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                   uint32_t* __restrict__ ranks, uint32_t* __restrict__ hist_out) {
  __shared__ uint32_t hist[HIST_WORDS];
  const uint32_t B = 1u << (c - 1);
  const uint32_t Bp = B < HIST_BUCKETS ? B : HIST_BUCKETS;
  const uint32_t w = blockIdx.y;
  const uint32_t lo_b = blockIdx.z * Bp;
  const int nbits = (int)w * c;
  uint32_t pat[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < (int)w; j++) {
    int pos = j * c + c - 1;
    if (pos < 256) pat[pos >> 5] |= 1u << (pos & 31);
  }
  const uint32_t hw = (Bp + 1) / 2;
  for (uint32_t b = threadIdx.x; b < hw; b += COUNT_T) hist[b] = 0;
  __syncthreads();
  const uint64_t i0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t i1 = min(i0 + chunk, (uint64_t)n);
  const uint32_t full = 1u << c;
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += COUNT_T) {
    uint32_t s[8];
    const uint4* q = reinterpret_cast<const uint4*>(sc + 8 * i);
    uint4 a = q[0], bq = q[1];
    s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
    s[4] = bq.x; s[5] = bq.y; s[6] = bq.z; s[7] = bq.w;
    uint32_t v = (nbits < 256 ? get_bits(s, nbits, c) : 0u) + window_carry(s, pat, nbits);
    bool owner = blockIdx.z == 0;
    uint32_t key = NOKEY, val = 0, r = 0;
    if (v != 0 && v != full) {
      bool neg = v > B;
      uint32_t bk = (neg ? full - v : v) - 1;
      owner = bk - lo_b < Bp;
      if (owner) {
        // windowed bases: bucket (w, bk), point i; fixed-base tables: bucket bk shared by
        // all windows, point w * n + i of the table
        const uint32_t lb = bk - lo_b;
        key = key_base + (fixed ? 0u : w * B) + bk;
        val = (fixed ? w * n + (uint32_t)i : (uint32_t)i) | (neg ? 0x80000000u : 0u);
        const uint32_t sh = 16 * (lb & 1u);
        r = (atomicAdd(&hist[lb >> 1], 1u << sh) >> sh) & 0xFFFFu;
      }
    }
    if (owner) {
      uint64_t e = (uint64_t)w * n + i;
      keys[e] = key;
      vals[e] = val;
      ranks[e] = r | (w << 24);
    }
  }
  __syncthreads();
  uint32_t* ho = hist_out + ((uint64_t)w * nchunks + blockIdx.x) * B + lo_b;
  for (uint32_t b = threadIdx.x; b < Bp; b += COUNT_T) ho[b] = (hist[b >> 1] >> (16 * (b & 1u))) & 0xFFFFu;
}

// For every (window, bucket) of one MSM: exclusive prefix of the chunk histograms in
// place (the chunk's base inside the bucket) and the bucket total into counts.
__global__ void k_column_scan(uint32_t* __restrict__ hist, uint32_t W, uint32_t B, uint32_t nchunks,
                              uint32_t* __restrict__ counts) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)W * B) return;
  uint32_t w = (uint32_t)(t / B), b = (uint32_t)(t % B);
  uint32_t* col = hist + (uint64_t)w * nchunks * B + b;
  uint32_t run = 0;
  uint32_t k = 0;
  for (; k + 8 <= nchunks; k += 8) {  // eight loads in flight
    uint32_t x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = col[(uint64_t)(k + u) * B];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      col[(uint64_t)(k + u) * B] = run;
      run += x[u];
    }
  }
  for (; k < nchunks; k++) {
    uint32_t x = col[(uint64_t)k * B];
    col[(uint64_t)k * B] = run;
    run += x;
  }
  counts[t] = run;
}

// exclusive scan, three kernels: 2048 elements per block
static constexpr int SCAN_T = 1024;
__global__ void k_scan_local(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t* __restrict__ bsum,
                             uint32_t n) {
  __shared__ uint32_t sh[2 * SCAN_T];
  uint32_t base = blockIdx.x * 2 * SCAN_T;
  uint32_t t = threadIdx.x;
  uint32_t i0 = base + 2 * t, i1 = i0 + 1;
  uint32_t a = i0 < n ? in[i0] : 0, b = i1 < n ? in[i1] : 0;
  sh[2 * t] = a;
  sh[2 * t + 1] = a + b;
  __syncthreads();
  // inclusive scan of the pair sums (Hillis-Steele over SCAN_T)
  uint32_t v = sh[2 * t + 1];
  for (int d = 1; d < SCAN_T; d <<= 1) {
    uint32_t x = t >= (uint32_t)d ? sh[2 * (t - d) + 1] : 0;
    __syncthreads();
    v += x;
    sh[2 * t + 1] = v;
    __syncthreads();
  }
  uint32_t excl = v - (a + b);
  if (i0 < n) out[i0] = excl;
  if (i1 < n) out[i1] = excl + a;
  if (t == SCAN_T - 1) bsum[blockIdx.x] = v;
}

// single block: exclusive scan of nblocks block sums in place; writes the total to *total
__global__ void k_scan_top(uint32_t* __restrict__ bsum, uint32_t nblocks, uint32_t* __restrict__ total) {
  __shared__ uint32_t sh[SCAN_T];
  uint32_t t = threadIdx.x;
  uint32_t per = (nblocks + SCAN_T - 1) / SCAN_T;
  uint32_t lo = t * per, hi = min(lo + per, nblocks);
  uint32_t s = 0;
  for (uint32_t i = lo; i < hi; i++) s += bsum[i];
  sh[t] = s;
  __syncthreads();
  uint32_t v = s;
  for (int d = 1; d < SCAN_T; d <<= 1) {
    uint32_t x = t >= (uint32_t)d ? sh[t - d] : 0;
    __syncthreads();
    v += x;
    sh[t] = v;
    __syncthreads();
  }
  uint32_t run = v - s;
  for (uint32_t i = lo; i < hi; i++) {
    uint32_t x = bsum[i];
    bsum[i] = run;
    run += x;
  }
  if (t == SCAN_T - 1) *total = v;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ bsum, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += bsum[i / (2 * SCAN_T)];
}

// k_scan_top and k_scan_add in one launch for at most SCAN_FOLD_MAX_BLOCKS local blocks:
// block b adds up the block sums before its own (at most 4 reads a thread) and adds them
// to its 2 SCAN_T elements; the last block writes the total.
static constexpr uint32_t SCAN_FOLD_MAX_BLOCKS = 4 * SCAN_T;
__global__ __launch_bounds__(SCAN_T) void k_scan_add_fold(uint32_t* __restrict__ out,
                                                          const uint32_t* __restrict__ bsum, uint32_t n,
                                                          uint32_t* __restrict__ total) {
  __shared__ uint32_t sh[SCAN_T];
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  uint32_t sum = 0;
  for (uint32_t j = t; j < b; j += SCAN_T) sum += bsum[j];
  sh[t] = sum;
  __syncthreads();
  for (uint32_t d = SCAN_T / 2; d > 0; d >>= 1) {
    if (t < d) sh[t] += sh[t + d];
    __syncthreads();
  }
  const uint32_t pre = sh[0];
  const uint64_t i0 = (uint64_t)b * 2 * SCAN_T + t, i1 = i0 + SCAN_T;
  if (i0 < n) out[i0] += pre;
  if (i1 < n) out[i1] += pre;
  if (b + 1 == gridDim.x && t == 0) *total = pre + bsum[b];
}

// zeroes two word ranges in one launch
__global__ void k_zero2_u32(uint32_t* __restrict__ a, uint64_t na, uint32_t* __restrict__ b, uint64_t nb) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < na) a[i] = 0;
  else if (i - na < nb) b[i - na] = 0;
}

__global__ void k_copy_u32(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

__global__ void k_zero_u32(uint32_t* __restrict__ out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = 0;
}

// hist[(w * nchunks + chunk) * B + b] += offsets of its bucket, making each chunk base an
// absolute position in the sorted array (one coalesced pass; the scatter then reads one
// table instead of two at random).  fixed: every window shares the bucket offsets.
__global__ void k_add_offsets(uint32_t* __restrict__ hist, uint64_t words, uint32_t B, uint32_t nchunks, int fixed,
                              const uint32_t* __restrict__ offsets) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= words) return;
  const uint32_t b = (uint32_t)(t & (B - 1));
  const uint64_t row = t / B;  // w * nchunks + chunk
  const uint64_t key = fixed ? b : (row / nchunks) * B + b;
  hist[t] += offsets[key];
}

// sorted[chunk position of its bucket + rank] = (key, val), one MSM's entries
__global__ void k_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                          const uint32_t* __restrict__ ranks, uint32_t n, uint32_t Wd, uint32_t B, int log2chunk,
                          uint32_t nchunks, uint32_t key_base, int fixed, const uint32_t* __restrict__ hist,
                          uint2* __restrict__ sorted) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (uint64_t)Wd * n) return;
  uint32_t key = keys[e];
  if (key == NOKEY) return;
  uint32_t val = vals[e], rk = ranks[e];
  uint32_t w = rk >> 24, idx = val & 0x7FFFFFFFu;
  uint32_t i = fixed ? idx - w * n : idx;
  uint32_t b = (key - key_base) & (B - 1);
  uint32_t pos = hist[((uint64_t)w * nchunks + (i >> log2chunk)) * B + b];
  sorted[pos + (rk & 0xFFFFFFu)] = make_uint2(key, val);
}

// ------------------------------------------------------------------ two-pass bucket sort
// The entries (key, val) of an MSM batch ordered by key in two counting passes over the
// key bits, each of whose scatters writes a block's entries of one bin to one contiguous
// range (the single-pass sort above writes every entry to a random 8-byte slot and reads a
// random histogram word for it):
//   pass 1, by the high key bits (at most 2^14 bins): a block takes 1024 scalars of one
//     MSM, forms all their signed digits, counts them in an LDS histogram, and after a
//     scan of the (bin, block) counts writes each entry to its bin's range;
//   pass 2, by the low 8 key bits inside each high-bits segment: tiles of 4096 entries of
//     one segment, an LDS histogram each, a scan of the (segment, bin, tile) counts, and
//     the scatter into the final order.
// Ranks within a bin come from LDS atomics, so the order of a bucket's entries is not
// fixed; a bucket's sum does not depend on it (the accumulation's special cases are exact).
static constexpr int SORT_T = 256;
static constexpr uint32_t P1_SCALARS = 1024;  // scalars per pass-1 block
static constexpr int P1_MAX_BITS = 14;        // pass-1 bins: a 64 KB LDS histogram
static constexpr int P2_MAX_BITS = 8;
static constexpr uint32_t P2_TILE = 4096;     // entries per pass-2 block

// Signed c-bit digits of a reduced scalar, low window first: digit_w = bits_w + carry_w,
// minus 2^c (carry 1 into window w+1) when it exceeds 2^(c-1); fn(w, bucket, negative) for
// each nonzero digit (the same digits as k_count's closed-form carry).
template <class Fn>
__device__ __forceinline__ void for_each_digit(const uint32_t* s, int c, int Wd, Fn fn) {
  const uint32_t B = 1u << (c - 1), full = 1u << c;
  uint32_t carry = 0;
  for (int w = 0; w < Wd; w++) {
    const int lo = w * c;
    const uint32_t v = (lo < 256 ? get_bits(s, lo, c) : 0u) + carry;
    carry = v > B ? 1u : 0u;
    if (v != 0 && v != full) {
      const bool neg = v > B;
      fn(w, (neg ? full - v : v) - 1u, neg);
    }
  }
}

struct SortArgs {
  const uint32_t* sc;  // reduced scalars of this MSM
  uint32_t n, key_base, blk_base, nb1_total, nbins1;
  int c, Wd, fixed, hb;
};

__device__ __forceinline__ void load_scalar(const uint32_t* __restrict__ sc, uint64_t j, uint32_t* s) {
  const uint4* q = reinterpret_cast<const uint4*>(sc + 8 * j);
  const uint4 a = q[0], b = q[1];
  s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
  s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
}

__global__ __launch_bounds__(SORT_T) void k_sort_p1_count(SortArgs a, uint32_t* __restrict__ hist1) {
  extern __shared__ uint32_t h1[];
  for (uint32_t b = threadIdx.x; b < a.nbins1; b += SORT_T) h1[b] = 0;
  __syncthreads();
  const uint32_t B = 1u << (a.c - 1);
  const uint64_t j0 = (uint64_t)blockIdx.x * P1_SCALARS, j1 = min(j0 + P1_SCALARS, (uint64_t)a.n);
  for (uint64_t j = j0 + threadIdx.x; j < j1; j += SORT_T) {
    uint32_t s[8];
    load_scalar(a.sc, j, s);
    for_each_digit(s, a.c, a.Wd, [&](int w, uint32_t bk, bool) {
      const uint32_t key = a.key_base + (a.fixed ? 0u : (uint32_t)w * B) + bk;
      atomicAdd(&h1[key >> a.hb], 1u);
    });
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < a.nbins1; b += SORT_T)
    hist1[(uint64_t)b * a.nb1_total + a.blk_base + blockIdx.x] = h1[b];
}

// base1: the exclusive scan of hist1 (absolute first positions of each (bin, block))
__global__ __launch_bounds__(SORT_T) void k_sort_p1_scatter(SortArgs a, const uint32_t* __restrict__ base1,
                                                            uint2* __restrict__ tmp) {
  extern __shared__ uint32_t cur1[];
  for (uint32_t b = threadIdx.x; b < a.nbins1; b += SORT_T)
    cur1[b] = base1[(uint64_t)b * a.nb1_total + a.blk_base + blockIdx.x];
  __syncthreads();
  const uint32_t B = 1u << (a.c - 1);
  const uint64_t j0 = (uint64_t)blockIdx.x * P1_SCALARS, j1 = min(j0 + P1_SCALARS, (uint64_t)a.n);
  for (uint64_t j = j0 + threadIdx.x; j < j1; j += SORT_T) {
    uint32_t s[8];
    load_scalar(a.sc, j, s);
    for_each_digit(s, a.c, a.Wd, [&](int w, uint32_t bk, bool neg) {
      const uint32_t key = a.key_base + (a.fixed ? 0u : (uint32_t)w * B) + bk;
      const uint32_t val = (a.fixed ? (uint32_t)w * a.n + (uint32_t)j : (uint32_t)j) | (neg ? 0x80000000u : 0u);
      tmp[atomicAdd(&cur1[key >> a.hb], 1u)] = make_uint2(key, val);
    });
  }
}

// One block: segment starts (column 0 of the scanned pass-1 counts; seg[nbins1] = total),
// tiles per segment and their exclusive scan (wbase[nbins1] = tiles in all).
__global__ __launch_bounds__(1024) void k_sort_tiles(const uint32_t* __restrict__ base1, uint32_t nb1_total,
                                                     uint32_t nbins1, const uint32_t* __restrict__ total,
                                                     uint32_t* __restrict__ seg, uint32_t* __restrict__ ntiles,
                                                     uint32_t* __restrict__ wbase) {
  __shared__ uint32_t sh[1024];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nbins1 + 1023) / 1024, lo = t * per, hi = min(lo + per, nbins1);
  const uint32_t tot = *total;
  uint32_t sum = 0;
  for (uint32_t b = lo; b < hi; b++) {
    const uint32_t st = base1[(uint64_t)b * nb1_total];
    const uint32_t en = b + 1 < nbins1 ? base1[(uint64_t)(b + 1) * nb1_total] : tot;
    seg[b] = st;
    const uint32_t nt = (en - st + P2_TILE - 1) / P2_TILE;
    ntiles[b] = nt;
    sum += nt;
  }
  sh[t] = sum;
  __syncthreads();
  uint32_t v = sum;
  for (uint32_t d = 1; d < 1024; d <<= 1) {
    const uint32_t x = t >= d ? sh[t - d] : 0u;
    __syncthreads();
    v += x;
    sh[t] = v;
    __syncthreads();
  }
  uint32_t run = v - sum;
  for (uint32_t b = lo; b < hi; b++) {
    wbase[b] = run;
    run += ntiles[b];
  }
  if (t == 1023) {
    wbase[nbins1] = v;
    seg[nbins1] = tot;
  }
}

// pass-2 block -> (segment, tile, entry range); false past the last tile
__device__ __forceinline__ bool p2_tile(const uint32_t* __restrict__ seg, const uint32_t* __restrict__ wbase,
                                        uint32_t nbins1, uint32_t& sg, uint32_t& tile, uint32_t& e0,
                                        uint32_t& e1) {
  const uint32_t b = blockIdx.x;
  if (b >= wbase[nbins1]) return false;
  uint32_t lo = 0, hi = nbins1;  // largest sg with wbase[sg] <= b (and a tile: wbase[sg + 1] > b)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (wbase[mid] <= b) lo = mid; else hi = mid;
  }
  sg = lo;
  tile = b - wbase[lo];
  e0 = seg[lo] + tile * P2_TILE;
  e1 = min(e0 + P2_TILE, seg[lo + 1]);
  return true;
}

__global__ __launch_bounds__(SORT_T) void k_sort_p2_count(const uint2* __restrict__ tmp,
                                                          const uint32_t* __restrict__ seg,
                                                          const uint32_t* __restrict__ ntiles,
                                                          const uint32_t* __restrict__ wbase, uint32_t nbins1, int hb,
                                                          uint32_t* __restrict__ hist2) {
  __shared__ uint32_t h2[1u << P2_MAX_BITS];
  uint32_t sg, tile, e0, e1;
  if (!p2_tile(seg, wbase, nbins1, sg, tile, e0, e1)) return;
  const uint32_t nbins2 = 1u << hb, mask = nbins2 - 1;
  for (uint32_t b = threadIdx.x; b < nbins2; b += SORT_T) h2[b] = 0;
  __syncthreads();
  for (uint32_t e = e0 + threadIdx.x; e < e1; e += SORT_T) atomicAdd(&h2[tmp[e].x & mask], 1u);
  __syncthreads();
  const uint32_t nt = ntiles[sg];
  const uint64_t base = (uint64_t)wbase[sg] * nbins2 + tile;
  for (uint32_t b = threadIdx.x; b < nbins2; b += SORT_T) hist2[base + (uint64_t)b * nt] = h2[b];
}

// The tile is ordered by bin in LDS first (local histogram, scan, ranks), then written
// out position by position: consecutive threads write consecutive slots of one bin's
// range, so the stores coalesce (ranked in place, every entry would land at a random
// 8-byte slot of its bin's range, about three times the bytes in partial lines).
__global__ __launch_bounds__(SORT_T) void k_sort_p2_scatter(const uint2* __restrict__ tmp,
                                                            const uint32_t* __restrict__ seg,
                                                            const uint32_t* __restrict__ ntiles,
                                                            const uint32_t* __restrict__ wbase, uint32_t nbins1,
                                                            int hb, const uint32_t* __restrict__ base2,
                                                            uint2* __restrict__ sorted) {
  __shared__ uint32_t lcount[1u << P2_MAX_BITS];  // counts, then local starts
  __shared__ uint32_t lcur[1u << P2_MAX_BITS];
  __shared__ uint32_t gbase[1u << P2_MAX_BITS];   // global start minus local start, per bin
  __shared__ uint2 stage[P2_TILE];
  uint32_t sg, tile, e0, e1;
  if (!p2_tile(seg, wbase, nbins1, sg, tile, e0, e1)) return;
  const uint32_t nbins2 = 1u << hb, mask = nbins2 - 1;
  const uint32_t nt = ntiles[sg];
  const uint64_t base = (uint64_t)wbase[sg] * nbins2 + tile;
  for (uint32_t b = threadIdx.x; b < nbins2; b += SORT_T) lcount[b] = 0;
  __syncthreads();
  constexpr uint32_t PER = P2_TILE / SORT_T;
  uint2 ent[PER];
#pragma unroll
  for (uint32_t k = 0; k < PER; k++) {
    const uint32_t e = e0 + k * SORT_T + threadIdx.x;
    ent[k] = e < e1 ? tmp[e] : make_uint2(0xFFFFFFFFu, 0u);
    if (e < e1) atomicAdd(&lcount[ent[k].x & mask], 1u);
  }
  __syncthreads();
  // exclusive scan of the (at most SORT_T) bin counts, one bin per thread
  static_assert((1u << P2_MAX_BITS) <= SORT_T, "one thread per pass-2 bin");
  const uint32_t b = threadIdx.x;
  const uint32_t own = b < nbins2 ? lcount[b] : 0u;
  uint32_t incl = own;
  for (uint32_t d = 1; d < nbins2; d <<= 1) {
    const uint32_t x = b >= d && b < nbins2 ? lcount[b - d] : 0u;
    __syncthreads();
    incl += x;
    if (b < nbins2) lcount[b] = incl;
    __syncthreads();
  }
  if (b < nbins2) {
    lcur[b] = incl - own;
    gbase[b] = base2[base + (uint64_t)b * nt] - (incl - own);
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < PER; k++)
    if (ent[k].x != 0xFFFFFFFFu) stage[atomicAdd(&lcur[ent[k].x & mask], 1u)] = ent[k];
  __syncthreads();
  const uint32_t cnt = e1 - e0;
  for (uint32_t q = threadIdx.x; q < cnt; q += SORT_T) {
    const uint2 v = stage[q];
    sorted[gbase[v.x & mask] + q] = v;
  }
}

// offsets[k] = first sorted position of key k (k <= nb; offsets[nb] = total)
__global__ void k_sort_offsets(const uint32_t* __restrict__ seg, const uint32_t* __restrict__ ntiles,
                               const uint32_t* __restrict__ wbase, const uint32_t* __restrict__ base2, uint32_t nb,
                               int hb, uint32_t* __restrict__ offsets) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nb) return;
  const uint32_t sg = k >> hb, b = k & ((1u << hb) - 1u);
  uint32_t v;
  const uint32_t nbins1 = (nb + (1u << hb) - 1) >> hb;
  if (k == nb || sg >= nbins1)
    v = seg[nbins1];
  else if (ntiles[sg] == 0)
    v = seg[sg];
  else
    v = base2[(uint64_t)wbase[sg] * (1u << hb) + (uint64_t)b * ntiles[sg]];
  offsets[k] = v;
}

__global__ void k_pack16(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t nout) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nout) out[i] = (in[2 * i] & 0xFFFFu) | (in[2 * i + 1] << 16);
}
__global__ void k_unpack16(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t nin) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nin) {
    uint32_t v = in[i];
    out[2 * i] = v & 0xFFFFu;
    out[2 * i + 1] = v >> 16;
  }
}

void launch_pack16(const uint32_t* in16, uint32_t* out32, uint64_t out_words, hipStream_t s) {
  if (!out_words) return;
  k_pack16<<<(unsigned)((out_words + 255) / 256), 256, 0, s>>>(in16, out32, out_words);
  BG_HIP(hipGetLastError());
}
void launch_unpack16(const uint32_t* in32, uint32_t* out16, uint64_t in_words, hipStream_t s) {
  if (!in_words) return;
  k_unpack16<<<(unsigned)((in_words + 255) / 256), 256, 0, s>>>(in32, out16, in_words);
  BG_HIP(hipGetLastError());
}

static unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

uint32_t msm_scan_blocks(uint32_t n) { return blocks_for(n, 2 * SCAN_T); }

void launch_zero_u32(uint32_t* p, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_zero_u32<<<blocks_for(n, 256), 256, 0, s>>>(p, n);
  BG_HIP(hipGetLastError());
}

// Up to SET_PTRS_MAX base-array addresses passed by value, so a recorded launch sequence
// carries them itself (a copy from a host staging buffer would read whatever the buffer
// holds when the sequence is replayed).
static constexpr uint32_t SET_PTRS_MAX = 8;
struct PtrTable {
  uint64_t p[SET_PTRS_MAX];
};
// the pointer table, and the two word ranges an MSM starts from zero (in one launch)
__global__ void k_set_ptrs(uint64_t* __restrict__ out, PtrTable t, uint32_t M, uint32_t* __restrict__ z1,
                           uint64_t n1, uint32_t* __restrict__ z2, uint64_t n2) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) out[i] = t.p[i];
  if (i < n1) z1[i] = 0;
  else if (i - n1 < n2) z2[i - n1] = 0;
}

void launch_set_ptrs(uint64_t* d_out, const uint64_t* h_ptrs, uint32_t M, hipStream_t s, uint32_t* z1, uint64_t n1,
                     uint32_t* z2, uint64_t n2) {
  if (M <= SET_PTRS_MAX) {
    PtrTable t{};
    for (uint32_t i = 0; i < M; i++) t.p[i] = h_ptrs[i];
    k_set_ptrs<<<blocks_for(std::max<uint64_t>(M, n1 + n2), 256), 256, 0, s>>>(d_out, t, M, z1, n1, z2, n2);
    BG_HIP(hipGetLastError());
  } else {
    BG_REQUIRE(!stream_capturing(s), BG_ERR_INTERNAL, "msm: a captured batch holds at most %u base arrays",
               SET_PTRS_MAX);
    BG_HIP(hipMemcpyAsync(d_out, h_ptrs, 8ull * M, hipMemcpyHostToDevice, s));
    if (n1 + n2) {
      k_zero2_u32<<<blocks_for(n1 + n2, 256), 256, 0, s>>>(z1, n1, z2, n2);
      BG_HIP(hipGetLastError());
    }
  }
}

__global__ void k_publish(const uint32_t* __restrict__ src, uint64_t n, uint32_t* __restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Runs after k_publish has completed (stream order), so its stores have reached the host.
__global__ void k_raise_flag(uint32_t* flag) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

void launch_publish(const uint32_t* src, uint64_t n, uint32_t* dst, uint32_t* flag, hipStream_t s) {
  if (n) k_publish<<<(unsigned)std::min<uint64_t>(blocks_for(n, 256), 64), 256, 0, s>>>(src, n, dst);
  k_raise_flag<<<1, 64, 0, s>>>(flag);
  BG_HIP(hipGetLastError());
}

// chunk of scalars per counting block: 2^14, or 2^15 when n/32 or half the bucket count
// is larger (the 16-bit LDS counters bound it at 2^15)
static int count_chunk_log2(uint64_t n, uint32_t B) {
  uint64_t want = std::max<uint64_t>((n + 31) / 32, B / 2);
  return want > (1ull << 14) ? MAX_CHUNK_LOG2 : 14;
}
static uint32_t count_chunk(uint64_t n, uint32_t B) { return 1u << count_chunk_log2(n, B); }

uint64_t msm_hist_words(const MsmPlan& p, const uint64_t* ns) {
  uint64_t t = 0;
  for (uint32_t i = 0; i < p.M; i++)
    if (ns[i]) {
      const uint32_t ch = count_chunk(ns[i], p.B);
      t += (uint64_t)p.Wd * ((ns[i] + ch - 1) / ch) * p.B;
    }
  return std::max<uint64_t>(t, 1);
}

// exclusive scan of in (n words) into out (in place when out == in); *total = the sum
static void scan_excl(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* bsum, uint32_t* total,
                      hipStream_t s) {
  const uint32_t nblk = msm_scan_blocks((uint32_t)n);
  k_scan_local<<<nblk, SCAN_T, 0, s>>>(in, out, bsum, (uint32_t)n);
  if (nblk <= SCAN_FOLD_MAX_BLOCKS) {
    k_scan_add_fold<<<nblk, SCAN_T, 0, s>>>(out, bsum, (uint32_t)n, total);
  } else {
    k_scan_top<<<1, SCAN_T, 0, s>>>(bsum, nblk, total);
    k_scan_add<<<blocks_for(n, 256), 256, 0, s>>>(out, bsum, (uint32_t)n);
  }
}

// digits + single-pass counting sort of (bucket, point) pairs by bucket (the key spaces
// above the two-pass sort's 2^22)
static void msm_sort_single(const uint32_t* const* d_scalars, const uint32_t* const* d_points, const uint64_t* ns,
                            const MsmPlan& p, hipStream_t s, int group, uint32_t* keys, uint32_t* vals,
                            uint32_t* ranks, uint32_t* counts, uint32_t* offsets, uint32_t* bsum, uint2* sorted,
                            uint32_t* fix, uint64_t fix_words, uint32_t* scal, uint32_t* hist) {
  const int fixed = p.fixed ? 1 : 0;
  {
    ScopedKTime kt(group == 1 ? "msm_g1_digits" : "msm_g2_digits", s);
    uint64_t off = 0, soff = 0, hoff = 0;
    for (uint32_t i = 0; i < p.M; i++) {
      const uint64_t n = ns[i];
      if (n) {
        const uint32_t ch = count_chunk(n, p.B), nch = (uint32_t)((n + ch - 1) / ch);
        k_prep_scalars<<<blocks_for(n, 256), 256, 0, s>>>(d_scalars[i], d_points[i], group == 1 ? 16 : 32,
                                                           (uint32_t)n, scal + soff);
        const uint32_t parts = p.B > HIST_BUCKETS ? p.B / HIST_BUCKETS : 1;
        k_count<<<dim3(nch, p.Wd, parts), COUNT_T, 0, s>>>(scal + soff, (uint32_t)n, p.c, ch, nch, i * p.kspan,
                                                           fixed, keys + off, vals + off, ranks + off, hist + hoff);
        // a bucket's column runs over every (window, chunk) histogram that feeds it
        if (fixed)
          k_column_scan<<<blocks_for(p.B, 256), 256, 0, s>>>(hist + hoff, 1, p.B, (uint32_t)p.Wd * nch,
                                                             counts + (uint64_t)i * p.kspan);
        else
          k_column_scan<<<blocks_for((uint64_t)p.W * p.B, 256), 256, 0, s>>>(hist + hoff, p.W, p.B, nch,
                                                                             counts + (uint64_t)i * p.kspan);
        hoff += (uint64_t)p.Wd * nch * p.B;
      }
      off += (uint64_t)p.Wd * n;
      soff += 8 * n;
    }
    BG_HIP(hipGetLastError());
  }
  {
    ScopedKTime kt(group == 1 ? "msm_g1_sort" : "msm_g2_sort", s);
    scan_excl(counts, offsets, p.nb, bsum, offsets + p.nb, s);
    uint64_t off = 0, hoff = 0;
    for (uint32_t i = 0; i < p.M; i++) {
      const uint64_t n = ns[i];
      if (n) {
        const uint32_t ch = count_chunk(n, p.B), nch = (uint32_t)((n + ch - 1) / ch);
        const uint64_t hw = (uint64_t)p.Wd * nch * p.B;
        k_add_offsets<<<blocks_for(hw, 256), 256, 0, s>>>(hist + hoff, hw, p.B, nch, fixed,
                                                          offsets + (uint64_t)i * p.kspan);
        k_scatter<<<blocks_for((uint64_t)p.Wd * n, 256), 256, 0, s>>>(keys + off, vals + off, ranks + off,
                                                                      (uint32_t)n, p.Wd, p.B, count_chunk_log2(n, p.B),
                                                                      nch, i * p.kspan, fixed, hist + hoff, sorted);
        hoff += (uint64_t)p.Wd * nch * p.B;
      }
      off += (uint64_t)p.Wd * n;
    }
    BG_HIP(hipGetLastError());
  }
}

// Key bits of an MSM batch and the split between the two passes (hb low bits in pass 2);
// false for the single-pass sort: below 2^21 entries (where its fewer dependent launches
// win: 2^10 .. 2^14 proofs 4-12 % faster with it, 2^16 and up slower), or when the high
// bits exceed the pass-1 histogram.  BG16_SORT=1 / 2 forces the single / two-pass sort.
static bool two_pass_bits(const MsmPlan& p, int& hb, uint32_t& nbins1) {
  const char* e = getenv("BG16_SORT");
  if (e && e[0] == '1') return false;
  if (!(e && e[0] == '2') && p.entries < (1ull << 21)) return false;
  int kb = 0;
  while ((1ull << kb) < (uint64_t)p.nb) kb++;
  hb = std::min(kb, P2_MAX_BITS);
  if (kb - hb > P1_MAX_BITS) return false;
  nbins1 = (uint32_t)(((uint64_t)p.nb + (1ull << hb) - 1) >> hb);
  return true;
}

void msm_sort(const uint32_t* const* d_scalars, const uint32_t* const* d_points, const uint64_t* ns,
              const MsmPlan& p, hipStream_t s, int group, MsmWork& w, uint32_t* offsets, uint2* sorted,
              uint32_t* fix, uint64_t fix_words) {
  int hb = 0;
  uint32_t nbins1 = 0;
  if (!two_pass_bits(p, hb, nbins1)) {
    uint32_t* counts = offsets - (p.nb + 1);
    msm_sort_single(d_scalars, d_points, ns, p, s, group, w.keys.as<uint32_t>(p.entries),
                    w.vals.as<uint32_t>(p.entries), w.ranks.as<uint32_t>(p.entries), counts, offsets,
                    w.scan_tmp.as<uint32_t>((uint64_t)msm_scan_blocks(p.nb) + 8), sorted, fix, fix_words,
                    w.scal.as<uint32_t>(std::max<uint64_t>(8 * p.n, 8)), w.hist.as<uint32_t>(msm_hist_words(p, ns)));
    return;
  }
  uint32_t nb1 = 0;  // pass-1 blocks over the batch
  for (uint32_t i = 0; i < p.M; i++) nb1 += (uint32_t)((ns[i] + P1_SCALARS - 1) / P1_SCALARS);
  const uint64_t h1n = std::max<uint64_t>((uint64_t)nbins1 * nb1, 1);
  const uint32_t g2 = (uint32_t)((p.entries + P2_TILE - 1) / P2_TILE) + nbins1;
  const uint32_t nbins2 = 1u << hb;
  const uint64_t h2n = (uint64_t)g2 * nbins2;
  BG_REQUIRE(h1n < (1ull << 31) && h2n < (1ull << 31), BG_ERR_ARG, "msm: sort histograms too large");
  uint32_t* scal = w.scal.as<uint32_t>(std::max<uint64_t>(8 * p.n, 8));
  uint32_t* hist1 = w.hist.as<uint32_t>(h1n);
  uint32_t* hist2 = w.ranks.as<uint32_t>(h2n);
  uint2* tmp = reinterpret_cast<uint2*>(w.keys.as<uint32_t>(2 * std::max<uint64_t>(p.entries, 1)));
  uint32_t* segw = w.vals.as<uint32_t>(3ull * (nbins1 + 1) + 8);
  uint32_t* seg = segw;
  uint32_t* ntiles = segw + nbins1 + 1;
  uint32_t* wbase = ntiles + nbins1 + 1;
  uint32_t* total = wbase + nbins1 + 1;
  uint32_t* bsum = w.scan_tmp.as<uint32_t>((uint64_t)msm_scan_blocks((uint32_t)std::max(h1n, h2n)) + 8);
  const int fixed = p.fixed ? 1 : 0;
  const size_t lds1 = (size_t)nbins1 * 4;
  {
    ScopedKTime kt(group == 1 ? "msm_g1_digits" : "msm_g2_digits", s);
    uint64_t soff = 0;
    uint32_t blk = 0;
    for (uint32_t i = 0; i < p.M; i++) {
      const uint64_t n = ns[i];
      if (n) {
        k_prep_scalars<<<blocks_for(n, 256), 256, 0, s>>>(d_scalars[i], d_points[i], group == 1 ? 16 : 32,
                                                           (uint32_t)n, scal + soff);
        const SortArgs a{scal + soff, (uint32_t)n, i * p.kspan, blk, nb1, nbins1, p.c, p.Wd, fixed, hb};
        k_sort_p1_count<<<(unsigned)((n + P1_SCALARS - 1) / P1_SCALARS), SORT_T, lds1, s>>>(a, hist1);
        blk += (uint32_t)((n + P1_SCALARS - 1) / P1_SCALARS);
      }
      soff += 8 * n;
    }
    BG_HIP(hipGetLastError());
  }
  {
    ScopedKTime kt(group == 1 ? "msm_g1_sort" : "msm_g2_sort", s);
    scan_excl(hist1, hist1, h1n, bsum, total, s);
    uint64_t soff = 0;
    uint32_t blk = 0;
    for (uint32_t i = 0; i < p.M; i++) {
      const uint64_t n = ns[i];
      if (n) {
        const SortArgs a{scal + soff, (uint32_t)n, i * p.kspan, blk, nb1, nbins1, p.c, p.Wd, fixed, hb};
        k_sort_p1_scatter<<<(unsigned)((n + P1_SCALARS - 1) / P1_SCALARS), SORT_T, lds1, s>>>(a, hist1, tmp);
        blk += (uint32_t)((n + P1_SCALARS - 1) / P1_SCALARS);
      }
      soff += 8 * n;
    }
    k_sort_tiles<<<1, 1024, 0, s>>>(hist1, nb1, nbins1, total, seg, ntiles, wbase);
    // every tile writes all its bins' counts, so hist2 needs no clearing: past the last
    // tile's counts the scan reads stale words that no later read depends on
    k_sort_p2_count<<<g2, SORT_T, 0, s>>>(tmp, seg, ntiles, wbase, nbins1, hb, hist2);
    scan_excl(hist2, hist2, h2n, bsum, total + 1, s);
    k_sort_p2_scatter<<<g2, SORT_T, 0, s>>>(tmp, seg, ntiles, wbase, nbins1, hb, hist2, sorted);
    k_sort_offsets<<<blocks_for((uint64_t)p.nb + 1, 256), 256, 0, s>>>(seg, ntiles, wbase, hist2, p.nb, hb, offsets);
    BG_HIP(hipGetLastError());
  }
}

void msm_enqueue_batch(int group, uint32_t M, const uint32_t* const* d_points, const uint32_t* const* d_scalars,
                       const uint64_t* ns, int window_bits, hipStream_t s, MsmWork& w) {
  if (group == 1)
    enqueue_g1(M, d_points, d_scalars, ns, window_bits, s, w, false);
  else
    enqueue_g2(M, d_points, d_scalars, ns, window_bits, s, w, false);
}

void msm_enqueue_fixed(int group, uint32_t M, const uint32_t* const* d_tables, const uint32_t* const* d_scalars,
                       const uint64_t* ns, int c, hipStream_t s, MsmWork& w) {
  if (group == 1)
    enqueue_g1(M, d_tables, d_scalars, ns, c, s, w, true);
  else
    enqueue_g2(M, d_tables, d_scalars, ns, c, s, w, true);
}

void msm_build_fixed_table(int group, const uint32_t* d_points, uint64_t n, int c, uint32_t* d_out, hipStream_t s) {
  if (group == 1)
    fixed_table_g1(d_points, n, c, d_out, s);
  else
    fixed_table_g2(d_points, n, c, d_out, s);
}

void msm_enqueue(int group, const uint32_t* d_points, const uint32_t* d_scalars, uint64_t n, int window_bits,
                 hipStream_t s, MsmWork& w) {
  msm_enqueue_batch(group, 1, &d_points, &d_scalars, &n, window_bits, s, w);
}

template <class E>
static void host_finish(const MsmPlan& p, const uint32_t* hout, std::vector<bgh::Pt<E>>& res) {
  using Pt = bgh::Pt<E>;
  const int xw = 4 * bgh::words<E>();
  int log2M = 0;
  while ((1u << log2M) < p.m * p.K) log2M++;
  const uint32_t MW = p.M * (uint32_t)p.W;
  // Window w's level-2 outputs (P_h, Q_h), h < H, give sum_h P_h + mK * sum_h h Q_h.  The
  // H outputs are cut into T chunks of Lc (powers of two) summed in parallel: chunk j
  // (from a = j Lc) returns S = sum P_h, U = sum Q_h and V = sum (h - a) Q_h, and
  // sum_h h Q_h = sum_j V_j + Lc * sum_j j U_j.
  uint32_t T = 1;
  while (T < 16 && p.H / (2 * T) >= 16) T *= 2;
  const uint32_t Lc = p.H / T;
  int log2Lc = 0;
  while ((1u << log2Lc) < Lc) log2Lc++;
  struct Part {
    Pt S, U, V;
  };
  std::vector<Part> parts((size_t)MW * T);
  auto chunk = [&](int id) {
    const uint32_t w = (uint32_t)id / T, j = (uint32_t)id % T;
    const uint32_t* base = hout + (uint64_t)2 * p.H * xw * w;
    const int a = (int)(j * Lc);
    Pt S = Pt::inf(), run = Pt::inf(), V = Pt::inf();
    for (int h = a + (int)Lc - 1; h >= a; h--) {
      S = S.add(bgh::load_xyzz<E>(base + (uint64_t)xw * (2 * h)));
      if (h > 0) run = run.add(bgh::load_xyzz<E>(base + (uint64_t)xw * (2 * h + 1)));
      if (h > a) V = V.add(run);
    }
    parts[id] = {S, run, V};
  };
  if ((uint64_t)p.H * MW < 64)
    for (uint32_t i = 0; i < MW * T; i++) chunk((int)i);
  else
    host_parallel((int)(MW * T), chunk);
  std::vector<Pt> wsum(MW);
  for (uint32_t w = 0; w < MW; w++) {
    Pt S = Pt::inf(), Vs = Pt::inf(), run = Pt::inf(), jU = Pt::inf();
    for (int j = (int)T - 1; j >= 0; j--) {
      const Part& q = parts[(size_t)w * T + j];
      S = S.add(q.S);
      Vs = Vs.add(q.V);
      if (j > 0) {
        run = run.add(q.U);
        jU = jU.add(run);
      }
    }
    wsum[w] = S.add(Vs.add(jU.dbl_n(log2Lc)).dbl_n(log2M));
  }
  res.assign(p.M, Pt::inf());
  auto horner = [&](int i) {
    Pt acc = Pt::inf();
    for (int w = p.W - 1; w >= 0; w--) acc = acc.dbl_n(p.c).add(wsum[i * p.W + w]);
    res[i] = acc;
  };
  if (p.M == 1)
    horner(0);
  else
    host_parallel((int)p.M, horner);
}

// Waits for the flag of a recorded MSM's outputs (msm_enqueue under capture); the graph's
// stream is polled so that a failed launch ends the wait with its error.
static void wait_published(const MsmWork& w) {
  const uint32_t* flag = static_cast<const uint32_t*>(w.host_pub.ptr);
  auto last = std::chrono::steady_clock::now();
  while (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == 0) {
    __builtin_ia32_pause();
    // the stream is asked only every 200 us: the runtime's queries take its locks
    const auto now = std::chrono::steady_clock::now();
    if (now - last < std::chrono::microseconds(200)) continue;
    last = now;
    const hipError_t q = hipStreamQuery(w.pub_stream);
    if (q == hipSuccess) {
      BG_REQUIRE(__atomic_load_n(flag, __ATOMIC_ACQUIRE) != 0, BG_ERR_INTERNAL,
                 "msm: a recorded launch finished without publishing its outputs");
      break;
    }
    if (q != hipErrorNotReady) BG_HIP(q);
  }
}

void msm_finish(MsmWork& w, uint32_t* out_affine, uint32_t* xyzz_out) {
  const uint32_t* hout;
  if (w.published) {
    wait_published(w);
    hout = static_cast<const uint32_t*>(w.host_pub.ptr) + MsmWork::PUB_HEAD;
  } else {
    BG_HIP(hipEventSynchronize(w.done));
    hout = static_cast<const uint32_t*>(w.host_out.ptr);
  }
  if (w.group == 1) {
    std::vector<bgh::P1> r;
    host_finish<bgh::Fq>(w.plan, hout, r);
    for (uint32_t i = 0; i < w.plan.M; i++) {
      if (out_affine) bgh::store_affine(r[i], out_affine + 16 * i);
      if (xyzz_out) {
        uint32_t* o = xyzz_out + 32 * i;
        r[i].X.to_u32(o); r[i].Y.to_u32(o + 8); r[i].ZZ.to_u32(o + 16); r[i].ZZZ.to_u32(o + 24);
      }
    }
  } else {
    std::vector<bgh::P2> r;
    host_finish<bgh::Fq2>(w.plan, hout, r);
    for (uint32_t i = 0; i < w.plan.M; i++) {
      if (out_affine) bgh::store_affine(r[i], out_affine + 32 * i);
      if (xyzz_out) {
        uint32_t* o = xyzz_out + 64 * i;
        r[i].X.to_u32(o); r[i].Y.to_u32(o + 16); r[i].ZZ.to_u32(o + 32); r[i].ZZZ.to_u32(o + 48);
      }
    }
  }
}

}  // namespace bg

using namespace bg;

// One-shot MSM through the C ABI (host or device buffers, either encoding).
static int msm_api(int group, const uint32_t* points, const uint32_t* scalars, uint64_t n, uint32_t* out, int flags,
                   int window_bits) {
  return guard([&] {
    BG_REQUIRE(out && (n == 0 || (points && scalars)), BG_ERR_ARG, "bg_msm: null pointer");
    BG_REQUIRE(window_bits == 0 || (window_bits >= 2 && window_bits <= 22), BG_ERR_ARG,
               "bg_msm: window_bits %d out of range (0 or 2..22)", window_bits);
    bool words16 = (flags & 1) == BG_FMT_WORDS16;
    bool dev = (flags & BG_MEM_DEVICE) != 0;
    BG_REQUIRE(!(dev && words16), BG_ERR_ARG, "bg_msm: device buffers must be packed");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    const uint64_t ew = 8;                    // packed words per field element
    const uint64_t pw = (group == 1 ? 2 : 4) * ew;  // packed words per affine point
    const uint32_t* dp = points;
    const uint32_t* ds = scalars;
    if (!dev && n > 0) {
      uint32_t* bp = c.scratch[10].as<uint32_t>(n * pw);
      uint32_t* bs = c.scratch[11].as<uint32_t>(n * ew);
      if (words16) {
        uint32_t* rp = c.scratch[12].as<uint32_t>(2 * n * pw);
        uint32_t* rs = c.scratch[13].as<uint32_t>(2 * n * ew);
        BG_HIP(hipMemcpyAsync(rp, points, 2 * n * pw * 4, hipMemcpyHostToDevice, c.stream));
        BG_HIP(hipMemcpyAsync(rs, scalars, 2 * n * ew * 4, hipMemcpyHostToDevice, c.stream));
        launch_pack16(rp, bp, n * pw, c.stream);
        launch_pack16(rs, bs, n * ew, c.stream);
      } else {
        BG_HIP(hipMemcpyAsync(bp, points, n * pw * 4, hipMemcpyHostToDevice, c.stream));
        BG_HIP(hipMemcpyAsync(bs, scalars, n * ew * 4, hipMemcpyHostToDevice, c.stream));
      }
      dp = bp;
      ds = bs;
    }
    static MsmWork work[2];
    MsmWork& w = work[group - 1];
    msm_enqueue(group, dp, ds, n, window_bits, c.stream, w);
    uint32_t res[32];
    msm_finish(w, res);
    if (dev) {
      BG_HIP(hipMemcpy(out, res, pw * 4, hipMemcpyHostToDevice));
    } else if (words16) {
      for (uint64_t i = 0; i < pw; i++) {
        out[2 * i] = res[i] & 0xFFFFu;
        out[2 * i + 1] = res[i] >> 16;
      }
    } else {
      memcpy(out, res, pw * 4);
    }
  });
}

extern "C" int bg_msm_batch(int group, uint32_t count, const uint32_t* const* points, const uint32_t* const* scalars,
                            const uint64_t* ns, uint32_t* outs, int flags, int window_bits) {
  return guard([&] {
    BG_REQUIRE((group == 1 || group == 2) && count >= 1 && points && scalars && ns && outs, BG_ERR_ARG,
               "bg_msm_batch: bad argument");
    BG_REQUIRE(window_bits == 0 || (window_bits >= 2 && window_bits <= 22), BG_ERR_ARG,
               "bg_msm_batch: window_bits %d out of range", window_bits);
    bool words16 = (flags & 1) == BG_FMT_WORDS16;
    bool dev = (flags & BG_MEM_DEVICE) != 0;
    BG_REQUIRE(!words16, BG_ERR_ARG, "bg_msm_batch: inputs must be packed");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    const uint64_t pw = group == 1 ? 16 : 32;
    std::vector<const uint32_t*> dp(count), ds(count);
    std::vector<uint64_t> n(ns, ns + count);
    if (dev) {
      for (uint32_t i = 0; i < count; i++) {
        dp[i] = points[i];
        ds[i] = scalars[i];
      }
    } else {
      uint64_t tot = 0;
      for (uint32_t i = 0; i < count; i++) tot += n[i];
      uint32_t* bp = c.scratch[10].as<uint32_t>(std::max<uint64_t>(1, tot * pw));
      uint32_t* bs = c.scratch[11].as<uint32_t>(std::max<uint64_t>(1, tot * 8));
      uint64_t off = 0;
      for (uint32_t i = 0; i < count; i++) {
        dp[i] = bp + off * pw;
        ds[i] = bs + off * 8;
        if (n[i]) {
          BG_HIP(hipMemcpyAsync(bp + off * pw, points[i], n[i] * pw * 4, hipMemcpyHostToDevice, c.stream));
          BG_HIP(hipMemcpyAsync(bs + off * 8, scalars[i], n[i] * 32, hipMemcpyHostToDevice, c.stream));
        }
        off += n[i];
      }
    }
    // Each MSM's own planned window width; MSMs sharing a width run as one batched
    // launch sequence (sizes 2^12 .. 2^20 in one call span four or five widths), the
    // width classes on separate streams.
    std::map<int, std::vector<uint32_t>> classes;
    for (uint32_t i = 0; i < count; i++) {
      int cw = window_bits > 0 ? window_bits : (n[i] ? msm_plan(n[i], group, 0).c : 0);
      classes[cw].push_back(i);
    }
    static std::map<std::pair<int, int>, std::unique_ptr<MsmWork>> works;
    hipStream_t streams[4] = {c.stream, c.aux[0], c.aux[1], c.aux[2]};
    struct Job {
      MsmWork* w;
      std::vector<uint32_t> idx;
    };
    std::vector<Job> jobs;
    std::vector<uint32_t> res(pw * count, 0);
    {
      ScopedKTime kt(group == 1 ? "msm_g1_batch" : "msm_g2_batch", c.stream);
      hipEvent_t start;
      BG_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
      BG_HIP(hipEventRecord(start, c.stream));  // host-to-device copies above are on c.stream
      int k = 0;
      for (auto& cl : classes) {
        if (cl.first == 0) continue;  // empty MSMs: result stays infinity
        auto& slot = works[{group, cl.first}];
        if (!slot) slot = std::make_unique<MsmWork>();
        std::vector<const uint32_t*> cp, cs;
        std::vector<uint64_t> cn;
        for (uint32_t i : cl.second) {
          cp.push_back(dp[i]);
          cs.push_back(ds[i]);
          cn.push_back(n[i]);
        }
        hipStream_t s = streams[k++ % 4];
        BG_HIP(hipStreamWaitEvent(s, start, 0));
        msm_enqueue_batch(group, (uint32_t)cp.size(), cp.data(), cs.data(), cn.data(), cl.first, s, *slot);
        jobs.push_back({slot.get(), cl.second});
      }
      for (auto& j : jobs) {
        std::vector<uint32_t> r(pw * j.idx.size());
        msm_finish(*j.w, r.data());
        for (size_t q = 0; q < j.idx.size(); q++) memcpy(&res[pw * j.idx[q]], &r[pw * q], pw * 4);
      }
      for (hipStream_t s : streams) {
        if (s == c.stream) continue;
        hipEvent_t e;
        BG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        BG_HIP(hipEventRecord(e, s));
        BG_HIP(hipStreamWaitEvent(c.stream, e, 0));
        (void)hipEventDestroy(e);
      }
      (void)hipEventDestroy(start);
    }
    if (dev)
      BG_HIP(hipMemcpy(outs, res.data(), res.size() * 4, hipMemcpyHostToDevice));
    else
      memcpy(outs, res.data(), res.size() * 4);
  });
}

extern "C" int bg_msm_plan_info(uint64_t n, int group, int window_bits, uint64_t* info8) {
  return guard([&] {
    BG_REQUIRE(info8 && (group == 1 || group == 2), BG_ERR_ARG, "bg_msm_plan_info: bad argument");
    MsmPlan p = msm_plan(n, group, window_bits);
    info8[0] = (uint64_t)p.c;
    info8[1] = (uint64_t)p.W;
    info8[2] = p.B;
    info8[3] = p.entries;
    info8[4] = p.L;
    info8[5] = p.acc_threads;
    info8[6] = (uint64_t)p.W * p.G;  // level-1 groups
    info8[7] = (uint64_t)p.W * p.H;  // level-2 outputs (host tail points / 2)
  });
}

extern "C" int bg_msm_g1(const uint32_t* points, const uint32_t* scalars, uint64_t n, uint32_t* out, int flags,
                         int window_bits) {
  return msm_api(1, points, scalars, n, out, flags, window_bits);
}

extern "C" int bg_msm_g2(const uint32_t* points, const uint32_t* scalars, uint64_t n, uint32_t* out, int flags,
                         int window_bits) {
  return msm_api(2, points, scalars, n, out, flags, window_bits);
}
