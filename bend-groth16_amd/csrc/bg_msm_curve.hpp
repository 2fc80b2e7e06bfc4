// Per-group MSM kernels (accumulation, partial merge, two-level bucket reduction) and
// the device pipeline driver, instantiated once per group in bg_msm_g1.hip /
// bg_msm_g2.hip so the two large instantiations compile in parallel.
#pragma once
#include "bg_msm.hpp"
#include "bg_curve.hpp"
#include "bg_curve_fd.hpp"

namespace bg {

inline unsigned blocks_of(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

static constexpr uint32_t NOKEY_C = 0xFFFFFFFFu;

// Each thread sums L consecutive sorted entries.  Runs wholly inside the slice are
// complete buckets; a run touching a slice edge whose key continues into the neighbour
// is a partial: slot 2t+1 if it continues into the next slice, else slot 2t.
// The hot loop uses only the generic mixed-add formula (every run starts from its first
// point and no addend is infinity); a run that meets x(acc) = x(point) is flagged, its
// bucket key is listed once in fix[1 + nb ..] (count in fix[0], per-bucket flag in
// fix[1 ..]) and k_fixup recomputes that bucket exactly.
template <class F>
__device__ __forceinline__ void flag_bucket(uint32_t key, uint32_t nb, uint32_t* __restrict__ fix) {
  if (atomicOr(&fix[1 + key], 1u) == 0u) {
    uint32_t idx = atomicAdd(&fix[0], 1u);
    fix[1 + nb + idx] = key;
  }
}

// Point base of an entry: the batch's MSM is key / kspan.
__device__ __forceinline__ const uint32_t* entry_points(const uint64_t* __restrict__ ptrs, uint32_t key,
                                                        uint32_t kspan) {
  return reinterpret_cast<const uint32_t*>(ptrs[key / kspan]);
}

// The running sum of a run lives in the digit domain (bg_curve_fd.hpp); points are read
// and buckets written in the word domain.
template <class F>
struct AccOf;
template <>
struct AccOf<Field<FqP>> {
  using type = fd::G1Acc<fd::FqDig>;
};
template <>
struct AccOf<Fq2> {
  using type = fd::G2Acc<fd::FqDig>;
};

constexpr bool same_digits(const uint32_t* a, const uint32_t* b) {
  for (int i = 0; i < 9; i++)
    if (a[i] != b[i]) return false;
  return true;
}
static_assert(same_digits(FqP::p29, fd::FqDig::p29) && FqP::n0_29 == fd::FqDig::n0_29 &&
                  FqP::p2[7] == fd::FqDig::top2p,
              "digit-domain constants of Fq");

// W words (a multiple of 4, 16-byte aligned) between global memory and registers (loads
// through a global-address-space pointer: bases reached through the pointer table are
// generic pointers, and flat loads would make every later wait a wait for all loads)
template <int W>
__device__ __forceinline__ void load_vec_global(const uint32_t* p, uint32_t* w) {
  typedef uint32_t U4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(1))) U4 GlobalU4;
  const GlobalU4* q = (const GlobalU4*)p;
#pragma unroll
  for (int i = 0; i < W / 4; i++) {
    const U4 v = q[i];
    w[4 * i] = v.x;
    w[4 * i + 1] = v.y;
    w[4 * i + 2] = v.z;
    w[4 * i + 3] = v.w;
  }
}
template <int W>
__device__ __forceinline__ void store_vec(uint32_t* __restrict__ p, const uint32_t* w) {
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int i = 0; i < W / 4; i++) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// A closed run's digit-form sum waits in a per-thread staging record (in LDS) until a
// wave-wide write-out converts it to the word domain (the conversion then runs once for
// every lane with a waiting run, instead of once per step at which any lane of the wave
// closes one).  Record of thread t, coordinate k (the k-th digit vector of A::Pt): digits
// 0..3 and 4..7 as 16-byte chunks 2k and 2k + 1 of column t, digit 8 in a word array after
// the chunks, then the run's key and destination kind (a word pair), so a wave's lanes
// touch contiguous spans and nothing of the waiting run is held in registers.
template <class A>
struct StageOf {
  static constexpr uint32_t coords = sizeof(typename A::Pt) / sizeof(fd::Fd);
  static constexpr uint64_t bytes_per_thread = 36ull * coords + 8;
  static_assert(sizeof(typename A::Pt) == coords * sizeof(fd::Fd), "digit-form point layout");
};
template <class A>
__device__ __forceinline__ void stage_put(uint32_t* __restrict__ st, uint32_t nthreads, uint32_t t,
                                          const typename A::Pt& p, uint32_t key, uint32_t kind) {
  constexpr uint32_t NC = StageOf<A>::coords;
  fd::Fd c[NC];
  __builtin_memcpy(c, &p, sizeof p);
  uint4* q = reinterpret_cast<uint4*>(st);
  uint32_t* d8 = st + 8ull * NC * nthreads;
#pragma unroll
  for (uint32_t k = 0; k < NC; k++) {
    q[(uint64_t)(2 * k) * nthreads + t] = make_uint4(c[k].d[0], c[k].d[1], c[k].d[2], c[k].d[3]);
    q[(uint64_t)(2 * k + 1) * nthreads + t] = make_uint4(c[k].d[4], c[k].d[5], c[k].d[6], c[k].d[7]);
    d8[(uint64_t)k * nthreads + t] = c[k].d[8];
  }
  reinterpret_cast<uint2*>(st + 9ull * NC * nthreads)[t] = make_uint2(key, kind);
}
template <class A>
__device__ __forceinline__ uint2 stage_meta(const uint32_t* __restrict__ st, uint32_t nthreads, uint32_t t) {
  return reinterpret_cast<const uint2*>(st + 9ull * StageOf<A>::coords * nthreads)[t];
}
// thread t's staged run, converted coordinate by coordinate, to the XYZZ words at dst
template <class A>
__device__ __forceinline__ void stage_write(const uint32_t* __restrict__ st, uint32_t nthreads, uint32_t t,
                                            uint32_t* __restrict__ dst) {
  constexpr uint32_t NC = StageOf<A>::coords;
  const uint4* q = reinterpret_cast<const uint4*>(st);
  const uint32_t* d8 = st + 8ull * NC * nthreads;
#pragma unroll
  for (uint32_t k = 0; k < NC; k++) {
    const uint4 a = q[(uint64_t)(2 * k) * nthreads + t];
    const uint4 b = q[(uint64_t)(2 * k + 1) * nthreads + t];
    const fd::Fd f = {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, d8[(uint64_t)k * nthreads + t]}};
    uint32_t w[8];
    A::store_fd(w, f);
    store_vec<8>(dst + 8 * k, w);
    // one coordinate's registers at a time (the write-out runs beside the live accumulator)
    __builtin_amdgcn_sched_barrier(0);
  }
}

// G1 accumulation keeps three waves per SIMD (at most 168 VGPRs), G2 two
template <class F>
struct AccWaves {
  static constexpr int value = sizeof(typename Curve<F>::P) > 128 ? 2 : 3;
};

template <class F>
__global__ __launch_bounds__(256, AccWaves<F>::value) void k_accumulate(const uint64_t* __restrict__ ptrs, uint32_t kspan,
                                                    const uint2* __restrict__ sorted,
                                                    const uint32_t* __restrict__ total_ptr, uint32_t L,
                                                    uint32_t nthreads, uint32_t* __restrict__ buckets,
                                                    uint32_t* __restrict__ pkeys, uint32_t* __restrict__ pflags,
                                                    uint32_t* __restrict__ ppts, uint32_t nb,
                                                    uint32_t* __restrict__ fix) {
  using C = Curve<F>;
  using A = typename AccOf<F>::type;
  static_assert(A::aff_words == C::aff_words && A::xyzz_words == C::xyzz_words, "point layout");
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nthreads) return;
  pkeys[2 * t] = NOKEY_C;
  pkeys[2 * t + 1] = NOKEY_C;
  // entry indices fit in 32 bits (the planner keeps a batch below 2^32 - 1 entries)
  const uint32_t T = *total_ptr;
  const uint64_t start64 = (uint64_t)t * L;
  if (start64 >= T) return;
  const uint32_t start = (uint32_t)start64;
  const uint32_t end = (uint32_t)min(start64 + L, (uint64_t)T);
  uint32_t prevkey = start > 0 ? sorted[start - 1].x : NOKEY_C;
  uint32_t nextkey = end < T ? sorted[end].x : NOKEY_C;
  // One point load, one run start, one addition and one store site in the loop: the
  // conversions at a run's start (two reductions into the digit domain) and at its write
  // (four packings into words) are long, so each is emitted once (a smaller kernel for the
  // instruction cache).  Step e == end closes the last run.  Loads run ahead of the
  // additions: the entry two steps ahead and the next point's words are requested before
  // the current addition, at addresses kept in bounds on every step (the last entry's
  // index, point 0 of the current MSM), so the loads are unconditional and every wait is
  // for the oldest loads only (G1 within three waves per SIMD, G2 within two).
  uint2 ent = sorted[start];                                                 // entry e
  uint2 ent1 = start + 1 < end ? sorted[start + 1] : make_uint2(NOKEY_C, 0u);  // entry e + 1
  const uint32_t* points = entry_points(ptrs, ent.x, kspan);  // base of entry e's MSM
  uint32_t kend = (ent.x / kspan + 1) * kspan;               // end of that MSM's keys
  uint32_t w[A::aff_words];                                   // entry e's point
  load_vec_global<A::aff_words>(points + (uint64_t)A::aff_words * (ent.y & 0x7FFFFFFFu), w);
  uint32_t cur = NOKEY_C;
  bool first = true;
  uint32_t bad = 0;
  typename A::Pt acc;
  // a run waits in the staging record (with its key and where it goes -- SK_BUCKET: a
  // complete bucket; SK_HEAD: head slot 2t, flag 1; SK_TAIL | flag: tail slot 2t + 1)
  constexpr uint32_t SK_BUCKET = 1, SK_HEAD = 2, SK_TAIL = 4;
  bool staged = false;
  // the block's staging records live in LDS (column threadIdx.x of 256): 38 KB for G1, 74 KB
  // for G2 (two G2 blocks per CU, as its registers allow)
  __shared__ uint32_t stage_lds[(uint32_t)(StageOf<A>::bytes_per_thread / 4) * 256];
  const uint32_t lt = threadIdx.x;
  // Step e == end closes the last run; step end + 1 writes out what is still staged.
  for (uint32_t e = start;; e++) {
    const bool drain = e > end;
    const bool last = e == end;
    const bool closes = !drain && e > start && ent.x != cur;
    // a lane about to stage a run while one waits: the wave writes its staged runs out
    if (__any((closes || drain) && staged)) {
      if (staged) {
        const uint2 meta = stage_meta<A>(stage_lds, 256, lt);  // key, kind
        uint32_t* dst;
        if (meta.y == SK_BUCKET) {
          dst = buckets + (uint64_t)A::xyzz_words * meta.x;
        } else {
          const uint32_t slot = meta.y == SK_HEAD ? 2 * t : 2 * t + 1;
          pkeys[slot] = meta.x;
          pflags[slot] = meta.y == SK_HEAD ? 1u : meta.y & 1u;
          dst = ppts + (uint64_t)A::xyzz_words * slot;
        }
        stage_write<A>(stage_lds, 256, lt, dst);
        staged = false;
      }
    }
    if (drain) break;
    if (closes) {
      // the run of key cur ends: a complete bucket, or a partial for the merge
      const bool cp = first && prevkey == cur;
      const bool cn = last && nextkey == cur;
      if (bad) flag_bucket<F>(cur, nb, fix);
      const uint32_t kind = !cp && !cn ? SK_BUCKET : (cn ? SK_TAIL | (cp ? 1u : 0u) : SK_HEAD);
      stage_put<A>(stage_lds, 256, lt, acc, cur, kind);
      staged = true;
      first = false;
    }
    if (last) continue;
    // entry e + 1's MSM base (rarely a new one) before the entry two ahead is requested
    // (keys ascend along the slice, so a key at or past kend belongs to a later MSM)
    if (e + 1 < end && ent1.x >= kend) {
      points = entry_points(ptrs, ent1.x, kspan);
      kend = (ent1.x / kspan + 1) * kspan;
    }
    uint2 ent2 = sorted[e + 2 < end ? e + 2 : end - 1];
    if (e + 2 >= end) ent2 = make_uint2(NOKEY_C, 0u);
    const typename A::Raw pt = A::load(w);
    load_vec_global<A::aff_words>(points + (uint64_t)A::aff_words * (ent1.y & 0x7FFFFFFFu), w);
    const bool neg = (ent.y >> 31) != 0;
    if (ent.x != cur) {
      cur = ent.x;
      bad = 0;
      acc = A::start(pt, neg);
    } else {
      acc = A::madd(acc, pt, neg, bad);
    }
    ent = ent1;
    ent1 = ent2;
  }
}

// Exact recomputation of the flagged buckets (doubling / cancellation inside a bucket:
// repeated or opposite bases), with the special-case-complete mixed add.  It also clears
// the reductions' exception counts rl1[0] and rl2[0] (the reductions run next on the same
// stream), which saves two launches per MSM.
template <class F>
__global__ __launch_bounds__(64) void k_fixup(const uint64_t* __restrict__ ptrs, uint32_t kspan,
                                              const uint2* __restrict__ sorted,
                                              const uint32_t* __restrict__ offsets, uint32_t nb,
                                              const uint32_t* __restrict__ fix, uint32_t* __restrict__ buckets,
                                              uint32_t* __restrict__ rl1, uint32_t* __restrict__ rl2) {
  using C = Curve<F>;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    rl1[0] = 0;
    rl2[0] = 0;
  }
  uint32_t count = fix[0];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    uint32_t key = fix[1 + nb + i];
    const uint32_t* points = entry_points(ptrs, key, kspan);
    typename C::P acc = C::infinity();
    for (uint32_t e = offsets[key]; e < offsets[key + 1]; e++) {
      uint32_t v = sorted[e].y;
      typename C::A pt = C::load_aff(points + (uint64_t)C::aff_words * (v & 0x7FFFFFFFu));
      if (v >> 31) pt = C::aff_neg(pt);
      acc = C::madd(acc, pt);
    }
    C::store(buckets + (uint64_t)C::xyzz_words * key, acc);
  }
}

// Partial runs, one item per accumulation thread (and, one level up, per merge block):
//   head slot 2t   : run that continues from item t-1 and ends in item t (NOKEY if none)
//   tail slot 2t+1 : run that continues into item t+1; flag 1 when it also continues from
//                    item t-1 (the run covers the whole item), else it starts in item t.
// A bucket spanning items t0..t1 is tail(t0) + tail(t0+1 .. t1-1) + head(t1).
// One merge level: a block of MERGE_T items sums each run of tail slots with a
// segmented tree in LDS (log2(MERGE_T) rounds of at most one add), adds the closing head
// and writes the finished bucket; a run entering from the previous block or leaving into
// the next becomes the block's head / tail slot of the next level, in the same format.
// Levels repeat until one block covers all items, so a bucket spread over m items costs
// O(log m) adds of depth instead of a serial walk (scalars that share digits -- small or
// repeated witness values, a short top window -- make such buckets).
// MERGE_T items per block: 256 for G1 (32 KB of LDS points), 128 for G2 (32 KB).
template <class F>
struct MergeT {
  static constexpr uint32_t value = sizeof(typename Curve<F>::P) > 128 ? 128u : 256u;
};

template <class F, uint32_t MERGE_T = MergeT<F>::value>
__global__ __launch_bounds__(MERGE_T) void k_merge_level(const uint32_t* __restrict__ ikeys,
                                                     const uint32_t* __restrict__ iflags,
                                                     const uint32_t* __restrict__ ipts, uint32_t nitems,
                                                     uint32_t* __restrict__ buckets, uint32_t* __restrict__ okeys,
                                                     uint32_t* __restrict__ oflags, uint32_t* __restrict__ opts) {
  using C = Curve<F>;
  using P = typename C::P;
  extern __shared__ __align__(16) unsigned char merge_smem[];
  P* shV = reinterpret_cast<P*>(merge_smem);
  __shared__ uint32_t shK[MERGE_T], shF[MERGE_T];
  __shared__ int shS[MERGE_T], shE[MERGE_T];
  const uint32_t i = threadIdx.x;
  const uint32_t b = blockIdx.x;
  const uint64_t u = (uint64_t)b * MERGE_T + i;
  const uint64_t xw = C::xyzz_words;
  uint32_t tk = NOKEY_C, tf = 0;
  if (u < nitems) {
    tk = ikeys[2 * u + 1];
    tf = tk != NOKEY_C ? iflags[2 * u + 1] : 0u;
  }
  shK[i] = tk;
  shF[i] = tf;
  P v;
  if (tk != NOKEY_C) {
    v = C::load(ipts + xw * (2 * u + 1));
    shV[i] = v;
  }
  __syncthreads();
  // segment bounds among tail slots: start = nearest start marker at or below i,
  // end = nearest end marker at or above i
  bool used = tk != NOKEY_C;
  int sm = (used && (tf == 0 || i == 0)) ? (int)i : -1;
  bool cont_next = i + 1 < MERGE_T && shK[i + 1] != NOKEY_C && shF[i + 1] != 0;
  int em = (used && !cont_next) ? (int)i : (int)MERGE_T;
  shS[i] = sm;
  shE[i] = em;
  __syncthreads();
  for (uint32_t d = 1; d < MERGE_T; d <<= 1) {
    int a = i >= d ? shS[i - d] : -1;
    int c = i + d < MERGE_T ? shE[i + d] : (int)MERGE_T;
    __syncthreads();
    sm = max(sm, a);
    em = min(em, c);
    shS[i] = sm;
    shE[i] = em;
    __syncthreads();
  }
  const int st = sm, en = em;
  const uint32_t rel = used ? i - (uint32_t)st : 1u;
  const uint32_t len = used ? (uint32_t)(en - st + 1) : 0u;
  for (uint32_t d = 1; d < MERGE_T; d <<= 1) {
    if (used && (rel & (2 * d - 1)) == 0 && rel + d < len) {
      v = C::add(v, shV[i + d]);
      shV[i] = v;
    }
    __syncthreads();
  }
  if (i == 0) {
    // head slot of this block at the next level: a run from the previous block that
    // ends here with no tail slot in this block passes through unchanged
    uint32_t hk = u < nitems ? ikeys[2 * u] : NOKEY_C;
    okeys[2 * b] = hk;
    if (hk != NOKEY_C) C::store(opts + xw * (2 * b), C::load(ipts + xw * (2 * u)));
  }
  if (i == MERGE_T - 1 && !used) okeys[2 * b + 1] = NOKEY_C;
  if (used && rel == 0) {
    const bool from_prev = i == 0 && tf != 0;
    if ((uint32_t)en + 1 < MERGE_T) {
      // the run closes at item u_end + 1 inside this block
      uint64_t uh = (uint64_t)b * MERGE_T + (uint32_t)en + 1;
      v = C::add(v, C::load(ipts + xw * (2 * uh)));
      if (from_prev) {
        okeys[2 * b] = tk;
        C::store(opts + xw * (2 * b), v);
      } else {
        C::store(buckets + xw * tk, v);
      }
    } else {
      okeys[2 * b + 1] = tk;
      oflags[2 * b + 1] = from_prev ? 1u : 0u;
      C::store(opts + xw * (2 * b + 1), v);
    }
  }
}

// k_merge_level with a quad per item (4 MERGE_T threads): the same segments, trees and
// outputs, each addition spread over the quad's four lanes (Quad); the lanes of a quad
// hold the same values, its lead lane writes.  For the levels with few blocks (latency-
// bound); a level that fills the device keeps k_merge_level.
static constexpr uint32_t MERGE_QT = 128;  // items per block of k_merge_levelq (512 threads)

template <class F, uint32_t MERGE_T = MERGE_QT>
__global__ __launch_bounds__(4 * MERGE_T) void k_merge_levelq(const uint32_t* __restrict__ ikeys,
                                                          const uint32_t* __restrict__ iflags,
                                                          const uint32_t* __restrict__ ipts, uint32_t nitems,
                                                          uint32_t* __restrict__ buckets, uint32_t* __restrict__ okeys,
                                                          uint32_t* __restrict__ oflags, uint32_t* __restrict__ opts) {
  using C = Curve<F>;
  using Q = Quad<F>;
  using P = typename C::P;
  extern __shared__ __align__(16) unsigned char merge_smem[];
  P* shV = reinterpret_cast<P*>(merge_smem);
  __shared__ uint32_t shK[MERGE_T], shF[MERGE_T];
  __shared__ int shS[MERGE_T], shE[MERGE_T];
  const uint32_t i = threadIdx.x >> 2;  // this quad's item
  const bool lead = (threadIdx.x & 3u) == 0;
  const uint32_t b = blockIdx.x;
  const uint64_t u = (uint64_t)b * MERGE_T + i;
  const uint64_t xw = C::xyzz_words;
  uint32_t tk = NOKEY_C, tf = 0;
  if (u < nitems) {
    tk = ikeys[2 * u + 1];
    tf = tk != NOKEY_C ? iflags[2 * u + 1] : 0u;
  }
  if (lead) {
    shK[i] = tk;
    shF[i] = tf;
  }
  P v;
  if (tk != NOKEY_C) {
    v = C::load(ipts + xw * (2 * u + 1));
    if (lead) shV[i] = v;
  }
  __syncthreads();
  bool used = tk != NOKEY_C;
  int sm = (used && (tf == 0 || i == 0)) ? (int)i : -1;
  bool cont_next = i + 1 < MERGE_T && shK[i + 1] != NOKEY_C && shF[i + 1] != 0;
  int em = (used && !cont_next) ? (int)i : (int)MERGE_T;
  if (lead) {
    shS[i] = sm;
    shE[i] = em;
  }
  __syncthreads();
  for (uint32_t d = 1; d < MERGE_T; d <<= 1) {
    int a = i >= d ? shS[i - d] : -1;
    int c = i + d < MERGE_T ? shE[i + d] : (int)MERGE_T;
    __syncthreads();
    sm = max(sm, a);
    em = min(em, c);
    if (lead) {
      shS[i] = sm;
      shE[i] = em;
    }
    __syncthreads();
  }
  const int st = sm, en = em;
  const uint32_t rel = used ? i - (uint32_t)st : 1u;
  const uint32_t len = used ? (uint32_t)(en - st + 1) : 0u;
  for (uint32_t d = 1; d < MERGE_T; d <<= 1) {
    if (used && (rel & (2 * d - 1)) == 0 && rel + d < len) {
      v = Q::add(v, shV[i + d]);
      if (lead) shV[i] = v;
    }
    __syncthreads();
  }
  if (i == 0 && lead) {
    uint32_t hk = u < nitems ? ikeys[2 * u] : NOKEY_C;
    okeys[2 * b] = hk;
    if (hk != NOKEY_C) C::store(opts + xw * (2 * b), C::load(ipts + xw * (2 * u)));
  }
  if (i == MERGE_T - 1 && lead && !used) okeys[2 * b + 1] = NOKEY_C;
  if (used && rel == 0) {
    const bool from_prev = i == 0 && tf != 0;
    if ((uint32_t)en + 1 < MERGE_T) {
      uint64_t uh = (uint64_t)b * MERGE_T + (uint32_t)en + 1;
      v = Q::add(v, C::load(ipts + xw * (2 * uh)));
      if (lead) {
        if (from_prev) {
          okeys[2 * b] = tk;
          C::store(opts + xw * (2 * b), v);
        } else {
          C::store(buckets + xw * tk, v);
        }
      }
    } else if (lead) {
      okeys[2 * b + 1] = tk;
      oflags[2 * b + 1] = from_prev ? 1u : 0u;
      C::store(opts + xw * (2 * b + 1), v);
    }
  }
}

// Level 1: group g sums the m = 2^logm buckets i = j*m + l (l < m) of window w:
//   S = sum_l b_l,  T = sum_l l * b_l.
// A block of four waves takes 64 groups, one role per wave (roles are wave-uniform, so
// no wave runs two of them): wave q < logm sums M_q = the b_l with bit q of l set, wave
// logm sums N = the b_l with l < m/2 (partial sums through the scratch mq); then wave 0
// forms T = sum_q 2^q M_q (Horner) and wave 1 forms S = M_{logm-1} + N.  Depth m/2 +
// 2(logm - 1) additions instead of a 2m-long running sum.
// CAREFUL = false uses the generic add with infinity selected around it and lists the
// group in rl[1 ..] (count rl[0]) when two partial sums share an x coordinate;
// CAREFUL = true recomputes only the listed groups with the exact add.
template <class F, bool CAREFUL>
__global__ __launch_bounds__(256) void k_reduce1(const uint32_t* __restrict__ buckets,
                                                 const uint32_t* __restrict__ offsets, uint32_t B, uint32_t logm,
                                                 uint32_t ngroups, uint32_t* __restrict__ lvlS,
                                                 uint32_t* __restrict__ lvlT, uint32_t* mq,
                                                 uint32_t* __restrict__ rl) {
  using C = Curve<F>;
  using P = typename C::P;
  __shared__ uint32_t shbad[64];
  const uint64_t xw = C::xyzz_words;
  const uint32_t m = 1u << logm, G = B >> logm;
  const uint32_t q = threadIdx.x >> 6, gl = threadIdx.x & 63;  // role = wave
  const uint32_t count = CAREFUL ? rl[0] : ngroups;
  const uint64_t mstride = (uint64_t)gridDim.x * 64;  // scratch slots per role
  for (uint32_t g0 = blockIdx.x * 64; g0 < count; g0 += gridDim.x * 64) {
    const uint32_t idx = g0 + gl;
    const bool valid = idx < count;
    const uint32_t g = valid ? (CAREFUL ? rl[1 + idx] : idx) : 0;
    const uint32_t w = g / G, j = g % G;
    const uint64_t base = (uint64_t)w * B + (uint64_t)j * m;
    const uint64_t slot = (uint64_t)blockIdx.x * 64 + gl;
    uint32_t bad = 0;
    if (q == 0) shbad[gl] = 0;
    if (valid && q <= logm) {
      P acc = C::infinity();
      for (uint32_t l = 0; l < m; l++) {
        const bool take = q < logm ? ((l >> q) & 1u) != 0 : l < m / 2;
        if (!take) continue;
        const uint64_t key = base + l;
        if (offsets[key + 1] > offsets[key]) {
          P bk = C::load(buckets + xw * key);
          acc = CAREFUL ? C::add(acc, bk) : C::add_lite(acc, bk, bad);
        }
      }
      C::store(mq + xw * (q * mstride + slot), acc);
    }
    __syncthreads();
    if (valid && q == 0) {
      P T = C::load(mq + xw * ((logm - 1) * mstride + slot));
      for (int qq = (int)logm - 2; qq >= 0; qq--) {
        T = C::dbl(T);
        P mqq = C::load(mq + xw * (qq * mstride + slot));
        T = CAREFUL ? C::add(T, mqq) : C::add_lite(T, mqq, bad);
      }
      C::store(lvlT + xw * g, T);
    } else if (valid && q == 1) {
      P hi = C::load(mq + xw * ((logm - 1) * mstride + slot));
      P lo = C::load(mq + xw * (logm * mstride + slot));
      C::store(lvlS + xw * g, CAREFUL ? C::add(hi, lo) : C::add_lite(hi, lo, bad));
    }
    if (bad) shbad[gl] = 1;
    __syncthreads();
    if (!CAREFUL && valid && q == 0 && shbad[gl]) rl[1 + atomicAdd(&rl[0], 1u)] = g;
    __syncthreads();
  }
}

// k_reduce1 (CAREFUL = false) with a quad per group: the same roles (one per wave), 16
// groups per block, each addition and doubling spread over the quad's four lanes (Quad).
template <class F>
__global__ __launch_bounds__(256) void k_reduce1q(const uint32_t* __restrict__ buckets,
                                                  const uint32_t* __restrict__ offsets, uint32_t B, uint32_t logm,
                                                  uint32_t ngroups, uint32_t* __restrict__ lvlS,
                                                  uint32_t* __restrict__ lvlT, uint32_t* mq, uint32_t* __restrict__ rl) {
  using C = Curve<F>;
  using Q = Quad<F>;
  using P = typename C::P;
  __shared__ uint32_t shbad[16];
  const uint64_t xw = C::xyzz_words;
  const uint32_t m = 1u << logm, G = B >> logm;
  const uint32_t q = threadIdx.x >> 6, gl = (threadIdx.x & 63) >> 2;  // role = wave, group = quad
  const bool lead = (threadIdx.x & 3u) == 0;
  const uint64_t mstride = (uint64_t)gridDim.x * 16;  // scratch slots per role
  for (uint32_t g0 = blockIdx.x * 16; g0 < ngroups; g0 += gridDim.x * 16) {
    const uint32_t g = g0 + gl;
    const bool valid = g < ngroups;
    const uint32_t w = valid ? g / G : 0, j = valid ? g % G : 0;
    const uint64_t base = (uint64_t)w * B + (uint64_t)j * m;
    const uint64_t slot = (uint64_t)blockIdx.x * 16 + gl;
    uint32_t bad = 0;
    if (q == 0 && lead) shbad[gl] = 0;
    if (valid && q <= logm) {
      P acc = C::infinity();
      for (uint32_t l = 0; l < m; l++) {
        const bool take = q < logm ? ((l >> q) & 1u) != 0 : l < m / 2;
        if (!take) continue;
        const uint64_t key = base + l;
        if (offsets[key + 1] > offsets[key]) acc = Q::add_lite(acc, C::load(buckets + xw * key), bad);
      }
      if (lead) C::store(mq + xw * (q * mstride + slot), acc);
    }
    __syncthreads();
    if (valid && q == 0) {
      P T = C::load(mq + xw * ((logm - 1) * mstride + slot));
      for (int qq = (int)logm - 2; qq >= 0; qq--) {
        T = Q::dbl(T);
        T = Q::add_lite(T, C::load(mq + xw * (qq * mstride + slot)), bad);
      }
      if (lead) C::store(lvlT + xw * g, T);
    } else if (valid && q == 1) {
      const P hi = C::load(mq + xw * ((logm - 1) * mstride + slot));
      const P lo = C::load(mq + xw * (logm * mstride + slot));
      const P sum = Q::add_lite(hi, lo, bad);
      if (lead) C::store(lvlS + xw * g, sum);
    }
    if (bad) shbad[gl] = 1;
    __syncthreads();
    if (valid && q == 0 && lead && shbad[gl]) rl[1 + atomicAdd(&rl[0], 1u)] = g;
    __syncthreads();
  }
}

// Level 1 with one lane per group (running sums from the top bucket, depth 2m): used
// for G2, whose additions hold a whole register file per lane, so four role waves per
// group would not fit the device in one round.
template <class F, bool CAREFUL>
__global__ __launch_bounds__(256) void k_reduce1_serial(const uint32_t* __restrict__ buckets,
                                                        const uint32_t* __restrict__ offsets, uint32_t B,
                                                        uint32_t m, uint32_t ngroups, uint32_t* __restrict__ lvlS,
                                                        uint32_t* __restrict__ lvlT, uint32_t* __restrict__ rl) {
  using C = Curve<F>;
  uint32_t count = CAREFUL ? rl[0] : ngroups;
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < count; idx += gridDim.x * blockDim.x) {
    uint32_t g = CAREFUL ? rl[1 + idx] : idx;
    uint32_t G = B / m;
    uint32_t w = g / G, j = g % G;
    uint64_t base = (uint64_t)w * B + (uint64_t)j * m;
    typename C::P R = C::infinity(), T = C::infinity();
    uint32_t bad = 0;
    for (int l = (int)m - 1; l >= 0; l--) {
      uint64_t key = base + l;
      if (offsets[key + 1] > offsets[key]) {
        typename C::P b = C::load(buckets + (uint64_t)C::xyzz_words * key);
        R = CAREFUL ? C::add(R, b) : C::add_lite(R, b, bad);
      }
      if (l > 0) T = CAREFUL ? C::add(T, R) : C::add_lite(T, R, bad);
    }
    C::store(lvlS + (uint64_t)C::xyzz_words * g, R);
    C::store(lvlT + (uint64_t)C::xyzz_words * g, T);
    if (!CAREFUL && bad) rl[1 + atomicAdd(&rl[0], 1u)] = g;
  }
}

// k_reduce1_serial (CAREFUL = false) with a quad per group (the running sums' additions
// spread over the quad's four lanes).
template <class F>
__global__ __launch_bounds__(256) void k_reduce1_serialq(const uint32_t* __restrict__ buckets,
                                                         const uint32_t* __restrict__ offsets, uint32_t B, uint32_t m,
                                                         uint32_t ngroups, uint32_t* __restrict__ lvlS,
                                                         uint32_t* __restrict__ lvlT, uint32_t* __restrict__ rl) {
  using C = Curve<F>;
  using Q = Quad<F>;
  const bool lead = (threadIdx.x & 3u) == 0;
  const uint32_t stride = (gridDim.x * blockDim.x) >> 2;
  for (uint32_t g = (blockIdx.x * blockDim.x + threadIdx.x) >> 2; g < ngroups; g += stride) {
    const uint32_t G = B / m;
    const uint32_t w = g / G, j = g % G;
    const uint64_t base = (uint64_t)w * B + (uint64_t)j * m;
    typename C::P R = C::infinity(), T = C::infinity();
    uint32_t bad = 0;
    for (int l = (int)m - 1; l >= 0; l--) {
      const uint64_t key = base + l;
      if (offsets[key + 1] > offsets[key])
        R = Q::add_lite(R, C::load(buckets + (uint64_t)C::xyzz_words * key), bad);
      if (l > 0) T = Q::add_lite(T, R, bad);
    }
    if (lead) {
      C::store(lvlS + (uint64_t)C::xyzz_words * g, R);
      C::store(lvlT + (uint64_t)C::xyzz_words * g, T);
      if (bad) rl[1 + atomicAdd(&rl[0], 1u)] = g;
    }
  }
}

// Level 2: a block of K = blockDim.x threads (power of two) takes K consecutive level-1
// groups of one window:
//   U = sum_k (T_k + S_k),  V = sum_k k*S_k (suffix scan + tree),  Z = sum_k S_k
//   out P = U + m*V, Q = Z
// CAREFUL as in k_reduce1, per block (list in rl2).
template <class F, bool CAREFUL>
__global__ __launch_bounds__(256) void k_reduce2(const uint32_t* __restrict__ lvlS, const uint32_t* __restrict__ lvlT,
                                                 int log2m, uint32_t* __restrict__ outPQ, uint32_t* __restrict__ rl2) {
  using C = Curve<F>;
  using P = typename C::P;
  extern __shared__ __align__(16) unsigned char red2_smem[];
  const uint32_t K = blockDim.x;
  P* shA = reinterpret_cast<P*>(red2_smem);
  P* shU = shA + K;
  __shared__ uint32_t blk_bad;
  uint32_t count = CAREFUL ? rl2[0] : gridDim.x;
  for (uint32_t bi = blockIdx.x; bi < count; bi += gridDim.x) {
    uint32_t blk = CAREFUL ? rl2[1 + bi] : bi;
    uint32_t k = threadIdx.x;
    uint32_t bad = 0;
    if (k == 0) blk_bad = 0;
    uint64_t g = (uint64_t)blk * K + k;
    P S = C::load(lvlS + (uint64_t)C::xyzz_words * g);
    P T = C::load(lvlT + (uint64_t)C::xyzz_words * g);
    P a = S;
    shA[k] = a;
    shU[k] = CAREFUL ? C::add(T, S) : C::add_lite(T, S, bad);
    __syncthreads();
    for (uint32_t d = 1; d < K; d <<= 1) {
      P x = (k + d < K) ? shA[k + d] : C::infinity();
      __syncthreads();
      a = CAREFUL ? C::add(a, x) : C::add_lite(a, x, bad);
      shA[k] = a;
      __syncthreads();
    }
    P Z = shA[0];
    __syncthreads();
    if (k == 0) shA[0] = C::infinity();
    __syncthreads();
    // the two tree sums side by side: the first half of the block reduces shA, the
    // second half shU (wave-uniform roles once K/2 >= 64)
    if (K >= 2) {
      const uint32_t half = K / 2, t = k % half;
      P* arr = k < half ? shA : shU;
      for (uint32_t s = half; s > 0; s >>= 1) {
        if (t < s) arr[t] = CAREFUL ? C::add(arr[t], arr[t + s]) : C::add_lite(arr[t], arr[t + s], bad);
        __syncthreads();
      }
    }
    if (k == 0) {
      P V = shA[0];
      for (int i = 0; i < log2m; i++) V = C::dbl(V);
      P out = CAREFUL ? C::add(shU[0], V) : C::add_lite(shU[0], V, bad);
      // canonical coordinates: the host finishes the sums
      C::store(outPQ + (uint64_t)C::xyzz_words * (2 * blk), C::canon(out));
      C::store(outPQ + (uint64_t)C::xyzz_words * (2 * blk + 1), C::canon(Z));
    }
    if (!CAREFUL && bad) atomicOr(&blk_bad, 1u);
    __syncthreads();
    if (!CAREFUL && k == 0 && blk_bad) rl2[1 + atomicAdd(&rl2[0], 1u)] = blk;
    __syncthreads();
  }
}

// k_reduce2 (CAREFUL = false) with a quad per level-1 group (blockDim.x = 4 K): the same
// scan, trees and outputs, each addition and doubling spread over the quad's four lanes
// (Quad), so the block's chain of dependent additions takes about half the time.
// Blocks of up to 512 threads for G1 (K = 128 groups; at 1024 a quad's G1 sum would
// spill past 128 VGPRs) and 256 for G2 (K = 64: a quad's G2 doubling holds ~210 VGPRs).
template <class F>
struct QuadThreads {
  static constexpr int value = sizeof(typename Curve<F>::P) > 128 ? 256 : 512;
};

template <class F>
__global__ __launch_bounds__(QuadThreads<F>::value) void k_reduce2q(const uint32_t* __restrict__ lvlS, const uint32_t* __restrict__ lvlT,
                                                   int log2m, uint32_t* __restrict__ outPQ, uint32_t* __restrict__ rl2) {
  using C = Curve<F>;
  using Q = Quad<F>;
  using P = typename C::P;
  extern __shared__ __align__(16) unsigned char red2_smem[];
  const uint32_t K = blockDim.x >> 2;
  P* shA = reinterpret_cast<P*>(red2_smem);
  P* shU = shA + K;
  __shared__ uint32_t blk_bad;
  const uint32_t blk = blockIdx.x;
  const uint32_t k = threadIdx.x >> 2;  // this quad's group
  const bool lead = (threadIdx.x & 3u) == 0;
  uint32_t bad = 0;
  if (threadIdx.x == 0) blk_bad = 0;
  const uint64_t g = (uint64_t)blk * K + k;
  const P S = C::load(lvlS + (uint64_t)C::xyzz_words * g);
  const P T = C::load(lvlT + (uint64_t)C::xyzz_words * g);
  P a = S;
  const P u = Q::add_lite(T, S, bad);
  if (lead) {
    shA[k] = a;
    shU[k] = u;
  }
  __syncthreads();
  for (uint32_t d = 1; d < K; d <<= 1) {
    P x = (k + d < K) ? shA[k + d] : C::infinity();
    __syncthreads();
    a = Q::add_lite(a, x, bad);
    if (lead) shA[k] = a;
    __syncthreads();
  }
  const P Z = shA[0];
  __syncthreads();
  if (threadIdx.x == 0) shA[0] = C::infinity();
  __syncthreads();
  if (K >= 2) {
    const uint32_t half = K / 2, t = k % half;
    P* arr = k < half ? shA : shU;
    for (uint32_t s = half; s > 0; s >>= 1) {
      // a quad reads arr[t] and arr[t + s] (t + s is written by no quad this round)
      // before its lead lane writes arr[t]
      if (t < s) {
        const P v = Q::add_lite(arr[t], arr[t + s], bad);
        if (lead) arr[t] = v;
      }
      __syncthreads();
    }
  }
  if (k == 0) {
    P V = shA[0];
    for (int i = 0; i < log2m; i++) V = Q::dbl(V);
    const P out = Q::add_lite(shU[0], V, bad);
    if (lead) {
      C::store(outPQ + (uint64_t)C::xyzz_words * (2 * blk), C::canon(out));
      C::store(outPQ + (uint64_t)C::xyzz_words * (2 * blk + 1), C::canon(Z));
    }
  }
  if (bad) atomicOr(&blk_bad, 1u);
  __syncthreads();
  if (threadIdx.x == 0 && blk_bad) rl2[1 + atomicAdd(&rl2[0], 1u)] = blk;
}

// Widest level-2 block (groups) the quad kernel can take: a quad per group, as many
// threads as its registers allow in one block (0 when BG16_QUAD=0).
template <class F>
uint32_t reduce2_quad_kmax() {
  const char* e = getenv("BG16_QUAD");  // read on every plan, so a process can switch
  if (e && atoi(e) == 0) return 0u;
  static const uint32_t cached = [] {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(&k_reduce2q<F>)) != hipSuccess ||
        attr.maxThreadsPerBlock < 8) {
      (void)hipGetLastError();
      return 0u;
    }
    uint32_t k = 1;
    while (8 * k <= (uint32_t)attr.maxThreadsPerBlock) k *= 2;
    return k;
  }();
  return cached;
}

// merge levels of at most this many blocks run k_merge_levelq
static constexpr uint32_t MERGE_QUAD_MAX_BLOCKS = 256;

template <class F>
uint64_t accumulate_round_threads() {
  static uint64_t cached = [] {
    int cus = 256;
    int wps = 2;
    hipFuncAttributes attr;
    try {
      cus = ctx().cu_count;
      if (hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(&k_accumulate<F>)) == hipSuccess &&
          attr.numRegs > 0) {
        int regs = (attr.numRegs + 7) / 8 * 8;
        wps = std::max(1, std::min(8, 512 / regs));
      }
    } catch (const Error&) {
    }
    return (uint64_t)cus * 4 * wps * 64;
  }();
  return cached;
}

// Fixed-base table: out[w * n + i] = 2^(w c) P_i (affine), one thread per point walking
// the windows (c doublings and one inversion per window, all inlined); the point at
// infinity (all-zero affine) stays all-zero in every window.
template <class F>
__global__ __launch_bounds__(256) void k_fixed_table(const uint32_t* __restrict__ pts, uint64_t n, int c, int Wd,
                                                     uint32_t* __restrict__ out) {
  using C = Curve<F>;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t aw = C::aff_words;
  typename C::A a = C::load_aff(pts + aw * i);
  C::store_aff(out + aw * i, a);
  const bool inf = C::aff_is_inf(a);
  typename C::P acc = C::from_affine(a);
  for (int w = 1; w < Wd; w++) {
    for (int k = 0; k < c; k++) acc = C::dbl(acc);
    typename C::A r = inf ? a : C::to_affine_inline(acc);
    C::store_aff(out + aw * ((uint64_t)w * n + i), r);
  }
}

template <class F>
void fixed_table_curve(const uint32_t* d_points, uint64_t n, int c, uint32_t* d_out, hipStream_t s) {
  if (!n) return;
  const int Wd = (255 + c - 1) / c;
  k_fixed_table<F><<<blocks_of(n, 256), 256, 0, s>>>(d_points, n, c, Wd, d_out);
  BG_HIP(hipGetLastError());
}

template <class F>
void enqueue_curve(uint32_t M, const uint32_t* const* d_points, const uint32_t* const* d_scalars, const uint64_t* ns,
                   int window_bits, hipStream_t s, MsmWork& w, int group, bool fixed) {
  using C = Curve<F>;
  BG_REQUIRE(M >= 1, BG_ERR_ARG, "msm: empty batch");
  for (uint32_t i = 0; i < M; i++)
    BG_REQUIRE(ns[i] < (1ull << 31), BG_ERR_ARG, "msm: n = %llu exceeds 2^31 - 1", (unsigned long long)ns[i]);
  if (fixed) BG_REQUIRE(window_bits >= 2 && window_bits <= 22, BG_ERR_ARG, "msm: fixed-base tables need c");
  MsmPlan p = fixed ? msm_plan_fixed(ns, M, group, window_bits) : msm_plan_batch(ns, M, group, window_bits);
  if (fixed)
    for (uint32_t i = 0; i < M; i++)
      BG_REQUIRE((uint64_t)p.Wd * ns[i] < (1ull << 31), BG_ERR_ARG, "msm: fixed-base table of %llu points too large",
                 (unsigned long long)ns[i]);
  const uint64_t n = p.n;
  BG_REQUIRE(p.entries < (1ull << 32) - 1 && (uint64_t)M * p.kspan < (1ull << 31), BG_ERR_ARG,
             "msm: batch too large (%llu entries)", (unsigned long long)p.entries);
  BG_REQUIRE(p.c >= 2 && p.c <= 22, BG_ERR_ARG, "msm: window_bits %d out of range", p.c);
  w.plan = p;
  w.group = group;
  const uint64_t xw = C::xyzz_words;
  uint32_t* counts = w.counts.as<uint32_t>(2 * ((uint64_t)p.nb + 1));  // counts | offsets
  uint32_t* offsets = counts + p.nb + 1;
  uint2* sorted = w.sorted.as<uint2>(p.entries);
  uint32_t* fix = w.fix.as<uint32_t>(1 + 2ull * p.nb);  // count | per-bucket flags | key list
  // reduction exception lists: level-1 count | groups, level-2 count | blocks
  uint64_t ngr = (uint64_t)p.M * p.W * p.G, nblk2 = ngr / p.K;
  uint32_t* rl1 = w.rfix.as<uint32_t>(2 + ngr + nblk2);
  uint32_t* rl2 = rl1 + 1 + ngr;
  uint32_t* buckets = w.buckets.as<uint32_t>((uint64_t)p.nb * xw);
  uint32_t* pkeys = w.pkeys.as<uint32_t>(2ull * p.acc_threads);
  uint32_t* pflags = w.pflags.as<uint32_t>(2ull * p.acc_threads);
  uint32_t* ppts = w.ppts.as<uint32_t>(2ull * p.acc_threads * xw);
  // merge levels ping-pong between two halves of mcap items each (sized for the narrower
  // of the two merge kernels' blocks)
  const uint64_t mcap = blocks_of(p.acc_threads, std::min<uint32_t>(MergeT<F>::value, MERGE_QT));
  uint32_t* mkeys = w.mkeys.as<uint32_t>(4 * mcap);
  uint32_t* mflags = w.mflags.as<uint32_t>(4 * mcap);
  uint32_t* mpts = w.mpts.as<uint32_t>(4 * mcap * xw);
  uint64_t ngroups = (uint64_t)p.M * p.W * p.G;
  uint32_t* lvlS = w.lvlS.as<uint32_t>(ngroups * xw);
  uint32_t* lvlT = w.lvlT.as<uint32_t>(ngroups * xw);
  uint32_t* mq = w.mq.as<uint32_t>(4 * std::max<uint64_t>((ngroups + 63) / 64 * 64, 256) * xw);
  uint64_t nout = (uint64_t)p.M * p.W * p.H;
  uint64_t* hptrs = static_cast<uint64_t*>(w.host_ptrs.get(8ull * M));
  for (uint32_t i = 0; i < M; i++) hptrs[i] = reinterpret_cast<uint64_t>(d_points[i]);
  uint64_t* dptrs = w.ptrs.as<uint64_t>(M);
  uint32_t* outPQ = w.outPQ.as<uint32_t>(2 * nout * xw);
  uint32_t* hout = static_cast<uint32_t*>(w.host_out.get(2 * nout * xw * 4));

  if (n == 0) {
    // empty MSM: every bucket empty -> zero outputs
    launch_zero_u32(outPQ, 2 * nout * xw, s);
  } else {
    // the sorts start from cleared bucket counts and fix-up flags (count | per-bucket flags)
    launch_set_ptrs(dptrs, hptrs, M, s, counts, (uint64_t)p.nb + 1, fix, 1 + (uint64_t)p.nb);
    msm_sort(d_scalars, d_points, ns, p, s, group, w, offsets, sorted, fix, 1 + (uint64_t)p.nb);
    {
      {
        ScopedKTime kt(group == 1 ? "msm_g1_accumulate" : "msm_g2_accumulate", s);
        k_accumulate<F><<<blocks_of(p.acc_threads, 256), 256, 0, s>>>(dptrs, p.kspan, sorted, offsets + p.nb, p.L,
                                                                       p.acc_threads, buckets, pkeys, pflags, ppts,
                                                                       p.nb, fix);
      }
      ScopedKTime kt(group == 1 ? "msm_g1_merge" : "msm_g2_merge", s);
      {
        const uint32_t* ik = pkeys;
        const uint32_t* ifl = pflags;
        const uint32_t* ip = ppts;
        uint32_t items = p.acc_threads;
        int side = 0;
        constexpr uint32_t MERGE_T = MergeT<F>::value;
        size_t msmem = MERGE_T * sizeof(typename C::P);
        do {
          // a level of few blocks is latency-bound: a quad per item (BG16_QUAD=0: never)
          const bool quad = reduce2_quad_kmax<F>() != 0 && blocks_of(items, MERGE_QT) <= MERGE_QUAD_MAX_BLOCKS;
          const uint32_t blocks = blocks_of(items, quad ? MERGE_QT : MERGE_T);
          uint32_t* ok = mkeys + (uint64_t)side * 2 * mcap;
          uint32_t* of = mflags + (uint64_t)side * 2 * mcap;
          uint32_t* op = mpts + (uint64_t)side * 2 * mcap * xw;
          if (quad)
            k_merge_levelq<F><<<blocks, 4 * MERGE_QT, MERGE_QT * sizeof(typename C::P), s>>>(ik, ifl, ip, items,
                                                                                          buckets, ok, of, op);
          else
            k_merge_level<F><<<blocks, MERGE_T, msmem, s>>>(ik, ifl, ip, items, buckets, ok, of, op);
          ik = ok;
          ifl = of;
          ip = op;
          items = blocks;
          side ^= 1;
        } while (items > 1);
      }
      k_fixup<F><<<4, 64, 0, s>>>(dptrs, p.kspan, sorted, offsets, p.nb, fix, buckets, rl1, rl2);
      BG_HIP(hipGetLastError());
    }
    if (!w.mid) BG_HIP(hipEventCreateWithFlags(&w.mid, hipEventDisableTiming));
    record_if_eager(w.mid, s);
    {
      int log2m = 0;
      while ((1u << log2m) < p.m) log2m++;
      unsigned nblocks = (unsigned)(ngroups / p.K);
      size_t smem = 2 * (size_t)p.K * sizeof(typename C::P);
      BG_REQUIRE(smem <= 65536, BG_ERR_INTERNAL, "msm: level-2 block %u needs %zu B of LDS", p.K, smem);
      {
        ScopedKTime kt(group == 1 ? "msm_g1_reduce1" : "msm_g2_reduce1", s);
        if (sizeof(typename C::P) > 128) {
          if (p.quad2)
            k_reduce1_serialq<F><<<blocks_of(4 * ngroups, 256), 256, 0, s>>>(buckets, offsets, p.B, p.m,
                                                                             (uint32_t)ngroups, lvlS, lvlT, rl1);
          else
            k_reduce1_serial<F, false><<<blocks_of(ngroups, 256), 256, 0, s>>>(buckets, offsets, p.B, p.m,
                                                                                (uint32_t)ngroups, lvlS, lvlT, rl1);
          k_reduce1_serial<F, true><<<16, 64, 0, s>>>(buckets, offsets, p.B, p.m, (uint32_t)ngroups, lvlS, lvlT,
                                                      rl1);
        } else {
          // one wave per role, 64 groups per block (16 with a quad per group)
          bool quad = false;
          if constexpr (sizeof(typename C::P) <= 128) {
            if (p.quad2) {
              k_reduce1q<F><<<(unsigned)((ngroups + 15) / 16), 256, 0, s>>>(buckets, offsets, p.B, (uint32_t)log2m,
                                                                            (uint32_t)ngroups, lvlS, lvlT, mq, rl1);
              quad = true;
            }
          }
          if (!quad) {
            const unsigned r1b = (unsigned)((ngroups + 63) / 64);
            k_reduce1<F, false><<<r1b, 256, 0, s>>>(buckets, offsets, p.B, (uint32_t)log2m, (uint32_t)ngroups, lvlS,
                                                    lvlT, mq, rl1);
          }
          k_reduce1<F, true><<<4, 256, 0, s>>>(buckets, offsets, p.B, (uint32_t)log2m, (uint32_t)ngroups, lvlS,
                                               lvlT, mq, rl1);
        }
      }
      ScopedKTime kt(group == 1 ? "msm_g1_reduce2" : "msm_g2_reduce2", s);
      bool quad = false;
      if (p.quad2 && p.K <= reduce2_quad_kmax<F>() && p.K >= 2) {
        k_reduce2q<F><<<nblocks, 4 * p.K, smem, s>>>(lvlS, lvlT, log2m, outPQ, rl2);
        quad = true;
      }
      if (!quad) k_reduce2<F, false><<<nblocks, p.K, smem, s>>>(lvlS, lvlT, log2m, outPQ, rl2);
      k_reduce2<F, true><<<8, p.K, smem, s>>>(lvlS, lvlT, log2m, outPQ, rl2);
      BG_HIP(hipGetLastError());
    }
  }
  // Outputs to the host: eagerly a copy and an event; in a recorded graph (which carries
  // no event nodes) a kernel writes them to coherent pinned memory and raises a flag that
  // msm_finish polls.  The landing is sized on eager runs too, so recording allocates
  // nothing.
  const uint64_t out_words = 2 * nout * xw;
  uint32_t* pub = static_cast<uint32_t*>(w.host_pub.get(4 * (MsmWork::PUB_HEAD + out_words), hipHostMallocCoherent));
  if (stream_capturing(s)) {
    launch_publish(outPQ, out_words, pub + MsmWork::PUB_HEAD, pub, s);
    w.published = true;
  } else {
    BG_HIP(hipMemcpyAsync(hout, outPQ, out_words * 4, hipMemcpyDeviceToHost, s));
    if (!w.done) BG_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
    BG_HIP(hipEventRecord(w.done, s));
    w.published = false;
  }
}


}  // namespace bg
