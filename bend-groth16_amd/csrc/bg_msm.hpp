// Pippenger MSM driver interface (device pipeline in bg_msm.hip).
//
// Replaces Msm.g1 / Msm.g2 (bend2/src/msm.bend:293-298, :525-530).  The reference uses
// unsigned c-bit windows with private per-leaf bucket arrays; here the scalars are cut
// into signed c-bit digits (2^(c-1) buckets per window), the (bucket, point) pairs are
// counting-sorted on the device, each thread accumulates a fixed-length slice of the
// sorted list (exact merging of buckets that straddle slices), and the per-window
// weighted bucket sums run as a two-level parallel reduction.  The last few points per
// window and the window combination run on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <array>
#include "bg_runtime.hpp"

namespace bg {

// Schedule of a batch of M MSMs of one group sharing one window width: the batch is
// handled as one MSM with M*W windows (window mw belongs to MSM mw / W), so every stage
// runs once over all MSMs.
struct MsmPlan {
  int c = 0;          // window width (bits)
  int W = 0;          // bucket windows per MSM (1 with fixed-base tables)
  int Wd = 0;         // digit windows per scalar, Wd*c >= 255 (= W without tables)
  bool fixed = false; // bases are fixed-base tables: entry (w, i) uses 2^(wc) P_i, every
                      // window adds into the same buckets
  uint32_t M = 1;     // MSMs in the batch
  uint32_t B = 0;     // buckets per window, 2^(c-1)
  uint32_t kspan = 0; // buckets per MSM, W*B
  uint32_t nb = 0;    // M*W*B
  uint32_t m = 0;     // buckets per level-1 thread
  uint32_t G = 0;     // level-1 groups per window
  uint32_t K = 0;     // level-2 block size
  uint32_t H = 0;     // level-2 outputs per window
  uint32_t L = 0;     // sorted entries per accumulation thread
  uint64_t n = 0;     // total points
  uint64_t entries = 0;  // Wd * total points
  uint32_t acc_threads = 0;
  bool quad2 = false;  // level 2 with a quad per group (k_reduce2q)
};

MsmPlan msm_plan(uint64_t n, int group, int window_bits);
MsmPlan msm_plan_batch(const uint64_t* ns, uint32_t M, int group, int window_bits);
// Plan over fixed-base tables of window width c (table w of MSM i at w * n_i).
MsmPlan msm_plan_fixed(const uint64_t* ns, uint32_t M, int group, int c);
// Window width for fixed-base tables of a batch of MSMs of these sizes.
int fixed_window_bits(const uint64_t* ns, uint32_t M, int group);

// Work buffers of one in-flight MSM (reused across calls).
struct MsmWork {
  DevBuf keys, vals, ranks, counts, sorted, buckets, pkeys, pflags, ppts, mkeys, mflags, mpts, lvlS, lvlT, outPQ,
      scan_tmp, fix, rfix, ptrs, scal, hist, mq;
  HostBuf host_out, host_ptrs;
  // Coherent pinned landing of the outputs of a recorded MSM (no event nodes in a graph):
  // PUB_HEAD words whose first is a flag the device raises once the outputs follow it.
  static constexpr uint64_t PUB_HEAD = 64;
  HostBuf host_pub;
  bool published = false;        // outputs come through host_pub (set for graph launches)
  hipStream_t pub_stream = nullptr;  // stream the recorded sequence runs on (polled for errors)
  // Clears the flag before the recorded sequence that holds this MSM is launched on s.
  void arm_publish(hipStream_t s) {
    published = true;
    pub_stream = s;
    __atomic_store_n(static_cast<uint32_t*>(host_pub.ptr), 0u, __ATOMIC_RELEASE);
  }
  MsmPlan plan;
  int group = 1;
  hipEvent_t done = nullptr;
  hipEvent_t mid = nullptr;  // recorded once the buckets are complete (before the reductions)
  // every device buffer: the byte count, the release and the buffer signature of recorded
  // launch sequences all walk this one list
  template <class Self>
  static auto buf_list(Self& w) {
    return std::array<decltype(&w.keys), 22>{&w.keys, &w.vals, &w.ranks, &w.counts, &w.sorted, &w.buckets,
                                             &w.pkeys, &w.pflags, &w.ppts, &w.mkeys, &w.mflags, &w.mpts,
                                             &w.lvlS, &w.lvlT, &w.outPQ, &w.scan_tmp, &w.fix, &w.rfix,
                                             &w.ptrs, &w.scal, &w.hist, &w.mq};
  }
  auto bufs() const { return buf_list(*this); }
  auto bufs() { return buf_list(*this); }
  uint64_t bytes() const {
    uint64_t t = 0;
    for (const DevBuf* b : bufs()) t += b->cap;
    return t;
  }
  void release() {
    for (DevBuf* b : bufs()) b->release();
    host_out.release();
    host_ptrs.release();
    host_pub.release();
    published = false;
    if (done) (void)hipEventDestroy(done);
    if (mid) (void)hipEventDestroy(mid);
    done = mid = nullptr;
  }
};

// Accumulation threads of one group resident on the device at once (CUs x SIMDs x the
// waves per SIMD that the accumulation kernel's registers allow).
uint64_t accumulate_threads_per_round(int group);

// Words of the per-chunk bucket histograms of the counting sort for this batch.
uint64_t msm_hist_words(const MsmPlan& p, const uint64_t* ns);

// Device bytes the work buffers of a batch with this plan occupy (upper estimate).
uint64_t msm_work_bytes(const MsmPlan& p, int group);

// Enqueue the device part of an MSM on stream s: points (packed affine, device),
// scalars (packed canonical, device), n.  Results land in w.host_out (pinned) when the
// recorded event completes.
void msm_enqueue(int group, const uint32_t* d_points, const uint32_t* d_scalars, uint64_t n, int window_bits,
                 hipStream_t s, MsmWork& w);

// Batch of M MSMs of one group: points[i] / scalars[i] / ns[i] per MSM (device).
void msm_enqueue_batch(int group, uint32_t M, const uint32_t* const* d_points, const uint32_t* const* d_scalars,
                       const uint64_t* ns, int window_bits, hipStream_t s, MsmWork& w);
// Batch over fixed-base tables (built by msm_build_fixed_table with the same c): d_tables[i]
// holds Wd * ns[i] affine points.
void msm_enqueue_fixed(int group, uint32_t M, const uint32_t* const* d_tables, const uint32_t* const* d_scalars,
                       const uint64_t* ns, int c, hipStream_t s, MsmWork& w);
// out[w * n + i] = 2^(w c) points[i] (affine), w < ceil(255 / c); out holds that many points.
void msm_build_fixed_table(int group, const uint32_t* d_points, uint64_t n, int c, uint32_t* d_out, hipStream_t s);

// Waits for the MSM(s) and finishes them on the host: out = affine results, one per MSM
// in the batch (packed words: G1 16, G2 32 each).  xyzz_out (optional) receives the XYZZ
// sums (G1 32, G2 64 words each).
void msm_finish(MsmWork& w, uint32_t* out_affine, uint32_t* xyzz_out = nullptr);

// Device kernels shared with the prover: FORMAT.md 16-bit words -> packed words.
void launch_pack16(const uint32_t* in16, uint32_t* out32, uint64_t out_words, hipStream_t s);
void launch_unpack16(const uint32_t* in32, uint32_t* out16, uint64_t in_words, hipStream_t s);


// pipeline pieces shared by the per-group translation units
uint32_t msm_scan_blocks(uint32_t n);
void launch_zero_u32(uint32_t* p, uint64_t n, hipStream_t s);
// d_out[i] = h_ptrs[i] for i < M on stream s (the MSM's table of base-array addresses,
// passed as kernel arguments: nothing on the host to copy from in a recorded sequence;
// h_ptrs may be reused once this returns), and z1[0, n1), z2[0, n2) cleared in the same
// launch.
void launch_set_ptrs(uint64_t* d_out, const uint64_t* h_ptrs, uint32_t M, hipStream_t s, uint32_t* z1, uint64_t n1,
                     uint32_t* z2, uint64_t n2);
// dst[i] = src[i] for i < n (dst in coherent pinned host memory), then *flag = 1 once the
// copy is complete and visible to the host.
void launch_publish(const uint32_t* src, uint64_t n, uint32_t* dst, uint32_t* flag, hipStream_t s);
// digits and the sort of the batch's (bucket, point) entries by bucket: sorted entries,
// offsets[0 .. nb] (offsets[nb] = entries); the work arrays come from w (offsets lies in
// w.counts, after nb + 1 words the single-pass sort uses for its counts).  Those counts
// and fix[0, fix_words) must be zero when it starts (launch_set_ptrs clears them).
void msm_sort(const uint32_t* const* d_scalars, const uint32_t* const* d_points, const uint64_t* ns,
              const MsmPlan& p, hipStream_t s, int group, MsmWork& w, uint32_t* offsets, uint2* sorted,
              uint32_t* fix, uint64_t fix_words);
void enqueue_g1(uint32_t M, const uint32_t* const* d_points, const uint32_t* const* d_scalars, const uint64_t* ns,
                int window_bits, hipStream_t s, MsmWork& w, bool fixed);
void enqueue_g2(uint32_t M, const uint32_t* const* d_points, const uint32_t* const* d_scalars, const uint64_t* ns,
                int window_bits, hipStream_t s, MsmWork& w, bool fixed);
void fixed_table_g1(const uint32_t* d_points, uint64_t n, int c, uint32_t* d_out, hipStream_t s);
void fixed_table_g2(const uint32_t* d_points, uint64_t n, int c, uint32_t* d_out, hipStream_t s);

}  // namespace bg
