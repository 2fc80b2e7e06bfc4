// Per-proof execution state and the prove driver shared by bg_prove and the proving
// engine (bg_engine.hip).  A resident key (bg_pk) is read-only while proofs run, so any
// number of slots may prove with one key at once.
#pragma once
#include "bg_runtime.hpp"
#include "bg_msm.hpp"

struct bg_pk;

namespace bg {

// One proof in flight: four streams (witness path + qap_h + H, G1 batch, spare, B2),
// phase events and the device work arrays, grown to the largest key proved on it.
struct ProveSlot {
  hipStream_t s[4] = {};
  hipEvent_t ev[16] = {};
  DevArray z, zc, ea, eb, ec, h, tmp;
  DevArray comb;  // the device combination's scratch (BG_COMBINE_DEVICE)
  HostBuf zpin;   // pinned staging of a host witness (packed words)
  HostBuf proofpin;  // pinned landing of a device-combined proof
  uint64_t cap_nz = 0, cap_d = 0;
  MsmWork work[3];  // G1 batch (A, B1, L), H, B2
  // Recorded launch sequences of single proofs (prove_on): three linear HIP graphs (one per
  // stream) per key, schedule, tuning table, BG16_ settings and buffer layout (`key`),
  // least recently used dropped beyond MAX_GRAPHS.  A graph is recorded only when an
  // earlier proof on this slot ran eagerly with the same key (same buffer addresses and
  // sizes), so every buffer it needs already has its final size and nothing is allocated
  // while recording.
  struct Graph {
    uint64_t key = 0;
    hipGraphExec_t exec[3] = {};  // launched on s[3], s[1], s[0]
    uint64_t used = 0;
    // host side of the recorded MSMs (how msm_finish reads their outputs); put back into
    // work[] on each launch, since other keys' proofs on this slot overwrite it
    MsmPlan plan[3];
    int group[3] = {};
  };
  static constexpr size_t MAX_GRAPHS = 8;
  std::vector<Graph> graphs;
  uint64_t graph_clock = 0;
  std::vector<uint64_t> warm;  // keys of recent eager proofs on this slot (newest last)
  uint64_t failed_key = 0;     // key whose recording failed (it runs eagerly)
  bool is_warm(uint64_t key) const;
  void add_warm(uint64_t key);
  hipEvent_t join[3] = {};  // join the side streams back into s[0] at the end of a recording
  // The graphs recorded for key with work[]'s host side restored to their recording, or
  // null.
  const Graph* find_graph(uint64_t key);
  const Graph* add_graph(uint64_t key, const hipGraphExec_t exec[3]);  // keeps work[]'s host side too
  void drop_graphs();
  uint64_t buffer_signature() const;  // addresses and sizes of every work buffer
  ProveSlot();
  ~ProveSlot();
  ProveSlot(const ProveSlot&) = delete;
  ProveSlot& operator=(const ProveSlot&) = delete;
  void ensure(uint64_t nz, uint64_t d);
  void release_buffers();
  uint64_t device_bytes() const;
};

// Device bytes one proof with this key needs in a fresh slot (work arrays + MSM work).
uint64_t prove_bytes_estimate(const bg_pk* pk);

// Checks a host witness of nz elements (either encoding) and r / s (host) as bg_prove
// does: 16-bit limbs and every value below r; throws BG_ERR_ARG otherwise.
void check_host_inputs(const uint32_t* witness, const uint32_t* rs, int flags, uint64_t nz);

// Proofs of a key that are queued or running: bg_pk_free refuses a key while any are.
void pk_acquire(bg_pk* pk);
void pk_release(bg_pk* pk);

// Final combination on the device (BG_COMBINE_DEVICE, bg_combine.hip): the key's points,
// r, s, rs (canonical) and the shapes of the three MSM batches' level-2 outputs.
struct CombineArgs {
  uint32_t alpha1[16], beta1[16], delta1[16];  // packed affine Montgomery
  uint32_t beta2[32], delta2[32];
  uint32_t sc[24];                              // r, s, rs canonical
  const uint32_t* dtab1;                        // j 2^(4w) delta (G1 XYZZ, w < 64, j < 16)
  const uint32_t* dtab2;                        // the same for delta2 (G2)
  uint32_t hn_g1, hn_h, hn_b2;                  // level-2 outputs per MSM (powers of two)
  int log2m_g1, log2m_h, log2m_b2;              // log2(m K) of each batch's plan
};
uint64_t comb_words();         // device words of one proof's combination scratch
uint32_t comb_proof_offset();  // where the proof (64 packed words) lands in it
void launch_delta_terms(const CombineArgs& a, uint32_t* cb, hipStream_t s);
void launch_combine_g1(const CombineArgs& a, const uint32_t* outPQ, uint32_t* cb, hipStream_t s);
void launch_affine_a(uint32_t* cb, hipStream_t s);
void launch_combine_b2(const CombineArgs& a, const uint32_t* outPQ, uint32_t* cb, hipStream_t s);
void launch_combine_final(const CombineArgs& a, const uint32_t* outPQ, uint32_t* cb, hipStream_t s);

// Runs one proof on slot sl (blocking until the proof is on the host).
void prove_on(bg_pk* pk, ProveSlot& sl, const uint32_t* witness, const uint32_t* rs, uint32_t* proof_out, int flags,
              double* t_ms);

// One proof of a batch: key, witness, r/s and output as bg_prove's arguments.
struct ProveJob {
  bg_pk* pk;
  const uint32_t* witness;
  const uint32_t* rs;
  uint32_t* proof_out;
  int flags;
};
// Keys whose proofs can share launches: fixed-base tables, identical headers and widths.
bool prove_batchable(const bg_pk* a, const bg_pk* b);
// Runs P proofs of batchable keys together on slot sl: one upload pass, one qap_h launch
// sequence over all P domains, and the five MSMs of every proof as three batched MSMs
// (B2 x P, the G1 batch A/B1/L x P, H x P).  Each proof equals the proof prove_on makes
// alone.  Blocks until every proof is on the host.
void prove_batch_on(const ProveJob* jobs, int P, ProveSlot& sl);

}  // namespace bg
