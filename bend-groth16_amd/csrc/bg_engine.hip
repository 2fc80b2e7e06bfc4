// Proving engine: continuous batching of Groth16 proofs on one device (C ABI:
// bg_engine_create, bg_engine_submit, bg_engine_wait, bg_engine_info, bg_engine_destroy,
// bg_prove_batch).
//
// Requests (key, witness, r/s) queue in arrival order.  Each worker thread owns one
// ProveSlot (four streams and the work arrays of its proofs) and takes the oldest request
// whose work arrays fit the device-memory budget, together with up to batch_limit - 1
// later requests whose keys have the same shape (prove_batch_on: one launch sequence for
// all of them, so small proofs share their latency-bound reductions and launches).
// Proofs of different sizes run concurrently on different slots: one proof's
// latency-bound reductions and host finishing overlap the others' accumulation.  Keys stay
// resident (bg_pk_load) and are shared read-only by all slots.  Every proof is computed
// exactly as bg_prove computes it alone.  The queue, admission and batching logic is the
// device-free Scheduler of bg_sched.hpp; this file supplies its prover-facing parts.
//
// The reference proves one instance per process run (bend2/prove.bend:129-162); this is
// the serving layer the batched and served configurations run on.
#include "bg_prover.hpp"
#include "bg_sched.hpp"
#include <chrono>
#include <algorithm>
#include <memory>
#include <string.h>
#include <vector>

namespace bg {

using Clock = sched::Clock;

struct Request {
  uint64_t ticket = 0;
  bg_pk* pk = nullptr;
  std::vector<uint32_t> witness;  // host copy (empty when the witness is on the device)
  const uint32_t* dev_witness = nullptr;
  uint32_t rs[32];
  int flags = 0;
  uint64_t need = 0;  // device bytes of its work arrays
  Clock::time_point t_submit, t_start, t_end;
  uint32_t proof[128];
  int status = 1;  // 1 pending, 0 ok, < 0 error
  std::string err;
};

// Proofs of one shape a slot takes together: 16 up to domain 2^14, 4 up to 2^16, 2 up to
// 2^18, and none from 2^19, where one proof fills the device.  Smaller batches as proofs
// grow keep several slots' batches in flight (a batch completes all at once): bursts of 32
// same-key proofs on 4 slots measured best at these sizes (2^15: 4 -> 1,175 proofs/s,
// 16 -> 798; 2^17: 2 -> 685, 8 -> 405; 2^19: 1 -> 212, 2 -> 200).  BG16_ENGINE_BATCH caps
// it (1 disables batching).
static size_t batch_limit_of(bg_pk* pk) {
  uint32_t hdr[8];
  if (bg_pk_header(pk, hdr) != 0) return 1;
  const uint32_t lg = hdr[3];  // log2 of the domain: constraints + 2 rounded up
  size_t lim = lg <= 14 ? 16 : lg <= 16 ? 4 : lg <= 18 ? 2 : 1;
  const char* env = getenv("BG16_ENGINE_BATCH");
  if (env && atoi(env) > 0) lim = std::min<size_t>(lim, (size_t)atoi(env));
  return lim;
}

// The scheduler's device-facing parts: slots are prover slots, a group of one request is
// proved by prove_on, a larger group (keys of one shape) by prove_batch_on.
struct EngineOps {
  using Request = bg::Request;
  using Slot = ProveSlot;
  std::unique_ptr<Slot> make_slot() {
    try {
      return std::make_unique<ProveSlot>();
    } catch (const Error&) {
      return nullptr;
    }
  }
  size_t batch_limit(const Request& r) { return batch_limit_of(r.pk); }
  bool batchable(const Request& head, const Request& other) { return prove_batchable(head.pk, other.pk); }
  // Longest a partial batch waits for more requests of its shape while the device is busy
  // (BG16_ENGINE_WAIT_US, default 2000 us).
  int64_t batch_wait_us() {
    static const int64_t w = [] {
      const char* e = getenv("BG16_ENGINE_WAIT_US");
      return e ? (int64_t)atoll(e) : (int64_t)2000;
    }();
    return w;
  }
  int run(Slot& slot, std::vector<Request*>& group, std::string& msg) {
    const int rc = guard([&] {
      if (group.size() == 1) {
        Request* r = group[0];
        prove_on(r->pk, slot, r->dev_witness ? r->dev_witness : r->witness.data(), r->rs, r->proof, r->flags,
                 nullptr);
      } else {
        std::vector<ProveJob> jobs;
        for (Request* r : group)
          jobs.push_back({r->pk, r->dev_witness ? r->dev_witness : r->witness.data(), r->rs, r->proof, r->flags});
        prove_batch_on(jobs.data(), (int)jobs.size(), slot);
      }
    });
    if (rc < 0) msg = bg_last_error();
    return rc;
  }
  void release(Slot& slot) { slot.release_buffers(); }
  uint64_t slot_bytes(const Slot& slot) { return slot.device_bytes(); }
  void done(Request& r) { pk_release(r.pk); }  // bg_pk_free refuses the key until here
};

}  // namespace bg

struct bg_engine {
  bg::EngineOps ops;
  std::unique_ptr<bg::sched::Scheduler<bg::EngineOps>> sched;
  int device = 0;
};

using namespace bg;

extern "C" int bg_engine_create(int slots, uint64_t hbm_budget_bytes, bg_engine** out) {
  return guard([&] {
    BG_REQUIRE(out && slots >= 1 && slots <= 64, BG_ERR_ARG, "bg_engine_create: slots %d out of 1..64", slots);
    *out = nullptr;
    Ctx& c = ctx();
    auto e = std::make_unique<bg_engine>();
    e->device = c.device;
    if (hbm_budget_bytes == 0) {
      size_t fr = 0, tot = 0;
      BG_HIP(hipMemGetInfo(&fr, &tot));
      hbm_budget_bytes = (uint64_t)(fr * 0.9);  // default: 90 % of the memory free now
    }
    e->sched = std::make_unique<sched::Scheduler<EngineOps>>(e->ops, slots, hbm_budget_bytes);
    *out = e.release();
  });
}

extern "C" int bg_engine_submit(bg_engine* e, bg_pk* pk, const uint32_t* witness, const uint32_t* rs, int flags,
                                uint64_t* ticket) {
  return guard([&] {
    BG_REQUIRE(e && pk && witness && rs && ticket, BG_ERR_ARG, "bg_engine_submit: null argument");
    bool words16 = (flags & 1) == BG_FMT_WORDS16;
    bool dev = (flags & BG_MEM_DEVICE) != 0;
    BG_REQUIRE(!(dev && words16), BG_ERR_ARG, "bg_engine_submit: a device witness must be packed");
    uint32_t hdr[8];
    BG_REQUIRE(bg_pk_header(pk, hdr) == 0, BG_ERR_ARG, "bg_engine_submit: bad key");
    check_host_inputs(witness, rs, flags, (uint64_t)hdr[1] + hdr[2]);
    auto r = std::make_unique<Request>();
    r->pk = pk;
    r->flags = flags;
    const uint64_t nz = (uint64_t)hdr[1] + hdr[2];
    const size_t ew = words16 ? 16 : 8;
    if (dev)
      r->dev_witness = witness;
    else
      r->witness.assign(witness, witness + nz * ew);
    memcpy(r->rs, rs, 2 * ew * 4);
    r->need = prove_bytes_estimate(pk);
    BG_REQUIRE(r->need <= e->sched->budget(), BG_ERR_NOMEM,
               "bg_engine_submit: one proof of this key needs %llu bytes, budget is %llu",
               (unsigned long long)r->need, (unsigned long long)e->sched->budget());
    pk_acquire(pk);  // released when the proof is done: bg_pk_free refuses the key until then
    *ticket = e->sched->submit(std::move(r));
  });
}

extern "C" int bg_engine_wait(bg_engine* e, uint64_t ticket, uint32_t* proof_out, double* t_ms) {
  std::unique_ptr<Request> r;
  int rc = guard([&] {
    BG_REQUIRE(e && proof_out, BG_ERR_ARG, "bg_engine_wait: null argument");
    r = e->sched->wait(ticket);
    BG_REQUIRE(r != nullptr, BG_ERR_ARG, "bg_engine_wait: unknown ticket %llu", (unsigned long long)ticket);
  });
  if (rc < 0) return rc;
  if (r->status < 0) {
    set_error("%s", r->err.c_str());
    return r->status;
  }
  const size_t words = (r->flags & 1) == BG_FMT_WORDS16 ? 128 : 64;
  memcpy(proof_out, r->proof, words * 4);
  if (t_ms) {
    using ms = std::chrono::duration<double, std::milli>;
    t_ms[0] = ms(r->t_start - r->t_submit).count();  // queued
    t_ms[1] = ms(r->t_end - r->t_start).count();     // proving
    t_ms[2] = ms(r->t_end - r->t_submit).count();    // latency
  }
  return 0;
}

extern "C" int bg_engine_info(bg_engine* e, uint64_t* info4) {
  return guard([&] {
    BG_REQUIRE(e && info4, BG_ERR_ARG, "bg_engine_info: null argument");
    e->sched->info(info4);
  });
}

extern "C" int bg_engine_destroy(bg_engine* e) {
  return guard([&] {
    if (!e) return;
    e->sched->shutdown();  // finishes what is queued or running, then stops the workers
    delete e;
  });
}

extern "C" int bg_prove_batch(int count, bg_pk* const* pks, const uint32_t* const* witnesses,
                              const uint32_t* const* rs, uint32_t* const* proofs, int flags, int slots) {
  bg_engine* e = nullptr;
  int rc = guard([&] {
    BG_REQUIRE(count >= 0 && (count == 0 || (pks && witnesses && rs && proofs)), BG_ERR_ARG,
               "bg_prove_batch: bad argument");
  });
  if (rc < 0 || count == 0) return rc;
  rc = bg_engine_create(slots > 0 ? slots : 4, 0, &e);
  if (rc < 0) return rc;
  std::vector<uint64_t> tickets(count, 0);
  int first_err = 0;
  std::string msg;
  for (int i = 0; i < count && !first_err; i++) {
    int r = bg_engine_submit(e, pks[i], witnesses[i], rs[i], flags, &tickets[i]);
    if (r < 0) {
      first_err = r;
      msg = bg_last_error();
    }
  }
  for (int i = 0; i < count; i++) {
    if (!tickets[i]) continue;
    int r = bg_engine_wait(e, tickets[i], proofs[i], nullptr);
    if (r < 0 && !first_err) {
      first_err = r;
      msg = bg_last_error();
    }
  }
  bg_engine_destroy(e);
  if (first_err) set_error("%s", msg.c_str());
  return first_err;
}
