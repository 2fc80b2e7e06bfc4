// The proving engine's scheduler (bend-groth16_amd/csrc/bg_sched.hpp) driven by a
// simulated prover on host threads, for ThreadSanitizer and AddressSanitizer builds: many
// submitting threads, requests of several shapes and sizes, a budget that admits only some
// at once.  Checks, while it runs and at the end: every request completes exactly once,
// with its own result; the slots' reservations never exceed the budget; a group holds one
// shape and at most its batch limit; a group's requests were queued in arrival order.
#include "bg_sched.hpp"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>

using namespace bg::sched;

struct SimRequest {
  uint64_t ticket = 0, need = 0;
  Clock::time_point t_submit, t_start, t_end;
  int status = 1;
  std::string err;
  int shape = 0;
  uint64_t input = 0, result = 0;
};

struct SimSlot {
  uint64_t bytes = 0;
};

static std::atomic<int> g_fail{0};
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      g_fail++;                                                         \
    }                                                                   \
  } while (0)

struct SimOps {
  using Request = SimRequest;
  using Slot = SimSlot;
  uint64_t budget = 0;
  Scheduler<SimOps>* sched = nullptr;
  std::atomic<int> runs{0}, released{0};
  std::atomic<uint64_t> done_count{0};
  std::unique_ptr<Slot> make_slot() { return std::make_unique<Slot>(); }
  size_t batch_limit(const Request& r) { return r.shape == 0 ? 8 : r.shape == 1 ? 3 : 1; }
  bool batchable(const Request& a, const Request& b) { return a.shape == b.shape; }
  int64_t batch_wait_us() { return 300; }
  int run(Slot& s, std::vector<Request*>& group, std::string& msg) {
    runs++;
    CHECK(!group.empty() && group.size() <= batch_limit(*group[0]));
    for (size_t i = 0; i < group.size(); i++) {
      CHECK(group[i]->shape == group[0]->shape);
      if (i) CHECK(group[i]->ticket > group[i - 1]->ticket);
    }
    uint64_t info[4];
    sched->info(info);
    CHECK(info[1] <= info[2]);  // reservations within the budget
    uint64_t need = 0;
    for (Request* r : group) need += r->need;
    s.bytes = std::max(s.bytes, need);
    std::this_thread::sleep_for(std::chrono::microseconds(50 + 40 * group[0]->shape));
    for (Request* r : group) r->result = r->input * 2654435761u + 17;  // the "proof"
    if (group[0]->shape == 3) {
      msg = "simulated failure";
      return -5;
    }
    return 0;
  }
  void release(Slot& s) {
    released++;
    s.bytes = 0;
  }
  uint64_t slot_bytes(const Slot& s) { return s.bytes; }
  void done(Request&) { done_count++; }
};

int main() {
  const uint64_t budget = 1000;
  SimOps ops;
  ops.budget = budget;
  const int submitters = 4, per = 150;
  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> tickets(submitters);  // (ticket, input)
  {
    Scheduler<SimOps> s(ops, 5, budget);
    ops.sched = &s;
    std::vector<std::thread> th;
    for (int t = 0; t < submitters; t++)
      th.emplace_back([&, t] {
        std::mt19937_64 rng(1000 + t);
        for (int i = 0; i < per; i++) {
          auto r = std::make_unique<SimRequest>();
          r->shape = (int)(rng() % 4);
          const uint64_t sizes[4] = {60, 200, 450, 30};
          r->need = sizes[r->shape];
          r->input = rng();
          const uint64_t in = r->input;
          tickets[t].push_back({s.submit(std::move(r)), in});
          if (rng() % 7 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 500));
        }
      });
    for (auto& x : th) x.join();
    // waiters on several threads, each collecting one submitter's tickets
    std::vector<std::thread> wt;
    std::atomic<int> ok{0}, failed{0};
    for (int t = 0; t < submitters; t++)
      wt.emplace_back([&, t] {
        for (auto& [tk, in] : tickets[t]) {
          std::unique_ptr<SimRequest> r = s.wait(tk);
          CHECK(r != nullptr);
          if (!r) continue;
          CHECK(r->ticket == tk && r->status != 1);
          if (r->shape == 3) {
            CHECK(r->status == -5 && r->err == "simulated failure");
            failed++;
          } else {
            CHECK(r->status == 0 && r->result == in * 2654435761u + 17);
            ok++;
          }
          CHECK(r->t_end >= r->t_start && r->t_start >= r->t_submit);
          CHECK(s.wait(tk) == nullptr);  // handed back once
        }
      });
    for (auto& x : wt) x.join();
    uint64_t info[4];
    s.info(info);
    CHECK(info[0] == 5 && info[2] == budget && info[3] == (uint64_t)submitters * per);
    CHECK(ok + failed == submitters * per);
    std::printf("runs %d, releases %d, completed %llu\n", ops.runs.load(), ops.released.load(),
                (unsigned long long)info[3]);
  }  // destructor: shutdown and join
  CHECK(ops.done_count == (uint64_t)submitters * per);
  if (g_fail) {
    std::fprintf(stderr, "%d checks failed\n", g_fail.load());
    return 1;
  }
  std::printf("scheduler simulation OK\n");
  return 0;
}
