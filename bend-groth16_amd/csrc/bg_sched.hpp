// Scheduler of the proving engine (host logic only, no device calls): a queue of requests
// in arrival order, worker threads that each own one slot, admission by a device-memory
// budget and grouping of queued requests of one shape into batches.
//
// bg_engine.hip instantiates it with the prover (slots = ProveSlot, runs = prove_on /
// prove_batch_on); the sanitizer tests instantiate it with a simulated prover, so the
// locking, admission and hand-back logic runs under ThreadSanitizer on the host.
//
// Ops supplies the device-facing parts:
//   using Request = ...;   has: uint64_t need, ticket; Clock::time_point t_submit, t_start,
//                          t_end; int status (1 pending, 0 ok, < 0 error); std::string err
//   using Slot = ...;
//   std::unique_ptr<Slot> make_slot();              // nullptr when it cannot be made
//   size_t batch_limit(const Request&);              // requests of this shape one slot takes
//   bool batchable(const Request& head, const Request& other);
//   int64_t batch_wait_us();                         // a partial batch's longest wait
//   int run(Slot&, std::vector<Request*>&, std::string& msg);  // proves the group; status
//   void release(Slot&);                             // gives the slot's work arrays back
//   uint64_t slot_bytes(const Slot&);                // device bytes the slot holds now
//   void done(Request&);                             // the request is finished
#pragma once
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace bg {
namespace sched {

using Clock = std::chrono::steady_clock;

template <class Ops>
class Scheduler {
 public:
  using Request = typename Ops::Request;
  using Slot = typename Ops::Slot;

  Scheduler(Ops& ops, int slots, uint64_t budget) : ops_(ops), budget_(budget), reserved_(slots, 0) {
    for (int i = 0; i < slots; i++) workers_.emplace_back([this, i] { worker_main(i); });
  }
  ~Scheduler() { shutdown(); }

  uint64_t budget() const { return budget_; }

  // Queues r (r->need must not exceed the budget); returns its ticket.
  uint64_t submit(std::unique_ptr<Request> r) {
    std::lock_guard<std::mutex> lk(mu_);
    r->ticket = next_ticket_++;
    r->t_submit = Clock::now();
    const uint64_t t = r->ticket;
    queue_.push_back(r.get());
    all_[t] = std::move(r);
    cv_work_.notify_all();
    return t;
  }

  // Blocks until the ticket's request is finished and hands it back (nullptr: unknown ticket).
  std::unique_ptr<Request> wait(uint64_t ticket) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = all_.find(ticket);
    if (it == all_.end()) return nullptr;
    Request* q = it->second.get();
    cv_done_.wait(lk, [&] { return q->status != 1; });
    std::unique_ptr<Request> r = std::move(it->second);
    all_.erase(it);
    return r;
  }

  // slots, device bytes reserved by the slots, budget, requests completed
  void info(uint64_t* info4) {
    std::lock_guard<std::mutex> lk(mu_);
    info4[0] = workers_.size();
    info4[1] = total_reserved();
    info4[2] = budget_;
    info4[3] = completed_;
  }

  // Finishes every queued or running request, then stops and joins the workers.
  void shutdown() {
    {
      std::unique_lock<std::mutex> lk(mu_);
      if (stop_) return;
      cv_done_.wait(lk, [&] {
        for (auto& kv : all_)
          if (kv.second->status == 1) return false;
        return true;
      });
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : workers_) t.join();
  }

 private:
  uint64_t total_reserved() const {
    uint64_t t = 0;
    for (uint64_t r : reserved_) t += r;
    return t;
  }

  // Takes the next group for slot_id (blocking); empty when the scheduler stops.
  std::vector<Request*> take(int slot_id, Slot* slot) {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      if (stop_) return {};
      if (!queue_.empty()) {
        Request* head = queue_.front();
        // the head and the next queued requests of the same shape
        std::vector<Request*> cand = {head};
        const size_t lim = ops_.batch_limit(*head);
        for (auto it = queue_.begin() + 1; it != queue_.end() && cand.size() < lim; ++it)
          if (ops_.batchable(*head, **it)) cand.push_back(*it);
        // a partial batch waits for more requests while another slot keeps the device busy,
        // up to the batch wait after the head arrived (a full batch, or an idle device, goes now)
        const auto ready_at = head->t_submit + std::chrono::microseconds(ops_.batch_wait_us());
        if (cand.size() < lim && running_ > 0 && Clock::now() < ready_at) {
          cv_work_.wait_until(lk, ready_at);
          continue;
        }
        const uint64_t mine = reserved_[slot_id];
        const uint64_t others = total_reserved() - mine;
        uint64_t need = head->need * cand.size();
        if (cand.size() > 1 && others + std::max(mine, need) > budget_) {
          cand.resize(1);  // the batch does not fit: the head alone
          need = head->need;
        }
        const uint64_t grown = std::max(mine, need);
        const bool fits = others + grown <= budget_;
        const bool fits_after_release = !fits && mine && others + need <= budget_;
        if (fits || fits_after_release) {
          // claim the requests before the lock may be dropped
          for (Request* r : cand) queue_.erase(std::find(queue_.begin(), queue_.end(), r));
          running_++;
          reserved_[slot_id] = fits ? grown : need;
          if (fits_after_release) {
            // the batch fits once this idle slot gives its arrays back
            lk.unlock();
            if (slot) ops_.release(*slot);
            lk.lock();
          }
          return cand;
        }
        if (mine) {
          // idle and holding memory that a waiting request needs: give it back
          reserved_[slot_id] = 0;
          lk.unlock();
          if (slot) ops_.release(*slot);
          lk.lock();
          cv_work_.notify_all();
          continue;
        }
      }
      cv_work_.wait(lk);
    }
  }

  void worker_main(int slot_id) {
    std::unique_ptr<Slot> slot = ops_.make_slot();
    for (;;) {
      std::vector<Request*> group = take(slot_id, slot.get());
      if (group.empty()) return;
      const auto t_start = Clock::now();
      int rc = 0;
      std::string msg;
      if (!slot) {
        rc = -2;  // BG_ERR_DEVICE
        msg = "engine worker could not create its slot";
      } else {
        rc = ops_.run(*slot, group, msg);
      }
      const auto t_end = Clock::now();
      {
        std::lock_guard<std::mutex> lk(mu_);
        running_--;
        for (Request* r : group) {
          r->t_start = t_start;
          r->t_end = t_end;
          r->status = rc < 0 ? rc : 0;
          r->err = msg;
          completed_++;
          ops_.done(*r);
        }
        // actual footprint after the run (grow-only buffers)
        if (slot) reserved_[slot_id] = std::max(reserved_[slot_id], ops_.slot_bytes(*slot));
      }
      cv_done_.notify_all();
      cv_work_.notify_all();
    }
  }

  Ops& ops_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Request*> queue_;
  std::map<uint64_t, std::unique_ptr<Request>> all_;
  std::vector<std::thread> workers_;
  const uint64_t budget_;
  std::vector<uint64_t> reserved_;  // per slot: device bytes its work arrays may hold
  uint64_t next_ticket_ = 1, completed_ = 0;
  int running_ = 0;  // slots running a group now
  bool stop_ = false;
};

}  // namespace sched
}  // namespace bg
