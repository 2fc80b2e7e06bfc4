// Library context, error reporting and kernel timing (C ABI: bg_init, bg_last_error,
// bg_set_timing, bg_timing_report, bg_timing_reset).
#include "bg_runtime.hpp"
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <thread>
#include <stdarg.h>
#include <string.h>
#include <map>
#include <string>

namespace bg {

static thread_local char g_err[512] = "";
static std::mutex g_init_mu;
static Ctx* g_ctx = nullptr;
static int g_device = 0;

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

void ensure_init() {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_ctx) return;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    set_error("no HIP device available (%s)", hipGetErrorString(e));
    throw Error{BG_ERR_DEVICE};
  }
  BG_REQUIRE(g_device >= 0 && g_device < n, BG_ERR_ARG, "device %d out of range (%d devices)", g_device, n);
  BG_HIP(hipSetDevice(g_device));
  Ctx* c = new Ctx();
  c->device = g_device;
  BG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& s : c->aux) BG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  BG_HIP(hipGetDeviceProperties(&prop, g_device));
  c->cu_count = prop.multiProcessorCount;
  g_ctx = c;
}

Ctx& ctx() {
  if (!g_ctx) ensure_init();
  BG_HIP(hipSetDevice(g_ctx->device));
  return *g_ctx;
}

// ---------------------------------------------------------------- timing
struct TimeAcc {
  double total_ms = 0;
  uint64_t count = 0;
};
static std::mutex g_tmu;
static std::map<std::string, TimeAcc> g_times;
struct Pending {
  std::string name;
  hipEvent_t a, b;
};
static std::vector<Pending> g_pending;

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  BG_HIP(hipStreamIsCapturing(s, &st));
  return st == hipStreamCaptureStatusActive;
}

bool record_if_eager(hipEvent_t e, hipStream_t s) {
  if (stream_capturing(s)) return false;
  BG_HIP(hipEventRecord(e, s));
  return true;
}

void timing_record(const char* name, hipStream_t, hipEvent_t start, hipEvent_t stop) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_pending.push_back({name, start, stop});
}

static void timing_flush() {
  std::lock_guard<std::mutex> lk(g_tmu);
  for (auto& p : g_pending) {
    float ms = 0;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto& t = g_times[p.name];
      t.total_ms += ms;
      t.count += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  g_pending.clear();
}

namespace {

class HostPool {
 public:
  explicit HostPool(int workers) {
    for (int i = 0; i < workers; i++) th_.emplace_back([this] { loop(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void run(int n, const std::function<void(int)>& f) {
    std::lock_guard<std::mutex> serial(run_mu_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &f;
      n_ = n;
      next_.store(0);
      open_ = true;
      gen_++;
    }
    cv_.notify_all();
    work(f, n);
    std::unique_lock<std::mutex> lk(mu_);
    open_ = false;  // late wakers skip this job
    done_cv_.wait(lk, [&] { return active_ == 0; });
  }

 private:
  void work(const std::function<void(int)>& f, int n) {
    for (int i; (i = next_.fetch_add(1)) < n;) f(i);
  }
  void loop() {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || (open_ && gen_ != seen); });
      if (stop_) return;
      seen = gen_;
      const std::function<void(int)>* f = job_;
      int n = n_;
      active_++;
      lk.unlock();
      work(*f, n);
      lk.lock();
      if (--active_ == 0) done_cv_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_, run_mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* job_ = nullptr;
  int n_ = 0, active_ = 0;
  std::atomic<int> next_{0};
  uint64_t gen_ = 0;
  bool open_ = false, stop_ = false;
};

}  // namespace

void host_parallel(int n, const std::function<void(int)>& f) {
  if (n <= 0) return;
  static HostPool* pool = [] {
    unsigned hc = std::thread::hardware_concurrency();
    int w = (int)std::min<unsigned>(hc > 1 ? hc - 1 : 0, 15);
    return new HostPool(w);  // lives until the process ends
  }();
  if (n == 1) {
    f(0);
    return;
  }
  pool->run(n, f);
}

}  // namespace bg

using namespace bg;

extern "C" int bg_init(int device) {
  return guard([&] {
    {
      std::lock_guard<std::mutex> lk(g_init_mu);
      if (g_ctx && g_ctx->device != device) {
        set_error("bg_init: already initialised on device %d", g_ctx->device);
        throw Error{BG_ERR_ARG};
      }
      g_device = device;
    }
    ensure_init();
  });
}

extern "C" const char* bg_last_error(void) { return g_err; }

extern "C" int bg_set_timing(int on) {
  return guard([&] { ctx().timing = on != 0; });
}

extern "C" int bg_timing_reset(void) {
  return guard([&] {
    timing_flush();
    std::lock_guard<std::mutex> lk(g_tmu);
    g_times.clear();
  });
}

extern "C" int bg_timing_report(char* names, double* avg_ms, uint64_t* counts, int cap) {
  int n = 0;
  int rc = guard([&] {
    timing_flush();
    std::lock_guard<std::mutex> lk(g_tmu);
    for (auto& kv : g_times) {
      if (n >= cap) break;
      if (names) {
        strncpy(names + 48 * n, kv.first.c_str(), 47);
        names[48 * n + 47] = 0;
      }
      if (avg_ms) avg_ms[n] = kv.second.total_ms / (double)kv.second.count;
      if (counts) counts[n] = kv.second.count;
      n++;
    }
  });
  return rc == BG_OK ? n : rc;
}
