// Device-tuned parameter table: see bg_tune.hpp.  Host-only.
#include "bg_tune.hpp"
#include "bg_runtime.hpp"
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

namespace bg {

namespace {

struct Table {
  std::map<std::tuple<int, int, int>, MsmTune> msm;  // (group, fixed, round(2 log2 n)) -> entry
  std::map<int, NttTune> ntt;
  std::string path;
  bool loaded = false;
};

std::mutex g_mu;
Table g_table;
std::atomic<uint64_t> g_generation{1};

std::string default_path() {
  const char* env = getenv("BG16_TUNE_FILE");
  if (env && env[0]) return env;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&default_path), &info) && info.dli_fname) {
    std::string so(info.dli_fname);
    size_t slash = so.rfind('/');
    return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tune_gfx950.txt";
  }
  return "tune_gfx950.txt";
}

int kv(const char* tok, const char* key) {
  size_t n = strlen(key);
  if (strncmp(tok, key, n) == 0 && tok[n] == '=') return atoi(tok + n + 1);
  return -1;
}

// Parses the table; a missing file is an empty table.  Malformed lines are errors.
Table parse(const std::string& path) {
  Table t;
  t.path = path;
  t.loaded = true;
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return t;
  char line[512];
  int lineno = 0;
  while (fgets(line, sizeof line, f)) {
    lineno++;
    if (char* h = strchr(line, '#')) *h = 0;
    char* save = nullptr;
    char* tok = strtok_r(line, " \t\r\n", &save);
    if (!tok) continue;
    if (strcmp(tok, "msm") == 0) {
      char* g = strtok_r(nullptr, " \t\r\n", &save);
      char* kind = strtok_r(nullptr, " \t\r\n", &save);
      char* lg = strtok_r(nullptr, " \t\r\n", &save);
      bool ok = g && kind && lg && (strcmp(kind, "windowed") == 0 || strcmp(kind, "fixed") == 0);
      int group = ok ? atoi(g) : 0;
      const double lv = ok ? strtod(lg, nullptr) : -1.0;
      const int l2 = (int)lround(2.0 * lv);  // half-octave key
      ok = ok && (group == 1 || group == 2) && lv >= 0.0 && lv <= 40.0;
      MsmTune e;
      while (ok && (tok = strtok_r(nullptr, " \t\r\n", &save))) {
        int v;
        if ((v = kv(tok, "c")) >= 0) e.c = v;
        else if ((v = kv(tok, "L")) >= 0) e.L = v;
        else if ((v = kv(tok, "m")) >= 0) e.m = v;
        else if ((v = kv(tok, "K")) >= 0) e.K = v;
        else ok = false;
      }
      ok = ok && (e.c == 0 || (e.c >= 2 && e.c <= 22)) && e.L >= 0 && e.L <= 64 &&
           (e.m == 0 || e.m == 2 || e.m == 4 || e.m == 8) && (e.K == 0 || (e.K >= 2 && e.K <= 256 && !(e.K & (e.K - 1))));
      if (!ok) {
        fclose(f);
        set_error("%s:%d: malformed msm line", path.c_str(), lineno);
        throw Error{BG_ERR_IO};
      }
      t.msm[std::make_tuple(group, strcmp(kind, "fixed") == 0 ? 1 : 0, l2)] = e;
    } else if (strcmp(tok, "ntt") == 0) {
      char* lg = strtok_r(nullptr, " \t\r\n", &save);
      int l2 = lg ? atoi(lg) : -1;
      bool ok = l2 >= 1 && l2 <= 28;
      NttTune e;
      while (ok && (tok = strtok_r(nullptr, " \t\r\n", &save))) {
        int v;
        if ((v = kv(tok, "tile")) >= 0) e.tile_log2 = v;
        else if ((v = kv(tok, "radix")) >= 0) e.radix = v;
        else ok = false;
      }
      ok = ok && (e.tile_log2 == 0 || (e.tile_log2 >= 9 && e.tile_log2 <= 11)) &&
           (e.radix == 0 || e.radix == 2 || e.radix == 4);
      if (!ok) {
        fclose(f);
        set_error("%s:%d: malformed ntt line", path.c_str(), lineno);
        throw Error{BG_ERR_IO};
      }
      t.ntt[l2] = e;
    } else {
      fclose(f);
      set_error("%s:%d: unknown line kind '%s'", path.c_str(), lineno, tok);
      throw Error{BG_ERR_IO};
    }
  }
  fclose(f);
  return t;
}

void ensure_loaded() {
  if (g_table.loaded) return;
  try {
    g_table = parse(default_path());
  } catch (const Error&) {
    g_table = Table();  // a malformed default table is ignored: the cost model plans
    g_table.loaded = true;
  }
}

}  // namespace

bool tune_msm(int group, bool fixed, uint64_t total_n, MsmTune& out) {
  if (total_n == 0) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  const int l2 = (int)lround(2.0 * log2((double)total_n));  // half-octave key
  auto it = g_table.msm.find(std::make_tuple(group, fixed ? 1 : 0, l2));
  if (it == g_table.msm.end()) return false;
  out = it->second;
  return true;
}

bool tune_ntt(int log2n, NttTune& out) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  auto it = g_table.ntt.find(log2n);
  if (it == g_table.ntt.end()) return false;
  out = it->second;
  return true;
}

uint64_t tune_generation() { return g_generation.load(); }

}  // namespace bg

using namespace bg;

extern "C" int bg_tune_load(const char* path) {
  return guard([&] {
    Table t = parse(path && path[0] ? std::string(path) : default_path());
    std::lock_guard<std::mutex> lk(g_mu);
    g_table = std::move(t);
    g_generation.fetch_add(1);
  });
}

extern "C" int bg_tune_lookup(int kind, int group, int fixed, uint64_t size, int* values4) {
  return guard([&] {
    BG_REQUIRE(values4 && (kind == 0 || kind == 1), BG_ERR_ARG, "bg_tune_lookup: bad argument");
    values4[0] = values4[1] = values4[2] = values4[3] = 0;
    if (kind == 0) {
      MsmTune e;
      if (!tune_msm(group, fixed != 0, size, e)) throw Error{1};
      values4[0] = e.c; values4[1] = e.L; values4[2] = e.m; values4[3] = e.K;
    } else {
      NttTune e;
      if (!tune_ntt((int)size, e)) throw Error{1};
      values4[0] = e.tile_log2; values4[1] = e.radix;
    }
  });
}
