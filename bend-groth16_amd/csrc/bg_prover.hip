// Resident-key Groth16 prover (C ABI: bg_pk_load, bg_pk_free, bg_pk_header,
// bg_pk_device_bytes, bg_prove, bg_proof_format).
//
// Device-side counterpart of bend2/prove.bend's Pr.run (:129-162): the FORMAT.md key is
// loaded once into HBM (packed 8 x u32 limbs), and each proof runs
//   upload z -> R1CS evaluation vectors (groth16.bend:162-208, with the libsnark
//   input-consistency rows) and canonical scalars (groth16.bend:210-219, 263-270)
//   -> qap_h (ntt.bend:444-459) -> five MSMs in the reference's order and slots
//   (prove.bend:64-116) -> proof combination on the host (groth16.bend:298-308).
#include "bg_runtime.hpp"
#include "bg_msm.hpp"
#include "bg_prover.hpp"
#include "bg_tune.hpp"
#include "bg_ntt.hpp"
#include "bg_field.hpp"
#include "bg_host_ec.hpp"
#include <atomic>
#include <chrono>
#include <memory>
#include <fstream>
#include <string>
#include <vector>
#include <unistd.h>

namespace bg {

// ------------------------------------------------------------------ FORMAT.md codec

static std::vector<uint32_t> read_words(const std::string& path, size_t expect_words = 0, bool exact = true) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  BG_REQUIRE(f.good(), BG_ERR_IO, "cannot open %s", path.c_str());
  std::streamsize sz = f.tellg();
  BG_REQUIRE(sz >= 0 && sz % 4 == 0, BG_ERR_IO, "%s: size %lld is not a multiple of 4", path.c_str(), (long long)sz);
  if (expect_words)
    BG_REQUIRE(exact ? (size_t)sz == expect_words * 4 : (size_t)sz >= expect_words * 4, BG_ERR_IO,
               "%s: %lld bytes, expected %zu", path.c_str(), (long long)sz, expect_words * 4);
  std::vector<uint32_t> w((size_t)sz / 4);
  f.seekg(0);
  f.read(reinterpret_cast<char*>(w.data()), sz);
  BG_REQUIRE(f.good() || sz == 0, BG_ERR_IO, "%s: read failed", path.c_str());
  return w;
}

// 16-bit-limb words -> packed words; rejects limbs >= 2^16 (bend2/tools/verify_bin.py:38)
static void pack_words(const uint32_t* in16, uint32_t* out32, size_t nout, const char* what) {
  uint32_t bad = 0;
  for (size_t i = 0; i < nout; i++) {
    uint32_t lo = in16[2 * i], hi = in16[2 * i + 1];
    bad |= (lo | hi) >> 16;
    out32[i] = lo | (hi << 16);
  }
  BG_REQUIRE(bad == 0, BG_ERR_IO, "%s: limb word >= 2^16", what);
}

static std::vector<uint32_t> packed_file(const std::string& path, size_t nwords16, const char* what) {
  auto w = read_words(path, nwords16);
  std::vector<uint32_t> out(nwords16 / 2);
  pack_words(w.data(), out.data(), out.size(), what);
  return out;
}

// Field moduli as little-endian 32-bit words: FORMAT.md stores x * 2^256 mod p, so every
// stored element is below its modulus.
static const uint32_t* const MOD_R = FrP::p;
static const uint32_t* const MOD_Q = FqP::p;

static bool below(const uint32_t* x, const uint32_t* p) {
  for (int i = 7; i >= 0; i--) {
    if (x[i] != p[i]) return x[i] < p[i];
  }
  return false;
}

// Index of the first of n packed elements (8 words each) that is not below p, or n.
static uint64_t first_unreduced(const uint32_t* x, uint64_t n, const uint32_t* p) {
  for (uint64_t i = 0; i < n; i++)
    if (!below(x + 8 * i, p)) return i;
  return n;
}

// y^2 = x^3 + b for packed affine Montgomery points (all-zero = infinity, accepted):
// G1 b = 3, G2 b = 3 / (9 + u) (the D-type twist).  Coordinates are checked below q
// first.  Returns the index of the first bad point, or n.
static uint64_t first_off_curve(int group, const uint32_t* pts, uint64_t n) {
  const int ew = group == 1 ? 8 : 16, pw = 2 * ew;
  const bgh::Fq b1 = bgh::Fq::to_mont(bgh::Fq::raw(3));
  const bgh::Fq2 xi = {bgh::Fq::to_mont(bgh::Fq::raw(9)), bgh::Fq::to_mont(bgh::Fq::raw(1))};
  const bgh::Fq2 b2 = xi.inv() * bgh::Fq2{b1, bgh::Fq::zero()};
  const uint64_t chunk = 1u << 13;
  const int nch = (int)((n + chunk - 1) / chunk);
  std::vector<uint64_t> bad(std::max(nch, 1), n);
  host_parallel(nch, [&](int t) {
    for (uint64_t i = (uint64_t)t * chunk; i < std::min<uint64_t>(n, (uint64_t)(t + 1) * chunk); i++) {
      const uint32_t* q = pts + (uint64_t)pw * i;
      bool zero = true;
      for (int k = 0; k < pw; k++) zero &= q[k] == 0;
      if (zero) continue;
      bool ok = true;
      for (int k = 0; k < pw / 8; k++) ok &= below(q + 8 * k, MOD_Q);
      if (ok) {
        if (group == 1) {
          const bgh::Fq x = bgh::Fq::from_u32(q), y = bgh::Fq::from_u32(q + 8);
          ok = y.sqr() == x.sqr() * x + b1;
        } else {
          const bgh::Fq2 x = bgh::Fq2::from_u32(q), y = bgh::Fq2::from_u32(q + 16);
          ok = y.sqr() == x.sqr() * x + b2;
        }
      }
      if (!ok) {
        bad[t] = i;
        return;
      }
    }
  });
  return *std::min_element(bad.begin(), bad.end());
}

struct Csr {
  uint32_t nnz = 0;
  DevArray offs, cols, coef;
};

}  // namespace bg

struct bg_pk {
  uint32_t hdr[8];  // nc, ni, nw, lg, na, nb, nh, nl
  bg::Csr A, B, C;
  bg::DevArray pa, pb1, pb2, ph, pl;
  uint32_t g1[3 * 16];  // alpha, beta, delta (packed affine)
  uint32_t g2[2 * 32];  // beta2, delta2
  // fixed-base tables: each query array holds Wd copies of its points, window w scaled by
  // 2^(w c); c for the G1 batch (A, B1, L), for H and for B2.  fixed = false keeps the
  // plain arrays (windowed MSMs).
  bool fixed = false;
  int c_g1 = 0, c_h = 0, c_b2 = 0;
  // 4-bit-window tables of delta (G1) and delta2 (G2): entry 16 w + j = j 2^(4w) delta,
  // XYZZ, for the device combination's r delta, s delta, rs delta and s delta2
  bg::DevArray dtab1, dtab2;
  uint64_t dev_bytes = 0;  // resident key bytes
  // slot used by bg_prove (created on first use; one bg_prove call per key at a time)
  std::unique_ptr<bg::ProveSlot> own;
  std::mutex own_mu;
  // proofs of this key queued in an engine or running (bg_pk_free refuses while > 0)
  std::atomic<int> inflight{0};
  // process-unique id (recorded launch sequences are keyed by it, not by the address,
  // which a later key may reuse)
  uint64_t serial = 0;
};

namespace bg {

// 64 x 16 XYZZ points j 2^(4w) P (w < 64, j < 16) of an affine point, packed words
template <class E>
static std::vector<uint32_t> window_table(const uint32_t* aff) {
  const int k = bgh::words<E>();
  std::vector<uint32_t> out((size_t)64 * 16 * 4 * k);
  bgh::Pt<E> base = bgh::load_affine<E>(aff);
  for (int w = 0; w < 64; w++) {
    bgh::Pt<E> acc = bgh::Pt<E>::inf();
    for (int j = 0; j < 16; j++) {
      uint32_t* o = &out[(size_t)(16 * w + j) * 4 * k];
      acc.X.to_u32(o);
      acc.Y.to_u32(o + k);
      acc.ZZ.to_u32(o + 2 * k);
      acc.ZZZ.to_u32(o + 3 * k);
      acc = acc.add(base);
    }
    base = base.dbl_n(4);
  }
  return out;
}

static void upload_array(DevArray& d, const void* src, size_t bytes, uint64_t& total) {
  d.alloc(bytes);
  if (bytes) BG_HIP(hipMemcpy(d.ptr, src, bytes, hipMemcpyHostToDevice));
  total += bytes;
}

// ------------------------------------------------------------------ device kernels

// Evaluation vector of one matrix over the 2^lg domain: row i < nc is the CSR row dot z
// (coefficient Montgomery-one skips the product, groth16.bend:108-116); for A the rows
// nc .. nc+ni-1 are the instance values z[i] (libsnark input consistency); zero elsewhere.
__global__ void k_evals(const uint32_t* __restrict__ offs, const uint32_t* __restrict__ cols,
                        const uint32_t* __restrict__ coef, const uint32_t* __restrict__ z, uint32_t nc, uint32_t ni,
                        int with_inputs, uint64_t d, uint32_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d) return;
  Fp acc = Fr::zero();
  if (i < nc) {
    const Fp one = Fr::one();
    for (uint32_t e = offs[i]; e < offs[i + 1]; e++) {
      Fp cf = load_fp(coef + 8ull * e);
      Fp zv = load_fp(z + 8ull * cols[e]);
      acc = Fr::add(acc, Fr::eq(cf, one) ? zv : Fr::mul(cf, zv));
    }
  } else if (with_inputs && i < (uint64_t)nc + ni) {
    acc = load_fp(z + 8 * (i - nc));
  }
  store_fp(out + 8 * i, acc);
}

__global__ void k_from_mont(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) store_fp(out + 8 * i, Fr::canon(Fr::from_mont(load_fp(in + 8 * i))));
}

static unsigned nb(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Order of qap_h and the MSMs inside one proof (see prove_on); BG16_PROVE_SCHEDULE=0..5,
// any other value keeps the default 0.
static int prove_schedule() {
  static int s = [] {
    const char* e = getenv("BG16_PROVE_SCHEDULE");
    const int v = e ? atoi(e) : 0;
    return v >= 0 && v <= 5 ? v : 0;
  }();
  return s;
}

// ------------------------------------------------------------------ host combination


static void canon_words(const uint32_t* mont_fr, uint64_t out[4]) {
  bgh::Fr x = bgh::Fr::from_u32(mont_fr).from_mont();
  memcpy(out, x.v, 32);
}

}  // namespace bg

using namespace bg;

// host side of a CSR file, validated
struct HostCsr {
  uint32_t nnz = 0;
  std::vector<uint32_t> offs, cols, coef;
};

static HostCsr read_csr(const std::string& path, uint32_t nc, uint64_t nz) {
  auto w = read_words(path);
  BG_REQUIRE(w.size() >= (size_t)nc + 2, BG_ERR_IO, "%s: truncated", path.c_str());
  HostCsr h;
  h.nnz = w[0];
  size_t expect = 1 + (size_t)nc + 1 + 17ull * h.nnz;
  BG_REQUIRE(w.size() == expect, BG_ERR_IO, "%s: %zu words, expected %zu for nnz=%u", path.c_str(), w.size(), expect,
             h.nnz);
  h.offs.assign(w.begin() + 1, w.begin() + 2 + nc);
  BG_REQUIRE(h.offs[0] == 0 && h.offs[nc] == h.nnz, BG_ERR_IO, "%s: bad row offsets", path.c_str());
  for (uint32_t i = 0; i < nc; i++)
    BG_REQUIRE(h.offs[i] <= h.offs[i + 1], BG_ERR_IO, "%s: row offsets decrease", path.c_str());
  h.cols.resize(h.nnz);
  h.coef.resize(8ull * h.nnz);
  const uint32_t* e = w.data() + 2 + nc;
  for (uint32_t k = 0; k < h.nnz; k++) {
    h.cols[k] = e[17ull * k];
    BG_REQUIRE(h.cols[k] < nz, BG_ERR_IO, "%s: column %u out of range (%llu wires)", path.c_str(), h.cols[k],
               (unsigned long long)nz);
    pack_words(e + 17ull * k + 1, &h.coef[8ull * k], 8, path.c_str());
    BG_REQUIRE(below(&h.coef[8ull * k], MOD_R), BG_ERR_IO, "%s: coefficient %u is not below r", path.c_str(), k);
  }
  return h;
}

extern "C" int bg_pk_load(const char* dir, bg_pk** out) {
  return guard([&] {
    BG_REQUIRE(dir && out, BG_ERR_ARG, "bg_pk_load: null argument");
    *out = nullptr;
    std::string d(dir);
    // 1. read and validate every file on the host (bend2/tools/verify_bin.py:24-64 checks)
    auto hdrw = read_words(d + "/header.bin", 8);
    auto pk = std::make_unique<bg_pk>();
    static std::atomic<uint64_t> next_serial{1};
    pk->serial = next_serial.fetch_add(1);
    memcpy(pk->hdr, hdrw.data(), 32);
    uint32_t nc = pk->hdr[0], ni = pk->hdr[1], nw = pk->hdr[2], lg = pk->hdr[3];
    uint32_t na = pk->hdr[4], nbq = pk->hdr[5], nh = pk->hdr[6], nl = pk->hdr[7];
    BG_REQUIRE(lg >= 1 && lg <= 28, BG_ERR_IO, "%s/header.bin: domain_log2 %u out of range", dir, lg);
    uint64_t dsz = 1ull << lg;
    uint64_t nz = (uint64_t)ni + nw;
    BG_REQUIRE(ni >= 1 && (uint64_t)nc + ni <= dsz && na <= nz && nbq <= nz && nh <= dsz && (uint64_t)nl + ni <= nz,
               BG_ERR_IO, "%s/header.bin: inconsistent sizes", dir);
    HostCsr ca = read_csr(d + "/r1cs_a.bin", nc, nz);
    HostCsr cb = read_csr(d + "/r1cs_b.bin", nc, nz);
    HostCsr cc = read_csr(d + "/r1cs_c.bin", nc, nz);
    auto g1 = packed_file(d + "/pk_g1.bin", 3 * 32, "pk_g1.bin");
    auto g2 = packed_file(d + "/pk_g2.bin", 2 * 64, "pk_g2.bin");
    BG_REQUIRE(first_off_curve(1, g1.data(), 3) == 3, BG_ERR_IO, "%s/pk_g1.bin: a point is not on G1", dir);
    BG_REQUIRE(first_off_curve(2, g2.data(), 2) == 2, BG_ERR_IO, "%s/pk_g2.bin: a point is not on the twist",
               dir);
    memcpy(pk->g1, g1.data(), sizeof pk->g1);
    memcpy(pk->g2, g2.data(), sizeof pk->g2);
    struct Q {
      const char* name;
      uint64_t n;
      int w16;
      DevArray* dst;
      std::vector<uint32_t> host;
    } qs[] = {{"pk_a.bin", na, 32, &pk->pa, {}},
              {"pk_b1.bin", nbq, 32, &pk->pb1, {}},
              {"pk_b2.bin", nbq, 64, &pk->pb2, {}},
              {"pk_h.bin", nh, 32, &pk->ph, {}},
              {"pk_l.bin", nl, 32, &pk->pl, {}}};
    for (auto& q : qs) {
      q.host = packed_file(d + "/" + q.name, q.n * q.w16, q.name);
      const int group = q.w16 == 32 ? 1 : 2;
      const uint64_t bad = first_off_curve(group, q.host.data(), q.n);
      BG_REQUIRE(bad == q.n, BG_ERR_IO, "%s/%s: point %llu is not on %s", dir, q.name, (unsigned long long)bad,
                 group == 1 ? "G1" : "the twist");
    }

    // 2. upload into HBM
    Ctx& c = ctx();
    std::lock_guard<std::mutex> lk(c.mu);
    uint64_t total = 0;
    for (auto pr : {std::make_pair(&ca, &pk->A), std::make_pair(&cb, &pk->B), std::make_pair(&cc, &pk->C)}) {
      pr.second->nnz = pr.first->nnz;
      upload_array(pr.second->offs, pr.first->offs.data(), 4ull * (nc + 1), total);
      upload_array(pr.second->cols, pr.first->cols.data(), 4ull * pr.first->nnz, total);
      upload_array(pr.second->coef, pr.first->coef.data(), 32ull * pr.first->nnz, total);
    }
    const char* fb = getenv("BG16_FIXED_BASE");
    pk->fixed = !(fb && fb[0] == '0');
    if (!pk->fixed) {
      for (auto& q : qs) upload_array(*q.dst, q.host.data(), q.host.size() * 4, total);
    } else {
      // tables built on the device from the uploaded points (window 0 is the point itself)
      const uint64_t g1n[3] = {na, nbq, nl};
      pk->c_g1 = fixed_window_bits(g1n, 3, 1);
      pk->c_h = fixed_window_bits(&qs[3].n, 1, 1);
      pk->c_b2 = fixed_window_bits(&qs[2].n, 1, 2);
      const int cs[5] = {pk->c_g1, pk->c_g1, pk->c_b2, pk->c_h, pk->c_g1};
      for (int k = 0; k < 5; k++) {
        auto& q = qs[k];
        const int group = k == 2 ? 2 : 1;
        const uint64_t wd = (255 + cs[k] - 1) / cs[k];
        DevArray plain;
        upload_array(plain, q.host.data(), q.host.size() * 4, total);
        total -= q.host.size() * 4;
        q.dst->alloc(std::max<uint64_t>(wd * q.host.size() * 4, 4));
        total += wd * q.host.size() * 4;
        msm_build_fixed_table(group, plain.as<uint32_t>(), q.n, cs[k], q.dst->as<uint32_t>(), c.stream);
        BG_HIP(hipStreamSynchronize(c.stream));
      }
    }
    {
      auto t1 = window_table<bgh::Fq>(pk->g1 + 32);
      auto t2 = window_table<bgh::Fq2>(pk->g2 + 32);
      upload_array(pk->dtab1, t1.data(), t1.size() * 4, total);
      upload_array(pk->dtab2, t2.data(), t2.size() * 4, total);
    }
    pk->dev_bytes = total;
    (void)ntt_domain((int)lg);  // twiddle / scale tables resident with the key
    *out = pk.release();
  });
}

extern "C" int bg_pk_free(bg_pk* pk) {
  return guard([&] {
    if (!pk) return;
    const int busy = pk->inflight.load();
    BG_REQUIRE(busy == 0, BG_ERR_ARG, "bg_pk_free: %d proofs of this key are queued or running", busy);
    {
      std::lock_guard<std::mutex> lk(pk->own_mu);
      pk->own.reset();
    }
    delete pk;
  });
}

extern "C" int bg_pk_header(const bg_pk* pk, uint32_t* header8) {
  return guard([&] {
    BG_REQUIRE(pk && header8, BG_ERR_ARG, "bg_pk_header: null argument");
    memcpy(header8, pk->hdr, 32);
  });
}

extern "C" int bg_pk_windows(const bg_pk* pk, int* c3) {
  return guard([&] {
    BG_REQUIRE(pk && c3, BG_ERR_ARG, "bg_pk_windows: null argument");
    c3[0] = pk->fixed ? pk->c_g1 : 0;
    c3[1] = pk->fixed ? pk->c_h : 0;
    c3[2] = pk->fixed ? pk->c_b2 : 0;
  });
}

extern "C" int bg_pk_device_bytes(const bg_pk* pk, uint64_t* bytes) {
  return guard([&] {
    BG_REQUIRE(pk && bytes, BG_ERR_ARG, "bg_pk_device_bytes: null argument");
    *bytes = pk->dev_bytes;
  });
}

namespace bg {

ProveSlot::ProveSlot() {
  (void)ctx();  // device current on this thread
  for (auto& x : s) BG_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
  for (auto& e : ev) BG_HIP(hipEventCreate(&e));
  for (auto& e : join) BG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  graphs.reserve(MAX_GRAPHS);  // entries never move while a pointer to one is in use
}

ProveSlot::~ProveSlot() {
  release_buffers();
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : join)
    if (e) (void)hipEventDestroy(e);
  for (auto& x : s)
    if (x) (void)hipStreamDestroy(x);
}

const ProveSlot::Graph* ProveSlot::find_graph(uint64_t key) {
  for (auto& g : graphs)
    if (g.key == key) {
      g.used = ++graph_clock;
      for (int i = 0; i < 3; i++) {
        work[i].plan = g.plan[i];
        work[i].group = g.group[i];
      }
      return &g;
    }
  return nullptr;
}

static void destroy_graph(ProveSlot::Graph& g) {
  for (auto& e : g.exec)
    if (e) (void)hipGraphExecDestroy(e);
}

const ProveSlot::Graph* ProveSlot::add_graph(uint64_t key, const hipGraphExec_t exec[3]) {
  if (graphs.size() >= MAX_GRAPHS) {
    auto lru = graphs.begin();
    for (auto it = graphs.begin(); it != graphs.end(); ++it)
      if (it->used < lru->used) lru = it;
    destroy_graph(*lru);
    graphs.erase(lru);
  }
  Graph g;
  g.key = key;
  for (int i = 0; i < 3; i++) g.exec[i] = exec[i];
  g.used = ++graph_clock;
  for (int i = 0; i < 3; i++) {
    g.plan[i] = work[i].plan;
    g.group[i] = work[i].group;
  }
  graphs.push_back(g);
  return &graphs.back();
}

void ProveSlot::drop_graphs() {
  for (auto& g : graphs) destroy_graph(g);
  graphs.clear();
  warm.clear();
  failed_key = 0;
}

bool ProveSlot::is_warm(uint64_t key) const {
  for (uint64_t k : warm)
    if (k == key) return true;
  return false;
}

void ProveSlot::add_warm(uint64_t key) {
  if (is_warm(key)) return;
  if (warm.size() >= 2 * MAX_GRAPHS) warm.erase(warm.begin());
  warm.push_back(key);
}

static uint64_t mix64(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
  return h * 0xff51afd7ed558ccdULL;
}

uint64_t ProveSlot::buffer_signature() const {
  uint64_t sig = 0x6a09e667f3bcc909ULL;
  for (const DevArray* a : {&z, &zc, &ea, &eb, &ec, &h, &tmp}) sig = mix64(mix64(sig, (uint64_t)a->ptr), a->bytes);
  for (const MsmWork& w : work) {
    for (const DevBuf* b : w.bufs()) sig = mix64(mix64(sig, (uint64_t)b->ptr), b->cap);
    for (const HostBuf* b : {&w.host_out, &w.host_ptrs, &w.host_pub}) sig = mix64(mix64(sig, (uint64_t)b->ptr), b->cap);
    sig = mix64(mix64(sig, (uint64_t)w.done), (uint64_t)w.mid);
  }
  return sig;
}

void ProveSlot::ensure(uint64_t nz, uint64_t d) {
  if (nz > cap_nz) {
    z.alloc(32 * nz);
    zc.alloc(32 * nz);
    cap_nz = nz;
  }
  if (d > cap_d) {
    for (DevArray* a : {&ea, &eb, &ec, &h, &tmp}) a->alloc(32 * d);
    cap_d = d;
  }
}

void ProveSlot::release_buffers() {
  drop_graphs();
  for (DevArray* a : {&z, &zc, &ea, &eb, &ec, &h, &tmp, &comb}) a->release();
  zpin.release();
  proofpin.release();
  cap_nz = cap_d = 0;
  for (auto& w : work) w.release();
}

uint64_t ProveSlot::device_bytes() const {
  uint64_t t = 64 * cap_nz + 160 * cap_d + comb.bytes;
  for (auto& w : work) t += w.bytes();
  return t;
}

void pk_acquire(bg_pk* pk) { pk->inflight.fetch_add(1); }
void pk_release(bg_pk* pk) { pk->inflight.fetch_sub(1); }

uint64_t prove_bytes_estimate(const bg_pk* pk) {
  const uint32_t ni = pk->hdr[1], nw = pk->hdr[2], lg = pk->hdr[3];
  const uint64_t na = pk->hdr[4], nbq = pk->hdr[5], nh = pk->hdr[6], nl = pk->hdr[7];
  const uint64_t nz = (uint64_t)ni + nw, d = 1ull << lg;
  uint64_t g1n[3] = {na, nbq, nl};
  uint64_t t = 64 * nz + 160 * d;
  if (pk->fixed) {
    t += msm_work_bytes(msm_plan_fixed(g1n, 3, 1, pk->c_g1), 1);
    t += msm_work_bytes(msm_plan_fixed(&nh, 1, 1, pk->c_h), 1);
    t += msm_work_bytes(msm_plan_fixed(&nbq, 1, 2, pk->c_b2), 2);
  } else {
    t += msm_work_bytes(msm_plan_batch(g1n, 3, 1, 0), 1);
    t += msm_work_bytes(msm_plan(nh, 1, 0), 1);
    t += msm_work_bytes(msm_plan(nbq, 2, 0), 2);
  }
  return t;
}

}  // namespace bg

extern "C" int bg_prove(bg_pk* pk, const uint32_t* witness, const uint32_t* rs, uint32_t* proof_out, int flags,
                        double* t_ms) {
  return guard([&] {
    BG_REQUIRE(pk && witness && rs && proof_out, BG_ERR_ARG, "bg_prove: null argument");
    pk_acquire(pk);
    struct Release {
      bg_pk* k;
      ~Release() { pk_release(k); }
    } rel{pk};
    std::lock_guard<std::mutex> lk(pk->own_mu);
    if (!pk->own) pk->own = std::make_unique<ProveSlot>();
    prove_on(pk, *pk->own, witness, rs, proof_out, flags, t_ms);
  });
}

namespace bg {

// Witness of nz packed elements into device memory z on stream s.  A host witness is
// packed (or copied) into the pinned staging buffer zp by the host worker pool, then
// moved by one DMA: a pageable source goes through the runtime's staged copy (~2.6 GB/s
// for the 16.8 MB words16 witness of a 2^18 proof).  zp must stay untouched until the
// copy has run.
static void upload_witness(const uint32_t* witness, int flags, uint64_t nz, uint32_t* z, uint32_t* zp,
                           hipStream_t s) {
  const bool words16 = (flags & 1) == BG_FMT_WORDS16;
  if (flags & BG_MEM_DEVICE) {
    BG_HIP(hipMemcpyAsync(z, witness, 32 * nz, hipMemcpyDeviceToDevice, s));
    return;
  }
  const uint64_t chunk = 1ull << 14;
  const int nch = (int)((nz + chunk - 1) / chunk);
  // FORMAT.md witness values are Montgomery Fr below r; a 16-bit word with high bits set
  // or a value at or above r is rejected (checked on the packed staging copy)
  std::vector<uint64_t> bad(std::max(nch, 1), nz);
  std::vector<uint32_t> wide(std::max(nch, 1), 0);
  host_parallel(nch, [&](int t) {
    const uint64_t lo = (uint64_t)t * chunk, hi = std::min<uint64_t>(nz, lo + chunk);
    if (words16) {
      uint32_t hb = 0;
      for (uint64_t i = 8 * lo; i < 8 * hi; i++) {
        const uint32_t a = witness[2 * i], b = witness[2 * i + 1];
        hb |= (a | b) >> 16;
        zp[i] = a | (b << 16);
      }
      wide[t] = hb;
    } else {
      memcpy(zp + 8 * lo, witness + 8 * lo, 32 * (hi - lo));
    }
    const uint64_t f = first_unreduced(zp + 8 * lo, hi - lo, MOD_R);
    if (f < hi - lo) bad[t] = lo + f;
  });
  for (uint32_t w : wide) BG_REQUIRE(w == 0, BG_ERR_ARG, "witness: limb word >= 2^16");
  const uint64_t b = *std::min_element(bad.begin(), bad.end());
  BG_REQUIRE(b == nz, BG_ERR_ARG, "witness: element %llu is not below r", (unsigned long long)b);
  BG_HIP(hipMemcpyAsync(z, zp, 32 * nz, hipMemcpyHostToDevice, s));
}

// r then s (FORMAT.md rs.bin) as packed words, each checked below r
static void read_rs(const uint32_t* rs, int flags, uint32_t* rs_packed) {
  if ((flags & 1) == BG_FMT_WORDS16)
    pack_words(rs, rs_packed, 16, "rs");
  else
    memcpy(rs_packed, rs, 64);
  BG_REQUIRE(first_unreduced(rs_packed, 2, MOD_R) == 2, BG_ERR_ARG, "rs: r or s is not below the modulus");
}

void check_host_inputs(const uint32_t* witness, const uint32_t* rs, int flags, uint64_t nz) {
  uint32_t tmp[16];
  read_rs(rs, flags, tmp);
  if (flags & BG_MEM_DEVICE) return;
  const bool words16 = (flags & 1) == BG_FMT_WORDS16;
  uint32_t e[8];
  for (uint64_t i = 0; i < nz; i++) {
    const uint32_t* x = witness + (words16 ? 16 : 8) * i;
    if (words16) {
      for (int k = 0; k < 8; k++) {
        BG_REQUIRE(((x[2 * k] | x[2 * k + 1]) >> 16) == 0, BG_ERR_ARG, "witness: limb word >= 2^16");
        e[k] = x[2 * k] | (x[2 * k + 1] << 16);
      }
      x = e;
    }
    BG_REQUIRE(below(x, MOD_R), BG_ERR_ARG, "witness: element %llu is not below r", (unsigned long long)i);
  }
}

// Single proofs replay recorded HIP graphs of their device work (BG16_GRAPH: 0 never,
// 1 always; unset: domains up to 2^GRAPH_DEFAULT_MAX_LOG2).  Measured per proof, eager vs
// replayed: K=10 0.99 / 0.98 ms, K=12 1.14 / 1.11, K=14 1.35 / 1.32, K=16 1.85 / 1.93,
// K=18 4.02 / 4.17 -- small proofs gain the host's launch work, large ones lose a little
// to the graphs' own dispatch.
static constexpr uint32_t GRAPH_DEFAULT_MAX_LOG2 = 15;
static bool graph_wanted(uint32_t lg) {
  const char* e = getenv("BG16_GRAPH");
  if (e && e[0]) return atoi(e) != 0;
  return lg <= GRAPH_DEFAULT_MAX_LOG2;
}

// Everything a recorded proof depends on besides its inputs: the key, the schedule, the
// tuning table, every BG16_ setting (window, slice and sort overrides change the launch
// sequence) and the addresses and sizes of the slot's buffers.
static uint64_t graph_key(const bg_pk* pk, const ProveSlot& sl, int sched) {
  uint64_t k = mix64(mix64(mix64(0x243f6a8885a308d3ULL, pk->serial), (uint64_t)sched), tune_generation());
  for (char** e = environ; e && *e; e++)
    if (strncmp(*e, "BG16_", 5) == 0)
      for (const char* c = *e; *c; c++) k = mix64(k, (uint64_t)(unsigned char)*c);
  return mix64(k, sl.buffer_signature());
}

// Records fn's launches on the slot's streams as one graph (s[0] forks the side streams;
// they are joined back before the recording ends) and instantiates it.  Nothing runs
// while recording.  Returns null, with the streams usable, if the recording fails; the
// proof then runs eagerly (and reports any real error from there).
static std::atomic<uint64_t> g_graphs_recorded{0}, g_graph_launches{0}, g_graph_failures{0};

// Records fn's launches on stream st as a linear graph and instantiates it (nothing runs
// while recording); null if recording fails, with st usable again.
template <class Fn>
static hipGraphExec_t record_linear(hipStream_t st, Fn&& fn) {
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  hipGraph_t g = nullptr;
  bool ok = true;
  try {
    fn();
  } catch (const Error&) {
    ok = false;
  }
  if (hipStreamEndCapture(st, &g) != hipSuccess) ok = false;
  hipGraphExec_t exec = nullptr;
  if (ok && g && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) exec = nullptr;
  if (g) (void)hipGraphDestroy(g);
  return ok ? exec : nullptr;
}

// Records schedule 0's three per-stream sequences (B2 on s[3], the G1 batch on s[1], the
// rest on s[0]) for key.  On failure nothing is kept, the key is not tried again on this
// slot, and the proof runs eagerly (which reports any real error).
template <class F3, class F1, class F0>
static const ProveSlot::Graph* record_graphs(ProveSlot& sl, uint64_t key, F3&& on_s3, F1&& on_s1, F0&& on_s0) {
  hipGraphExec_t ex[3] = {record_linear(sl.s[3], on_s3), record_linear(sl.s[1], on_s1), record_linear(sl.s[0], on_s0)};
  (void)hipGetLastError();
  if (!ex[0] || !ex[1] || !ex[2]) {
    for (auto e : ex)
      if (e) (void)hipGraphExecDestroy(e);
    sl.failed_key = key;
    g_graph_failures.fetch_add(1);
    return nullptr;
  }
  g_graphs_recorded.fetch_add(1);
  return sl.add_graph(key, ex);
}

void prove_on(bg_pk* pk, ProveSlot& sl, const uint32_t* witness, const uint32_t* rs, uint32_t* proof_out, int flags,
              double* t_ms) {
  {
    BG_REQUIRE(pk && witness && rs && proof_out, BG_ERR_ARG, "bg_prove: null argument");
    bool words16 = (flags & 1) == BG_FMT_WORDS16;
    bool dev_witness = (flags & BG_MEM_DEVICE) != 0;
    BG_REQUIRE(!(dev_witness && words16), BG_ERR_ARG, "bg_prove: a device witness must be packed");
    (void)ctx();  // device current on this thread
    auto t0 = std::chrono::steady_clock::now();
    uint32_t rs_packed[16];
    read_rs(rs, flags, rs_packed);
    const uint32_t nc = pk->hdr[0], ni = pk->hdr[1], nw = pk->hdr[2], lg = pk->hdr[3];
    const uint32_t na = pk->hdr[4], nbq = pk->hdr[5], nh = pk->hdr[6], nl = pk->hdr[7];
    const uint64_t d = 1ull << lg, nz = (uint64_t)ni + nw;
    // Streams: s0 runs the witness path (upload, evaluation vectors, canonical scalars),
    // qap_h and then MSM H; MSMs A+L, B1 and B2 each run on their own stream once the
    // canonical scalars exist, so the latency-bound tails of one MSM overlap with the
    // accumulation of the others.
    hipStream_t s0 = sl.s[0], s1 = sl.s[1], s2 = sl.s[2], s3 = sl.s[3];
    const int sched = prove_schedule();
    if (sched == 5) s1 = s2 = s3 = s0;  // one stream: every kernel alone (profiling)
    hipEvent_t* ev = sl.ev;
    sl.ensure(nz, d);
    // r, s and r s canonical: the scalars of the combination's delta terms and products
    uint64_t rc[4], sc[4], rsc[4];
    canon_words(rs_packed, rc);
    canon_words(rs_packed + 8, sc);
    {
      bgh::Fr rs_m = bgh::Fr::from_u32(rs_packed) * bgh::Fr::from_u32(rs_packed + 8);
      bgh::Fr x = rs_m.from_mont();
      memcpy(rsc, x.v, 32);
    }
    // the combination on the device (BG_COMBINE_DEVICE): its delta terms need only r and s,
    // so they run from the start beside everything else
    const bool dev_comb = (flags & BG_COMBINE_DEVICE) != 0;
    BG_REQUIRE(!dev_comb || pk->fixed, BG_ERR_ARG, "bg_prove: the device combination needs fixed-base tables");
    CombineArgs ca{};
    uint32_t* cb = nullptr;
    if (dev_comb) {
      if (!sl.comb.ptr) sl.comb.alloc(4 * comb_words());
      cb = sl.comb.as<uint32_t>();
      memcpy(ca.alpha1, pk->g1, 64);
      memcpy(ca.beta1, pk->g1 + 16, 64);
      memcpy(ca.delta1, pk->g1 + 32, 64);
      memcpy(ca.beta2, pk->g2, 128);
      memcpy(ca.delta2, pk->g2 + 32, 128);
      memcpy(ca.sc, rc, 32);
      memcpy(ca.sc + 8, sc, 32);
      memcpy(ca.sc + 16, rsc, 32);
      ca.dtab1 = pk->dtab1.as<uint32_t>();
      ca.dtab2 = pk->dtab2.as<uint32_t>();
      launch_delta_terms(ca, cb, s2);
      BG_HIP(hipEventRecord(ev[12], s2));
    }
    BG_HIP(hipEventRecord(ev[0], s0));
    // 1. witness upload (a host witness is packed by the worker pool into pinned staging
    //    and sent in one copy; a device witness is copied in place)
    uint32_t* z = sl.z.as<uint32_t>();
    upload_witness(witness, flags, nz, z, static_cast<uint32_t*>(sl.zpin.get(32 * nz)), s0);
    BG_HIP(hipEventRecord(ev[1], s0));
    // 2.-4. the device work of the proof after the upload: evaluation vectors, canonical
    //    scalars, qap_h and the five MSMs, each MSM's result copied to pinned host memory
    //    and its event recorded.  Run eagerly, or recorded once as a HIP graph and replayed
    //    (one launch instead of ~100 kernel launches and the events between streams).
    // 2.-4. The device work after the upload.  `head` makes the canonical scalars and forks
    //    the side streams from them; then each stream has its own launch sequence: B2 on s3,
    //    the G1 batch A, B1, L on s1, and the evaluation vectors, qap_h and H on s0.  The
    //    G1 MSMs other than H run as one batch (one sort, accumulation and reduction over
    //    all their windows); H needs qap_h's output.
    //    Schedules (BG16_PROVE_SCHEDULE; every one gives the same proof):
    //   0: B2 on s3 and the G1 batch A, B1, L on s1 start at once; qap_h then H on s0
    //      (default: the shortest from 2^12 to 2^18)
    //   1: B2 on s3 at once; qap_h then the G1 batch A, B1, L, H on s0
    //   2: as 0, but the G1 batch waits until B2's buckets are complete
    //   3: as 0, but the G1 batch waits for qap_h
    //   4: qap_h first; then B2 on s3, the G1 batch A, B1, L on s1 and H on s0
    //   5: B2, qap_h, the G1 batch and H on a single stream (no overlap: per-kernel times
    //      for profiling)
    //    Schedule 0's three per-stream sequences are recorded as linear HIP graphs on the
    //    second proof of a key on this slot and replayed after that (one launch per stream
    //    instead of ~100 kernel launches).  A graph that forks and joins streams itself
    //    was measured slower: this runtime's launch of it waits on the host for its
    //    branches (~1.2 ms per launch of a 2^10 proof).
    uint32_t* zc = sl.zc.as<uint32_t>();
    const uint32_t* g1p[4] = {pk->pa.as<uint32_t>(), pk->pb1.as<uint32_t>(), pk->pl.as<uint32_t>(),
                              pk->ph.as<uint32_t>()};
    const uint32_t* g1s[4] = {zc, zc, zc + 8 * ni, sl.h.as<uint32_t>()};
    const uint64_t g1n[4] = {na, nbq, nl, nh};
    auto head = [&] {
      ScopedKTime kt("evals", s0);
      k_from_mont<<<nb(nz, 256), 256, 0, s0>>>(z, zc, nz);
      BG_HIP(hipGetLastError());
      BG_HIP(hipEventRecord(ev[2], s0));  // canonical scalars ready
      for (hipStream_t sx : {s1, s2, s3}) BG_HIP(hipStreamWaitEvent(sx, ev[2], 0));
    };
    auto run_evals = [&] {
      {
        ScopedKTime kt("evals", s0);
        k_evals<<<nb(d, 256), 256, 0, s0>>>(pk->A.offs.as<uint32_t>(), pk->A.cols.as<uint32_t>(),
                                            pk->A.coef.as<uint32_t>(), z, nc, ni, 1, d, sl.ea.as<uint32_t>());
        k_evals<<<nb(d, 256), 256, 0, s0>>>(pk->B.offs.as<uint32_t>(), pk->B.cols.as<uint32_t>(),
                                            pk->B.coef.as<uint32_t>(), z, nc, ni, 0, d, sl.eb.as<uint32_t>());
        k_evals<<<nb(d, 256), 256, 0, s0>>>(pk->C.offs.as<uint32_t>(), pk->C.cols.as<uint32_t>(),
                                            pk->C.coef.as<uint32_t>(), z, nc, ni, 0, d, sl.ec.as<uint32_t>());
        BG_HIP(hipGetLastError());
      }
      record_if_eager(ev[3], s0);
    };
    auto run_qap = [&] {
      ScopedKTime kt("qap_h", s0);
      qap_h_run(sl.ea.as<uint32_t>(), sl.eb.as<uint32_t>(), sl.ec.as<uint32_t>(), sl.h.as<uint32_t>(),
                sl.tmp.as<uint32_t>(), (int)lg, true, s0);
      record_if_eager(ev[4], s0);
    };
    auto run_b2 = [&] {
      record_if_eager(ev[9], s3);
      if (pk->fixed) {
        const uint32_t* tb = pk->pb2.as<uint32_t>();
        const uint64_t nn = nbq;
        msm_enqueue_fixed(2, 1, &tb, &zc, &nn, pk->c_b2, s3, sl.work[2]);
      } else {
        msm_enqueue(2, pk->pb2.as<uint32_t>(), zc, nbq, 0, s3, sl.work[2]);
      }
      record_if_eager(ev[10], s3);
    };
    auto run_g1_batch3 = [&] {
      record_if_eager(ev[5], s1);
      if (pk->fixed)
        msm_enqueue_fixed(1, 3, g1p, g1s, g1n, pk->c_g1, s1, sl.work[0]);
      else
        msm_enqueue_batch(1, 3, g1p, g1s, g1n, 0, s1, sl.work[0]);
      record_if_eager(ev[6], s1);
    };
    auto run_h = [&] {
      if (pk->fixed) {
        const uint32_t* tb = pk->ph.as<uint32_t>();
        const uint32_t* hs = sl.h.as<uint32_t>();
        const uint64_t nn = nh;
        msm_enqueue_fixed(1, 1, &tb, &hs, &nn, pk->c_h, s0, sl.work[1]);
      } else {
        msm_enqueue(1, pk->ph.as<uint32_t>(), sl.h.as<uint32_t>(), nh, 0, s0, sl.work[1]);
      }
    };
    auto run_s0 = [&] {  // schedule 0's sequence on s0
      run_evals();
      run_qap();
      run_h();
      record_if_eager(ev[11], s0);
    };
    auto enqueue_device = [&] {
      head();
      if (sched == 0) {
        run_b2();
        run_g1_batch3();
        run_s0();
        return;
      }
      run_evals();
      if (sched == 4) {
        run_qap();
        BG_HIP(hipStreamWaitEvent(s3, ev[4], 0));
        BG_HIP(hipStreamWaitEvent(s1, ev[4], 0));
        run_b2();
        run_g1_batch3();
      } else {
        run_b2();
        if (sched == 2) {
          BG_HIP(hipStreamWaitEvent(s1, sl.work[2].mid, 0));
          run_g1_batch3();
        }
        run_qap();
        if (sched == 3 || sched == 5) {
          BG_HIP(hipStreamWaitEvent(s1, ev[4], 0));
          run_g1_batch3();
        }
      }
      if (sched != 1) {
        run_h();
      } else {
        record_if_eager(ev[5], s0);
        if (pk->fixed) {
          // schedule 1 batches H with A, B1, L: its table has its own width, so H runs alone
          msm_enqueue_fixed(1, 3, g1p, g1s, g1n, pk->c_g1, s0, sl.work[0]);
          const uint32_t* tb = pk->ph.as<uint32_t>();
          const uint32_t* hs = sl.h.as<uint32_t>();
          const uint64_t nn = nh;
          msm_enqueue_fixed(1, 1, &tb, &hs, &nn, pk->c_h, s0, sl.work[1]);
        } else {
          msm_enqueue_batch(1, 4, g1p, g1s, g1n, 0, s0, sl.work[0]);
        }
        record_if_eager(ev[6], s0);
      }
      record_if_eager(ev[11], s0);
    };
    const bool graph_ok = !dev_comb && sched == 0 && !ctx().timing && graph_wanted(lg);
    bool graphed = false;
    if (graph_ok) {
      const uint64_t key = graph_key(pk, sl, sched);
      const ProveSlot::Graph* g = sl.find_graph(key);
      if (!g && sl.is_warm(key) && sl.failed_key != key) g = record_graphs(sl, key, run_b2, run_g1_batch3, run_s0);
      if (g) {
        head();
        sl.work[0].arm_publish(s1);  // G1 batch
        sl.work[1].arm_publish(s0);  // H
        sl.work[2].arm_publish(s3);  // B2
        BG_HIP(hipGraphLaunch(g->exec[0], s3));
        BG_HIP(hipGraphLaunch(g->exec[1], s1));
        BG_HIP(hipGraphLaunch(g->exec[2], s0));
        g_graph_launches.fetch_add(1);
        graphed = true;
      }
    }
    if (!graphed) {
      enqueue_device();
      if (graph_ok) sl.add_warm(graph_key(pk, sl, sched));  // the buffers now fit this key's proofs
    }

    if (dev_comb) {
      // A, B1, L finished and combined once the G1 batch is reduced (its stream), B2 on its
      // stream, C after MSM H on s0; only the affine proof comes back
      auto shape = [](const MsmWork& w, uint32_t& hn, int& log2m) {
        hn = w.plan.H;
        log2m = 0;
        while ((1u << log2m) < w.plan.m * w.plan.K) log2m++;
      };
      shape(sl.work[0], ca.hn_g1, ca.log2m_g1);
      shape(sl.work[1], ca.hn_h, ca.log2m_h);
      shape(sl.work[2], ca.hn_b2, ca.log2m_b2);
      BG_REQUIRE(sl.work[0].plan.W == 1 && sl.work[1].plan.W == 1 && sl.work[2].plan.W == 1 &&
                     sl.work[0].plan.M == 3, BG_ERR_INTERNAL, "bg_prove: unexpected MSM shape for the device combination");
      const hipStream_t sg1 = sched == 1 ? s0 : s1;  // the G1 batch's stream
      BG_HIP(hipStreamWaitEvent(sg1, ev[12], 0));
      launch_combine_g1(ca, static_cast<const uint32_t*>(sl.work[0].outPQ.ptr), cb, sg1);
      BG_HIP(hipEventRecord(ev[13], sg1));  // C0 ready
      launch_affine_a(cb, sg1);             // A's normalisation, off C's path
      BG_HIP(hipEventRecord(ev[7], sg1));
      BG_HIP(hipStreamWaitEvent(s3, ev[12], 0));
      launch_combine_b2(ca, static_cast<const uint32_t*>(sl.work[2].outPQ.ptr), cb, s3);
      BG_HIP(hipEventRecord(ev[14], s3));
      BG_HIP(hipStreamWaitEvent(s0, ev[13], 0));
      BG_HIP(hipStreamWaitEvent(s0, ev[14], 0));
      launch_combine_final(ca, static_cast<const uint32_t*>(sl.work[1].outPQ.ptr), cb, s0);
      BG_HIP(hipStreamWaitEvent(s0, ev[7], 0));
      uint32_t* pin = static_cast<uint32_t*>(sl.proofpin.get(256));
      BG_HIP(hipMemcpyAsync(pin, cb + comb_proof_offset(), 256, hipMemcpyDeviceToHost, s0));
      BG_HIP(hipEventRecord(ev[15], s0));
      BG_HIP(hipEventSynchronize(ev[15]));
      uint32_t proof[64];
      memcpy(proof, pin, sizeof proof);
      if (words16) {
        for (int i = 0; i < 64; i++) {
          proof_out[2 * i] = proof[i] & 0xFFFFu;
          proof_out[2 * i + 1] = proof[i] >> 16;
        }
      } else {
        memcpy(proof_out, proof, sizeof proof);
      }
      if (t_ms) {
        const int pairs[8][2] = {{0, 1}, {1, 3}, {3, 4}, {5, 6}, {5, 6}, {9, 10}, {4, 11}, {5, 6}};
        for (int i = 0; i < 8; i++) {
          float e = 0;
          BG_HIP(hipEventElapsedTime(&e, ev[pairs[i][0]], ev[pairs[i][1]]));
          t_ms[i] = e;
        }
        t_ms[8] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      }
      return;
    }

    // host work that depends only on r, s overlaps with the device
    bgh::P1 alpha = bgh::load_affine<bgh::Fq>(pk->g1), beta1 = bgh::load_affine<bgh::Fq>(pk->g1 + 16),
       delta1 = bgh::load_affine<bgh::Fq>(pk->g1 + 32);
    bgh::P2 beta2 = bgh::load_affine<bgh::Fq2>(pk->g2), delta2 = bgh::load_affine<bgh::Fq2>(pk->g2 + 32);
    bgh::P1 r_d1 = delta1.mul(rc), s_d1 = delta1.mul(sc), rs_d1 = delta1.mul(rsc);
    bgh::P2 s_d2 = delta2.mul(sc);

    // 5. finish the MSMs and combine (groth16.bend:298-308); slots as prove.bend:64-116.
    //    Everything but the H term is combined while MSM H still runs on the device: B2
    //    first (its stream finishes first), then the G1 batch, whose s A and r B1 run on
    //    two workers.  MSM results stay in XYZZ form; only A, B2 and C are made affine.
    uint32_t g1x[4 * 32], b2x[64], hx[32];
    uint32_t proof[64];
    msm_finish(sl.work[2], nullptr, b2x);
    bgh::P2 B2 = bgh::load_xyzz<bgh::Fq2>(b2x).add(s_d2).add(beta2);
    bgh::store_affine(B2, proof + 16);
    msm_finish(sl.work[0], nullptr, g1x);
    const uint32_t *xA = g1x, *xB1 = g1x + 32, *xL = g1x + 64;
    const uint32_t* xH = g1x + 96;
    bgh::P1 A = bgh::load_xyzz<bgh::Fq>(xA).add(r_d1).add(alpha);
    bgh::P1 B1 = bgh::load_xyzz<bgh::Fq>(xB1).add(s_d1).add(beta1);
    bgh::P1 sA, rB1;
    host_parallel(2, [&](int i) {
      if (i == 0)
        sA = A.mul(sc);
      else
        rB1 = B1.mul(rc);
    });
    bgh::P1 C0 = sA.add(rB1).add(rs_d1.neg()).add(bgh::load_xyzz<bgh::Fq>(xL));
    bgh::store_affine(A, proof);
    if (sched != 1 || pk->fixed) {
      msm_finish(sl.work[1], nullptr, hx);
      xH = hx;
    }
    bgh::P1 C = C0.add(bgh::load_xyzz<bgh::Fq>(xH));
    bgh::store_affine(C, proof + 48);
    if (words16) {
      for (int i = 0; i < 64; i++) {
        proof_out[2 * i] = proof[i] & 0xFFFFu;
        proof_out[2 * i + 1] = proof[i] >> 16;
      }
    } else {
      memcpy(proof_out, proof, sizeof proof);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (t_ms) {
      // device time of each phase on its own stream (the MSMs overlap), then the wall
      // time of the whole call
      // upload, evals, qap, G1 batch, (unused), B2, H alone (schedule 0) / G1 batch end
      // (a replayed graph records no events: its device phases read -1)
      const int pairs[8][2] = {{0, 1}, {1, 3}, {3, 4}, {5, 6}, {5, 6}, {9, 10}, {4, 11}, {5, 6}};
      for (int i = 0; i < 8; i++) {
        float e = -1;
        if (!graphed || i == 0) BG_HIP(hipEventElapsedTime(&e, ev[pairs[i][0]], ev[pairs[i][1]]));
        t_ms[i] = e;
      }
      t_ms[8] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
  }
}

bool prove_batchable(const bg_pk* a, const bg_pk* b) {
  return a && b && a->fixed && b->fixed && memcmp(a->hdr, b->hdr, sizeof a->hdr) == 0 && a->c_g1 == b->c_g1 &&
         a->c_h == b->c_h && a->c_b2 == b->c_b2;
}

void prove_batch_on(const ProveJob* jobs, int P, ProveSlot& sl) {
  BG_REQUIRE(P >= 1 && jobs, BG_ERR_ARG, "prove batch: no jobs");
  const bg_pk* pk0 = jobs[0].pk;
  for (int j = 0; j < P; j++) {
    const ProveJob& q = jobs[j];
    BG_REQUIRE(q.pk && q.witness && q.rs && q.proof_out, BG_ERR_ARG, "prove batch: null argument");
    BG_REQUIRE(!((q.flags & BG_MEM_DEVICE) && (q.flags & 1) == BG_FMT_WORDS16), BG_ERR_ARG,
               "prove batch: a device witness must be packed");
    BG_REQUIRE(prove_batchable(pk0, q.pk), BG_ERR_ARG, "prove batch: keys of different shapes");
    uint32_t tmp[16];
    read_rs(q.rs, q.flags, tmp);
  }
  (void)ctx();
  const uint32_t nc = pk0->hdr[0], ni = pk0->hdr[1], nw = pk0->hdr[2], lg = pk0->hdr[3];
  const uint32_t na = pk0->hdr[4], nbq = pk0->hdr[5], nh = pk0->hdr[6], nl = pk0->hdr[7];
  const uint64_t d = 1ull << lg, nz = (uint64_t)ni + nw;
  hipStream_t s0 = sl.s[0], s1 = sl.s[1], s3 = sl.s[3];
  hipEvent_t* ev = sl.ev;
  // every per-proof array holds the P proofs back to back
  sl.ensure(P * nz, P * d);
  uint32_t *z = sl.z.as<uint32_t>(), *zc = sl.zc.as<uint32_t>(), *ea = sl.ea.as<uint32_t>(),
           *eb = sl.eb.as<uint32_t>(), *ec = sl.ec.as<uint32_t>(), *h = sl.h.as<uint32_t>();
  uint32_t* zp = static_cast<uint32_t*>(sl.zpin.get(32 * nz * P));
  for (int j = 0; j < P; j++) upload_witness(jobs[j].witness, jobs[j].flags, nz, z + 8 * nz * j, zp + 8 * nz * j, s0);
  {
    ScopedKTime kt("evals", s0);
    k_from_mont<<<nb(P * nz, 256), 256, 0, s0>>>(z, zc, P * nz);
    BG_HIP(hipEventRecord(ev[2], s0));  // canonical scalars of every proof ready
    for (int j = 0; j < P; j++) {
      const bg_pk* pk = jobs[j].pk;
      const uint32_t* zj = z + 8 * nz * j;
      k_evals<<<nb(d, 256), 256, 0, s0>>>(pk->A.offs.as<uint32_t>(), pk->A.cols.as<uint32_t>(),
                                          pk->A.coef.as<uint32_t>(), zj, nc, ni, 1, d, ea + 8 * d * j);
      k_evals<<<nb(d, 256), 256, 0, s0>>>(pk->B.offs.as<uint32_t>(), pk->B.cols.as<uint32_t>(),
                                          pk->B.coef.as<uint32_t>(), zj, nc, ni, 0, d, eb + 8 * d * j);
      k_evals<<<nb(d, 256), 256, 0, s0>>>(pk->C.offs.as<uint32_t>(), pk->C.cols.as<uint32_t>(),
                                          pk->C.coef.as<uint32_t>(), zj, nc, ni, 0, d, ec + 8 * d * j);
    }
    BG_HIP(hipGetLastError());
  }
  // MSMs as schedule 0, each over all P proofs: B2 (P MSMs) on s3 and the G1 batch (A, B1,
  // L of every proof: 3P MSMs) on s1 at once; qap_h of all proofs, then H (P MSMs) on s0
  std::vector<const uint32_t*> tb2(P), sb2(P), tg1(3 * P), sg1(3 * P), th(P), sh(P);
  std::vector<uint64_t> nb2(P, nbq), ng1(3 * P), nhv(P, nh);
  for (int j = 0; j < P; j++) {
    const bg_pk* pk = jobs[j].pk;
    const uint32_t* zcj = zc + 8 * nz * j;
    tb2[j] = pk->pb2.as<uint32_t>();
    sb2[j] = zcj;
    tg1[3 * j] = pk->pa.as<uint32_t>();
    tg1[3 * j + 1] = pk->pb1.as<uint32_t>();
    tg1[3 * j + 2] = pk->pl.as<uint32_t>();
    sg1[3 * j] = zcj;
    sg1[3 * j + 1] = zcj;
    sg1[3 * j + 2] = zcj + 8 * ni;
    ng1[3 * j] = na;
    ng1[3 * j + 1] = nbq;
    ng1[3 * j + 2] = nl;
    th[j] = pk->ph.as<uint32_t>();
    sh[j] = h + 8 * d * j;
  }
  for (hipStream_t sx : {s1, s3}) BG_HIP(hipStreamWaitEvent(sx, ev[2], 0));
  msm_enqueue_fixed(2, P, tb2.data(), sb2.data(), nb2.data(), pk0->c_b2, s3, sl.work[2]);
  msm_enqueue_fixed(1, 3 * P, tg1.data(), sg1.data(), ng1.data(), pk0->c_g1, s1, sl.work[0]);
  {
    ScopedKTime kt("qap_h", s0);
    qap_h_run(ea, eb, ec, h, sl.tmp.as<uint32_t>(), (int)lg, true, s0, (unsigned)P);
  }
  msm_enqueue_fixed(1, P, th.data(), sh.data(), nhv.data(), pk0->c_h, s0, sl.work[1]);

  // host: per proof r, s scalars and the delta / alpha / beta terms while the device runs
  struct Comb {
    uint64_t rc[4], sc[4], rsc[4];
    bgh::P1 alpha, beta1, r_d1, s_d1, rs_d1, A, C0;
    bgh::P2 beta2, s_d2;
    uint32_t proof[64];
  };
  std::vector<Comb> cb(P);
  host_parallel(P, [&](int j) {
    const ProveJob& q = jobs[j];
    Comb& c = cb[j];
    uint32_t rs_packed[16];
    read_rs(q.rs, q.flags, rs_packed);  // checked before any launch
    canon_words(rs_packed, c.rc);
    canon_words(rs_packed + 8, c.sc);
    bgh::Fr x = (bgh::Fr::from_u32(rs_packed) * bgh::Fr::from_u32(rs_packed + 8)).from_mont();
    memcpy(c.rsc, x.v, 32);
    const bg_pk* pk = q.pk;
    c.alpha = bgh::load_affine<bgh::Fq>(pk->g1);
    c.beta1 = bgh::load_affine<bgh::Fq>(pk->g1 + 16);
    const bgh::P1 delta1 = bgh::load_affine<bgh::Fq>(pk->g1 + 32);
    c.beta2 = bgh::load_affine<bgh::Fq2>(pk->g2);
    const bgh::P2 delta2 = bgh::load_affine<bgh::Fq2>(pk->g2 + 32);
    c.r_d1 = delta1.mul(c.rc);
    c.s_d1 = delta1.mul(c.sc);
    c.rs_d1 = delta1.mul(c.rsc);
    c.s_d2 = delta2.mul(c.sc);
  });
  // combination (groth16.bend:298-308) as in prove_on, the H term last
  std::vector<uint32_t> g1res(3 * P * 16), b2res(P * 32), hres(P * 16);
  msm_finish(sl.work[0], g1res.data());
  host_parallel(P, [&](int j) {
    Comb& c = cb[j];
    const uint32_t* r = &g1res[48 * j];
    c.A = bgh::load_affine<bgh::Fq>(r).add(c.r_d1).add(c.alpha);
    const bgh::P1 B1 = bgh::load_affine<bgh::Fq>(r + 16).add(c.s_d1).add(c.beta1);
    c.C0 = c.A.mul(c.sc).add(B1.mul(c.rc)).add(c.rs_d1.neg()).add(bgh::load_affine<bgh::Fq>(r + 32));
    bgh::store_affine(c.A, c.proof);
  });
  msm_finish(sl.work[2], b2res.data());
  msm_finish(sl.work[1], hres.data());
  host_parallel(P, [&](int j) {
    Comb& c = cb[j];
    const bgh::P2 B2 = bgh::load_affine<bgh::Fq2>(&b2res[32 * j]).add(c.s_d2).add(c.beta2);
    bgh::store_affine(B2, c.proof + 16);
    bgh::store_affine(c.C0.add(bgh::load_affine<bgh::Fq>(&hres[16 * j])), c.proof + 48);
    uint32_t* out = jobs[j].proof_out;
    if ((jobs[j].flags & 1) == BG_FMT_WORDS16) {
      for (int i = 0; i < 64; i++) {
        out[2 * i] = c.proof[i] & 0xFFFFu;
        out[2 * i + 1] = c.proof[i] >> 16;
      }
    } else {
      memcpy(out, c.proof, sizeof c.proof);
    }
  });
}

}  // namespace bg

static void hex_be(const uint32_t* packed_mont, char* out) {
  // canonical big-endian "0x" + 64 hex digits (bend2/src/fq.bend:2482-2489)
  bgh::Fq x = bgh::Fq::from_u32(packed_mont).from_mont();
  static const char* H = "0123456789abcdef";
  out[0] = '0';
  out[1] = 'x';
  for (int i = 0; i < 64; i++) {
    int nib = 63 - i;
    uint64_t w = x.v[nib / 16];
    out[2 + i] = H[(w >> (4 * (nib % 16))) & 0xF];
  }
  out[66] = 0;
}

extern "C" int bg_proof_format(const uint32_t* p, char* buf, int cap) {
  int n = 0;
  int rc = guard([&] {
    BG_REQUIRE(p && buf && cap > 0, BG_ERR_ARG, "bg_proof_format: bad argument");
    char h[8][67];
    const int off[8] = {0, 8, 16, 24, 32, 40, 48, 56};
    for (int i = 0; i < 8; i++) hex_be(p + off[i], h[i]);
    n = snprintf(buf, (size_t)cap, "A %s %s\nB %s %s %s %s\nC %s %s\n", h[0], h[1], h[2], h[3], h[4], h[5], h[6],
                 h[7]);
    BG_REQUIRE(n < cap, BG_ERR_ARG, "bg_proof_format: buffer of %d bytes too small", cap);
  });
  return rc == BG_OK ? n : rc;
}

extern "C" int bg_graph_info(uint64_t* info3) {
  return guard([&] {
    BG_REQUIRE(info3, BG_ERR_ARG, "bg_graph_info: null argument");
    info3[0] = bg::g_graphs_recorded.load();
    info3[1] = bg::g_graph_launches.load();
    info3[2] = bg::g_graph_failures.load();
  });
}
