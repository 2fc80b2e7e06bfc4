// Known-answer kernels, element-wise field/curve kernels for the parity tests, and
// throughput probes (C ABI: bg_kat_field_mul_xor, bg_kat_g1_add_chain, bg_field_op,
// bg_curve_op, bg_probe_rate, bg_probe_launch).
#include "bg_curve.hpp"
#include "bg_runtime.hpp"
#include <string.h>
#include <chrono>
#include <functional>

namespace bg {

// ---------------------------------------------------------------- field_mul KAT
// bend2/bench/field_mul.bend:9-31: h(i) = ((i * 2654435761) ^ (i >> 15)) * 2246822519,
// operand s has 16-bit limbs h(s + off_k) (off_0 = 0, off_k = 40503 k + 1), top limb
// masked to 0x2FFF; product i multiplies operands 2i and 2i+1; checksum = xor of the
// low 16 bits of every Montgomery product.
__device__ __forceinline__ uint32_t kat_h(uint32_t i) { return ((i * 2654435761u) ^ (i >> 15)) * 2246822519u; }

__device__ __forceinline__ Fp kat_rnd(uint32_t s) {
  Fp r;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    uint32_t k0 = 2 * j, k1 = 2 * j + 1;
    uint32_t o0 = k0 == 0 ? 0u : 40503u * k0 + 1u;
    uint32_t o1 = 40503u * k1 + 1u;
    uint32_t lo = kat_h(s + o0) & 0xFFFFu;
    uint32_t hi = kat_h(s + o1) & (k1 == 15 ? 12287u : 0xFFFFu);
    r.v[j] = lo | (hi << 16);
  }
  return r;
}

__global__ void k_kat_field_mul(uint64_t total, uint32_t* out) {
  uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t x = 0;
  for (uint64_t i = tid; i < total; i += stride) {
    Fp a = kat_rnd((uint32_t)(2 * i));
    Fp b = kat_rnd((uint32_t)(2 * i + 1));
    Fp c = Fr::canon(Fr::mul(a, b));
    x ^= c.v[0] & 0xFFFFu;
  }
  // wave-level xor reduction, then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) x ^= __shfl_xor(x, off);
  if ((threadIdx.x & 63) == 0) atomicXor(out, x);
}

// ---------------------------------------------------------------- g1_add chain KAT
__global__ void k_kat_g1_chain(uint64_t k, uint32_t* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  using J = JacRef<Fq>;
  Aff<Fq> g = {Fq::one(), Fq::add(Fq::one(), Fq::one())};
  J::J acc = {g.x, g.y, Fq::one()};
  for (uint64_t i = 0; i < k; i++) acc = J::madd(acc, g);
  out[0] = Fq::canon(acc.x).v[0] & 0xFFFFu;
}

// ---------------------------------------------------------------- element-wise ops
template <class F>
__device__ typename F::E field_apply(int op, const typename F::E& a, const typename F::E& b);

template <class P>
__device__ Fp fp_apply(int op, const Fp& a, const Fp& b) {
  using F = Field<P>;
  switch (op) {
    case 0: return F::add(a, b);
    case 1: return F::sub(a, b);
    case 2: return F::mul(a, b);
    case 3: return F::neg(a);
    case 4: return F::inv(a);
    case 5: return F::sqr(a);
    case 6: return F::to_mont(a);
    default: return F::from_mont(a);
  }
}

__device__ Fp2 fp2_apply(int op, const Fp2& a, const Fp2& b) {
  switch (op) {
    case 0: return Fq2::add(a, b);
    case 1: return Fq2::sub(a, b);
    case 2: return Fq2::mul(a, b);
    case 3: return Fq2::neg(a);
    case 4: return Fq2::inv(a);
    default: return Fq2::sqr(a);
  }
}

__global__ void k_field_op(int field, int op, const uint32_t* a, const uint32_t* b, uint32_t* out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (field == 2) {
    Fp2 x = load_fp2(a + 16 * i), y = load_fp2(b + 16 * i);
    store_fp2(out + 16 * i, Fq2::canon(fp2_apply(op, x, y)));
  } else {
    Fp x = load_fp(a + 8 * i), y = load_fp(b + 8 * i);
    Fp r = field == 0 ? Fr::canon(fp_apply<FrP>(op, x, y)) : Fq::canon(fp_apply<FqP>(op, x, y));
    store_fp(out + 8 * i, r);
  }
}

template <class F>
__global__ void k_curve_op(int op, const uint32_t* p, const uint32_t* q, uint32_t* out, uint64_t n) {
  using C = Curve<F>;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int aw = C::aff_words;
  auto a = C::load_aff(p + aw * i);
  auto b = C::load_aff(q + aw * i);
  typename C::P r;
  switch (op) {
    case 0: r = C::madd(C::from_affine(a), b); break;
    case 1: r = C::dbl(C::from_affine(a)); break;
    case 2: r = C::from_affine(C::aff_neg(a)); break;
    default: r = C::add(C::from_affine(a), C::madd(C::infinity(), b)); break;
  }
  C::store_aff(out + aw * i, C::aff_canon(C::to_affine(r)));
}

// ---------------------------------------------------------------- throughput probes
__global__ void k_probe_mad(uint64_t iters, uint32_t seed, uint64_t* sink) {
  // 16 independent chains of acc = lo(acc) * y + acc: one v_mad_u64_u32 per step
  uint32_t t = threadIdx.x + blockIdx.x * blockDim.x;
  uint64_t acc[16];
#pragma unroll
  for (int k = 0; k < 16; k++) acc[k] = ((uint64_t)(seed * (k + 1)) << 32) | (t + k);
  uint32_t y = seed ^ t;
  for (uint64_t it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 16; k++) acc[k] = (uint64_t)(uint32_t)acc[k] * y + acc[k];
  }
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) s ^= acc[k];
  if (s == 0x1234567ull) sink[0] = s;
}

__global__ void k_probe_fqmul(uint64_t iters, uint32_t seed, uint32_t* sink) {
  uint32_t t = threadIdx.x + blockIdx.x * blockDim.x;
  Fp a[4], b;
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int j = 0; j < 8; j++) a[k].v[j] = (seed * (j + 7 * k + 1) + t) & (j == 7 ? 0x0fffffffu : 0xffffffffu);
#pragma unroll
  for (int j = 0; j < 8; j++) b.v[j] = (seed ^ (j * 977 + t)) & (j == 7 ? 0x0fffffffu : 0xffffffffu);
  for (uint64_t it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 4; k++) a[k] = Fq::mul(a[k], b);
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) s ^= a[k].v[0];
  if (s == 0x1234567u) sink[0] = s;
}

__global__ void k_probe_madd(uint64_t iters, uint32_t seed, uint32_t* sink) {
  uint32_t t = threadIdx.x + blockIdx.x * blockDim.x;
  // pseudo-random (not on-curve) coordinates: the formula cost is what is measured
  Xyzz<Fq> acc;
  Aff<Fq> q;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    uint32_t m = j == 7 ? 0x0fffffffu : 0xffffffffu;
    acc.X.v[j] = (seed * (j + 1) + t) & m;
    acc.Y.v[j] = (seed * (j + 3) ^ t) & m;
    acc.ZZ.v[j] = (seed * (j + 5) + 3 * t) & m;
    acc.ZZZ.v[j] = (seed * (j + 9) ^ 5 * t) & m;
    q.x.v[j] = (seed * (j + 11) + 7 * t) & m;
    q.y.v[j] = (seed * (j + 13) ^ 9 * t) & m;
  }
  for (uint64_t it = 0; it < iters; it++) acc = G1::madd(acc, q);
  if ((acc.X.v[0] ^ acc.ZZZ.v[1]) == 0x1234567u) sink[0] = 1;
}

// latency probes: ILP independent dependency chains per lane
template <int ILP>
__global__ void k_probe_mad_chain(uint64_t iters, uint32_t seed, uint64_t* sink) {
  uint32_t t = threadIdx.x + blockIdx.x * blockDim.x;
  uint64_t acc[ILP];
#pragma unroll
  for (int k = 0; k < ILP; k++) acc[k] = ((uint64_t)(seed * (k + 1)) << 32) | (t + k);
  uint32_t y = seed ^ t;
  for (uint64_t it = 0; it < iters; it++) {
#pragma unroll
    for (int r = 0; r < 16 / ILP; r++)
#pragma unroll
      for (int k = 0; k < ILP; k++) acc[k] = (uint64_t)(uint32_t)acc[k] * y + acc[k];
  }
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < ILP; k++) s ^= acc[k];
  if (s == 0x1234567ull) sink[0] = s;
}

template <int ILP>
__global__ void k_probe_fqmul_ilp(uint64_t iters, uint32_t seed, uint32_t* sink) {
  uint32_t t = threadIdx.x + blockIdx.x * blockDim.x;
  Fp a[ILP], b;
#pragma unroll
  for (int k = 0; k < ILP; k++)
#pragma unroll
    for (int j = 0; j < 8; j++) a[k].v[j] = (seed * (j + 7 * k + 1) + t) & (j == 7 ? 0x0fffffffu : 0xffffffffu);
#pragma unroll
  for (int j = 0; j < 8; j++) b.v[j] = (seed ^ (j * 977 + t)) & (j == 7 ? 0x0fffffffu : 0xffffffffu);
  for (uint64_t it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < ILP; k++) a[k] = Fq::mul(a[k], b);
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < ILP; k++) s ^= a[k].v[0];
  if (s == 0x1234567u) sink[0] = s;
}

__global__ void k_probe_fq2mul(uint64_t iters, uint32_t seed, uint32_t* sink) {
  uint32_t t = threadIdx.x + blockIdx.x * blockDim.x;
  Fp2 a, b;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    uint32_t m = j == 7 ? 0x0fffffffu : 0xffffffffu;
    a.c0.v[j] = (seed * (j + 1) + t) & m;
    a.c1.v[j] = (seed * (j + 3) ^ t) & m;
    b.c0.v[j] = (seed * (j + 5) + 3 * t) & m;
    b.c1.v[j] = (seed * (j + 9) ^ 5 * t) & m;
  }
  for (uint64_t it = 0; it < iters; it++) a = Fq2::mul(a, b);
  if ((a.c0.v[0] ^ a.c1.v[1]) == 0x1234567u) sink[0] = 1;
}

}  // namespace bg

using namespace bg;

static void upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
  BG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}

extern "C" int bg_kat_field_mul_xor(int log2_count, uint32_t* xor_out) {
  return guard([&] {
    BG_REQUIRE(log2_count >= 0 && log2_count <= 31 && xor_out, BG_ERR_ARG, "bg_kat_field_mul_xor: bad argument");
    Ctx& c = ctx();
    uint32_t* d = c.scratch[0].as<uint32_t>(1);
    BG_HIP(hipMemsetAsync(d, 0, 4, c.stream));
    uint64_t total = 1ull << log2_count;
    int blocks = (int)std::min<uint64_t>((total + 255) / 256, (uint64_t)c.cu_count * 16);
    k_kat_field_mul<<<blocks, 256, 0, c.stream>>>(total, d);
    BG_HIP(hipGetLastError());
    BG_HIP(hipMemcpyAsync(xor_out, d, 4, hipMemcpyDeviceToHost, c.stream));
    BG_HIP(hipStreamSynchronize(c.stream));
  });
}

extern "C" int bg_kat_g1_add_chain(uint64_t k, uint32_t* x16_out) {
  return guard([&] {
    BG_REQUIRE(x16_out && k <= (1ull << 20), BG_ERR_ARG, "bg_kat_g1_add_chain: bad argument");
    Ctx& c = ctx();
    uint32_t* d = c.scratch[0].as<uint32_t>(1);
    k_kat_g1_chain<<<1, 64, 0, c.stream>>>(k, d);
    BG_HIP(hipGetLastError());
    BG_HIP(hipMemcpyAsync(x16_out, d, 4, hipMemcpyDeviceToHost, c.stream));
    BG_HIP(hipStreamSynchronize(c.stream));
  });
}

extern "C" int bg_field_op(int field, int op, const uint32_t* a, const uint32_t* b, uint32_t* out, uint64_t n) {
  return guard([&] {
    BG_REQUIRE(field >= 0 && field <= 2 && op >= 0 && op <= 7 && a && out, BG_ERR_ARG, "bg_field_op: bad argument");
    BG_REQUIRE(!(field == 2 && op > 5), BG_ERR_ARG, "bg_field_op: Fq2 has no Montgomery conversion op");
    if (n == 0) return;
    Ctx& c = ctx();
    size_t w = field == 2 ? 16 : 8;
    size_t bytes = n * w * 4;
    uint32_t* da = c.scratch[0].as<uint32_t>(n * w);
    uint32_t* db = c.scratch[1].as<uint32_t>(n * w);
    uint32_t* dout = c.scratch[2].as<uint32_t>(n * w);
    upload(da, a, bytes, c.stream);
    upload(db, b ? b : a, bytes, c.stream);
    k_field_op<<<(unsigned)((n + 127) / 128), 128, 0, c.stream>>>(field, op, da, db, dout, n);
    BG_HIP(hipGetLastError());
    BG_HIP(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c.stream));
    BG_HIP(hipStreamSynchronize(c.stream));
  });
}

extern "C" int bg_curve_op(int group, int op, const uint32_t* p, const uint32_t* q, uint32_t* out, uint64_t n) {
  return guard([&] {
    BG_REQUIRE((group == 1 || group == 2) && op >= 0 && op <= 3 && p && out, BG_ERR_ARG, "bg_curve_op: bad argument");
    // The G2 variant of this diagnostic kernel took an illegal memory access on gfx950
    // with no cause found in its code; it stays disabled until one is shown.
    BG_REQUIRE(group == 1, BG_ERR_ARG, "bg_curve_op: the G2 diagnostic kernel is disabled");
    if (n == 0) return;
    Ctx& c = ctx();
    size_t w = group == 1 ? 16 : 32;
    size_t bytes = n * w * 4;
    uint32_t* dp = c.scratch[0].as<uint32_t>(n * w);
    uint32_t* dq = c.scratch[1].as<uint32_t>(n * w);
    uint32_t* dout = c.scratch[2].as<uint32_t>(n * w);
    upload(dp, p, bytes, c.stream);
    upload(dq, q ? q : p, bytes, c.stream);
    unsigned blocks = (unsigned)((n + 63) / 64);
    k_curve_op<Fq><<<blocks, 64, 0, c.stream>>>(op, dp, dq, dout, n);
    BG_HIP(hipGetLastError());
    BG_HIP(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c.stream));
    BG_HIP(hipStreamSynchronize(c.stream));
  });
}

// Throughput / latency probes.  kind: 0 v_mad_u64_u32 (16 independent chains, 16 waves
// per CU), 1 Fq products (4 independent per lane), 2 G1 mixed adds, 3/4 v_mad chains with
// ILP 1/4 at one wave per SIMD, 5/6 Fq products ILP 1 at one / four waves per SIMD,
// 7 Fq products ILP 4 at one wave per SIMD, 8 Fq2 products at one wave per SIMD,
// 9 Fq2 products at two waves per SIMD.  result[0] = operations/s, result[1] = ms.
extern "C" int bg_probe_rate(int kind, uint64_t iters, double* result) {
  return guard([&] {
    BG_REQUIRE(kind >= 0 && kind <= 9 && result && iters > 0, BG_ERR_ARG, "bg_probe_rate: bad argument");
    Ctx& c = ctx();
    uint32_t* sink = c.scratch[0].as<uint32_t>(4);
    uint64_t* sink64 = reinterpret_cast<uint64_t*>(sink);
    const int threads = 256;  // 4 waves: one per SIMD of a CU
    int wps = 4;              // waves per SIMD
    if (kind == 3 || kind == 4 || kind == 5 || kind == 7 || kind == 8) wps = 1;
    if (kind == 9) wps = 2;
    if (kind == 6) wps = 4;
    int blocks = c.cu_count * wps;
    double per_thread = 1.0;
    hipEvent_t e0, e1;
    BG_HIP(hipEventCreate(&e0));
    BG_HIP(hipEventCreate(&e1));
    auto launch = [&](uint64_t it) {
      switch (kind) {
        case 0: k_probe_mad<<<blocks, threads, 0, c.stream>>>(it, 12345u, sink64); per_thread = 16; break;
        case 1: k_probe_fqmul<<<blocks, threads, 0, c.stream>>>(it, 12345u, sink); per_thread = 4; break;
        case 2: k_probe_madd<<<blocks, threads, 0, c.stream>>>(it, 12345u, sink); per_thread = 1; break;
        case 3: k_probe_mad_chain<1><<<blocks, threads, 0, c.stream>>>(it, 12345u, sink64); per_thread = 16; break;
        case 4: k_probe_mad_chain<4><<<blocks, threads, 0, c.stream>>>(it, 12345u, sink64); per_thread = 16; break;
        case 5:
        case 6: k_probe_fqmul_ilp<1><<<blocks, threads, 0, c.stream>>>(it, 12345u, sink); per_thread = 1; break;
        case 7: k_probe_fqmul_ilp<4><<<blocks, threads, 0, c.stream>>>(it, 12345u, sink); per_thread = 4; break;
        default: k_probe_fq2mul<<<blocks, threads, 0, c.stream>>>(it, 12345u, sink); per_thread = 1; break;
      }
    };
    launch(1);  // warm-up
    BG_HIP(hipEventRecord(e0, c.stream));
    launch(iters);
    BG_HIP(hipEventRecord(e1, c.stream));
    BG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    BG_HIP(hipEventElapsedTime(&ms, e0, e1));
    result[0] = (double)blocks * threads * per_thread * (double)iters / (ms * 1e-3);
    result[1] = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

// ---------------------------------------------------------------- launch probe
// The cost of launching `kernels` short kernels (256 blocks of 64 threads, `spin` dependent
// integer steps each) eagerly or from HIP graphs, the shapes the prover's recorded
// sequences choose between: us6 = microseconds per repetition (launch to completion) of
//   0 eager, one stream            1 one linear graph
//   2 eager, fork to 4 streams      3 one graph that forks and joins the 4 streams
//   4 as 3 with a host-visible event node (hipEventRecordExternal) on each side stream
//   5 one linear graph per stream, launched after an eager fork
namespace bg {
static constexpr int PROBE_NB = 256;
__global__ void k_probe_work(uint32_t* __restrict__ p, int i, int spin) {
  const uint32_t k = blockIdx.x * 64 + threadIdx.x;
  uint32_t v = p[k];
  for (int j = 0; j < spin; j++) v = v * 1664525u + 1013904223u;
  p[k] = v + (uint32_t)i;
}
}  // namespace bg

extern "C" int bg_probe_launch(int kernels, int spin, int reps, double* us6) {
  return guard([&] {
    BG_REQUIRE(kernels >= 8 && kernels <= 4096 && spin >= 0 && spin <= 1000000 && reps >= 1 && reps <= 10000 && us6,
               BG_ERR_ARG, "bg_probe_launch: bad argument");
    Ctx& c = ctx();
    uint32_t* d = c.scratch[1].as<uint32_t>(4ull * PROBE_NB * 64);
    BG_HIP(hipMemsetAsync(d, 0, 16ull * PROBE_NB * 64, c.stream));
    BG_HIP(hipStreamSynchronize(c.stream));
    hipStream_t s[4] = {};
    hipEvent_t fork = nullptr, join[3] = {}, ext[3] = {};
    hipGraphExec_t ex[6] = {};
    struct Cleanup {
      hipStream_t* s;
      hipEvent_t* fork;
      hipEvent_t* join;
      hipEvent_t* ext;
      hipGraphExec_t* ex;
      ~Cleanup() {
        for (int i = 0; i < 6; i++)
          if (ex[i]) (void)hipGraphExecDestroy(ex[i]);
        if (*fork) (void)hipEventDestroy(*fork);
        for (int i = 0; i < 3; i++) {
          if (join[i]) (void)hipEventDestroy(join[i]);
          if (ext[i]) (void)hipEventDestroy(ext[i]);
        }
        for (int i = 0; i < 4; i++)
          if (s[i]) (void)hipStreamDestroy(s[i]);
      }
    } cleanup{s, &fork, join, ext, ex};
    for (auto& x : s) BG_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    BG_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (int i = 0; i < 3; i++) {
      BG_HIP(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
      BG_HIP(hipEventCreateWithFlags(&ext[i], hipEventDisableTiming));
    }
    const int per = kernels / 4;
    auto linear = [&](hipStream_t st, uint32_t* p, int n) {
      for (int i = 0; i < n; i++) k_probe_work<<<PROBE_NB, 64, 0, st>>>(p, i, spin);
    };
    auto forked = [&](bool ext_events) {
      k_probe_work<<<PROBE_NB, 64, 0, s[0]>>>(d, 0, spin);
      BG_HIP(hipEventRecord(fork, s[0]));
      for (int b = 1; b < 4; b++) {
        BG_HIP(hipStreamWaitEvent(s[b], fork, 0));
        linear(s[b], d + (uint64_t)b * PROBE_NB * 64, per);
        if (ext_events) BG_HIP(hipEventRecordWithFlags(ext[b - 1], s[b], hipEventRecordExternal));
      }
      linear(s[0], d, per);
      for (int b = 1; b < 4; b++) {
        BG_HIP(hipEventRecord(join[b - 1], s[b]));
        BG_HIP(hipStreamWaitEvent(s[0], join[b - 1], 0));
      }
    };
    auto record = [&](hipStream_t st, const std::function<void()>& fn) {
      hipGraph_t g = nullptr;
      hipGraphExec_t e = nullptr;
      BG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      fn();
      BG_HIP(hipStreamEndCapture(st, &g));
      const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      BG_HIP(r);
      return e;
    };
    ex[0] = record(s[0], [&] { linear(s[0], d, kernels); });
    ex[1] = record(s[0], [&] { forked(false); });
    ex[2] = record(s[0], [&] { forked(true); });
    for (int b = 0; b < 3; b++) ex[3 + b] = record(s[b + 1], [&] { linear(s[b + 1], d + (uint64_t)(b + 1) * PROBE_NB * 64, per); });
    auto split = [&] {
      k_probe_work<<<PROBE_NB, 64, 0, s[0]>>>(d, 0, spin);
      BG_HIP(hipEventRecord(fork, s[0]));
      for (int b = 1; b < 4; b++) {
        BG_HIP(hipStreamWaitEvent(s[b], fork, 0));
        BG_HIP(hipGraphLaunch(ex[2 + b], s[b]));
      }
      linear(s[0], d, per);
      for (int b = 1; b < 4; b++) {
        BG_HIP(hipEventRecord(join[b - 1], s[b]));
        BG_HIP(hipStreamWaitEvent(s[0], join[b - 1], 0));
      }
    };
    const std::function<void()> variants[6] = {
        [&] { linear(s[0], d, kernels); },       [&] { BG_HIP(hipGraphLaunch(ex[0], s[0])); },
        [&] { forked(false); },                   [&] { BG_HIP(hipGraphLaunch(ex[1], s[0])); },
        [&] { BG_HIP(hipGraphLaunch(ex[2], s[0])); }, split};
    for (int v = 0; v < 6; v++) {
      variants[v]();  // warm-up
      BG_HIP(hipStreamSynchronize(s[0]));
      const auto t0 = std::chrono::steady_clock::now();
      for (int r = 0; r < reps; r++) {
        variants[v]();
        BG_HIP(hipStreamSynchronize(s[0]));
      }
      us6[v] = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / reps;
    }
  });
}
