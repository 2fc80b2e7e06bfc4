// Final combination and normalisation of a proof on the device (the BG_COMBINE_DEVICE path
// of bg_prove): bend2/src/groth16.bend:281-308 (acc_g1 / acc_g2 and combine) and the MSM
// finishing of bend2/src/msm.bend:210-235 (window sums), with no host round trip between
// the MSMs and the affine proof.
//
//   A  = alpha + MSM_A + r delta
//   B  = beta2 + MSM_B2 + s delta2           (G2)
//   B1 = beta + MSM_B1 + s delta
//   C  = s A + r B1 - r s delta + MSM_L + MSM_H
//
// The kernels follow the proof's streams: k_delta_terms (r delta, s delta, rs delta,
// s delta2 from the key's 4-bit-window tables of delta) starts with the proof, since it
// needs only r and s; once the G1 batch is reduced, k_finish sums its three MSMs (A, B1, L)
// from their level-2 outputs and k_combine_g1 forms A and C0 = s A + r B1 - rs delta + L
// while MSM H still runs; k_combine_b2 finishes B2 and normalises it; k_combine_final adds
// MSM H and normalises C.  Every addition is the exact XYZZ law, so the result is the
// group element the host path computes, and its affine bits are identical.  These kernels
// are single dependent chains (one or a few waves) running beside the MSMs' accumulation
// waves, so each raises its waves' issue priority (s_setprio 3).
#include "bg_curve.hpp"
#include "bg_prover.hpp"

namespace bg {

// Scratch of one proof's combination (comb_words() words, device): G1 XYZZ points at 32
// words each, s delta2 (G2 XYZZ, 64 words), then the proof (A, B, C affine, 64 words).
enum : uint32_t {
  CB_RD1 = 0, CB_SD1 = 1, CB_RSD1 = 2,  // r delta, s delta, rs delta
  CB_MA = 3, CB_MB1 = 4, CB_ML = 5,     // MSM A, B1, L
  CB_A = 6, CB_C0 = 7, CB_G1_COUNT = 8
};
static constexpr uint32_t CB_G2_OFF = CB_G1_COUNT * 32;
static constexpr uint32_t CB_PROOF_OFF = CB_G2_OFF + 64;
static constexpr uint32_t CB_WORDS = CB_PROOF_OFF + 64;

uint64_t comb_words() { return CB_WORDS; }
uint32_t comb_proof_offset() { return CB_PROOF_OFF; }

// k * P (k: 8 canonical words), 4-bit fixed windows from the top: the multiples j P
// (j < 16) in this lane's LDS table, then 4 doublings and one exact addition per window.
template <class F>
__device__ __forceinline__ typename Curve<F>::P scalar_mul(const typename Curve<F>::P& p, const uint32_t* k,
                                           typename Curve<F>::P* tab) {
  using C = Curve<F>;
  tab[0] = C::infinity();
  tab[1] = p;
#pragma unroll 1
  for (int j = 2; j < 16; j++) tab[j] = C::add(tab[j - 1], p);
  typename C::P acc = C::infinity();
#pragma unroll 1
  for (int w = 63; w >= 0; w--) {
#pragma unroll 1
    for (int q = 0; q < 4; q++) acc = C::dbl(acc);
    acc = C::add(acc, tab[(k[w >> 3] >> (4 * (w & 7))) & 15u]);
  }
  return acc;
}

// Tree sum over the 64 lanes of one wave (sh: 64 points of LDS); lane 0 returns the sum.
template <class F>
__device__ __forceinline__ typename Curve<F>::P wave_sum(typename Curve<F>::P v, typename Curve<F>::P* sh) {
  using C = Curve<F>;
  const uint32_t lane = threadIdx.x & 63;
  sh[lane] = v;
  __syncthreads();
  for (uint32_t d = 1; d < 64; d <<= 1) {
    if ((lane & (2 * d - 1)) == 0) sh[lane] = C::add(sh[lane], sh[lane + d]);
    __syncthreads();
  }
  return sh[0];
}

// One MSM's sum from its level-2 outputs (one window, fixed-base tables): the outputs
// (P_h, Q_h), h < Hn, give sum_h P_h + 2^log2M sum_h h Q_h.  The 64 lanes of a wave take
// Hn/64 outputs each (Hn a power of two), then merge neighbours in a tree: a run of n
// outputs carries (S = sum P, U = sum Q, V = sum (h - first) Q), and merging left and
// right gives V = V_l + V_r + n_l U_r.  Lanes beyond Hn hold infinity.  The running sums
// live in LDS (sh: 3 x 64 points) and each step keeps at most two points in registers, so
// the kernels need no scratch memory.  Lane 0 returns the sum.
template <class F>
__device__ __forceinline__ typename Curve<F>::P finish_msm(const uint32_t* __restrict__ base, uint32_t Hn, int log2M,
                                           typename Curve<F>::P* sh) {
  using C = Curve<F>;
  using P = typename C::P;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t xw = C::xyzz_words;
  const uint32_t per = Hn > 64 ? Hn / 64 : 1;
  P* sS = sh;
  P* sU = sh + 64;
  P* sV = sh + 128;
  sS[lane] = C::infinity();
  sU[lane] = C::infinity();
  sV[lane] = C::infinity();
  if (lane * per < Hn) {
    const uint32_t a = lane * per;
#pragma unroll 1
    for (int h = (int)(a + per) - 1; h >= (int)a; h--) {
      P x = sS[lane], y = C::load(base + xw * (2 * (uint64_t)h));
      sS[lane] = C::add(x, y);
      x = sU[lane];
      y = C::load(base + xw * (2 * (uint64_t)h + 1));
      const P u = C::add(x, y);
      sU[lane] = u;
      if (h > (int)a) {
        x = sV[lane];
        sV[lane] = C::add(x, u);
      }
    }
  }
  __syncthreads();
  int len_log = 0;  // log2 of the outputs per run at this level
  while ((1u << len_log) < per) len_log++;
#pragma unroll 1
  for (uint32_t d = 1; d < 64; d <<= 1, len_log++) {
    if ((lane & (2 * d - 1)) == 0) {
      P x = sS[lane], y = sS[lane + d];
      sS[lane] = C::add(x, y);
      P ur = sU[lane + d];
      x = sU[lane];
      sU[lane] = C::add(x, ur);
#pragma unroll 1
      for (int q = 0; q < len_log; q++) ur = C::dbl(ur);
      x = sV[lane];
      y = C::add(x, ur);
      x = sV[lane + d];
      sV[lane] = C::add(y, x);
    }
    __syncthreads();
  }
  P out = sV[0];
#pragma unroll 1
  for (int q = 0; q < log2M; q++) out = C::dbl(out);
  out = C::add(sS[0], out);
  __syncthreads();
  return out;
}

template <class F>
__device__ __forceinline__ typename Curve<F>::P from_aff(const uint32_t* a) {
  using C = Curve<F>;
  return C::from_affine(C::load_aff(a));
}

// k * delta from the key's table T[w][j] = j 2^(4w) delta (XYZZ, w < 64, j < 16): lane w
// looks up its window's digit, the wave sums.
template <class F>
__device__ __forceinline__ typename Curve<F>::P fixed_mul(const uint32_t* __restrict__ tab, const uint32_t* k,
                                                          typename Curve<F>::P* sh) {
  using C = Curve<F>;
  const uint32_t w = threadIdx.x & 63;
  const uint32_t digit = (k[w >> 3] >> (4 * (w & 7))) & 15u;
  return wave_sum<F>(C::load(tab + (uint64_t)C::xyzz_words * (16 * w + digit)), sh);
}

// blocks 0, 1, 2: r delta, s delta, rs delta (G1); block 3: s delta2 (G2)
__global__ __launch_bounds__(64) void k_delta_terms(CombineArgs args, uint32_t* __restrict__ cb) {
  __builtin_amdgcn_s_setprio(3);  // a latency chain beside the MSMs' throughput waves
  __shared__ G2::P sh[64];
  const uint32_t b = blockIdx.x;
  if (b < 3) {
    const G1::P v = fixed_mul<Fq>(args.dtab1, args.sc + 8 * b, reinterpret_cast<G1::P*>(sh));
    if (threadIdx.x == 0) G1::store(cb + 32 * (CB_RD1 + b), v);
  } else {
    const G2::P v = fixed_mul<Fq2>(args.dtab2, args.sc + 8, sh);
    if (threadIdx.x == 0) G2::store(cb + CB_G2_OFF, v);
  }
}

// block i: MSM i of the G1 batch (A, B1, L) from its level-2 outputs
__global__ __launch_bounds__(64) void k_finish_g1(CombineArgs args, const uint32_t* __restrict__ outPQ,
                                                  uint32_t* __restrict__ cb) {
  __builtin_amdgcn_s_setprio(3);  // a latency chain beside the MSMs' throughput waves
  __shared__ G1::P sh[3 * 64];
  const uint32_t i = blockIdx.x;
  const G1::P m = finish_msm<Fq>(outPQ + (uint64_t)G1::xyzz_words * 2 * args.hn_g1 * i, args.hn_g1, args.log2m_g1, sh);
  if (threadIdx.x == 0) G1::store(cb + 32 * (CB_MA + i), m);
}

// A = alpha + MSM_A + r delta and C0 = s A + r B1 - rs delta + L (B1 = beta + MSM_B1 +
// s delta); lane 0 forms s A, lane 1 r B1.
__global__ __launch_bounds__(64) void k_combine_g1(CombineArgs args, uint32_t* __restrict__ cb) {
  __builtin_amdgcn_s_setprio(3);  // a latency chain beside the MSMs' throughput waves
  __shared__ G1::P tab[2][16];
  __shared__ G1::P res[2];
  const uint32_t lane = threadIdx.x;
  if (lane < 2) {
    const G1::P base = from_aff<Fq>(lane == 0 ? args.alpha1 : args.beta1);
    const G1::P dt = G1::load(cb + 32 * (lane == 0 ? CB_RD1 : CB_SD1));
    const G1::P x = G1::add(G1::add(G1::load(cb + 32 * (CB_MA + lane)), dt), base);
    if (lane == 0) G1::store(cb + 32 * CB_A, x);
    res[lane] = scalar_mul<Fq>(x, args.sc + 8 * (1 - lane), tab[lane]);
  }
  __syncthreads();
  if (lane == 0) {
    G1::P c0 = G1::add(res[0], res[1]);
    c0 = G1::add(c0, G1::neg(G1::load(cb + 32 * CB_RSD1)));
    c0 = G1::add(c0, G1::load(cb + 32 * CB_ML));
    G1::store(cb + 32 * CB_C0, c0);
  }
}

// A affine into the proof (beside MSM H, on the G1 batch's stream).
__global__ __launch_bounds__(64) void k_affine_a(uint32_t* __restrict__ cb) {
  __builtin_amdgcn_s_setprio(3);  // a latency chain beside the MSMs' throughput waves
  if (threadIdx.x == 0)
    G1::store_aff(cb + CB_PROOF_OFF, G1::aff_canon(G1::to_affine_inline(G1::load(cb + 32 * CB_A))));
}

// B = beta2 + MSM_B2 + s delta2, affine, into the proof.
__global__ __launch_bounds__(64) void k_combine_b2(CombineArgs args, const uint32_t* __restrict__ outPQ,
                                                   uint32_t* __restrict__ cb) {
  __builtin_amdgcn_s_setprio(3);  // a latency chain beside the MSMs' throughput waves
  __shared__ G2::P sh[3 * 64];
  const G2::P m = finish_msm<Fq2>(outPQ, args.hn_b2, args.log2m_b2, sh);
  if (threadIdx.x == 0) {
    const G2::P b = G2::add(G2::add(m, G2::load(cb + CB_G2_OFF)), from_aff<Fq2>(args.beta2));
    G2::store_aff(cb + CB_PROOF_OFF + 16, G2::aff_canon(G2::to_affine_inline(b)));
  }
}

// C = C0 + MSM_H, affine into the proof.
__global__ __launch_bounds__(64) void k_combine_final(CombineArgs args, const uint32_t* __restrict__ outPQ,
                                                      uint32_t* __restrict__ cb) {
  __builtin_amdgcn_s_setprio(3);  // a latency chain beside the MSMs' throughput waves
  __shared__ G1::P sh[3 * 64];
  const G1::P h = finish_msm<Fq>(outPQ, args.hn_h, args.log2m_h, sh);
  if (threadIdx.x == 0) {
    const G1::P c = G1::add(G1::load(cb + 32 * CB_C0), h);
    G1::store_aff(cb + CB_PROOF_OFF + 48, G1::aff_canon(G1::to_affine_inline(c)));
  }
}

void launch_delta_terms(const CombineArgs& a, uint32_t* cb, hipStream_t s) {
  k_delta_terms<<<4, 64, 0, s>>>(a, cb);
  BG_HIP(hipGetLastError());
}
void launch_combine_g1(const CombineArgs& a, const uint32_t* outPQ, uint32_t* cb, hipStream_t s) {
  k_finish_g1<<<3, 64, 0, s>>>(a, outPQ, cb);
  k_combine_g1<<<1, 64, 0, s>>>(a, cb);
  BG_HIP(hipGetLastError());
}
void launch_affine_a(uint32_t* cb, hipStream_t s) {
  k_affine_a<<<1, 64, 0, s>>>(cb);
  BG_HIP(hipGetLastError());
}
void launch_combine_b2(const CombineArgs& a, const uint32_t* outPQ, uint32_t* cb, hipStream_t s) {
  k_combine_b2<<<1, 64, 0, s>>>(a, outPQ, cb);
  BG_HIP(hipGetLastError());
}
void launch_combine_final(const CombineArgs& a, const uint32_t* outPQ, uint32_t* cb, hipStream_t s) {
  k_combine_final<<<1, 64, 0, s>>>(a, outPQ, cb);
  BG_HIP(hipGetLastError());
}

}  // namespace bg
