// Radix-2 NTT over BN254 Fr and the QAP quotient (C ABI: bg_ntt, bg_qap_h).
//
// Replaces Ntt.fft / ifft / coset_fft / coset_ifft and Ntt.qap_h (bend2/src/ntt.bend:
// 431-459).  Natural order in and out.  Per domain size the library keeps twiddle and
// scale tables resident in HBM (per-stage omega^(j 2^k), n^-1 g^j, n^-1 g^-j), built on the
// device from two short host-computed geometric tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bg_runtime.hpp"

namespace bg {

struct NttDomain {
  int log2 = 0;
  uint64_t n = 0;
  DevArray tw_fwd;       // per-stage twiddles: stage k's omega^(j 2^k), j < n/2^(k+1), at n - n/2^k
  DevArray tw_inv;       // the same for omega^-1
  DevArray sc_g;         // g^j (Montgomery)                    -- coset_fft input scale
  DevArray sc_ninv_g;    // n^-1 g^j (Montgomery)               -- ifft then coset_fft
  DevArray sc_ninv_gi;   // n^-1 g^-j (Montgomery)              -- coset_ifft output scale
  DevArray sc_ninv_gi_c; // n^-1 g^-j as raw integers: Montgomery product gives canonical out
  DevArray sc_ninv_g_rev; // n^-1 g^rev(j) (Montgomery)          -- coset scale in bit-reversed order
  uint32_t ninv[8];      // n^-1 (Montgomery)
  uint32_t zinv[8];      // (g^n - 1)^-1 (Montgomery)
};

// Cached per log2 (thread-safe); tables are built on the context stream.
NttDomain& ntt_domain(int log2);

// In-place transforms on device data (packed Montgomery Fr).  tmp: n elements scratch.
void ntt_run(uint32_t* d, uint32_t* tmp, int log2, int direction, hipStream_t s);

// qap_h on device buffers a, b, c (clobbered); h receives n coefficients, canonical if
// canonical_out (the MSM scalar form, bend2/prove.bend:56), else Montgomery.  batch > 1:
// a, b, c, h each hold that many contiguous n-element instances, transformed together.
void qap_h_run(uint32_t* a, uint32_t* b, uint32_t* c, uint32_t* h, uint32_t* tmp, int log2, bool canonical_out,
               hipStream_t s, unsigned batch = 1);

}  // namespace bg
