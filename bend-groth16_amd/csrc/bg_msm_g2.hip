// G2 (over Fq2) instantiation of the MSM pipeline.
#include "bg_msm_curve.hpp"

namespace bg {
void enqueue_g2(uint32_t M, const uint32_t* const* d_points, const uint32_t* const* d_scalars, const uint64_t* ns,
                int window_bits, hipStream_t s, MsmWork& w, bool fixed) {
  enqueue_curve<Fq2>(M, d_points, d_scalars, ns, window_bits, s, w, 2, fixed);
}
void fixed_table_g2(const uint32_t* d_points, uint64_t n, int c, uint32_t* d_out, hipStream_t s) {
  fixed_table_curve<Fq2>(d_points, n, c, d_out, s);
}
uint64_t accumulate_threads_g2() { return accumulate_round_threads<Fq2>(); }
uint32_t reduce2_quad_kmax_g2() { return reduce2_quad_kmax<Fq2>(); }
}  // namespace bg
