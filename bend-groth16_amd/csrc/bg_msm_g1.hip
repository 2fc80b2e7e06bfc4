// G1 (over Fq) instantiation of the MSM pipeline.
#include "bg_msm_curve.hpp"

namespace bg {
void enqueue_g1(uint32_t M, const uint32_t* const* d_points, const uint32_t* const* d_scalars, const uint64_t* ns,
                int window_bits, hipStream_t s, MsmWork& w, bool fixed) {
  enqueue_curve<Fq>(M, d_points, d_scalars, ns, window_bits, s, w, 1, fixed);
}
void fixed_table_g1(const uint32_t* d_points, uint64_t n, int c, uint32_t* d_out, hipStream_t s) {
  fixed_table_curve<Fq>(d_points, n, c, d_out, s);
}
uint64_t accumulate_threads_g1() { return accumulate_round_threads<Fq>(); }
uint32_t reduce2_quad_kmax_g1() { return reduce2_quad_kmax<Fq>(); }
}  // namespace bg
