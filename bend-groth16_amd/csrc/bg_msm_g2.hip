// G2 (over Fq2) instantiation of the MSM pipeline.
#include "bg_msm_curve.hpp"

namespace bg {
void enqueue_g2(const uint32_t* d_points, const uint32_t* d_scalars, uint64_t n, int window_bits, hipStream_t s,
                MsmWork& w) {
  enqueue_curve<Fq2>(d_points, d_scalars, n, window_bits, s, w, 2);
}
}  // namespace bg
