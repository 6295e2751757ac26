// Float32 SGM path launcher (the fused pipeline's post-processing chain).
#include "sgm_scan.hpp"

namespace cbmv {

int launch_sgm_path_f32(const float* x0, const float* x1, const float* in, float* out, int H,
                        int W, int D, VolStrides st, float pi1, float pi2, float tau_so,
                        float alpha1, float q1, float q2, int direction, int path, int mode,
                        double scale, hipStream_t s) {
  return sgm_path<float>(x0, x1, in, out, H, W, D, st, pi1, pi2, tau_so, alpha1, q1, q2,
                         direction, path, mode, scale, s);
}

}  // namespace cbmv
