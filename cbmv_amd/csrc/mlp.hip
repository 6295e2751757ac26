// MLP classifier (placeholder until the MFMA kernels land).
#include "mlp.hpp"

namespace cbmv {

int mlp_upload(MlpDevice& m, const float*, const float*, const float*, const float*,
               const float*, const float*, const float*, const float*) {
  m.ready = false;
  set_error("MLP not available in this build");
  return CBMV_ESTATE;
}

int mlp_classify(const MlpDevice&, const MlpRun&, hipStream_t) {
  set_error("MLP not available in this build");
  return CBMV_ESTATE;
}

}  // namespace cbmv
