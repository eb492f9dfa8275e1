// RK5 kernels of the models with two fixed ports (rk5_models.h):
// aeif_psc_exp, aeif_cond_alpha, aeif_cond_beta, aeif_psc_alpha.
#include "rk5_device.h"

namespace ngpu {

void launch_rk5_two_port(int model, const RK5Args& a, hipStream_t s) {
  switch (model) {
    case MODEL_AEIF_PSC_EXP: launch_rk5_t<AeifPscExp>(a, s); break;
    case MODEL_AEIF_COND_ALPHA: launch_rk5_t<AeifCondAlpha2>(a, s); break;
    case MODEL_AEIF_COND_BETA: launch_rk5_t<AeifCondBeta2>(a, s); break;
    case MODEL_AEIF_PSC_ALPHA: launch_rk5_t<AeifPscAlpha>(a, s); break;
    default: throw std::runtime_error("RK5: not a two-port model");
  }
}

bool layout_two_port(int model, int iv, int is, int wp, int ws) {
  switch (model) {
    case MODEL_AEIF_PSC_EXP: return fold_layout_is<AeifPscExp>(iv, is, wp, ws);
    case MODEL_AEIF_COND_ALPHA: return fold_layout_is<AeifCondAlpha2>(iv, is, wp, ws);
    case MODEL_AEIF_COND_BETA: return fold_layout_is<AeifCondBeta2>(iv, is, wp, ws);
    case MODEL_AEIF_PSC_ALPHA: return fold_layout_is<AeifPscAlpha>(iv, is, wp, ws);
    default: return false;
  }
}

}  // namespace ngpu
