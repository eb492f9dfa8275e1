// RK5 kernels of the AeifCondBeta<1..kMaxPort> model family (rk5_models.h),
// instantiated in a file of their own so the families compile in parallel.
#include "rk5_device.h"

namespace ngpu {

void launch_rk5_cond_beta_ms(int n_port, const RK5Args& a, hipStream_t s) {
  dispatch_ports<AeifCondBeta>(n_port, a, s);
}

bool layout_cond_beta_ms(int n_port, int iv, int is, int wp, int ws) {
  return fold_layout_ports<AeifCondBeta>(n_port, iv, is, wp, ws);
}

}  // namespace ngpu
