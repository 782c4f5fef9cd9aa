#include "itnbp_ops.hpp"
extern "C" {
int itnbp_energy_grad(itnbp_net* net, int, int, const double*, const double*, double*, double*) {
  return itnbp_fail(net ? net->ctx : nullptr, ITNBP_ERR_UNSUPPORTED, "not built yet");
}
int itnbp_update_batched(int, itnbp_net**, const itnbp_update_opts*, itnbp_update_result*) {
  return ITNBP_ERR_UNSUPPORTED;
}
}
