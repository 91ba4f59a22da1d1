// MLP continuation regressor for LSMC (placeholder until the MFMA kernels land).
#include "lsmc.hpp"
namespace kg {
void american_mlp(kg_context*, LsmcHost&, const kg_lsmc_options&, kg_lsmc_report*) {
  fail(KG_UNSUPPORTED_OPERATION, "not built yet");
}
}  // namespace kg
