// Autocallable engine (placeholder until the path kernel lands).
#include "host_common.hpp"
using namespace kg;
extern "C" kg_status kg_autocallable_price(kg_context* ctx, const kg_settings*,
                                           const kg_autocall_terms*, const kg_bsm_market*,
                                           kg_estimate*) {
  return guarded(ctx, [&] { fail(KG_UNSUPPORTED_OPERATION, "not built yet"); });
}
