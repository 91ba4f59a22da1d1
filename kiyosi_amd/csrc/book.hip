// Barrier, Asian, accumulator engines and the persistent book kernel
// (placeholders until the book kernel lands).
#include "host_common.hpp"
using namespace kg;
extern "C" {
kg_status kg_barrier_price(kg_context* ctx, const kg_settings*, const kg_barrier_terms*,
                           const kg_bsm_market*, kg_estimate*) {
  return guarded(ctx, [&] { fail(KG_UNSUPPORTED_OPERATION, "not built yet"); });
}
kg_status kg_asian_price(kg_context* ctx, const kg_settings*, const kg_asian_terms*,
                         const kg_bsm_market*, kg_estimate*) {
  return guarded(ctx, [&] { fail(KG_UNSUPPORTED_OPERATION, "not built yet"); });
}
kg_status kg_accumulator_price(kg_context* ctx, const kg_settings*, const kg_accumulator_terms*,
                               const kg_bsm_market*, kg_estimate*) {
  return guarded(ctx, [&] { fail(KG_UNSUPPORTED_OPERATION, "not built yet"); });
}
kg_status kg_price_book(kg_context* ctx, const kg_book_contract*, size_t, int32_t, kg_book_tuning*,
                        kg_estimate*) {
  return guarded(ctx, [&] { fail(KG_UNSUPPORTED_OPERATION, "not built yet"); });
}
}
