// forage.hip -- foraging (Agent.update) step.  Placeholder until the kernel lands.
#include "../../include/abm_amd.h"
#include "forage_batch.hpp"

namespace abm { int set_error(int code, const char* fmt, ...); }
using namespace abm;

extern "C" {
int abm_forage_batch_create(const abm_forage_replica*, int32_t, abm_forage_batch** out) {
    if (out) *out = nullptr;
    return set_error(ABM_ERR_UNSUPPORTED, "foraging step not built yet");
}
int abm_forage_batch_destroy(abm_forage_batch*) { return ABM_OK; }
int abm_forage_batch_get_info(const abm_forage_batch*, abm_batch_info*) {
    return set_error(ABM_ERR_UNSUPPORTED, "foraging step not built yet");
}
int abm_forage_batch_get_offsets(const abm_forage_batch*, int32_t*, int64_t*) {
    return set_error(ABM_ERR_UNSUPPORTED, "foraging step not built yet");
}
int abm_forage_step_f64(abm_forage_batch*, const abm_forage_state*, const double*, const int64_t*,
                        const int32_t*, int32_t*, uint32_t*, void*) {
    return set_error(ABM_ERR_UNSUPPORTED, "foraging step not built yet");
}
int abm_forage_projection(int32_t, double, double, int32_t, double, int32_t, double, double, double,
                          double, const double*, const double*, int32_t, int32_t, double*, void*) {
    return set_error(ABM_ERR_UNSUPPORTED, "foraging step not built yet");
}
}
