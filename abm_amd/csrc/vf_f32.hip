// vf_f32.hip -- fused float32 step.  Placeholder until the kernel lands.
#include "vf_batch.hpp"
namespace abm {
int launch_vf_step_f32(abm_vf_batch*, const abm_vf_state_f32*, uint32_t*, float*, uint32_t, hipStream_t) {
    return set_error(ABM_ERR_UNSUPPORTED, "abm_vf_step_f32 not built yet");
}
}
