// forage_batch.hpp -- the batch object behind the opaque abm_forage_batch handle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/abm_amd.h"

struct abm_forage_batch {
    int32_t n_replicas = 0;
    int32_t n_agents_total = 0;
    int64_t field_words = 0;
    int32_t max_agents = 0;
    int32_t max_rays = 0;
    int device = 0;
    void* reps_dev = nullptr;
    int32_t* agent_offset_host = nullptr;
    int64_t* field_offset_host = nullptr;
};
