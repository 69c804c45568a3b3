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
    void* order_dev = nullptr;        // replicas by descending cost: the blocks' update order
    int32_t* agent_offset_host = nullptr;
    int64_t* field_offset_host = nullptr;
};

// The resource patches' per-replica constants and pending-placement scratch
// behind the opaque abm_forage_world handle.
struct abm_forage_world {
    abm_forage_batch* batch = nullptr;
    int32_t n_replicas = 0;
    int64_t patch_total = 0;
    size_t member_lds = 0;           // bytes of the membership masks per block
    void* wreps_dev = nullptr;
    int32_t* kill_rank = nullptr;    // [patch_total], -1 = not killed
    int32_t* pend = nullptr;         // [2 * n_replicas]: kind, count
    int32_t* pend_id = nullptr;      // [patch_total]
    int64_t* patch_offset_host = nullptr;
};
