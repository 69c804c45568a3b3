// forage_batch.hpp -- the batch object behind the opaque abm_forage_batch handle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

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
    // Replicas the step kernel updates (the first n_small of order_dev) and
    // the ones it cannot hold in one workgroup's LDS, updated by sweeps.
    int32_t n_small = 0;
    int32_t small_max_agents = 0, small_max_rays = 0;
    int32_t path = 0;                  // ABM_FORAGE_PATH_*
    std::vector<int32_t> rep_n, rep_r; // agents and rays of every replica
    std::vector<int32_t> order_host;   // every replica, step-kernel ones first by cost
    std::vector<uint8_t> too_big;      // [b]: beyond one workgroup's LDS
    std::vector<uint8_t> exclusion;    // [b]: visual exclusion
    std::vector<int32_t> sweep;        // replicas updated by sweeps
    int32_t sweep_max_agents = 0;      // the sweep buffers' capacity
    void* sweep_mem = nullptr;         // two tentative states, uniforms, candidates, control
    int32_t* sweep_ctrl_host = nullptr;   // pinned copy of the control words
    int64_t sweep_steps = 0, sweeps_done = 0;   // replica updates by sweeps, sweeps they took
};

// The resource patches' per-replica constants and pending-placement scratch
// behind the opaque abm_forage_world handle.
struct abm_forage_world {
    abm_forage_batch* batch = nullptr;
    int32_t n_replicas = 0;
    int64_t patch_total = 0;
    size_t member_lds = 0;           // bytes of the membership masks per block
    size_t collide_lds = 0;          // LDS of the collision kernel (0: no replica collides)
    void* wreps_dev = nullptr;
    int32_t* kill_rank = nullptr;    // [patch_total], -1 = not killed
    int32_t* pend = nullptr;         // [2 * n_replicas]: kind, count
    int32_t* pend_id = nullptr;      // [patch_total]
    int64_t* patch_offset_host = nullptr;
};
