// vf_batch.hpp -- device-side replica records and the batch object behind the
// opaque abm_vf_batch handle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>


#include "../../include/abm_amd.h"

namespace abm {

// Uniform grid of the wide (large-flock) step over the periodic arena: cells
// at least the visibility cut-off wide, so every pair that can be drawn lies
// in the focal agent's 3 x 3 cells.
constexpr int kWideGridMaxSide = 64;
constexpr int kWideGridMaxCells = kWideGridMaxSide * kWideGridMaxSide;
struct WideGrid {
    int32_t ok;        // the grid path applies (torus, >= 3 cells a side, W/2 and H/2 beyond reach)
    int32_t nx, ny;    // cells a side
    int32_t pad_;
    float cw, ch;      // cell width, height
    float x0, y0;      // arena origin (pad)
    float W, H;
};

// One replica as the kernels read it (built and validated on the host by
// abm_vf_batch_create from an abm_vf_replica).
struct VFRep {
    int32_t agent_offset;   // first agent in the SoA state
    int32_t n_agents;       // N
    int32_t n_rays;         // R
    int32_t words;          // ceil(R / 32)
    int32_t table_offset;   // first ray of this R in the ray tables
    int32_t boundary;       // ABM_BOUNDARY_*
    int32_t fov_px0, fov_px1;
    int32_t phi_ok;         // len(arange(-pi, pi, 2pi/R)) == R
    int32_t limit_movement;
    int64_t field_offset;   // first uint32 word of agent 0's field
    double width, height, pad;
    double gam, v0, alp0, alp1, bet0, bet1;
    double max_vel, max_th;
    double step;            // linspace(-pi, pi, R) step
    double cull_c;          // visibility cut-off d > radius * cull_c: cot(2 pi / R) (1 + 2e-4)
};

// float32 ray-table entry used by the fused kernel, one per ray.
struct RayTab32 {
    float wcos, wsin;   // trapz weight * cos(PHI_k), * sin(PHI_k)
    float cos, sin;     // cos(PHI_k), sin(PHI_k)
};

// Replica groups of the grouped fused kernel (vf_group.hip) for one group size G.
struct GroupPlan {
    int32_t* groups = nullptr;   // [n_groups][G] replica ids, -1 = empty slot
    int32_t n_groups = 0;
    float2* nib = nullptr;       // nibble tables (blob, then spike) in the kernel's LDS layout
    int32_t tab_entries = 0;     // float2 entries per table
};

// Replicas above this many agents form the large size class of a batch:
// the push kernel (or, where their fields do not fit in LDS, the 16-wave
// kernel) updates them, the per-replica kernel the rest.
constexpr int kBigClass = 64;

// The push kernel (one workgroup of 8 waves per replica): rays per lane of
// its summing wave for a launch whose largest R is r (the one-wave
// per-replica kernel's rule), and its LDS for a replica of n agents and r
// rays: start and new poses, speeds and parameters, heading vectors, the
// pushing waves' queues, and every agent's field (one spare word per row).
constexpr int kPushWaves = 8;
constexpr int kPushQueue = 128;   // each pushing wave's ring of queued pairs
inline int push_rpl(int r) {
    const int need = (r + 63) / 64;
    return need <= 4 ? 4 : (need <= 8 ? 8 : (need <= 16 ? 16 : 32));
}
inline size_t lds_bytes_push(int n, int r) {
    const int w1 = ((r + 31) >> 5) + 1;
    const size_t b = sizeof(float) * 4 * 3 * (size_t)n + sizeof(float) * 2 * (size_t)n +
                     sizeof(uint32_t) * kPushQueue * kPushWaves + sizeof(uint32_t) * (size_t)n * w1;
    return (b + 15) & ~(size_t)15;
}

// Largest replica the grouped kernel takes, and the largest R.
constexpr int kGroupMaxAgents = 256;
constexpr int kGroupMaxRays = 1024;

}  // namespace abm

// The opaque handle.
struct abm_vf_batch {
    int32_t n_replicas = 0;
    int32_t n_agents_total = 0;
    int64_t field_words = 0;
    int32_t max_agents = 0;
    int32_t max_rays = 0;
    int32_t n_tables = 0;
    int32_t ray_total = 0;          // entries in each ray table
    int32_t waves_per_group = 0;    // per-replica fused kernel: waves per replica (0 = default)
    int32_t group_ok = 0;           // the grouped kernel can run this batch
    int32_t group_g = 0;            // 0 = automatic, -1 = per-replica kernel, 1/2/4 = replicas per wave
    int device = 0;
    int dev_max_lds = 0;            // device limits, read once at creation
    int n_cu = 0;
    int occ_f32[2][2][2][3] = {};   // [size classes][sync][plain][nw 1,2,4]: groups per CU
    abm::VFRep* reps_dev = nullptr;
    double* tab_cos = nullptr;      // float64 tables: cos(PHI), sin(PHI), diff(PHI)
    double* tab_sin = nullptr;
    double* tab_dphi = nullptr;
    abm::RayTab32* tab32 = nullptr;
    unsigned long long* counters = nullptr;   // [0] pairs, [1] float64 re-checks, [2] disagreements, [3] float64 near ties
    int32_t* order_dev = nullptr;   // replica launch order: the large size class first, each
                                    // class by descending work
    int32_t n_big_class = 0;        // replicas above kBigClass agents (first in order_dev)
    uint32_t last_kernels = 0;      // ABM_KERNEL_* bits of the last step's launches
    size_t push_lds_all = 0;        // push kernel LDS of the largest replica (bytes)
    size_t push_lds_big = 0;        // ... of the largest in the large size class
    int32_t small_max_agents = 0, small_max_rays = 0;   // of the other replicas
    int32_t big_max_agents = 0, big_max_rays = 0;       // of the large class
    int64_t big_sum_agents = 0;     // agents in the large class
    float4* wide_out = nullptr;     // new poses of the wide (large-flock) step
    float4* grid_sorted = nullptr;  // wide step: the replica's agents sorted by grid cell
    int32_t* grid_start = nullptr;  // wide step: first sorted agent of each cell (+ end)
    void* grid_param = nullptr;     // wide step: the grid of the current replica (WideGrid)
    uint32_t* chain_fields = nullptr;   // chain step: every agent's field (max_agents x words)
    void* chain_u = nullptr;        // chain step: start heading vectors (float2 per agent)
    int32_t min_agents = 0;         // of the smallest replica
    abm::GroupPlan gplan[3];        // G = 1, 2, 4
    // host copies
    abm::VFRep* reps_host = nullptr;
    int32_t* agent_offset_host = nullptr;
    int64_t* field_offset_host = nullptr;
};

namespace abm {
// launchers (vf_f64.hip, vf_f32.hip)
int launch_vf_step_f64(abm_vf_batch* b, const abm_vf_state_f64* st, uint32_t* bits,
                       double* diag, uint32_t flags, hipStream_t s);
int launch_vf_step_f32(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits,
                       float* diag, uint32_t flags, hipStream_t s);
// Whether the push kernel holds every replica (all) or those of the large
// size class.
bool push_fits(const abm_vf_batch* b, bool all);
int launch_vf_group_f32(abm_vf_batch* b, int G, const abm_vf_state_f32* st, uint32_t* bits,
                        float* diag, uint32_t flags, hipStream_t s);
int group_rpl(int G, int R);
int selected_group_size(const abm_vf_batch* b);
int set_error(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);
}  // namespace abm
