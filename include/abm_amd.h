/*
 * abm_amd.h -- C ABI of the MI355X (gfx950) perception/motion step for
 * scioip34/ABM agents.
 *
 * Plain C: device pointers, sizes and POD descriptors only.  Every entry
 * point returns 0 on success or a negative ABM_ERR_* code; the message of the
 * last failure on the calling thread is returned by abm_last_error().  No
 * entry point throws across the ABI.  Kernels are launched on the caller's
 * stream (a hipStream_t passed as void*; NULL = the null stream) and are
 * stream ordered; a call allocates nothing except abm_*_create and the first
 * abm_vf_state_variables call for an n_rays (its cached ray tables).
 *
 * Each step entry point replaces one call of the reference's per-step agent
 * update (pygame Group.update over all agents):
 *   abm_vf_step_f64 / abm_vf_step_f32
 *       -> VFSimulation.step_sim: self.agents.update(self.agents)
 *          (abm/projects/visual_flocking/vf_simulation/vf_sims.py:302), i.e.
 *          VFAgent.update for every agent in id order
 *          (abm/projects/visual_flocking/vf_agent/vf_agent.py:52-77)
 *   abm_forage_step_f64
 *       -> Simulation.start: self.agents.update(self.agents)
 *          (abm/simulation/sims.py:861), i.e. Agent.update in id order
 *          (abm/agent/agent.py:212-283)
 *   abm_forage_world_run
 *       -> the whole loop body of Simulation.start with agent collisions off:
 *          agent-resource interaction and patch regeneration
 *          (sims.py:789-858, 321-368), then the update above
 * The probe entry points replace single functions:
 *   abm_vf_projection_rows   -> vf_supcalc.projection_field
 *                               (abm/projects/visual_flocking/vf_agent/vf_supcalc.py:20-138)
 *   abm_vf_state_variables   -> vf_supcalc.VSWRM_flocking_state_variables and
 *                               dPhi_V_of (vf_supcalc.py:161-277)
 *   abm_forage_projection    -> Agent.projection_field (abm/agent/agent.py:457-597)
 *   abm_cs_projection        -> cooperative-signalling projection_field and the
 *                               two combines of CSAgent (see below)
 */
#ifndef ABM_AMD_H
#define ABM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define ABM_OK               0
#define ABM_ERR_INVALID     -1  /* bad argument, size or descriptor          */
#define ABM_ERR_HIP         -2  /* a HIP runtime call failed                  */
#define ABM_ERR_UNSUPPORTED -3  /* valid request this build cannot run        */

/* ------------------------------------------------------------------ enums */
#define ABM_BOUNDARY_WALLS    0   /* Agent.reflect_from_walls (agent.py:347-395)         */
#define ABM_BOUNDARY_INFINITE 1   /* torus images + VFAgent.teleport_infinite_arena      */

/* step flags */
#define ABM_ORDER_SEQUENTIAL  0u  /* reference order: agent k sees agents <k moved      */
#define ABM_ORDER_SYNCHRONOUS 1u  /* every field from the pre-step poses (not reference) */
#define ABM_WANT_DIAG         2u  /* write per-agent (dv, dpsi, a_blob, a_edge, b_blob, b_edge) */
#define ABM_COUNT_RECHECKS    4u  /* fused step: count pairs and float64 re-checks       */
#define ABM_VERIFY_FAST_PATH  8u  /* fused step: also decide every pair in float64 and
                                     count float32 decisions that differ (testing)     */
#define ABM_KAHAN_SUMS       16u  /* fused step, grouped kernel: Kahan-compensated ray
                                     sums, group sums in float64 (ABM_ERR_UNSUPPORTED
                                     for batches the grouped kernel does not take)      */

/* ----------------------------------------------------- replica description */
/*
 * One independent simulation (a "replica") of the visual-flocking model.
 * Mirrors the VFSimulation / VFAgent / vf_params attributes the step reads:
 *   n_agents      N                          (Simulation.__init__ N)
 *   n_rays        R = v_field_res after the 1/FOV scaling (vf_sims.py:40-44)
 *   fov0, fov1    (-f*pi, f*pi)              (sims.py:161)
 *   width, height arena size W, H            (agent.py:146-150)
 *   window_pad    arena offset (30 in the apps, app.py:19)
 *   boundary      ABM_BOUNDARY_*             (vf_params.BOUNDARY)
 *   gam .. max_th vf_params constants        (vf_params.py:12-23); ALP2 and BET2
 *                 multiply an identically zero field (vf_supcalc.py:199) and
 *                 are therefore not inputs.
 */
typedef struct abm_vf_replica {
    int32_t n_agents;
    int32_t n_rays;
    int32_t boundary;
    int32_t limit_movement;
    double  fov0, fov1;
    double  width, height, window_pad;
    double  gam, v0, alp0, alp1, bet0, bet1;
    double  max_vel, max_th;
} abm_vf_replica;

/* Offsets the library assigns when a batch is created (replicas packed in order). */
typedef struct abm_batch_info {
    int32_t n_replicas;
    int32_t n_agents_total;   /* length of every per-agent state array            */
    int64_t field_words;      /* uint32 words of one packed field set (all agents) */
    int32_t max_agents;
    int32_t max_rays;
} abm_batch_info;

typedef struct abm_vf_batch abm_vf_batch;   /* opaque */

/* Per-agent state, structure of arrays over all agents of all replicas
 * (replica b owns agents [agent_offset(b), agent_offset(b)+N_b)).  Positions
 * are sprite top-left corners as in the reference (centre = position + radius).
 * alp0/bet0/v0 are optional per-agent overrides (VFAgent.ALP0/BET0/V0,
 * vf_agent.py:24-26); NULL means the replica value. */
typedef struct abm_vf_state_f64 {
    double* x; double* y; double* theta; double* v;
    const int32_t* radius;
    const double* alp0; const double* bet0; const double* v0;
} abm_vf_state_f64;

typedef struct abm_vf_state_f32 {
    float* x; float* y; float* theta; float* v;
    const int32_t* radius;
    const float* alp0; const float* bet0; const float* v0;
} abm_vf_state_f32;

/* Validate host replica descriptors, upload them with the ray tables.
 * Allocates device memory (the only allocating calls are *_create). */
int abm_vf_batch_create(const abm_vf_replica* replicas_host, int32_t n_replicas,
                        abm_vf_batch** out);
int abm_vf_batch_destroy(abm_vf_batch* batch);
int abm_vf_batch_get_info(const abm_vf_batch* batch, abm_batch_info* out);
/* Host copies of the per-replica agent offsets [n_replicas+1] and field word
 * offsets [n_replicas+1] (agent i of replica b has its field at
 * field_offset[b] + i*ceil(R_b/32)). */
int abm_vf_batch_get_offsets(const abm_vf_batch* batch, int32_t* agent_offset,
                             int64_t* field_offset);

/* One step of every replica, plain float64: the reference arithmetic on the
 * GPU, one workgroup per replica, per-ray fields in LDS.
 * field_bits (optional, device, info.field_words uint32): each agent's
 *   binary field in raw raster order (bit k = ray k, before the reference's
 *   flip), as used by the integrals; soc_v_field = flip(bits).
 * diag (optional, device, 6*n_agents_total): with ABM_WANT_DIAG. */
int abm_vf_step_f64(abm_vf_batch* batch, const abm_vf_state_f64* state,
                    uint32_t* field_bits, double* diag, uint32_t flags, void* stream);

/* One step of every replica, fused float32 path: bit-packed fields, float32
 * integrals from float64-built ray tables, float32 poses.  Ray-bin and half-width
 * decisions are made in float32 with a guard band and re-made with the exact
 * float64 arithmetic whenever the float32 value falls inside it, so the
 * rasterised field equals the float64 one on the same poses. */
int abm_vf_step_f32(abm_vf_batch* batch, const abm_vf_state_f32* state,
                    uint32_t* field_bits, float* diag, uint32_t flags, void* stream);

/* Cumulative counters of the fused step since the batch was created, kept
 * only by steps run with ABM_COUNT_RECHECKS / ABM_VERIFY_FAST_PATH: pairs
 * decided, pairs re-decided in float64, and (verify mode) float32 decisions
 * that differed from the float64 ones.  Synchronises the stream. */
int abm_vf_f32_fallback_count(abm_vf_batch* batch, int64_t* pairs_total,
                              int64_t* pairs_rechecked, int64_t* pairs_disagree, void* stream);

/* Pairs the plain float64 step decided within 1e-12 of a tie -- the nearest
 * linspace bin within 1e-12 rad of the second nearest, or proj_size / 2
 * within 1e-12 relative of an integer -- where the device's libm (acos, atan)
 * and the host's may round the deciding value to opposite sides; cumulative
 * over the float64 steps run with ABM_COUNT_RECHECKS.  Synchronises the
 * stream.  No counterpart in the reference (a diagnostic of the parity
 * claim). */
int abm_vf_f64_near_ties(abm_vf_batch* batch, int64_t* near_ties, void* stream);

/* Launch shapes of the fused step (no counterpart in the reference, which
 * has no kernels).  Two kernels implement it: the per-replica kernel, with
 * 1, 2 or 4 waves cooperating on one replica, and the grouped kernel, with
 * 1, 2 or 4 replicas advanced in lockstep by one wave (for batches with one
 * n_rays <= 1024 and at most 256 agents per replica; replicas that differ in
 * more than N and the ALP0/BET0/V0 defaults never share a wave).  Two more
 * give one workgroup to each replica: the 16-wave kernel (every pose in LDS,
 * the pairs of each agent decided in its turn) and the push kernel (every
 * agent's field in LDS, the pairs decided off the id-order chain as soon as
 * their object has moved), which by default runs the large size class of a
 * batch (replicas above 64 agents) when their fields fit and the class's
 * agents, spread over the CUs, are no more than its largest replica's chain
 * (one push workgroup per CU), next to the per-replica kernel for the rest;
 * the 16-wave kernel takes the class when its fields do not fit.
 * Batches of flocks above 2048 agents in the reference order take the chain
 * step: each agent's pairs with later agents at their start poses over the
 * whole GPU, then segments of 256 agents moved in id order by one workgroup,
 * each segment's new poses decided in the later agents' fields over the
 * whole GPU (the push kernel's sums and moves, so the same results).
 * abm_vf_set_launch selects the per-replica kernel with waves_per_group waves
 * per replica (1, 2, 4), the 16-wave kernel for every replica (16), the push
 * kernel for every replica (32; ABM_ERR_UNSUPPORTED at step time when the
 * fields do not fit in LDS), the chain step for every replica in the
 * reference order (64; allocates its buffers when the batch has none: call it
 * outside graph capture), or the automatic choice (0).  The 16-wave and push
 * kernels and the chain step run the reference order only: a step with
 * ABM_ORDER_SYNCHRONOUS under shapes 16, 32 or 64 returns
 * ABM_ERR_UNSUPPORTED, and the automatic choice runs it on the per-replica
 * (or grouped, or for flocks above 2048 agents the wide) kernel. */
int abm_vf_set_launch(abm_vf_batch* batch, int32_t waves_per_group);
/* The kernels the batch's last step launched (ABM_KERNEL_* bits), for
 * callers that log or check the launch choice.  No counterpart in the
 * reference. */
#define ABM_KERNEL_GROUPED 1u
#define ABM_KERNEL_PER_REPLICA 2u
#define ABM_KERNEL_WIDE 4u
#define ABM_KERNEL_BIG 8u
#define ABM_KERNEL_PUSH 16u
#define ABM_KERNEL_F64 32u
#define ABM_KERNEL_CHAIN 64u
int abm_vf_last_kernels(const abm_vf_batch* batch, uint32_t* mask);
/* replicas_per_wave: 1, 2 or 4 = grouped kernel with that many replicas per
 * wave (ABM_ERR_UNSUPPORTED when the batch does not qualify), -1 = per-replica
 * kernel, 0 = automatic choice. */
int abm_vf_set_grouping(abm_vf_batch* batch, int32_t replicas_per_wave);
/* Time every candidate launch shape on the given state (the state is
 * advanced by the trial steps) and keep the fastest.  *chosen: waves per
 * replica (1, 2, 4) of the per-replica kernel, 16 for the 16-wave kernel,
 * 0 for the size classes, or -G for the grouped kernel with G replicas per
 * wave. */
int abm_vf_autotune(abm_vf_batch* batch, const abm_vf_state_f32* state, int32_t trial_steps,
                    int32_t* chosen_waves, float* chosen_ms, void* stream);

/* -------------------------------------------------------------- probes */
/* Per-object projection rows of one focal agent (vf_supcalc.projection_field).
 * All pointers are device pointers.  focal = {x, y, theta, radius};
 * objects: ox[n], oy[n], osize[n].  rows: n_obj * ceil(R/32) uint32, raw
 * raster order.  rec (optional): n_obj * 5 int32 (phi_target, half, start,
 * end, drawn) with drawn = -1 for a skipped (coincident) object. */
int abm_vf_projection_rows(int32_t n_rays, double fov0, double fov1, int32_t boundary,
                           double width, double height,
                           double fx, double fy, double ftheta, double fradius,
                           const double* ox, const double* oy, const double* osize,
                           int32_t n_obj, uint32_t* rows, int32_t* rec, void* stream);

/* VSWRM_flocking_state_variables for n independent (velocity, field) pairs:
 * bits: n * ceil(R/32) uint32 raw raster order (V_now); vel[n]; out: n*6
 * doubles (dvel, dpsi, a_blob, a_edge, b_blob, b_edge).  alp0/bet0/v0 per
 * row.  dphi (optional): n*R int8 values of dPhi_V_of. */
int abm_vf_state_variables(int32_t n_rays, int32_t n, const uint32_t* bits, const double* vel,
                           const double* alp0, const double* bet0, const double* v0,
                           double gam, double alp1, double bet1,
                           double* out, int8_t* dphi, void* stream);

/* abm_vf_state_variables for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference: abm_vf_state_variables
 * allocates its float64 ray tables on first use per R; this variant calls no
 * allocating routine and keeps the tables in a static device array of the
 * library (3 * 8192 doubles = 192 KiB per GPU, held until the process ends),
 * rebuilt by a small kernel on every call.  Same arguments and results as
 * abm_vf_state_variables (cos/sin of the grid from the device's libm, within
 * an ulp of the host's); n_rays above 8192 get ABM_ERR_INVALID.  Calls that
 * overlap on one GPU (other streams or host threads) race on the tables. */
int abm_vf_state_variables_noalloc(int32_t n_rays, int32_t n, const uint32_t* bits,
                                   const double* vel, const double* alp0, const double* bet0,
                                   const double* v0, double gam, double alp1, double bet1,
                                   double* out, int8_t* dphi, void* stream);

/* ------------------------------------------------ cooperative signalling */
/* Field parameters of cs_supcalc.projection_field
 * (abm/projects/cooperative_signaling/cs_agent/cs_supcalc.py:204-295). */
typedef struct abm_cs_params {
    int32_t n_rays;              /* v_field_resolution                          */
    int32_t has_max_proj_size;   /* 0: max_proj_size=None (every size is valid) */
    double  fov0, fov1;          /* fov                                         */
    double  max_proj_size;       /* objects with proj_size > this are not drawn */
} abm_cs_params;

#define ABM_CS_ROWS 0   /* the per-object rows projection_field returns        */
#define ABM_CS_SUM  1   /* rows summed over objects, / norm[f]
                           (CSAgent.calc_crowing_density_proj, cs_agent.py:83-104) */
#define ABM_CS_MAX  2   /* maximum over rows, 0 without objects
                           (CSAgent.calc_others_signaling_density_proj,
                            cs_agent.py:106-131)                               */

/* projection_field of the cooperative-signalling agent for n_focal focal
 * agents, each with its own object list, in float64 with the reference's
 * operation order.  Replaces cs_supcalc.projection_field (rows) and the
 * visual_field.sum(axis=0) / len(agents) and visual_field.max(axis=0) steps of
 * CSAgent.  All arrays are device arrays:
 *   focal[n_focal][4]: x, y, radius, orientation (position = top-left);
 *   obj_offset[n_focal + 1]: objects of focal f are obj[obj_offset[f] ..
 *     obj_offset[f + 1]), each {x, y} (object_positions);
 *   meters (optional): one per object (object_meters; NULL = None);
 *   norm (optional, ABM_CS_SUM): divisor per focal agent (len(agents));
 *   out: ABM_CS_ROWS: [total objects][n_rays] rows in the returned order
 *     (flipped, FOV-masked); ABM_CS_SUM / ABM_CS_MAX: [n_focal][n_rays].
 * n_rays in [2, 7936]. */
int abm_cs_projection(const abm_cs_params* params, int32_t n_focal, const double* focal,
                      const int64_t* obj_offset, const double* obj, const double* meters,
                      const double* norm, int32_t mode, double* out, void* stream);

/* ------------------------------------------------- ensemble statistics */
/* Per-replica statistics of the current state, as abm/loader/data_loader.py
 * computes them offline from the logged trajectories (positions logged as
 * int(position), abm/monitoring/ifdb.py:83-84):
 *   out[4 b + 0] polarisation |sum (cos th, sin th)| / N (calculate_polarization,
 *                data_loader.py:1761-1836; NaN orientations skipped as nansum does)
 *   out[4 b + 1] mean inter-individual distance over pairs i < j
 *                (calculate_interindividual_distance, :1367-1459)
 *   out[4 b + 2] mean over agents of the nearest-neighbour distance
 *   out[4 b + 3] the value calculate_mean_NN_dist (:1461-1486) returns for the
 *                step: min_j>0 d(0, j) / N, because its nanmin runs over an iid
 *                matrix that holds distances only above the diagonal
 * agent_offset[n_replicas + 1] (device): agents of replica b are
 * [agent_offset[b], agent_offset[b + 1]).  x, y, theta: device arrays of
 * float32 (state_f32 = 1) or float64.  periodic: distances on a torus of
 * int(width) x int(height) (supcalc.distance_torus).  A replica with one
 * agent gives NaN distances. */
int abm_ensemble_stats(int32_t n_replicas, const int32_t* agent_offset, const void* x, const void* y,
                       const void* theta, int32_t state_f32, int32_t periodic, double width,
                       double height, double* out, void* stream);

/* ------------------------------------------------------- agent records */
/* Device arrays of a run's per-step agent records, indexed [t][agent]
 * (edges [t][agent][ceil(R/32)] bits).  The optional arrays may be NULL. */
typedef struct abm_log_buffers {
    int32_t* posx; int32_t* posy;            /* int(position): INT32_MIN if not finite */
    double* orientation; double* velocity;
    double* w; double* u; double* I_priv; double* collected_r;   /* optional */
    int32_t* mode; int32_t* expl_patch_id;                        /* optional */
    uint32_t* up; uint32_t* down;            /* optional: vfield_up / vfield_down as bits */
} abm_log_buffers;

/* One step of ifdb.save_agent_data_RAM (abm/monitoring/ifdb.py:55-96) for
 * n_agents agents into slot t (0 <= t < capacity) of buf.  x, y, theta, v:
 * float32 (state_f32 = 1) or float64 device arrays; the foraging arrays are
 * optional.  field_bits (optional): ceil(n_rays/32) words per agent, the
 * agent's soc_v_field in output order (raw_raster = 0, the foraging field) or
 * its raw raster (raw_raster = 1, the VSWRM field, soc_v_field =
 * flip(raster)); bit k of up / down is set where np.roll(f, 1)[k] < f[k] /
 * > f[k]. */
int abm_log_record(const abm_log_buffers* buf, int64_t t, int64_t capacity, int32_t n_agents,
                   const void* x, const void* y, const void* theta, const void* v, int32_t state_f32,
                   const double* w, const double* u, const double* I_priv, const double* collected_r,
                   const int32_t* mode, const int32_t* expl_patch_id, const uint32_t* field_bits,
                   int32_t n_rays, int32_t raw_raster, void* stream);

/* ---------------------------------------------------------- foraging */
/* One foraging replica (Simulation + Agent parameters, abm/simulation/sims.py:60-161,
 * abm/contrib/decision_params.py:13-42, movement_params.py:13-24). */
typedef struct abm_forage_replica {
    int32_t n_agents;
    int32_t n_rays;
    int32_t visual_exclusion;     /* VISUAL_EXCLUSION                       */
    int32_t patchwise_exclusion;  /* PATCHWISE_SOCIAL_EXCLUSION             */
    int32_t tau;                  /* novelty ring length (DEC_TAU)          */
    int32_t pad_;
    double  fov0, fov1;
    double  width, height, window_pad;
    double  vision_range;
    /* decision process */
    double  T_w, Eps_w, g_w, B_w, w_max;
    double  T_u, Eps_u, g_u, B_u, u_max;
    double  S_wu, S_uw, F_N, F_R;
    /* movement */
    double  exp_vel_max, exp_theta_min, exp_theta_max;
    double  reloc_theta_max, exp_stop_ratio;
} abm_forage_replica;

/* Agent modes as stored in abm_forage_state.mode (the reference's strings). */
#define ABM_MODE_EXPLORE  0
#define ABM_MODE_EXPLOIT  1
#define ABM_MODE_RELOCATE 2
#define ABM_MODE_COLLIDE  3
#define ABM_MODE_POOL     4

typedef struct abm_forage_batch abm_forage_batch;   /* opaque */

/* Foraging agent state (SoA over all agents of all replicas).
 * overriding_mode: -1 = None, else ABM_MODE_*.  novelty: tau doubles per agent
 * (ring as a shifted array, index 0 newest).  is_moved_with_cursor is not
 * modelled (always 0 in headless runs). */
typedef struct abm_forage_state {
    double* x; double* y; double* theta; double* v;
    double* w; double* u; double* I_priv;
    double* collected_r; double* collected_r_before;
    double* novelty;
    int32_t* mode; int32_t* overriding_mode;
    int32_t* env_status; int32_t* exploited_patch_id;
    const int32_t* radius;
} abm_forage_state;

int abm_forage_batch_create(const abm_forage_replica* replicas_host, int32_t n_replicas,
                            abm_forage_batch** out);
int abm_forage_batch_destroy(abm_forage_batch* batch);
int abm_forage_batch_get_info(const abm_forage_batch* batch, abm_batch_info* out);
int abm_forage_batch_get_offsets(const abm_forage_batch* batch, int32_t* agent_offset,
                                 int64_t* field_offset);

/* One Agent.update for every agent, reference order, float64.  One
 * workgroup per replica (256 threads, or 512 when a replica has more than
 * 128 agents) holds the replica in LDS: about 172 bytes per agent plus one
 * byte per ray, so up to about 920 agents per replica at R = 1200 on gfx950
 * (160 KB).  Larger replicas (up to 65536 agents) are updated by sweeps (see
 * abm_forage_set_path); with visual exclusion they get ABM_ERR_UNSUPPORTED
 * from abm_forage_batch_create. */
/* Arguments:
 * uniforms: per replica, a device array of the next numpy legacy-MT19937
 *   doubles (random_sample) of the global stream, uniform_offset[b] its start
 *   in the flat array and uniform_count[b] how many are available; the
 *   random walk draws lo + (hi-lo)*u from it in agent order
 *   (abm/agent/supcalc.py:38-47).
 * uniforms_used[b] (device, int32): how many were consumed, so the host can
 *   advance its global generator by exactly that many draws.
 * field_bits (optional): soc_v_field in output order (bit k = soc_v_field[k]). */
int abm_forage_step_f64(abm_forage_batch* batch, const abm_forage_state* state,
                        const double* uniforms, const int64_t* uniform_offset,
                        const int32_t* uniform_count, int32_t* uniforms_used,
                        uint32_t* field_bits, void* stream);

/* Update paths of abm_forage_step_f64 and abm_forage_world_run* (no
 * counterpart in the reference, which has one Python loop).  A replica that
 * fits one workgroup's LDS is updated by the step kernel; a larger one
 * (visual exclusion off, at most 8192 rays) by sweeps: every agent's update
 * recomputed at once from the latest tentative states of the agents before
 * it, repeated until nothing changes, which is the reference's id-order
 * result exactly.  A call with such a replica reads a few control words per
 * round of sweeps, so it synchronises the stream.  ABM_FORAGE_PATH_SWEEP
 * sweeps every replica without visual exclusion (same results). */
#define ABM_FORAGE_PATH_AUTO  0
#define ABM_FORAGE_PATH_SWEEP 1
int abm_forage_set_path(abm_forage_batch* batch, int32_t path);
/* Replica updates made by sweeps since the batch was created, and the sweeps
 * they took (the last one of each confirms the fixed point). */
int abm_forage_sweep_counts(const abm_forage_batch* batch, int64_t* updates, int64_t* sweeps);

/* Agent.projection_field for one focal agent (abm/agent/agent.py:457-597).
 * Obstacles are the social cues, followed by n_nonsocial non-social
 * occluders when non_social >= 0 (the non_expl_agents argument); fov given
 * explicitly.  Output: out_field[R] doubles (the returned array, after flip and
 * FOV mask); keep_distance selects keep_distance_info. */
int abm_forage_projection(int32_t n_rays, double fov0, double fov1, int32_t visual_exclusion,
                          double vision_range, int32_t keep_distance,
                          double fx, double fy, double ftheta, double fradius,
                          const double* ox, const double* oy, int32_t n_social,
                          int32_t n_nonsocial, double* out_field, void* stream);

/* numpy's legacy MT19937 state (numpy/random/src/mt19937/mt19937.h,
 * mt19937_state): the address a RandomState's bit generator reports as
 * ctypes.state_address points at exactly this struct. */
typedef struct abm_mt19937_state {
    uint32_t key[624];
    int32_t pos;
} abm_mt19937_state;

/* ------------------------------------------- agent-resource interaction */
/* The resource patches of every foraging replica and the loop body around
 * Agent.update that uses them (abm/simulation/sims.py:736-861): agent-agent
 * collisions when on (sims.py:736-787: pygame.sprite.groupcollide with
 * collide_circle on the agents' rect centres, int(position) + radius, with
 * every radius grown by 2; the proximity response of
 * agent_agent_collision_proximity, sims.py:421-467, from the
 * keep_distance_info projection field of the agents within 2 r + 20; the
 * collide modes ended or notified, replicas of up to 2048 agents), patch
 * membership of agent centres (groupcollide +
 * refine_ar_overlap_group, sims.py:45-56, 790-799), bias towards the patch
 * centre (bias_agent_towards_res_center, sims.py:544-552), notify_agent
 * (sims.py:29-42), teleport, Rescource.deplete (abm/environment/rescource.py:
 * 118-133), kill_resource and its regeneration (add_new_resource_patch,
 * sims.py:321-368) and the off-patch notifications (sims.py:847-855).
 * Per-replica parameters (Simulation kwargs, sims.py:60-68, 160-177). */
typedef struct abm_forage_world_params {
    int32_t n_patches;            /* N_RESOURCES: patches created, the list's capacity (<= 64) */
    int32_t patch_radius;         /* RADIUS_RESOURCE                                  */
    int32_t min_units, max_units; /* np.random.randint(min_units, max_units), max already
                                     fixed up as Simulation does when negative          */
    int32_t pooling_time;         /* POOLING_TIME                                     */
    int32_t regenerate;           /* REGENERATE_PATCHES                               */
    int32_t teleport_exploit;     /* TELEPORT_TO_MIDDLE                               */
    int32_t allow_border_overlap; /* ALLOW_BORDER_PATCH_OVERLAP                       */
    int32_t collide_agents;       /* AGENT_AGENT_COLLISION                            */
    int32_t ghost_mode;           /* GHOST_WHILE_EXPLOIT                              */
    double  consumption;          /* AGENT_CONSUMPTION                                */
    double  min_quality, max_quality;  /* np.random.uniform bounds (max fixed up)     */
} abm_forage_world_params;

/* Patches of all replicas, in each replica's Group order.  Replica b owns the
 * slots [patch_offset[b], patch_offset[b] + n_patches_b) of every array, of
 * which the first count[b] are live. */
typedef struct abm_patch_state {
    int32_t* count;               /* [n_replicas] */
    int32_t* id;
    double* x; double* y;         /* position (top-left corner), centre = position + radius */
    double* radius;
    double* units;                /* resc_units   */
    double* left;                 /* resc_left    */
    double* quality;              /* unit_per_timestep */
} abm_patch_state;

typedef struct abm_forage_world abm_forage_world;   /* opaque */

int abm_forage_world_create(abm_forage_batch* batch, const abm_forage_world_params* params_host,
                            abm_forage_world** out);
int abm_forage_world_destroy(abm_forage_world* world);
/* patch_offset[n_replicas + 1] (host): each replica's first patch slot. */
int abm_forage_world_get_offsets(const abm_forage_world* world, int64_t* patch_offset);

/* Phases of abm_forage_world_run. */
#define ABM_WORLD_STEP   0   /* interaction, regeneration, then Agent.update      */
#define ABM_WORLD_RESUME 1   /* finish what a phase left waiting for words        */
#define ABM_WORLD_FILL   2   /* add patches up to n_patches (create_resources,
                                sims.py:538-542, ids max + 1)                     */
/* Per-replica status after a run. */
#define ABM_WORLD_DONE       0
#define ABM_WORLD_NEED_WORDS 1   /* the word buffer ran out: nothing drawn, resume  */
#define ABM_WORLD_TIMEOUT    2   /* more than 10000 placement retries (the reference
                                    raises "Reached timeout while trying to create
                                    resources without overlap!")                   */
#define ABM_WORLD_BAD_COUNT  4   /* count[b] outside [0, n_patches]: nothing done    */
#define ABM_WORLD_CROWDED    5   /* a colliding agent has more than 256 agents within
                                    2 r + 20 (the proximity field's limit): nothing
                                    done                                            */

/* One phase for every replica.  words: per replica the next raw 32-bit outputs
 * of its numpy legacy MT19937 stream (abm_mt19937_peek32), word_offset[b] /
 * word_count[b] locating them; patch regeneration draws from them first
 * (randint by masked rejection, uniform from two words, as numpy's legacy
 * RandomState does), then the random walk of Agent.update.  A STEP keeps
 * 2 * n_agents words for the walk.  words_used[b] (device): words consumed;
 * status[b] (device): ABM_WORLD_*.  A replica left NEED_WORDS has consumed
 * nothing of this buffer; a RESUME with more words finishes it (and its
 * Agent.update when the phase was a STEP) and leaves every other replica
 * untouched.  pool_success: device int32 per agent (0 until notified).
 * field_bits as in abm_forage_step_f64. */
int abm_forage_world_run(abm_forage_world* world, int32_t phase, const abm_forage_state* state,
                         int32_t* pool_success, const abm_patch_state* patches,
                         const uint32_t* words, const int64_t* word_offset,
                         const int32_t* word_count, int32_t* words_used, int32_t* status,
                         uint32_t* field_bits, void* stream);

/* The same loop body with each replica's numpy legacy MT19937 state kept in
 * device memory (mt: one abm_mt19937_state per replica, the layout numpy
 * keeps behind a RandomState): the placements draw from it in the interaction
 * kernel and the random walk in the update, which advances it by exactly the
 * words used, so no host peek, skip or synchronisation is needed per step and
 * no replica ever waits for words.  Phases ABM_WORLD_STEP and ABM_WORLD_FILL.
 * status[b] and words_used[b] (placements + walk) as above, read when the
 * caller wishes; ABM_WORLD_TIMEOUT leaves the replica's generator untouched. */
int abm_forage_world_run_dev(abm_forage_world* world, int32_t phase, const abm_forage_state* state,
                             int32_t* pool_success, const abm_patch_state* patches,
                             abm_mt19937_state* mt_states, int32_t* words_used, int32_t* status,
                             uint32_t* field_bits, void* stream);

/* ------------------------------------------------ host random numbers */

/* Host functions (no GPU).  Replace the per-replica draws of
 * np.random.uniform in the random walk (abm/agent/supcalc.py:38-47):
 * abm_mt19937_peek writes, for each state b, the next counts[b] doubles that
 * random_sample would return, consecutively into out, leaving the states
 * unchanged; abm_mt19937_skip advances each state b by counts[b] draws. */
int abm_mt19937_peek(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts,
                     double* out);
int abm_mt19937_skip(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts);
/* The same for raw 32-bit outputs (next_uint32), the unit the resource loop's
 * randint and uniform draws are made of (sims.py:349-359). */
int abm_mt19937_peek32(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts,
                       uint32_t* out);
int abm_mt19937_skip32(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts);

/* --------------------------------------------------------------- misc */
const char* abm_last_error(void);
/* Library version string and the gfx target it was built for. */
const char* abm_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* ABM_AMD_H */
