// forage.hip -- collective-foraging agent step (Agent.update for every agent
// in id order) in float64, plus the Agent.projection_field probe.
//
// One workgroup per replica; the replica's agents live in LDS and are
// updated in place in id order (Group.update, abm/simulation/sims.py:861).
// Per focal agent k:
//   1. lanes over the other agents: cue class (agent.py:396-419) and the
//      projection record of each obstacle (agent.py:501-556): closed angle with
//      the strict 0 < a' < pi rule, the focal radius for both centres, int()
//      truncation of t -+ proj_size/2, strict FOV test;
//   2. visual exclusion (agent.py:421-445): records ranked by (distance,
//      obstacle-list position); each record then walks the nearer ones in rank
//      order and applies the three interval rules sequentially;
//   3. raster of the social records with Python slice semantics
//      (agent.py:569-590) into a per-ray byte field; flip and FOV mask
//      (agent.py:593-595) when counting;
//   4. one lane: I_priv, w/u integrators, mode table, random walk from the
//      host-supplied uniforms of numpy's global stream, relocation force,
//      Euler step and walls (agent.py:168-283, supcalc.py:38-47, 81-92).
// Built with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/abm_amd.h"
#include "forage_batch.hpp"
#include "vf_exact.hpp"
#include "vf_f32_common.hpp"

namespace abm {
int set_error(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

namespace {

// Obstacle classes, in the order the reference lists them to projection_field.
constexpr int kSocial = 0;      // exploiting agents of other patches (the cues)
constexpr int kNonExpl = 1;     // in-range agents that are not exploiting (self included)
constexpr int kSamePatch = 2;   // exploiting agents of the focal's own patch
constexpr int kNone = -1;

struct PairFG {
    int skipped, visible, t, start, end;
    double dist, vis;
};

// One obstacle of Agent.projection_field (agent.py:501-556).
ABM_HD PairFG pair_forage_exact(const Focal& f, double ox, double oy, int R, double step,
                                double fov0, double fov1) {
#pragma clang fp contract(off)
    PairFG p;
    p.skipped = 0; p.visible = 0; p.t = 0; p.start = 0; p.end = 0; p.dist = 0.0; p.vis = 0.0;
    if (ox == f.px && oy == f.py) { p.skipped = 1; return p; }
    const double v2ex = ox + f.radius;          // the focal radius for the obstacle too
    const double v2ey = oy + f.radius;
    const double v2x = v2ex - f.cx, v2y = v2ey - f.cy;
    const double n2 = norm2(v2x, v2y);
    const double u2x = v2x / n2, u2y = v2y / n2;
    double d = dot2(f.u1x, f.u1y, u2x, u2y);
    if (d < -1.0) d = -1.0;
    else if (d > 1.0) d = 1.0;
    double a = acos(d);
    const double cross = f.u1x * u2y - f.u1y * u2x;
    if (cross < 0.0) a = -a;
    a = remainder_2pi(a);
    const double ca = (0.0 < a && a < kPi) ? -a : kTwoPi - a;   // strict (agent.py:520)
    p.dist = n2;
    p.vis = 2.0 * atan(f.radius / (1.0 * n2));
    p.t = linspace_argmin(ca, R, step);
    p.visible = (fov0 < ca && ca < fov1) ? 1 : 0;
    const double ps = (p.vis / kTwoPi) * (double)R;
    p.start = (int)trunc((double)p.t - ps / 2.0);
    p.end = (int)trunc((double)p.t + ps / 2.0);
    return p;
}

// Indices of range(R)[a:b] (Python slice semantics), as a half-open range.
ABM_HD void py_slice(int a, int b, int R, int* lo, int* hi) {
    if (a < 0) { a += R; if (a < 0) a = 0; }
    if (a > R) a = R;
    if (b < 0) { b += R; if (b < 0) b = 0; }
    if (b > R) b = R;
    *lo = a; *hi = b > a ? b : a;
}

struct Rec {
    int valid;        // visible record (obstacle passed the FOV test)
    int social;
    int order;        // position in the reference's obstacle list
    int start, end;   // int(t -+ proj_size/2)
    int sx, ex;       // after exclusion
    double dist;
    double vis;
};

// Paint one record with Python slice semantics: wrap pieces get `wrap_code`,
// the main slice `main_code` (agent.py:577-590).
__device__ inline void paint_record(uint8_t* V, const Rec& rc, int R, uint8_t wrap_code,
                                    uint8_t main_code) {
    int s = rc.sx, e = rc.ex, lo, hi;
    if (s < 0) {
        py_slice(R + s, R, R, &lo, &hi);
        for (int r = lo; r < hi; ++r) V[r] = wrap_code;
        s = 0;
    }
    if (e >= R) {
        py_slice(0, e - R, R, &lo, &hi);
        for (int r = lo; r < hi; ++r) V[r] = wrap_code;
        e = R - 1;
    }
    py_slice(s, e, R, &lo, &hi);
    for (int r = lo; r < hi; ++r) V[r] = main_code;
}

// Per-replica constants in device memory.
struct FGRep {
    int32_t agent_offset, n_agents, n_rays, words, visual_exclusion, patchwise, tau, pad_;
    int32_t keep_lo, keep_hi;   // soc_v_field indices the FOV mask keeps (a contiguous range)
    int64_t field_offset;
    double fov0, fov1, width, height, pad, vision_range, step;
    double T_w, Eps_w, g_w, B_w, w_max, T_u, Eps_u, g_u, B_u, u_max, S_wu, S_uw, F_N, F_R;
    double exp_vel_max, exp_theta_min, exp_theta_max, reloc_theta_max, exp_stop_ratio;
};

__device__ inline int get_mode(int overriding, double w, double T_w) {
    // agent.py:659-669
    if (overriding >= 0) return overriding;
    return w > T_w ? ABM_MODE_RELOCATE : ABM_MODE_EXPLORE;
}

struct FGSmem {
    Focal* focal;   // step kernel: focal geometry of the next two agents, by parity
    double *x, *y, *th, *v, *w, *u, *cr, *crb, *nov, *rad;
    int *mode, *over, *env, *pid, *cls;
    Rec* rec;
    int* sorted;
    int* lst;       // step kernel: obstacles needing records, two lists of maxN by parity
    int* cls2;      // step kernel: classes for the agent after next (cls holds the next one's)
    double* uni;    // step kernel: the step's uniforms (or pairs of words), staged
    uint32_t* mtk;  // step kernel: the replica's generator key (device generators)
    uint8_t* V;
    int* cnt;       // step kernel: [2p] total, [2p+1] left half by parity; [4 + p] list sizes
    double* lastd;  // keep_distance: distance of the last processed obstacle
    double* spec;   // step kernel: [12 p ..] an agent's headings after a walk turn or none, with sin and
                    // cos, by parity p; [24 + p] uniforms drawn before an agent of parity p
};

__device__ inline FGSmem fg_carve(void* base, int maxN, int maxR) {
    FGSmem s;
    char* p = reinterpret_cast<char*>(base);
    s.focal = reinterpret_cast<Focal*>(p); p += 128;
    double** dptrs[] = {&s.x, &s.y, &s.th, &s.v, &s.w, &s.u, &s.cr, &s.crb, &s.nov, &s.rad};
    for (double** d : dptrs) { *d = reinterpret_cast<double*>(p); p += sizeof(double) * maxN; }
    s.lastd = reinterpret_cast<double*>(p); p += sizeof(double) * 2;
    s.spec = reinterpret_cast<double*>(p); p += sizeof(double) * 28;
    s.uni = reinterpret_cast<double*>(p); p += sizeof(double) * maxN;
    s.rec = reinterpret_cast<Rec*>(p); p += sizeof(Rec) * maxN;
    int** iptrs[] = {&s.mode, &s.over, &s.env, &s.pid, &s.cls};
    for (int** q : iptrs) { *q = reinterpret_cast<int*>(p); p += sizeof(int) * maxN; }
    s.sorted = reinterpret_cast<int*>(p); p += sizeof(int) * maxN;
    s.lst = reinterpret_cast<int*>(p); p += sizeof(int) * 2 * maxN;
    s.cls2 = reinterpret_cast<int*>(p); p += sizeof(int) * maxN;
    s.cnt = reinterpret_cast<int*>(p); p += sizeof(int) * 8;
    s.mtk = reinterpret_cast<uint32_t*>(p); p += sizeof(uint32_t) * 624;
    s.V = reinterpret_cast<uint8_t*>(p);
    return s;
}

static_assert(sizeof(Focal) <= 64, "two Focal slots of 64 bytes");

size_t fg_lds_bytes(int maxN, int maxR) {
    return 128 + sizeof(double) * ((size_t)11 * maxN + 30) + sizeof(Rec) * (size_t)maxN +
           sizeof(int) * ((size_t)9 * maxN + 8) + sizeof(uint32_t) * 624 + (size_t)maxR + 16;
}

// Records of one focal agent against the obstacle set marked in S.cls
// (kNone = not an obstacle), visual exclusion, and the raster into S.V.
// Every thread of the block calls it; it ends synchronised.
__device__ void project_focal(const FGSmem& S, int N, int R, double step, double fov0, double fov1,
                              bool exclusion, bool with_nonsocial, const Focal& f,
                              bool keep_distance = false) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int r = tid; r < R; r += nthr) S.V[r] = 0;
    for (int j = tid; j < N; j += nthr) {
        Rec rc;
        rc.valid = 0; rc.social = 0; rc.order = 0; rc.start = rc.end = rc.sx = rc.ex = 0;
        rc.dist = 0.0;
        rc.vis = 0.0;
        const int c = S.cls[j];
        if (c == kSocial || (with_nonsocial && (c == kNonExpl || c == kSamePatch))) {
            const PairFG p = pair_forage_exact(f, S.x[j], S.y[j], R, step, fov0, fov1);
            if (!p.skipped && p.visible) {
                rc.valid = 1;
                rc.social = c == kSocial;
                rc.order = c * N + j;
                rc.start = rc.sx = p.start;
                rc.end = rc.ex = p.end;
                rc.dist = p.dist;
                rc.vis = p.vis;
            }
        }
        S.rec[j] = rc;
    }
    __syncthreads();
    if (exclusion) {
        // rank = position in the stable ascending sort by distance
        for (int j = tid; j < N; j += nthr) {
            const Rec rf = S.rec[j];
            if (!rf.valid) continue;
            int rank = 0;
            for (int i = 0; i < N; ++i) {
                const Rec ri = S.rec[i];
                if (ri.valid && (ri.dist < rf.dist || (ri.dist == rf.dist && ri.order < rf.order)))
                    ++rank;
            }
            S.sorted[rank] = j;
        }
        __syncthreads();
        for (int j = tid; j < N; j += nthr) {
            Rec rf = S.rec[j];
            if (!rf.valid) continue;
            // walk the nearer records in rank order; rank 0 keeps its interval
            for (int q = 0; q < N; ++q) {
                const int i = S.sorted[q];
                if (i == j) break;
                const Rec ri = S.rec[i];
                if (ri.dist < rf.dist) {
                    if (rf.sx <= ri.start && ri.start <= rf.ex) rf.ex = ri.start;
                    if (rf.sx <= ri.end && ri.end <= rf.ex) rf.sx = ri.end;
                    if (ri.start <= rf.sx && ri.end >= rf.ex) { rf.sx = 0; rf.ex = 0; }
                }
            }
            S.rec[j].sx = rf.sx;
            S.rec[j].ex = rf.ex;
        }
        __syncthreads();
    }
    if (!keep_distance) {
        // every record paints 1, so the order of the records does not matter
        for (int j = tid; j < N; j += nthr) {
            const Rec rc = S.rec[j];
            if (!rc.valid || !rc.social) continue;
            paint_record(S.V, rc, R, 1, 1);
        }
    } else if (tid == 0) {
        // keep_distance_info: wrap pieces paint 1 and main slices the distance
        // value (code 2); later records overwrite, in descending vis_angle
        // order with ties in obstacle order (agent.py:567, 569-590)
        int last_vis_idx = -1;
        for (int n = 0; n < N; ++n) {
            int best = -1;
            for (int j = 0; j < N; ++j) {
                const Rec rc = S.rec[j];
                if (!rc.valid || !rc.social) continue;
                if (last_vis_idx >= 0) {
                    const Rec lr = S.rec[last_vis_idx];
                    const bool after = rc.vis < lr.vis || (rc.vis == lr.vis && rc.order > lr.order);
                    if (!after) continue;
                }
                if (best < 0) { best = j; continue; }
                const Rec br = S.rec[best];
                if (rc.vis > br.vis || (rc.vis == br.vis && rc.order < br.order)) best = j;
            }
            if (best < 0) break;
            paint_record(S.V, S.rec[best], R, 1, 2);
            last_vis_idx = best;
        }
    }
    __syncthreads();
}

// soc_v_field[k] = V[R-1-k] unless linspace_k is outside [fov0, fov1].
__device__ inline int soc_bit(const uint8_t* V, int k, int R, double step, double fov0, double fov1) {
    const double ph = linspace_at(k, R, step);
    if (ph < fov0 || ph > fov1) return 0;
    return V[R - 1 - k];
}

// Cue class of agent j for focal agent f (agent.py:396-413): in range by
// supcalc.distance; exploiting agents of another patch are the cues, those
// of the focal's own patch occlude (patchwise exclusion), exploiting agents
// without a patch are dropped; the rest in range (the focal included) occlude.
__device__ inline int cue_class(const FGSmem& S, const FGRep& P, int j, int f, double cfx, double cfy,
                                int pid_f) {
#pragma clang fp contract(off)
    const double dx = (S.x[j] + S.rad[j]) - cfx, dy = (S.y[j] + S.rad[j]) - cfy;
    const bool in_range = norm2(dx, dy) <= P.vision_range;
    const bool expl = in_range && j != f && get_mode(S.over[j], S.w[j], P.T_w) == ABM_MODE_EXPLOIT;
    int c = kNone;
    if (expl) {
        if (P.patchwise && S.pid[j] == pid_f) c = kSamePatch;
        else if (S.pid[j] != -1) c = kSocial;
    } else if (in_range) {
        c = kNonExpl;
    }
    return c;
}

// Whether the raster needs obstacle j's record: the cues, and with visual
// exclusion every obstacle in range (agent.py:415-419).
__device__ inline bool needs_record(int c, bool excl) { return c == kSocial || (excl && c != kNone); }

// Per-replica float32 constants of the guarded record.
struct FastRec {
    float inv_step;   // bins of linspace(-pi, pi, R) per radian
    float hs;         // R / (2 pi): proj_size / 2 per radian of atan(r / d)
    float g_bins;     // bin-tie guard in bins (angle guard + float32 rounding of the coordinate)
    float fov0, fov1;
};

__device__ inline FastRec fast_rec_const(int R, double step, double fov0, double fov1) {
    FastRec q;
    q.inv_step = (float)(1.0 / step);
    q.hs = (float)((double)R / kTwoPi);
    q.g_bins = (float)(f32::kAngleGuardF / step + 1.5e-7 * R);
    q.fov0 = (float)fov0;
    q.fov1 = (float)fov1;
    return q;
}

// Record of obstacle j (class c) for focal geometry f.  The bin, the
// interval and the FOV test come from a float32 decision when it is outside
// every guard band, else from the exact float64 record (agent.py:501-556):
//  * the centre difference (focal radius for both centres) in float64,
//    rounded once, and the focal's float64 unit heading: with atan2_f32 the
//    signed angle a from heading to object is within 1e-6 rad of the exact
//    one, and the reference's acos-based angle within 5e-8 rad of that;
//  * the reference's closed angle is -a for 0 < |a| < pi (a' = a mod 2 pi,
//    ca = -a' or 2 pi - a'); a = 0 gives 2 pi and |a| = pi gives pi: 4e-6 rad
//    bands around 0 and pi, and around the strict FOV ends;
//  * t = nearest bin of ca (bin-tie band as in the flocking kernels);
//  * start / end = int(t -+ proj_size / 2), truncated toward zero: with half
//    = proj_size / 2 outside a 4e-6 relative band around an integer, start =
//    t - floor(half) - 1 when that stays positive, else t - floor(half),
//    and end = t + floor(half).
// The float64 distance, which only visual exclusion ranks by, is exact.
__device__ inline Rec record_at(double ox, double oy, int c, int order, const Focal& f, int R, double step,
                                double fov0, double fov1, const FastRec& FR, bool need_dist) {
#pragma clang fp contract(off)
    Rec rc;
    rc.valid = 0; rc.social = 0; rc.order = 0; rc.start = rc.end = rc.sx = rc.ex = 0;
    rc.dist = 0.0; rc.vis = 0.0;
    if (ox == f.px && oy == f.py) return rc;                 // skipped (agent.py:502)
    const double v2x = (ox + f.radius) - f.cx, v2y = (oy + f.radius) - f.cy;
    {
        const float dx = (float)v2x, dy = (float)v2y;
        const float u1x = (float)f.u1x, u1y = (float)f.u1y;
        const float d2 = fmaf(dy, dy, dx * dx);
        const float a = f32::atan2_f32(fmaf(u1x, dy, -u1y * dx), fmaf(u1x, dx, u1y * dy));
        const float aa = fabsf(a);
        const float ca = -a;
        const float xb = (f32::kPiF - a) * FR.inv_step;     // bin coordinate of ca
        const float tr = rintf(xb);
        const float dfr = xb - tr;
        const float half = f32::atan_pos_f32((float)f.radius, __builtin_amdgcn_sqrtf(d2)) * FR.hs;
        const float hfl = floorf(half);
        const float hfrac = half - hfl;
        const float hg = 4e-6f * half + 1e-6f;
        const float g = f32::kAngleGuardF;
        const bool ok = (d2 > 1e-30f) & (aa > g) & (aa < f32::kPiF - g) &
                        (fabsf(dfr) <= 0.5f - FR.g_bins) & (hfrac >= hg) & (hfrac <= 1.0f - hg) &
                        (fabsf(ca - FR.fov0) >= g) & (fabsf(ca - FR.fov1) >= g);
        if (ok) {
            if (FR.fov0 < ca && ca < FR.fov1) {
                const int t = (int)tr, h = (int)hfl;
                rc.valid = 1;
                rc.social = c == kSocial;
                rc.order = order;
                rc.start = rc.sx = t - h >= 1 ? t - h - 1 : t - h;
                rc.end = rc.ex = t + h;
                if (need_dist) rc.dist = norm2(v2x, v2y);
            }
            return rc;
        }
    }
    const PairFG p = pair_forage_exact(f, ox, oy, R, step, fov0, fov1);
    if (!p.skipped && p.visible) {
        rc.valid = 1;
        rc.social = c == kSocial;
        rc.order = order;
        rc.start = rc.sx = p.start;
        rc.end = rc.ex = p.end;
        rc.dist = p.dist;
        rc.vis = p.vis;
    }
    return rc;
}

__device__ inline Rec record_of(const FGSmem& S, int j, int c, int N, const Focal& f, int R, double step,
                                double fov0, double fov1, const FastRec& FR, bool need_dist) {
    return record_at(S.x[j], S.y[j], c, c * N + j, f, R, step, fov0, fov1, FR, need_dist);
}

// Classes of the agents j != skip for focal agent f into S.cls; those whose
// records the raster needs are appended to lst (in any order: records are
// kept per agent and ranked by (distance, list position)).  Thread t0 of
// `stride` threads, both multiples of 64 apart per wave.
__device__ void classify(const FGSmem& S, const FGRep& P, int N, int f, int skip1, int skip2, bool excl,
                         int* cls, int* lst, int* lst_n, int t0, int stride) {
    const double cfx = S.x[f] + S.rad[f], cfy = S.y[f] + S.rad[f];
    const int pid_f = S.pid[f];
    const int lane = t0 & 63;
    for (int base = t0 - lane; base < N; base += stride) {
        const int j = base + lane;
        bool need = false;
        if (j < N && j != skip1 && j != skip2) {
            // without exclusion only a listed obstacle's class is read, and
            // only an exploiting agent (overriding mode exploit) is listed
            int c = kNone;
            if (excl || S.over[j] == ABM_MODE_EXPLOIT) c = cue_class(S, P, j, f, cfx, cfy, pid_f);
            cls[j] = c;
            need = needs_record(c, excl);
        }
        const uint64_t m = __ballot(need);
        if (m) {
            int at = 0;
            if (lane == 0) at = atomicAdd(lst_n, (int)__popcll(m));
            at = __builtin_amdgcn_readfirstlane(at);
            if (need) lst[at + (int)__popcll(m & ((1ull << lane) - 1ull))] = j;
        }
    }
}

// make_focal with one sincos for the heading (the same values as cos, sin).
__device__ inline Focal focal_of(const FGSmem& S, int k) {
#pragma clang fp contract(off)
    const double px = S.x[k], py = S.y[k], radius = S.rad[k];
    double sn, cs;
    sincos(S.th[k], &sn, &cs);
    Focal f;
    f.px = px; f.py = py; f.radius = radius;
    f.cx = px + radius;
    f.cy = py + radius;
    const double ex = px + (1.0 + cs) * radius;
    const double ey = py + (1.0 - sn) * radius;
    const double v1x = ex - f.cx, v1y = ey - f.cy;
    const double n1 = norm2(v1x, v1y);
    f.u1x = v1x / n1;
    f.u1y = v1y / n1;
    return f;
}

// The heading agent f takes after a random-walk turn drawn from uniform
// d + c (c = 0, 1) or after no turn (c = 2), with its sin and cos, into
// buf[3c .. 3c + 2] (agent.py:187-200, 595-603: prove_orientation of the
// turned heading, then cos and sin for the move); buf[9] = d.  The same
// arithmetic as the move, computed by a spare wave before agent f's
// decision: the walk's draw index is d or d + 1 (the agent before draws at
// most one uniform), so the decision only picks a slot.
__device__ inline void spec_heading(const FGSmem& S, const FGRep& P, int N, int f, int d, int c, double* buf) {
#pragma clang fp contract(off)
    double dth = 0.0;
    if (c < 2) {
        const double uu = d + c < N ? S.uni[d + c] : 0.0;
        dth = P.exp_theta_min + (P.exp_theta_max - P.exp_theta_min) * uu;
    }
    const double th = prove_orientation(S.th[f] + dth);
    double sn, cs;
    sincos(th, &sn, &cs);
    buf[3 * c] = th;
    buf[3 * c + 1] = sn;
    buf[3 * c + 2] = cs;
    if (c == 0) buf[9] = (double)d;
}

// Set bits [lo, hi) of a bit-packed field.
__device__ inline void or_range(uint32_t* Vb, int lo, int hi) {
    while (lo < hi) {
        const int w = lo >> 5, b0 = lo & 31;
        const int end = min(hi, (w + 1) * 32);
        const int nb = end - lo;
        atomicOr(&Vb[w], nb == 32 ? 0xffffffffu : ((1u << nb) - 1u) << b0);
        lo = end;
    }
}

// paint_record for the step kernel's bit-packed raw field: every social
// record paints 1, so the pieces are ORed in any order (agent.py:577-590).
__device__ inline void paint_bits(uint32_t* Vb, const Rec& rc, int R) {
    int s = rc.sx, e = rc.ex, lo, hi;
    if (s < 0) {
        py_slice(R + s, R, R, &lo, &hi);
        or_range(Vb, lo, hi);
        s = 0;
    }
    if (e >= R) {
        py_slice(0, e - R, R, &lo, &hi);
        or_range(Vb, lo, hi);
        e = R - 1;
    }
    py_slice(s, e, R, &lo, &hi);
    or_range(Vb, lo, hi);
}

// Bits of word w inside the inclusive index range [a, b].
__device__ inline uint32_t word_range_mask(int w, int a, int b) {
    const int lo = max(a, w * 32), hi = min(b, w * 32 + 31);
    if (lo > hi) return 0u;
    const int nb = hi - lo + 1;
    return nb == 32 ? 0xffffffffu : ((1u << nb) - 1u) << (lo & 31);
}

// Visual exclusion (agent.py:421-445) over the listed records, then their
// raster (agent.py:569-590) into the bit-packed raw field S.V (uint32
// words, bit r = V[r]).  Every thread; ends synchronised.
__device__ void raster_list(const FGSmem& S, int R, bool excl, const int* lst, int n0, int extra) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    // the listed records, and one more (extra >= 0) kept outside the list
    const int n = n0 + (extra >= 0 ? 1 : 0);
    auto item = [&](int q) { return q < n0 ? lst[q] : extra; };
    if (excl) {
        // rank = position in the stable ascending sort by distance
        for (int q = tid; q < n; q += nthr) {
            const int j = item(q);
            const Rec rf = S.rec[j];
            if (!rf.valid) continue;
            int rank = 0;
            for (int p = 0; p < n; ++p) {
                const Rec ri = S.rec[item(p)];
                if (ri.valid && (ri.dist < rf.dist || (ri.dist == rf.dist && ri.order < rf.order))) ++rank;
            }
            S.sorted[rank] = j;
        }
        __syncthreads();
        for (int q = tid; q < n; q += nthr) {
            const int j = item(q);
            Rec rf = S.rec[j];
            if (!rf.valid) continue;
            // walk the nearer records in rank order; rank 0 keeps its interval
            for (int p = 0; p < n; ++p) {
                const int i = S.sorted[p];
                if (i == j) break;
                const Rec ri = S.rec[i];
                if (ri.dist < rf.dist) {
                    if (rf.sx <= ri.start && ri.start <= rf.ex) rf.ex = ri.start;
                    if (rf.sx <= ri.end && ri.end <= rf.ex) rf.sx = ri.end;
                    if (ri.start <= rf.sx && ri.end >= rf.ex) { rf.sx = 0; rf.ex = 0; }
                }
            }
            S.rec[j].sx = rf.sx;
            S.rec[j].ex = rf.ex;
        }
        __syncthreads();
    }
    uint32_t* Vb = reinterpret_cast<uint32_t*>(S.V);
    for (int q = tid; q < n; q += nthr) {
        const Rec rc = S.rec[item(q)];
        if (rc.valid && rc.social) paint_bits(Vb, rc, R);
    }
    __syncthreads();
}

// Bits [lo, hi) of a field that fall in word w.
__device__ inline uint32_t word_slice_mask(int w, int lo, int hi) {
    return hi > lo ? word_range_mask(w, lo, hi - 1) : 0u;
}

// Sum over the 64 lanes of wave 0 with DPP moves (no LDS round trips):
// prefix sums inside each row of 16 lanes (row_shr 1, 2, 4, 8), then row
// 15's and row 31's lane carried into the rows above (row_bcast 15, 31), so
// lane 63 holds the total.  Lanes a move has no source for add 0.
__device__ inline int wave_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ inline void fg_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Raw-field word w of one social interval [s, e] of a record: the pieces
// paint_bits ORs (agent.py:577-590).
__device__ inline uint32_t interval_word(int w, int s, int e, int R) {
    if (s >= 0 && e < R) return word_slice_mask(w, s, e > s ? e : s);   // one piece: py_slice(s, e)
    uint32_t bits = 0u;
    int lo, hi;
    if (s < 0) {
        py_slice(R + s, R, R, &lo, &hi);
        bits |= word_slice_mask(w, lo, hi);
        s = 0;
    }
    if (e >= R) {
        py_slice(0, e - R, R, &lo, &hi);
        bits |= word_slice_mask(w, lo, hi);
        e = R - 1;
    }
    py_slice(s, e, R, &lo, &hi);
    return bits | word_slice_mask(w, lo, hi);
}

// Raster without visual exclusion, by one wave: raw-field word w of the
// live social records lst[0..n).  Lane l loads records l, l + 64, ...; each
// live interval reaches every lane by readlane, so there is no atomic.
__device__ inline uint32_t list_word(const FGSmem& S, int R, const int* lst, int n, int w) {
    const int lane = threadIdx.x & 63;
    uint32_t bits = 0u;
    for (int base = 0; base < n; base += 64) {
        const int q = base + lane;
        int sx = 0, ex = 0;
        bool live = false;
        if (q < n) {
            const Rec& rc = S.rec[lst[q]];
            live = rc.valid && rc.social;
            sx = rc.sx;
            ex = rc.ex;
        }
        uint64_t m = __ballot(live);
        while (m) {
            const int l = __builtin_ctzll(m);
            m &= m - 1;
            bits |= interval_word(w, __builtin_amdgcn_readlane(sx, l), __builtin_amdgcn_readlane(ex, l), R);
        }
    }
    return bits;
}

// Counts of the field without visual exclusion, by wave 0: the words the
// listed records made (S.V), with the interval [s, e] of the record kept
// outside the list ORed in when live (wave-uniform): the rays set on the
// kept raw range [ta, tb] and those in its left half [la, tb], to every lane.
// m_tot / m_left: the lane's word masks of the two ranges when W <= 64.
__device__ void count_wave(const FGSmem& S, int R, int W, bool live, int s, int e, int ta, int tb, int la,
                           uint32_t m_tot, uint32_t m_left, int* n_set, int* n_left) {
    const int lane = threadIdx.x & 63;
    const uint32_t* Vb = reinterpret_cast<const uint32_t*>(S.V);
    int tot = 0, left = 0;
    if (W <= 64) {   // one word per lane
        uint32_t bits = lane < W ? Vb[lane] : 0u;
        if (live) bits |= interval_word(lane, s, e, R);
        tot = __popc(bits & m_tot);
        left = __popc(bits & m_left);
    } else {
        for (int w = lane; w < W; w += 64) {
            uint32_t bits = Vb[w];
            if (live) bits |= interval_word(w, s, e, R);
            tot += __popc(bits & word_range_mask(w, ta, tb));
            left += __popc(bits & word_range_mask(w, la, tb));
        }
    }
    if (R < 65536) {   // one sum of both counts, each below 2^16
        const int both = wave_sum(tot << 16 | left);
        *n_set = both >> 16;
        *n_left = both & 0xffff;
    } else {
        *n_set = wave_sum(tot);
        *n_left = wave_sum(left);
    }
}

// The returned field of agent k: counts for mean() and the left/right
// halves into cnt; soc_v_field[r] = V[R-1-r] on the rays the FOV mask keeps.
__device__ void count_field(const FGSmem& S, const FGRep& P, int R, int* cnt, uint32_t* bits_out, int k) {
    const int tid = threadIdx.x, nthr = blockDim.x, nl = R / 2, W = P.words;
    const uint32_t* Vb = reinterpret_cast<const uint32_t*>(S.V);
    // soc_v_field[q] = V[R-1-q] for keep_lo <= q <= keep_hi: raw indices
    // [R-1-keep_hi, R-1-keep_lo]; the left half q < R/2 is raw r >= R - R/2
    const int ta = R - 1 - P.keep_hi, tb = R - 1 - P.keep_lo, la = max(R - nl, ta);
    for (int w = tid; w < W; w += nthr) {
        const uint32_t bits = Vb[w];
        const int tot = __popc(bits & word_range_mask(w, ta, tb));
        const int left = __popc(bits & word_range_mask(w, la, tb));
        if (tot) atomicAdd(&cnt[0], tot);
        if (left) atomicAdd(&cnt[1], left);
    }
    if (bits_out) {
        uint32_t* dst = bits_out + P.field_offset + (int64_t)k * W;
        for (int wd = tid; wd < W; wd += nthr) {
            uint32_t m = 0;
            for (int b = 0; b < 32; ++b) {
                const int q = wd * 32 + b, r = R - 1 - q;
                if (q >= P.keep_lo && q <= P.keep_hi && ((Vb[r >> 5] >> (r & 31)) & 1u)) m |= 1u << b;
            }
            dst[wd] = m;
        }
    }
}

// ---- numpy's legacy MT19937 on the device (the published generator of
// Matsumoto and Nishimura, in numpy's mt19937_state layout: 624 key words and
// a position; numpy twists lazily, when a word is drawn at position 624)
constexpr int kMtN = 624, kMtM = 397;

__device__ inline uint32_t mt_temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ inline uint32_t mt_twist_word(uint32_t cur, uint32_t nxt, uint32_t far) {
    const uint32_t y = (cur & 0x80000000u) | (nxt & 0x7fffffffu);
    return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

// key[i] for i = 0..623 in order, in place, by one thread.
__device__ inline void mt_twist_serial(uint32_t* key) {
    int i = 0;
    for (; i < kMtN - kMtM; ++i) key[i] = mt_twist_word(key[i], key[i + 1], key[i + kMtM]);
    for (; i < kMtN - 1; ++i) key[i] = mt_twist_word(key[i], key[i + 1], key[i + kMtM - kMtN]);
    key[kMtN - 1] = mt_twist_word(key[kMtN - 1], key[0], key[kMtM - 1]);
}

// The same by a block of at least 227 threads: in the chunks [0, 227),
// [227, 454), [454, 623) and {623} every element reads key[i + 1] not yet
// rewritten and key[(i + 397) % 624] already final, so each chunk is read
// in full, then written.  Every thread of the block calls it.
__device__ inline void mt_twist_block(uint32_t* key) {
    const int tid = threadIdx.x;
    constexpr int lo[4] = {0, kMtN - kMtM, 2 * (kMtN - kMtM), kMtN - 1};
    constexpr int hi[4] = {kMtN - kMtM, 2 * (kMtN - kMtM), kMtN - 1, kMtN};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int i = lo[c] + tid;
        const bool mine = i < hi[c];
        uint32_t v = 0u;
        if (mine) {
            const int far = i + kMtM < kMtN ? i + kMtM : i + kMtM - kMtN;
            v = mt_twist_word(key[i], key[i + 1 < kMtN ? i + 1 : 0], key[far]);
        }
        __syncthreads();
        if (mine) key[i] = v;
        __syncthreads();
    }
}

// Draws from a generator state held in LDS by one thread; never runs out.
struct MtCursor {
    uint32_t* key;
    int pos, used;
    bool out;
    __device__ uint32_t next() {
        if (pos >= kMtN) {
            mt_twist_serial(key);
            pos = 0;
        }
        ++used;
        return mt_temper(key[pos++]);
    }
};

// Status of a replica between the interaction kernel and the update.
constexpr int kWorldReady = 3;

// numpy's legacy random_sample from two raw MT19937 outputs (mt19937_next_double).
__device__ inline double words_to_double(uint32_t w0, uint32_t w1) {
    const uint32_t a = w0 >> 5, b = w1 >> 6;
    return ((double)a * 67108864.0 + (double)b) / 9007199254740992.0;
}

// The random walk's uniforms come either as doubles (uniforms) or as raw
// 32-bit outputs (words, two per draw, starting after the words the resource
// loop consumed: u_used[b] on entry).  With status, only replicas marked
// kWorldReady are updated, and they are marked ABM_WORLD_DONE.  FIELDS false:
// a step that writes no fields (bits_out is ignored); EXCL false: no replica
// of the launch has visual exclusion.  Either takes the other paths' tests
// out of the agent loop.
template <bool FIELDS, bool EXCL>
__global__ void __launch_bounds__(512)
forage_step_kernel(const FGRep* __restrict__ reps, abm_forage_state st,
                   const double* __restrict__ uniforms, const uint32_t* __restrict__ words,
                   const int64_t* __restrict__ u_off, const int32_t* __restrict__ u_cnt,
                   int32_t* __restrict__ u_used, int32_t* __restrict__ status,
                   uint32_t* __restrict__ bits_out, const int32_t* __restrict__ order,
                   abm_mt19937_state* __restrict__ mt, int maxN, int maxR) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) double smem[];
    if (!FIELDS) bits_out = nullptr;
    // blocks take the replicas largest first, so the longest id-order chains
    // start at once and the small replicas fill in behind them
    const int rb = order[blockIdx.x];
    if (status && status[rb] != kWorldReady) return;   // uniform over the block
    const int tid = threadIdx.x, nthr = blockDim.x;
    // the replica's parameters in LDS: read where used, rather than held in
    // scalar registers over the whole agent loop (which spills them)
    __shared__ FGRep sP;
    static_assert(sizeof(FGRep) % 8 == 0 && sizeof(FGRep) / 8 <= 64, "FGRep copy");
    if (tid < (int)(sizeof(FGRep) / 8))
        reinterpret_cast<uint64_t*>(&sP)[tid] = reinterpret_cast<const uint64_t*>(reps + rb)[tid];
    __syncthreads();
    const FGRep& P = sP;
    const int N = P.n_agents, R = P.n_rays, a0 = P.agent_offset;
    FGSmem S = fg_carve(smem, maxN, maxR);
    for (int i = tid; i < N; i += nthr) {
        const int g = a0 + i;
        S.x[i] = st.x[g]; S.y[i] = st.y[g]; S.th[i] = st.theta[g]; S.v[i] = st.v[g];
        S.w[i] = st.w[g]; S.u[i] = st.u[g]; S.cr[i] = st.collected_r[g];
        S.crb[i] = st.collected_r_before[g];
        double m = st.novelty[(int64_t)g * P.tau];
        for (int q = 1; q < P.tau; ++q) m = fmax(m, st.novelty[(int64_t)g * P.tau + q]);
        S.nov[i] = m;
        S.rad[i] = (double)st.radius[g];
        S.mode[i] = st.mode[g]; S.over[i] = st.overriding_mode[g];
        S.env[i] = st.env_status[g]; S.pid[i] = st.exploited_patch_id[g];
    }
    // the step's uniforms, staged in LDS: agent i's draw is the i-th (at most
    // one per agent); doubles, or pairs of words after those the resource
    // loop used (u_used[b] on entry); NaN past the end of the buffer
    const int ucount = mt ? 0 : u_cnt[rb];
    const int used0 = (words || mt) ? u_used[rb] : 0;
    int mt_pos0 = 0;
    if (mt) {
        // the replica's generator on the device: its next 2N words (a copy of
        // the key twisted in LDS as needed), as N uniforms
        for (int i = tid; i < kMtN; i += nthr) S.mtk[i] = mt[rb].key[i];
        mt_pos0 = mt[rb].pos;
        __syncthreads();
        uint32_t* wb = reinterpret_cast<uint32_t*>(S.uni);
        int made = 0, p = mt_pos0;
        while (made < 2 * N) {
            if (p >= kMtN) {
                mt_twist_block(S.mtk);
                p = 0;
            }
            const int take = min(kMtN - p, 2 * N - made);
            for (int j = tid; j < take; j += nthr) wb[made + j] = mt_temper(S.mtk[p + j]);
            made += take;
            p += take;
        }
        __syncthreads();
        for (int i = tid; i < N; i += nthr) {   // each thread rewrites its own 8 bytes
            const uint32_t w0 = wb[2 * i], w1 = wb[2 * i + 1];
            S.uni[i] = words_to_double(w0, w1);
        }
    } else {
        for (int i = tid; i < N; i += nthr) {
            double uu = NAN;
            if (words) {
                const uint32_t* wb = words + u_off[rb];
                const int at = used0 + 2 * i;
                if (at + 2 <= ucount) uu = words_to_double(wb[at], wb[at + 1]);
            } else if (i < ucount) {
                uu = uniforms[u_off[rb] + i];
            }
            S.uni[i] = uu;
        }
    }
    int draws = 0;   // meaningful in thread 0
    int xlive = 0, xs = 0, xe = 0;   // thread 0: agent k's social interval for agent k+1
    const int nl = R / 2;
    const bool excl = EXCL && P.visual_exclusion != 0;
    const FastRec FR = fast_rec_const(R, P.step, P.fov0, P.fov1);
    // without exclusion and field output, wave 0 rasters and counts alone
    const bool wave_raster = !excl && !bits_out;
    const int ta = R - 1 - P.keep_hi, tb = R - 1 - P.keep_lo, la = max(R - nl, ta);
    // wave 0's lane masks of the kept range and its left half (one word per lane)
    const int wl0 = tid & 63;
    const uint32_t m_tot = wl0 < P.words ? word_range_mask(wl0, ta, tb) : 0u;
    const uint32_t m_left = wl0 < P.words ? word_range_mask(wl0, la, tb) : 0u;
    int* const lst[2] = {S.lst, S.lst + maxN};
    int* const clsb[2] = {S.cls, S.cls2};
    int* const lst_n = S.cnt + 4;
    // Pipeline over the agents in id order.  Agent m's field depends on the
    // agents moved before it only through their records.  So while thread 0
    // moves agent k and then adds agent k's class (and record) for agent
    // k+1, wave 1 builds agent k+1's records of every other obstacle
    // (without exclusion, also their raster), and the other waves classify
    // agent k+2's obstacles but k and k+1; one of them then builds agent
    // k+1's speculative headings and the next agent k+2's focal geometry
    // (agents move only in their own turn, so it is the start-of-step
    // pose).  After one barrier come exclusion, raster and
    // counts of agent k+1 (without exclusion: agent k's interval ORed in and
    // the counts, by wave 0 alone), during which wave 2 adds agent k's class
    // for agent k+2.
    // Classes, lists, focal geometry and counts alternate between two
    // buffers by parity of the focal agent.
    const int rbase = 64, cbase = 128, cstride = nthr - 128;   // blockDim 256 or 512
    if (tid == 0) {
        for (int q = 0; q < 8; ++q) S.cnt[q] = 0;
    }
    for (int w = tid; w < P.words; w += nthr) reinterpret_cast<uint32_t*>(S.V)[w] = 0u;
    __syncthreads();
    // ---- agent 0's field; agent 1's classes but agent 0's; agent 0's
    // speculative headings (slot 0, draw index 0)
    // two waves that classify once per focal agent also build the
    // speculative headings and the focal geometry: waves 4 and 5 of eight
    // (waves 2 and 3 classify twice for 512 agents), waves 2 and 3 of four
    const int spec_tid = nthr == 512 ? 256 : 128, focal_tid = spec_tid + 64;
    if (tid >= spec_tid && tid < spec_tid + 3) spec_heading(S, P, N, 0, 0, tid - spec_tid, S.spec);
    if (tid == 0) S.spec[24] = 0.0;    // uniforms drawn before agent 0
    if (tid == 0) S.focal[0] = focal_of(S, 0);
    if (tid == cbase && N > 1) S.focal[1] = focal_of(S, 1);
    classify(S, P, N, 0, -1, -1, excl, clsb[0], lst[0], &lst_n[0], tid, nthr);
    if (N > 1) classify(S, P, N, 1, 0, -1, excl, clsb[1], lst[1], &lst_n[1], tid, nthr);
    __syncthreads();
    for (int q = tid; q < lst_n[0]; q += nthr) {
        const int j = lst[0][q];
        S.rec[j] = record_of(S, j, clsb[0][j], N, S.focal[0], R, P.step, P.fov0, P.fov1, FR, excl);
    }
    __syncthreads();
    raster_list(S, R, excl, lst[0], lst_n[0], -1);
    if (tid == 0) lst_n[0] = 0;   // rebuilt for agent 2
    count_field(S, P, R, S.cnt, bits_out, 0);
    __syncthreads();

    // wave 0 carries the chain of moves: first issue on its SIMD.  It holds
    // the focal agent's counts in registers (from LDS after count_field).
    if (tid < 64) __builtin_amdgcn_s_setprio(3);
    int n_set = S.cnt[0], n_left = S.cnt[1];
    for (int k = 0; k < N; ++k) {
        const int f1 = k + 1, par = f1 & 1, f2 = k + 2, par2 = par ^ 1;
        if (!wave_raster && k > 0) { n_set = S.cnt[2 * (k & 1)]; n_left = S.cnt[2 * (k & 1) + 1]; }
        // ---- decision and motion (agent.py:168-283)
        if (tid == 0) {
            // every LDS operand first, in one batch: the replica's constants
            // and agent k's state (the loads then wait once, not in turn)
            const double T_w = P.T_w, Eps_w = P.Eps_w, g_w = P.g_w, B_w = P.B_w, w_max = P.w_max;
            const double T_u = P.T_u, Eps_u = P.Eps_u, g_u = P.g_u, B_u = P.B_u, u_max = P.u_max;
            const double S_wu = P.S_wu, S_uw = P.S_uw, F_N = P.F_N, F_R = P.F_R;
            const double vmax = P.exp_vel_max, rtmax = P.reloc_theta_max, stop = P.exp_stop_ratio;
            const double pad = P.pad, width = P.width, height = P.height;
            double w = S.w[k], u = S.u[k];
            const double vel0 = S.v[k], cr = S.cr[k], crb = S.crb[k], nov = S.nov[k];
            const double x0 = S.x[k], y0 = S.y[k], th0 = S.th[k], rad = S.rad[k];
            int over = S.over[k], mode = S.mode[k];
            const int env = S.env[k];
            const double* spec = S.spec + 12 * (k & 1);
            const double spec_d = spec[9];
            asm volatile("" ::: "memory");
            const double count = (double)n_set, cleft = (double)n_left;
            const double collected = cr - crb;
            const double ipriv = F_N * nov + F_R * collected;
            const double w_p = w > T_w ? w : 0.0;
            const double u_p = u > T_u ? u : 0.0;
            const double mean = count / (double)R;
            const double dw = Eps_w * mean - g_w * (w - B_w) - u_p * S_uw;
            const double du = Eps_u * ipriv - g_u * (u - B_u) - w_p * S_wu;
            w = w + dw;
            u = u + du;
            if (w > w_max) w = w_max;
            if (w < -w_max) w = -w_max;
            if (u > u_max) u = u_max;
            if (u < -u_max) u = -u_max;
            double dvel = 0.0, dth = 0.0;
            int slot = 2;   // speculative heading: 2 = no turn, 0 / 1 = walk, -1 = relocation
            const bool tw = w > T_w, tu = u > T_w;   // tr_u compares with T_w
            auto walk = [&]() {
                slot = draws - (int)spec_d;   // the uniform drawn: S.uni[draws]
                ++draws;
                dvel = vmax;
                over = -1; mode = ABM_MODE_EXPLORE;
            };
            auto reloc = [&]() {
                const double lm = cleft / (double)nl;
                const double rm = (count - cleft) / (double)(R - nl);
                dvel = vmax - vel0;
                dth = (lm - rm) * rtmax;
                slot = -1;
                over = -1; mode = ABM_MODE_RELOCATE;
            };
            auto exploit = [&]() {
                over = ABM_MODE_EXPLOIT; mode = ABM_MODE_EXPLOIT;
                dvel = -vel0 * stop;
                dth = 0.0;
            };
            if (over != ABM_MODE_COLLIDE) {   // get_mode() == "collide" only via the override
                if (!tw && !tu) walk();
                else if (tw && tu) { if (env == 1) exploit(); else reloc(); }
                else if (tw && !tu) reloc();
                else { if (env == 1) exploit(); else walk(); }
            }
            Pose p{x0, y0, th0, vel0};
            double sn, cs;
            if (slot < 0) {
                p.theta = prove_orientation(p.theta + dth);
                sincos(p.theta, &sn, &cs);
            } else {   // prove_orientation(theta + turn) and its sin, cos from the spare wave
                p.theta = spec[3 * slot];
                sn = spec[3 * slot + 1];
                cs = spec[3 * slot + 2];
            }
            p.v = p.v + dvel;
            // prove_velocity (agent.py:612-620): only in explore mode
            if (get_mode(over, w, T_w) == ABM_MODE_EXPLORE && fabs(p.v) > 1.0) p.v = vmax;
            p.x = p.x + p.v * cs;
            p.y = p.y - p.v * sn;
            p = reflect_walls(p, rad, pad, width, height);
            S.x[k] = p.x; S.y[k] = p.y; S.th[k] = p.theta; S.v[k] = p.v;
            S.w[k] = w; S.u[k] = u; S.over[k] = over; S.mode[k] = mode;
            S.crb[k] = cr;
            S.nov[k] = ipriv;   // read only by agent k's decision: now its I_priv
            S.spec[24 + (f1 & 1)] = (double)draws;   // uniforms drawn before agent k+1
        }
        // ---- thread 0, after the move: agent k's class and record for agent k+1
        if (tid == 0 && f1 < N) {
            int extra = -1;
            xlive = 0;
            // without exclusion only an exploiting agent k (overriding mode
            // exploit) can be a cue for agent k+1: skip the distance otherwise
            int c = kNone;
            if (excl || S.over[k] == ABM_MODE_EXPLOIT) {
                const double cfx = S.x[f1] + S.rad[f1], cfy = S.y[f1] + S.rad[f1];
                c = cue_class(S, P, k, f1, cfx, cfy, S.pid[f1]);
            }
            if (needs_record(c, excl)) {
                const Rec rk = record_of(S, k, c, N, S.focal[par], R, P.step, P.fov0, P.fov1, FR, excl);
                S.rec[k] = rk;
                extra = k;
                xlive = rk.valid && rk.social;
                xs = rk.sx;
                xe = rk.ex;
            }
            S.cnt[6] = extra;
            S.cnt[2 * par] = 0; S.cnt[2 * par + 1] = 0;   // agent k+1's counts (last read by k-1)
        }
        // ---- wave 1: agent k+1's records but agent k's
        if (f1 < N && tid >= rbase && tid < cbase) {
            uint32_t* Vb = reinterpret_cast<uint32_t*>(S.V);
            if (!wave_raster)
                for (int w = tid - rbase; w < P.words; w += 64) Vb[w] = 0u;
            const Focal f = S.focal[par];
            const int n = lst_n[par], wl = tid - rbase;
            if (wave_raster && n <= 64 && P.words <= 64) {
                // one record and one field word per lane: the live intervals
                // go from registers to every lane by readlane (the records
                // are not needed after the raster)
                int sx = 0, ex = 0;
                bool live = false;
                if (wl < n) {
                    const int j = lst[par][wl];
                    const Rec rc = record_of(S, j, clsb[par][j], N, f, R, P.step, P.fov0, P.fov1, FR, excl);
                    live = rc.valid && rc.social;
                    sx = rc.sx;
                    ex = rc.ex;
                }
                uint64_t m = __ballot(live);
                uint32_t bits = 0u;
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    bits |= interval_word(wl, __builtin_amdgcn_readlane(sx, l), __builtin_amdgcn_readlane(ex, l), R);
                }
                if (wl < P.words) Vb[wl] = bits;
            } else {
                for (int q = wl; q < n; q += 64) {
                    const int j = lst[par][q];
                    S.rec[j] = record_of(S, j, clsb[par][j], N, f, R, P.step, P.fov0, P.fov1, FR, excl);
                }
            }
            if (wave_raster && !(n <= 64 && P.words <= 64)) {   // the listed records' raster; agent k's comes in phase 2
                fg_wave_sync();
                for (int w = wl; w - wl < P.words; w += 64) {
                    const uint32_t bits = list_word(S, R, lst[par], lst_n[par], w);
                    if (w < P.words) Vb[w] = bits;
                }
            }
        }
        // ---- waves 2 to 7: agent k+2's classes but agent k's and k+1's
        if (f2 < N && tid >= cbase)
            classify(S, P, N, f2, k, f1, excl, clsb[par2], lst[par2], &lst_n[par2], tid - cbase, cstride);
        // ---- one classifying wave, then: agent k+1's speculative headings,
        // from the uniforms drawn before agent k
        if (f1 < N && tid >= spec_tid && tid < spec_tid + 3)
            spec_heading(S, P, N, f1, (int)S.spec[24 + (k & 1)], tid - spec_tid, S.spec + 12 * par);
        // ---- the next wave, then: agent k+2's focal geometry (its
        // start-of-step pose: agents move only in their own turn); that
        // buffer was last read for agent k
        if (tid == focal_tid && f2 < N) S.focal[par2] = focal_of(S, f2);
        __syncthreads();
        if (f1 >= N) break;
        // ---- exclusion, raster and counts of agent k+1; meanwhile wave 2
        // adds agent k's class for agent k+2
        if (wave_raster) {   // agent k's interval from thread 0's registers
            if (tid < 64)
                count_wave(S, R, P.words, __builtin_amdgcn_readlane(xlive, 0) != 0, __builtin_amdgcn_readlane(xs, 0),
                           __builtin_amdgcn_readlane(xe, 0), ta, tb, la, m_tot, m_left, &n_set, &n_left);
        } else {
            raster_list(S, R, excl, lst[par], lst_n[par], S.cnt[6]);
        }
        if (tid == 0) lst_n[par] = 0;   // rebuilt for agent k+3
        if (tid == cbase && f2 < N) {   // agent k's class for agent k+2; its record comes with the others
            // without exclusion only a listed obstacle's class is read, and
            // only an exploiting agent k can be listed: skip the distance otherwise
            int c2 = kNone;
            if (excl || S.over[k] == ABM_MODE_EXPLOIT) {
                const double gx = S.x[f2] + S.rad[f2], gy = S.y[f2] + S.rad[f2];
                c2 = cue_class(S, P, k, f2, gx, gy, S.pid[f2]);
            }
            clsb[par2][k] = c2;
            if (needs_record(c2, excl)) {
                lst[par2][lst_n[par2]] = k;
                lst_n[par2] += 1;
            }
        }
        if (!wave_raster) count_field(S, P, R, S.cnt + 2 * par, bits_out, f1);
        __syncthreads();
    }
    for (int i = tid; i < N; i += nthr) {
        const int g = a0 + i;
        st.x[g] = S.x[i]; st.y[g] = S.y[i]; st.theta[g] = S.th[i]; st.v[g] = S.v[i];
        st.w[g] = S.w[i]; st.u[g] = S.u[i]; st.collected_r_before[g] = S.crb[i];
        st.mode[g] = S.mode[i]; st.overriding_mode[g] = S.over[i];
        st.I_priv[g] = S.nov[i];
    }
    if (tid == 0 && u_used) u_used[rb] = used0 + (words || mt ? 2 : 1) * draws;
    if (tid == 0 && status) status[rb] = ABM_WORLD_DONE;
    if (mt) {
        // advance the generator by the 2 * draws words used: numpy's state
        // after them is the key twisted v times at position pos0 + c - 624 v
        if (tid == 0) S.cnt[7] = draws;
        __syncthreads();
        const int c = 2 * S.cnt[7];
        int v = 0, pos = mt_pos0 + c;
        if (pos > kMtN) {
            v = (pos - 1) / kMtN;
            pos -= kMtN * v;
        }
        if (v > 0) {
            for (int i = tid; i < kMtN; i += nthr) S.mtk[i] = mt[rb].key[i];
            __syncthreads();
            for (int t = 0; t < v; ++t) mt_twist_block(S.mtk);
            for (int i = tid; i < kMtN; i += nthr) mt[rb].key[i] = S.mtk[i];
        }
        if (tid == 0) mt[rb].pos = pos;
    }
}

// ------------------------------------------------ replicas beyond one workgroup
// Agent.update in id order (sims.py:861) for a replica whose agents do not fit
// one workgroup's LDS, without visual exclusion.  Agent k's update reads the
// agents j < k after their own update and j > k before it, and only through
// what makes j a social cue for k: j's overriding mode (exploit), its patch
// id (never changed by the update) and its position.  Sweeps recompute every
// agent from the latest tentative states X^m (j < k) and the step's start
// (j >= k) at once:
//   X^{m+1}_k = F_k(X^m_0, .., X^m_{k-1}, start_k, .., start_{N-1}),
// the walk's draw index being the number of draws in X^{m+1}_j, j < k.
// Agent k is final after k + 1 sweeps, and the only fixed point X = F(X) is
// the reference's sequential result, so the sweeps stop at the first
// X^{m+1} == X^m (every value bitwise), which is that result; in practice a
// few sweeps, as only exploiting agents carry dependencies.
struct SweepX {
    double *x, *y, *th, *v, *w, *u, *ip, *dth, *dvel;
    int32_t *over, *mode, *kind;   // kind: 0 = draws a walk turn, 1 = fixed increments, -1 = unset
};

// Sweep control words.
constexpr int kSwConverged = 0, kSwSkip = 1, kSwCand = 2, kSwIters = 3, kSwDraws = 4,
              kSwUsed0 = 5, kSwPos0 = 6, kSwBad = 7, kSwTwists = 8, kSwWords = 9;

constexpr int kSweepMaxWords = 256;   // R <= 8192: four field words per lane

// Exploiting agents with a patch (at the step's start or in X), the only
// candidates for a social cue.
__device__ inline void sweep_candidates(const abm_forage_state& st, int a0, int N, const int32_t* xover,
                                        int32_t* cand, int32_t* ctrl) {
    if (threadIdx.x == 0) ctrl[kSwCand] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int base = threadIdx.x - lane; base < N; base += blockDim.x) {
        const int j = base + lane;
        bool c = false;
        if (j < N) {
            const int g = a0 + j;
            c = st.exploited_patch_id[g] != -1 &&
                (st.overriding_mode[g] == ABM_MODE_EXPLOIT || (xover && xover[j] == ABM_MODE_EXPLOIT));
        }
        const uint64_t m = __ballot(c);
        if (m) {
            int at = 0;
            if (lane == 0) at = atomicAdd(&ctrl[kSwCand], (int)__popcll(m));
            at = __builtin_amdgcn_readfirstlane(at);
            if (c) cand[at + (int)__popcll(m & ((1ull << lane) - 1ull))] = j;
        }
    }
    __syncthreads();
}

// One block of 1024 threads: status, the step's uniforms, X^0 = the start
// state, the first candidates.
__global__ void __launch_bounds__(1024)
sweep_prep_kernel(const FGRep* __restrict__ reps, int rb, abm_forage_state st,
                  const double* __restrict__ uniforms, const uint32_t* __restrict__ words,
                  const int64_t* __restrict__ u_off, const int32_t* __restrict__ u_cnt,
                  const int32_t* __restrict__ u_used, const int32_t* __restrict__ status,
                  const abm_mt19937_state* __restrict__ mt, SweepX X0, double* __restrict__ uni,
                  uint32_t* __restrict__ keys, int32_t* __restrict__ cand, int32_t* __restrict__ ctrl) {
    __shared__ uint32_t key[kMtN];
    const FGRep& P = reps[rb];
    const int N = P.n_agents, a0 = P.agent_offset, tid = threadIdx.x, nthr = blockDim.x;
    const bool skip = status && status[rb] != kWorldReady;
    if (tid == 0) {
        ctrl[kSwConverged] = 0; ctrl[kSwSkip] = skip ? 1 : 0; ctrl[kSwIters] = 0;
        ctrl[kSwDraws] = 0; ctrl[kSwBad] = 0; ctrl[kSwTwists] = 0;
        ctrl[kSwUsed0] = (words || mt) ? u_used[rb] : 0;
        ctrl[kSwPos0] = mt ? mt[rb].pos : 0;
    }
    if (skip) return;   // uniform over the block
    for (int i = tid; i < N; i += nthr) {
        const int g = a0 + i;
        X0.x[i] = st.x[g]; X0.y[i] = st.y[g]; X0.th[i] = st.theta[g]; X0.v[i] = st.v[g];
        X0.w[i] = st.w[g]; X0.u[i] = st.u[g]; X0.ip[i] = st.I_priv[g];
        X0.dth[i] = 0.0; X0.dvel[i] = 0.0;
        X0.over[i] = st.overriding_mode[g]; X0.mode[i] = st.mode[g]; X0.kind[i] = -1;
    }
    // agent i's walk draws the i-th uniform at most: doubles, pairs of words
    // after those the resource loop used, or the device generator's next 2N
    // words; NaN past the end of a buffer
    if (mt) {
        for (int i = tid; i < kMtN; i += nthr) key[i] = mt[rb].key[i];
        __syncthreads();
        uint32_t* wb = reinterpret_cast<uint32_t*>(uni);
        int made = 0, p = mt[rb].pos, twists = 0;
        while (made < 2 * N) {
            if (p >= kMtN) {
                // each twisted key is kept, so the commit takes the state
                // after the words drawn without twisting again
                mt_twist_block(key);
                for (int i = tid; i < kMtN; i += nthr) keys[(int64_t)twists * kMtN + i] = key[i];
                ++twists;
                p = 0;
            }
            const int take = min(kMtN - p, 2 * N - made);
            for (int j = tid; j < take; j += nthr) wb[made + j] = mt_temper(key[p + j]);
            made += take;
            p += take;
            __syncthreads();
        }
        for (int i = tid; i < N; i += nthr) {   // each thread rewrites its own 8 bytes
            const uint32_t w0 = wb[2 * i], w1 = wb[2 * i + 1];
            uni[i] = words_to_double(w0, w1);
        }
        if (tid == 0) ctrl[kSwTwists] = twists;
    } else {
        const int ucount = u_cnt[rb];
        const int used0 = words ? u_used[rb] : 0;
        for (int i = tid; i < N; i += nthr) {
            double uu = NAN;
            if (words) {
                const int at = used0 + 2 * i;
                if (at + 2 <= ucount) uu = words_to_double(words[u_off[rb] + at], words[u_off[rb] + at + 1]);
            } else if (i < ucount) {
                uu = uniforms[u_off[rb] + i];
            }
            uni[i] = uu;
        }
    }
    __syncthreads();
    sweep_candidates(st, a0, N, nullptr, cand, ctrl);
}

// One wave per focal agent k: its social field from the view (X^m for j < k,
// the start for j > k), the counts, and the decision (agent.py:168-265)
// into X^{m+1}; the field bits when asked for.
__global__ void __launch_bounds__(256)
sweep_field_kernel(const FGRep* __restrict__ reps, int rb, abm_forage_state st, SweepX Xc, SweepX Xn,
                   const int32_t* __restrict__ cand, const int32_t* __restrict__ ctrl,
                   uint32_t* __restrict__ bits_out) {
#pragma clang fp contract(off)
    __shared__ uint32_t raw[4][kSweepMaxWords];
    if (ctrl[kSwConverged] | ctrl[kSwSkip]) return;
    const FGRep& P = reps[rb];
    const int N = P.n_agents, R = P.n_rays, W = P.words, a0 = P.agent_offset;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + wv;
    if (k >= N) return;   // uniform over the wave
    const int gk = a0 + k;
    const double rk = (double)st.radius[gk];
    const double pxk = st.x[gk], pyk = st.y[gk];
    const int pid_k = st.exploited_patch_id[gk];
    Focal f;
    {
        double sn, cs;
        sincos(st.theta[gk], &sn, &cs);
        f.px = pxk; f.py = pyk; f.radius = rk;
        f.cx = pxk + rk;
        f.cy = pyk + rk;
        const double ex = pxk + (1.0 + cs) * rk;
        const double ey = pyk + (1.0 - sn) * rk;
        const double v1x = ex - f.cx, v1y = ey - f.cy;
        const double n1 = norm2(v1x, v1y);
        f.u1x = v1x / n1;
        f.u1y = v1y / n1;
    }
    const FastRec FR = fast_rec_const(R, P.step, P.fov0, P.fov1);
    const int ncand = ctrl[kSwCand];
    uint32_t bits[4] = {0u, 0u, 0u, 0u};
    for (int base = 0; base < ncand; base += 64) {
        const int q = base + lane;
        bool live = false;
        int sx = 0, ex = 0;
        if (q < ncand) {
            const int j = cand[q];
            const int g = a0 + j;
            const bool moved = j < k;
            const int over = moved ? Xc.over[j] : st.overriding_mode[g];
            const int pid = st.exploited_patch_id[g];
            // agent.py:396-413: a cue is an exploiting agent in range with a
            // patch other than the focal's own (patchwise exclusion)
            if (j != k && over == ABM_MODE_EXPLOIT && pid != -1 && !(P.patchwise && pid == pid_k)) {
                const double ox = moved ? Xc.x[j] : st.x[g];
                const double oy = moved ? Xc.y[j] : st.y[g];
                const double rj = (double)st.radius[g];
                if (norm2((ox + rj) - f.cx, (oy + rj) - f.cy) <= P.vision_range) {
                    const Rec rc = record_at(ox, oy, kSocial, j, f, R, P.step, P.fov0, P.fov1, FR, false);
                    live = rc.valid != 0;
                    sx = rc.sx;
                    ex = rc.ex;
                }
            }
        }
        uint64_t m = __ballot(live);
        while (m) {
            const int l = __builtin_ctzll(m);
            m &= m - 1;
            const int s = __builtin_amdgcn_readlane(sx, l), e = __builtin_amdgcn_readlane(ex, l);
#pragma unroll
            for (int qw = 0; qw < 4; ++qw) {
                const int w = lane + 64 * qw;
                if (w < W) bits[qw] |= interval_word(w, s, e, R);
            }
        }
    }
    // counts of the returned field: the kept raw range [ta, tb] and its left half
    const int nl = R / 2;
    const int ta = R - 1 - P.keep_hi, tb = R - 1 - P.keep_lo, la = max(R - nl, ta);
    int tot = 0, left = 0;
#pragma unroll
    for (int qw = 0; qw < 4; ++qw) {
        const int w = lane + 64 * qw;
        if (w < W) {
            tot += __popc(bits[qw] & word_range_mask(w, ta, tb));
            left += __popc(bits[qw] & word_range_mask(w, la, tb));
        }
    }
    const int n_set = wave_sum(tot), n_left = wave_sum(left);
    if (bits_out) {
        // soc_v_field[q] = V[R-1-q] on the kept rays
#pragma unroll
        for (int qw = 0; qw < 4; ++qw) {
            const int w = lane + 64 * qw;
            if (w < W) raw[wv][w] = bits[qw];
        }
        fg_wave_sync();
        uint32_t* dst = bits_out + P.field_offset + (int64_t)k * W;
        for (int wd = lane; wd < W; wd += 64) {
            uint32_t mm = 0u;
            for (int b = 0; b < 32; ++b) {
                const int q = wd * 32 + b, r = R - 1 - q;
                if (q >= P.keep_lo && q <= P.keep_hi && ((raw[wv][r >> 5] >> (r & 31)) & 1u)) mm |= 1u << b;
            }
            dst[wd] = mm;
        }
    }
    if (lane != 0) return;
    // ---- the decision of agent k (agent.py:168-265), as in the step kernel
    double w = st.w[gk], u = st.u[gk];
    const double vel0 = st.v[gk], cr = st.collected_r[gk], crb = st.collected_r_before[gk];
    double nov = st.novelty[(int64_t)gk * P.tau];
    for (int q = 1; q < P.tau; ++q) nov = fmax(nov, st.novelty[(int64_t)gk * P.tau + q]);
    int over = st.overriding_mode[gk], mode = st.mode[gk];
    const int env = st.env_status[gk];
    const double count = (double)n_set, cleft = (double)n_left;
    const double collected = cr - crb;
    const double ipriv = P.F_N * nov + P.F_R * collected;
    const double w_p = w > P.T_w ? w : 0.0;
    const double u_p = u > P.T_u ? u : 0.0;
    const double mean = count / (double)R;
    const double dw = P.Eps_w * mean - P.g_w * (w - P.B_w) - u_p * P.S_uw;
    const double du = P.Eps_u * ipriv - P.g_u * (u - P.B_u) - w_p * P.S_wu;
    w = w + dw;
    u = u + du;
    if (w > P.w_max) w = P.w_max;
    if (w < -P.w_max) w = -P.w_max;
    if (u > P.u_max) u = P.u_max;
    if (u < -P.u_max) u = -P.u_max;
    double dvel = 0.0, dth = 0.0;
    int kind = 1;
    const bool tw = w > P.T_w, tu = u > P.T_w;   // tr_u compares with T_w
    auto walk = [&]() {
        kind = 0;
        dvel = P.exp_vel_max;
        over = -1; mode = ABM_MODE_EXPLORE;
    };
    auto reloc = [&]() {
        const double lm = cleft / (double)nl;
        const double rm = (count - cleft) / (double)(R - nl);
        dvel = P.exp_vel_max - vel0;
        dth = (lm - rm) * P.reloc_theta_max;
        over = -1; mode = ABM_MODE_RELOCATE;
    };
    auto exploit = [&]() {
        over = ABM_MODE_EXPLOIT; mode = ABM_MODE_EXPLOIT;
        dvel = -vel0 * P.exp_stop_ratio;
        dth = 0.0;
    };
    if (over != ABM_MODE_COLLIDE) {
        if (!tw && !tu) walk();
        else if (tw && tu) { if (env == 1) exploit(); else reloc(); }
        else if (tw && !tu) reloc();
        else { if (env == 1) exploit(); else walk(); }
    }
    Xn.w[k] = w; Xn.u[k] = u; Xn.ip[k] = ipriv;
    Xn.over[k] = over; Xn.mode[k] = mode; Xn.kind[k] = kind;
    Xn.dth[k] = dth; Xn.dvel[k] = dvel;
}

// One block of 1024 threads: the draw indices (a scan of the walk flags in
// id order), the moves of X^{m+1}, the comparison with X^m and the next
// candidates.  The agents go in chunks of one per thread, so every load is
// coalesced; a chunk's draw indices come from wave ballots and the waves'
// totals in LDS.
__global__ void __launch_bounds__(1024)
sweep_motion_kernel(const FGRep* __restrict__ reps, int rb, abm_forage_state st, SweepX Xc, SweepX Xn,
                    const double* __restrict__ uni, int32_t* __restrict__ cand, int32_t* __restrict__ ctrl) {
#pragma clang fp contract(off)
    __shared__ int wtot[2][16];
    __shared__ int changed;
    if (ctrl[kSwConverged] | ctrl[kSwSkip]) return;   // uniform over the block
    const FGRep& P = reps[rb];
    const int N = P.n_agents, a0 = P.agent_offset, tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    if (tid == 0) changed = 0;
    auto bits_ne = [](double a, double b) { return __double_as_longlong(a) != __double_as_longlong(b); };
    bool diff = false;
    int d_base = 0;   // draws before the chunk (uniform over the block)
    for (int c0 = 0, par = 0; c0 < N; c0 += nthr, par ^= 1) {
        const int k = c0 + tid;
        const bool in = k < N;
        const int kind = in ? Xn.kind[k] : 1;
        const uint64_t m = __ballot(kind == 0);
        if (lane == 0) wtot[par][wave] = (int)__popcll(m);
        __syncthreads();   // wtot[par] is rewritten two chunks later, after this chunk's barrier
        int d = d_base + (int)__popcll(m & ((1ull << lane) - 1ull)), tot = 0;
        for (int w = 0; w < nw; ++w) {
            const int t = wtot[par][w];
            if (w < wave) d += t;
            tot += t;
        }
        d_base += tot;
        if (!in) continue;
        const int g = a0 + k;
        double dth = Xn.dth[k];
        if (kind == 0) {
            const double uu = d < N ? uni[d] : NAN;
            dth = P.exp_theta_min + (P.exp_theta_max - P.exp_theta_min) * uu;
        }
        Pose p{st.x[g], st.y[g], st.theta[g], st.v[g]};
        p.theta = prove_orientation(p.theta + dth);
        double sn, cs;
        sincos(p.theta, &sn, &cs);
        p.v = p.v + Xn.dvel[k];
        // prove_velocity (agent.py:612-620): only in explore mode
        if (get_mode(Xn.over[k], Xn.w[k], P.T_w) == ABM_MODE_EXPLORE && fabs(p.v) > 1.0) p.v = P.exp_vel_max;
        p.x = p.x + p.v * cs;
        p.y = p.y - p.v * sn;
        p = reflect_walls(p, (double)st.radius[g], P.pad, P.width, P.height);
        Xn.x[k] = p.x; Xn.y[k] = p.y; Xn.th[k] = p.theta; Xn.v[k] = p.v;
        diff = diff || bits_ne(p.x, Xc.x[k]) || bits_ne(p.y, Xc.y[k]) || bits_ne(p.theta, Xc.th[k]) ||
               bits_ne(p.v, Xc.v[k]) || bits_ne(Xn.w[k], Xc.w[k]) || bits_ne(Xn.u[k], Xc.u[k]) ||
               bits_ne(Xn.ip[k], Xc.ip[k]) || Xn.over[k] != Xc.over[k] || Xn.mode[k] != Xc.mode[k] ||
               kind != Xc.kind[k];
    }
    if (diff) changed = 1;   // benign race: every writer stores 1
    __syncthreads();
    if (tid == 0) {
        ctrl[kSwDraws] = d_base;
        ctrl[kSwIters] += 1;
    }
    if (!changed) {
        if (tid == 0) ctrl[kSwConverged] = 1;
        return;   // uniform over the block
    }
    sweep_candidates(st, a0, N, Xn.over, cand, ctrl);
}

// Writes the converged sweep into the state (one block of 1024 threads):
// poses, decision state, collected_r_before, the words used and status, and
// the device generator advanced by the words drawn (the key the prep
// twisted to).
__global__ void __launch_bounds__(1024)
sweep_commit_kernel(const FGRep* __restrict__ reps, int rb, abm_forage_state st, SweepX X,
                    int32_t* __restrict__ u_used, int32_t* __restrict__ status, int words_or_mt,
                    abm_mt19937_state* __restrict__ mt, const uint32_t* __restrict__ keys,
                    int32_t* __restrict__ ctrl) {
    __shared__ uint32_t key[kMtN];
    if (ctrl[kSwSkip]) return;
    if (!ctrl[kSwConverged]) {
        if (threadIdx.x == 0) ctrl[kSwBad] = 1;
        return;
    }
    const FGRep& P = reps[rb];
    const int N = P.n_agents, a0 = P.agent_offset, tid = threadIdx.x, nthr = blockDim.x;
    for (int i = tid; i < N; i += nthr) {
        const int g = a0 + i;
        st.x[g] = X.x[i]; st.y[g] = X.y[i]; st.theta[g] = X.th[i]; st.v[g] = X.v[i];
        st.w[g] = X.w[i]; st.u[g] = X.u[i]; st.I_priv[g] = X.ip[i];
        st.collected_r_before[g] = st.collected_r[g];
        st.mode[g] = X.mode[i]; st.overriding_mode[g] = X.over[i];
    }
    const int draws = ctrl[kSwDraws];
    if (tid == 0 && u_used) u_used[rb] = ctrl[kSwUsed0] + (words_or_mt ? 2 : 1) * draws;
    if (tid == 0 && status) status[rb] = ABM_WORLD_DONE;
    if (mt) {
        // numpy's state after the 2 * draws words: the key twisted v times at
        // position pos0 + c - 624 v (as the step kernel)
        const int c = 2 * draws;
        int v = 0, pos = ctrl[kSwPos0] + c;
        if (pos > kMtN) {
            v = (pos - 1) / kMtN;
            pos -= kMtN * v;
        }
        // the prep kept every key it twisted, and c <= 2N words never need
        // more; twisting here again is only the fallback
        if (v > ctrl[kSwTwists]) {   // uniform over the block
            for (int i = tid; i < kMtN; i += nthr) key[i] = mt[rb].key[i];
            __syncthreads();
            for (int t = 0; t < v; ++t) mt_twist_block(key);
            for (int i = tid; i < kMtN; i += nthr) mt[rb].key[i] = key[i];
        } else if (v > 0) {
            for (int i = tid; i < kMtN; i += nthr) mt[rb].key[i] = keys[(int64_t)(v - 1) * kMtN + i];
        }
        if (tid == 0) mt[rb].pos = pos;
    }
}

// ------------------------------------------------ agent-resource interaction
constexpr int kMaxPatches = 64;     // patches per replica (one 64-bit membership word)
constexpr int kMaxWorldChunks = 1024;   // 64-agent words of a replica's membership (65536 agents)
constexpr int kMaxRetries = 10000;  // add_new_resource_patch's max_retries (sims.py:335)

// Pending patch placements of a replica, kept between a phase that ran out of
// words and its resume.
constexpr int kPendNone = 0, kPendKill = 1, kPendFill = 2;

// Per-replica world constants on the device.
struct WRep {
    int32_t n_patches, patch_radius, pooling_time, regenerate, teleport, collide, ghost, pad_;
    int64_t patch_offset;
    int64_t x_lo, x_hi, y_lo, y_hi, units_lo, units_hi;   // randint bounds [lo, hi)
    double consumption, q_lo, q_hi;
};

// Raw MT19937 outputs of one replica; reading past the end sets `out`.
struct WordCursor {
    const uint32_t* w;
    int used, limit;
    bool out;
    __device__ uint32_t next() {
        if (used >= limit) { out = true; return 0u; }
        return w[used++];
    }
};

// RandomState.randint(lo, hi) for a range below 2^32 - 1: numpy draws
// next_uint32 & mask until the value is <= hi - 1 - lo, mask being the
// smallest all-ones word covering it (random_bounded_uint64_fill with
// use_masked; no draw when the range holds one value).
template <typename Cursor>
__device__ inline int64_t np_randint(Cursor& c, int64_t lo, int64_t hi) {
    const uint32_t rng = (uint32_t)(hi - 1 - lo);
    if (rng == 0u) return lo;
    uint32_t mask = rng;
    mask |= mask >> 1; mask |= mask >> 2; mask |= mask >> 4; mask |= mask >> 8; mask |= mask >> 16;
    uint32_t v;
    do { v = c.next() & mask; } while (v > rng && !c.out);
    return lo + (int64_t)v;
}

// RandomState.uniform(lo, hi) = lo + (hi - lo) * random_sample.
template <typename Cursor>
__device__ inline double np_uniform(Cursor& c, double lo, double hi) {
#pragma clang fp contract(off)
    const uint32_t w0 = c.next(), w1 = c.next();
    return lo + (hi - lo) * words_to_double(w0, w1);
}

// Python's float x % (2*pi) (and numpy's remainder): fmod, shifted into
// [0, 2*pi) when negative, +0 for a zero.
__device__ inline double py_mod_2pi(double x) {
#pragma clang fp contract(off)
    double m = fmod(x, kTwoPi);
    if (m != 0.0) {
        if (m < 0.0) m = m + kTwoPi;
    } else {
        m = 0.0;
    }
    return m;
}

// notify_agent (sims.py:29-42): env status, novelty ring, pool_success, patch id.
__device__ inline void notify(const abm_forage_state& st, int32_t* pool_success, int tau, int g,
                              int status, int patch_id) {
    const int before = st.env_status[g];
    st.env_status[g] = status;
    double* nov = st.novelty + (int64_t)g * tau;
    // np.roll(novelty, 1), eight entries at a time from the top: a chunk's
    // loads are all issued before its stores, and read below what it writes
    for (int top = tau - 1; top > 0; top -= 8) {
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (top - i > 0) v[i] = nov[top - i - 1];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (top - i > 0) nov[top - i] = v[i];
    }
    nov[0] = (status - before) > 0 ? 1.0 : 0.0;
    pool_success[g] = 1;
    st.exploited_patch_id[g] = patch_id;
}

// The condition under which the loop notifies an agent (sims.py:816-817, 852).
__device__ inline bool pooled(int mode, int pool_success, int pooling_time) {
    return ((mode == ABM_MODE_POOL || mode == ABM_MODE_RELOCATE) && pool_success) || pooling_time == 0;
}

// pygame.sprite.collide_circle of two patches: integer rect centres,
// squared distance <= (r1 + r2)^2.
__device__ inline bool patches_collide(int64_t cx1, int64_t cy1, int64_t r1, int64_t cx2, int64_t cy2,
                                       int64_t r2) {
    const int64_t dx = cx1 - cx2, dy = cy1 - cy2, rr = r1 + r2;
    return dx * dx + dy * dy <= rr * rr;
}

struct FreshPatch {
    int32_t id;
    int64_t x, y;
    double units, quality;
};

// Place the replica's pending patches (regenerations of the patches killed
// this step, in kill order, or create_resources' new patches) from the word
// buffer, then compact the list: live patches in order, new ones appended.
// One thread.  Commits nothing unless every placement fits in the buffer.
template <typename Cursor>
__device__ int place_patches(const WRep& W, abm_patch_state ps, int b, int kind, int n_pending,
                             const int32_t* __restrict__ pend_id, int32_t* __restrict__ kill_rank,
                             Cursor& cur, FreshPatch* fresh) {
    const int64_t po = W.patch_offset;
    const int count = ps.count[b];
    int n_new = 0;
    if (kind == kPendKill) n_new = W.regenerate ? n_pending : 0;
    else if (kind == kPendFill) n_new = W.n_patches - count;
    const int64_t r = W.patch_radius;
    for (int k = 0; k < n_new; ++k) {
        int32_t id;
        if (kind == kPendKill) {
            id = pend_id[po + k];                 // force_id: the killed patch's id
        } else {
            bool any = false;                     // max(ids) + 1, or 1 for an empty list
            int32_t mx = 0;
            for (int t = 0; t < count; ++t) {
                const int32_t v = ps.id[po + t];
                if (!any || v > mx) mx = v;
                any = true;
            }
            for (int f = 0; f < k; ++f) {
                if (!any || fresh[f].id > mx) mx = fresh[f].id;
                any = true;
            }
            id = (any ? mx : 0) + 1;
        }
        int retries = 0;
        for (;;) {
            if (retries > kMaxRetries) return ABM_WORLD_TIMEOUT;
            const int64_t x = np_randint(cur, W.x_lo, W.x_hi);
            const int64_t y = np_randint(cur, W.y_lo, W.y_hi);
            const int64_t units = np_randint(cur, W.units_lo, W.units_hi);
            const double q = np_uniform(cur, W.q_lo, W.q_hi);
            if (cur.out) return ABM_WORLD_NEED_WORDS;
            const int64_t cx = x + r, cy = y + r;
            bool ok = true;
            // the Group at the time: live patches, those killed later this
            // step, and the ones placed before this one
            for (int t = 0; t < count && ok; ++t) {
                const int64_t s = po + t;
                if (kill_rank[s] >= 0 && kill_rank[s] <= k) continue;
                const double pr = ps.radius[s];
                const int64_t pcx = (int64_t)trunc(ps.x[s] + pr), pcy = (int64_t)trunc(ps.y[s] + pr);
                if (patches_collide(pcx, pcy, (int64_t)pr, cx, cy, r)) ok = false;
            }
            for (int f = 0; f < k && ok; ++f)
                if (patches_collide(fresh[f].x + r, fresh[f].y + r, r, cx, cy, r)) ok = false;
            ++retries;
            if (ok) {
                fresh[k].id = id; fresh[k].x = x; fresh[k].y = y;
                fresh[k].units = (double)units; fresh[k].quality = q;
                break;
            }
        }
    }
    // commit: drop the killed patches, keep the order, append the new ones
    int wpos = 0;
    for (int t = 0; t < count; ++t) {
        const int64_t s = po + t;
        if (kill_rank[s] >= 0) { kill_rank[s] = -1; continue; }
        if (wpos != t) {
            const int64_t d = po + wpos;
            ps.id[d] = ps.id[s]; ps.x[d] = ps.x[s]; ps.y[d] = ps.y[s]; ps.radius[d] = ps.radius[s];
            ps.units[d] = ps.units[s]; ps.left[d] = ps.left[s]; ps.quality[d] = ps.quality[s];
        }
        ++wpos;
    }
    for (int k = 0; k < n_new; ++k) {
        const int64_t d = po + wpos + k;
        ps.id[d] = fresh[k].id; ps.x[d] = (double)fresh[k].x; ps.y[d] = (double)fresh[k].y;
        ps.radius[d] = (double)r; ps.units[d] = fresh[k].units; ps.left[d] = fresh[k].units;
        ps.quality[d] = fresh[k].quality;
    }
    ps.count[b] = wpos + n_new;
    return ABM_WORLD_DONE;
}

// ------------------------------------------------ agent-agent collisions
constexpr int kCollideMaxAgents = 2048;   // replicas with collisions on
constexpr int kCollideMaxNear = 256;      // agents within 2 r + 20 of a colliding agent

size_t collide_lds_bytes(int maxN, int maxR) {
    return ((sizeof(int32_t) * 5 * (size_t)maxN + sizeof(int32_t) * kCollideMaxNear + 15) & ~(size_t)15) +
           fg_lds_bytes(kCollideMaxNear, maxR) + 16;
}

// numpy's pairwise float64 sum (np.add.reduce of a contiguous array,
// pairwise_sum in numpy/_core/src/umath/loops_utils.h.src): fewer than 8
// elements in sequence from 0.0; up to 128 in eight strided partial sums
// combined as ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7)), then the
// remainder in sequence; longer runs split at n / 2 rounded down to a multiple
// of 8, each half summed the same way.  val(i) is element i.
template <class Val>
__device__ double np_block_sum(const Val& val, int lo, int n) {
#pragma clang fp contract(off)
    if (n < 8) {
        double res = 0.0;
        for (int i = 0; i < n; ++i) res += val(lo + i);
        return res;
    }
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = val(lo + j);
    int i = 8;
    for (; i < n - (n % 8); i += 8)
        for (int j = 0; j < 8; ++j) r[j] += val(lo + i + j);
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += val(lo + i);
    return res;
}

template <class Val>
__device__ double np_pairwise_sum(const Val& val, int lo, int n) {
#pragma clang fp contract(off)
    // post-order walk of the split tree on an explicit stack
    int s_lo[32], s_n[32], s_state[32];
    double s_left[32];
    int sp = 0;
    s_lo[0] = lo; s_n[0] = n; s_state[0] = 0;
    for (;;) {
        if (s_n[sp] > 128) {
            const int n2 = s_n[sp] / 2 - (s_n[sp] / 2) % 8;
            s_state[sp] = 1;
            s_lo[sp + 1] = s_lo[sp]; s_n[sp + 1] = n2; s_state[sp + 1] = 0;
            ++sp;
            continue;
        }
        double ret = np_block_sum(val, s_lo[sp], s_n[sp]);
        for (;;) {
            if (sp == 0) return ret;   // the node just finished is the root
            --sp;
            if (s_state[sp] == 1) {    // its left half: start the right half
                const int n2 = s_n[sp] / 2 - (s_n[sp] / 2) % 8;
                s_left[sp] = ret;
                s_state[sp] = 2;
                s_lo[sp + 1] = s_lo[sp] + n2; s_n[sp + 1] = s_n[sp] - n2; s_state[sp + 1] = 0;
                ++sp;
                break;
            }
            ret = s_left[sp] + ret;    // its right half: this node is done
        }
    }
}

// Agent-agent collisions of the loop body (sims.py:736-787), one block per
// replica before the resource loop.  Collisions are those of
// pygame.sprite.groupcollide(agents, agents, ..., within_group_collision)
// with every radius grown by 2 (sims.py:738-747): collide_circle on the rect
// centres, int(position) + radius (the rect holds int() of the position and
// is 2 r wide), squared integer distance <= (r1 + 2 + r2 + 2)^2.  Each agent2
// of a colliding pair (agent1, agent2) takes the proximity response of
// agent_agent_collision_proximity (sims.py:421-467) once per partner agent1
// that passes the ghost-mode test; the response reads only agent2's own
// heading and everyone's positions, which it does not change, so the agents
// are handled one at a time with their responses in sequence.  Then the
// collided list (sims.py:760-777) ends or notifies the collide modes
// (sims.py:779-784).  A replica with an agent beyond kCollideMaxNear
// neighbours within 2 r + 20 is reported ABM_WORLD_CROWDED and not touched.
__global__ void __launch_bounds__(256)
forage_collide_kernel(const FGRep* __restrict__ reps, const WRep* __restrict__ wreps,
                      abm_forage_state st, int32_t* __restrict__ pool_success,
                      const int32_t* __restrict__ pend, int32_t* __restrict__ status) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) double smem[];
    const int b = blockIdx.x, tid = threadIdx.x, nthr = blockDim.x;
    const FGRep& P = reps[b];
    const WRep& W = wreps[b];
    if (!W.collide || pend[2 * b] != kPendNone) return;   // uniform over the block
    const int N = P.n_agents, a0 = P.agent_offset, R = P.n_rays;
    char* base = reinterpret_cast<char*>(smem);
    int32_t* cx = reinterpret_cast<int32_t*>(base);
    int32_t* cy = cx + N;
    int32_t* rr = cy + N;
    int32_t* cnt = rr + N;     // responses of each agent as agent2
    int32_t* flag = cnt + N;   // bit 0 collided, bit 1 exploiting
    int32_t* near = flag + N;
    FGSmem S = fg_carve(base + ((sizeof(int32_t) * 5 * (size_t)N + sizeof(int32_t) * kCollideMaxNear + 15) &
                                ~(size_t)15),
                        kCollideMaxNear, R);
    __shared__ int s_crowded, s_nn;
    if (tid == 0) s_crowded = 0;
    for (int j = tid; j < N; j += nthr) {
        const int g = a0 + j;
        const int r = st.radius[g];
        cx[j] = (int)trunc(st.x[g]) + r;
        cy[j] = (int)trunc(st.y[g]) + r;
        rr[j] = r;
        flag[j] = get_mode(st.overriding_mode[g], st.w[g], P.T_w) == ABM_MODE_EXPLOIT ? 2 : 0;
    }
    __syncthreads();
    // ---- pairs, responses per agent and the collided list
    for (int j = tid; j < N; j += nthr) {
        const bool ex_j = flag[j] & 2;
        int n_resp = 0;
        bool any = false, both = false;
        for (int i = 0; i < N; ++i) {
            if (i == j) continue;
            const long dx = cx[i] - cx[j], dy = cy[i] - cy[j], rs = rr[i] + rr[j] + 4;
            if (dx * dx + dy * dy > rs * rs) continue;
            const bool ne = !(flag[i] & 2) && !ex_j;
            any = true;
            both = both || ne;
            if (!W.ghost || ne) ++n_resp;
        }
        // sims.py:760-777: collided agents (the exploiting status never changes here)
        const bool collided = W.teleport ? (any && !ex_j) : (!W.ghost ? any : both);
        cnt[j] = n_resp;
        flag[j] |= collided ? 1 : 0;
        if (n_resp > 0) {
            // neighbours within 2 r + 20 (supcalc.distance, each agent's own radius)
            const int g = a0 + j;
            const double rj = (double)rr[j];
            const double cxj = st.x[g] + rj, cyj = st.y[g] + rj;
            int nn = 0;
            for (int i = 0; i < N; ++i) {
                if (i == j) continue;
                const int gi = a0 + i;
                const double ri = (double)rr[i];
                if (norm2((st.x[gi] + ri) - cxj, (st.y[gi] + ri) - cyj) < 2.0 * rj + 20.0) ++nn;
            }
            if (nn > kCollideMaxNear) s_crowded = 1;   // benign race: every writer stores 1
        }
    }
    __syncthreads();
    if (s_crowded) {
        if (tid == 0) status[b] = ABM_WORLD_CROWDED;
        return;   // uniform over the block
    }
    // ---- the proximity responses, agent by agent
    const double fov0 = -kPi, fov1 = kPi;
    for (int j = 0; j < N; ++j) {
        if (cnt[j] == 0) continue;   // uniform over the block
        const int g = a0 + j;
        const bool ex_j = flag[j] & 2;
        const double rj = (double)rr[j];
        const double pxj = st.x[g], pyj = st.y[g];
        if (tid == 0) {
            if (!ex_j) { st.overriding_mode[g] = ABM_MODE_COLLIDE; st.mode[g] = ABM_MODE_COLLIDE; }
            // the agents within 2 r + 20, in id order
            int nn = 0;
            for (int i = 0; i < N; ++i) {
                if (i == j) continue;
                const int gi = a0 + i;
                const double ri = (double)rr[i];
                if (norm2((st.x[gi] + ri) - (pxj + rj), (st.y[gi] + ri) - (pyj + rj)) < 2.0 * rj + 20.0) {
                    if (nn < kCollideMaxNear) near[nn] = gi;   // the count was checked above
                    ++nn;
                }
            }
            s_nn = nn < kCollideMaxNear ? nn : kCollideMaxNear;
        }
        __syncthreads();
        const int nn = s_nn;
        for (int q = tid; q < nn; q += nthr) {
            S.x[q] = st.x[near[q]];
            S.y[q] = st.y[near[q]];
            S.cls[q] = kSocial;
        }
        __syncthreads();
        for (int c = 0; c < cnt[j]; ++c) {
            const Focal f = make_focal(pxj, pyj, st.theta[g], rj);
            if (tid == 0) {
                // the distance left in projection_field's loop variable: its
                // last non-coincident obstacle (agent.py:501-556)
                S.lastd[0] = 0.0;
                for (int q = nn - 1; q >= 0; --q) {
                    const PairFG p = pair_forage_exact(f, S.x[q], S.y[q], R, P.step, fov0, fov1);
                    if (!p.skipped) { S.lastd[0] = p.dist; break; }
                }
            }
            project_focal(S, nn, R, P.step, fov0, fov1, P.visual_exclusion != 0, false, f, true);
            if (tid == 0) {
                // left / right excitation and the front (sims.py:444-462)
                const double val2 = 1.0 - S.lastd[0] / P.vision_range;
                auto val = [&](int r) {
                    const int code = soc_bit(S.V, r, R, P.step, fov0, fov1);
                    return code == 0 ? 0.0 : (code == 2 ? val2 : 1.0);
                };
                const int h = R / 2;
                const double left = np_pairwise_sum(val, 0, h) / (double)h;
                const double right = np_pairwise_sum(val, h, R - h) / (double)(R - h);
                const double D = left > right ? 1.0 : -1.0;   // np.sign, 0 taken as -1
                int lo, hi;
                py_slice(h - 100, h + 100, R, &lo, &hi);
                const bool front = hi > lo && np_pairwise_sum(val, lo, hi - lo) / (double)(hi - lo) > 0.0;
                if (!ex_j) st.theta[g] = st.theta[g] - D * 0.2;
                if (front) st.v[g] = 0.0;
                else if (!ex_j) st.v[g] = P.exp_vel_max;
            }
            __syncthreads();
        }
    }
    // ---- collide modes ended, or notified (sims.py:779-784)
    for (int j = tid; j < N; j += nthr) {
        const int g = a0 + j;
        if (get_mode(st.overriding_mode[g], st.w[g], P.T_w) != ABM_MODE_COLLIDE) continue;
        if (!(flag[j] & 1)) {
            st.overriding_mode[g] = -1;
            st.mode[g] = ABM_MODE_EXPLORE;
        } else {
            notify(st, pool_success, P.tau, g, -1, -1);
        }
    }
    if (tid == 0) status[b] = kWorldReady;
}

// One replica per block.  STEP: agent-resource interaction (sims.py:789-855)
// then the placement of regenerated patches; FILL: create_resources; RESUME:
// the placements a previous phase left pending.  Membership is computed for
// every (patch, agent) pair in parallel from the poses at the start of the
// step; the loop over patches and their agents runs per agent in parallel,
// with only each patch's depletion chain in id order (one thread per patch).
__global__ void __launch_bounds__(1024)
forage_world_kernel(const FGRep* __restrict__ reps, const WRep* __restrict__ wreps, int phase,
                    abm_forage_state st, int32_t* __restrict__ pool_success, abm_patch_state ps,
                    int32_t* __restrict__ kill_rank, int32_t* __restrict__ pend,
                    int32_t* __restrict__ pend_id, const uint32_t* __restrict__ words,
                    const int64_t* __restrict__ w_off, const int32_t* __restrict__ w_cnt,
                    int32_t* __restrict__ w_used, int32_t* __restrict__ status,
                    abm_mt19937_state* __restrict__ mt) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) uint64_t member[];   // [patch][64-agent chunk]
    __shared__ FreshPatch fresh[kMaxPatches];
    __shared__ uint32_t mkey[kMtN];
    __shared__ int s_kind, s_pending, s_pos;
    const int b = blockIdx.x, tid = threadIdx.x, nthr = blockDim.x;
    // the replica's parameters in LDS, read where used
    __shared__ FGRep sP;
    __shared__ WRep sW;
    static_assert(sizeof(WRep) % 8 == 0 && sizeof(WRep) / 8 <= 64, "WRep copy");
    if (tid < (int)(sizeof(FGRep) / 8))
        reinterpret_cast<uint64_t*>(&sP)[tid] = reinterpret_cast<const uint64_t*>(reps + b)[tid];
    else if (tid >= 64 && tid < 64 + (int)(sizeof(WRep) / 8))
        reinterpret_cast<uint64_t*>(&sW)[tid - 64] = reinterpret_cast<const uint64_t*>(wreps + b)[tid - 64];
    __syncthreads();
    const FGRep& P = sP;
    const WRep& W = sW;
    const int N = P.n_agents, a0 = P.agent_offset, C = (N + 63) / 64;
    const int64_t po = W.patch_offset;
    int kind = pend[2 * b], n_pending = pend[2 * b + 1];
    const int n_live = ps.count[b];
    if (n_live < 0 || n_live > W.n_patches) {
        if (tid == 0) { w_used[b] = 0; status[b] = ABM_WORLD_BAD_COUNT; }
        return;
    }
    if (phase == ABM_WORLD_RESUME) {
        if (status[b] != ABM_WORLD_NEED_WORDS || kind == kPendNone) {
            if (tid == 0) w_used[b] = 0;
            return;
        }
    } else if (kind != kPendNone) {
        // placements of an earlier phase still wait for words: resume first
        if (tid == 0) { w_used[b] = 0; status[b] = ABM_WORLD_NEED_WORDS; }
        return;
    }
    if (phase == ABM_WORLD_STEP) {
        // a replica the collision kernel found crowded is left as it is
        if (W.collide && status[b] == ABM_WORLD_CROWDED) {
            if (tid == 0) w_used[b] = 0;
            return;
        }
        const int np_ = n_live;
        // the live patches' geometry, quality and units staged in LDS
        __shared__ double s_pr[kMaxPatches], s_px[kMaxPatches], s_py[kMaxPatches];
        __shared__ double s_pq[kMaxPatches], s_pl[kMaxPatches];
        __shared__ int32_t s_pid[kMaxPatches];
        for (int p = tid; p < np_; p += nthr) {
            const int64_t s = po + p;
            s_pr[p] = ps.radius[s]; s_px[p] = ps.x[s]; s_py[p] = ps.y[s];
            s_pq[p] = ps.quality[s]; s_pl[p] = ps.left[s]; s_pid[p] = ps.id[s];
        }
        __syncthreads();
        // ---- membership: supcalc.distance(resc, agent) < resc.radius.  A
        // wave takes 64 agents, loads their centres once and tests them
        // against every patch; the norm (a sqrt) only where both centre
        // differences are within the radius (with a margin far above the
        // rounding of the norm, so no pair inside is skipped)
        for (int base = tid & ~63; base < C * 64; base += nthr) {
            const int j = base + (tid & 63);
            double cx = 0.0, cy = 0.0;
            if (j < N) {
                const int g = a0 + j;
                const double ra = (double)st.radius[g];
                cx = st.x[g] + ra;
                cy = st.y[g] + ra;
            }
            for (int p = 0; p < np_; ++p) {
                const double pr = s_pr[p];
                const double dx = cx - (s_px[p] + pr), dy = cy - (s_py[p] + pr);
                const double lim = pr * 1.0000001;
                bool in = false;
                if (j < N && fabs(dx) <= lim && fabs(dy) <= lim) in = norm2(dx, dy) < pr;
                const uint64_t m = __ballot(in);
                if ((tid & 63) == 0) member[p * C + (j >> 6)] = m;
            }
        }
        __syncthreads();
        // ---- the patches' agent groups (sims.py:803-845).  The reference
        // walks the patches in Group order and their agents in id order; an
        // agent's centre lies in at most one patch (patches never overlap),
        // so only the depletion chain of each patch is sequential (patches a
        // caller loaded that overlap take the serial walk).  Phase A
        // (every thread): per agent on a patch, the bias towards the patch
        // centre and its mode and pooling flags, read before any notify of
        // the walk (a notify never changes them; a later notify of the same
        // agent only follows its own pooling test).  Phase B (one thread per
        // patch): the exploiting agents deplete the patch in id order until
        // its last unit; the agent that took it is the destroyer.  Phase C
        // (every thread): each agent's notifies in the walk's order: a
        // pooled agent up to the destroyer hears of the patch (and teleports),
        // then every agent of a destroyed patch hears -1 from the
        // destroyer's notify-all, and agents after the destroyer hear -1
        // once more in their own turn.
        __shared__ uint64_t s_on[kMaxWorldChunks], s_expl[kMaxWorldChunks], s_pool[kMaxWorldChunks];
        __shared__ int32_t s_destroyer[kMaxPatches];
        __shared__ int s_multi;
        if (tid == 0) s_multi = 0;
        __syncthreads();
        for (int c = tid; c < C; c += nthr) {
            uint64_t on = 0, dup = 0;
            for (int p = 0; p < np_; ++p) {
                dup |= on & member[p * C + c];
                on |= member[p * C + c];
            }
            s_on[c] = on;
            if (dup) s_multi = 1;
        }
        for (int p = tid; p < np_; p += nthr) s_destroyer[p] = N;
        __syncthreads();
        if (s_multi) {
            // patches loaded by the caller that overlap: the walk in the
            // reference's order, in one thread
            if (tid == 0) {
                int n_kill = 0;
                for (int p = 0; p < np_; ++p) {
                    const int64_t s = po + p;
                    const double pr = s_pr[p];
                    const double pcx = s_px[p] + pr, pcy = s_py[p] + pr;   // resc.center
                    const int32_t pid = s_pid[p];
                    bool destroy = false;
                    for (int c = 0; c < C; ++c) {
                        uint64_t m = member[p * C + c];
                        while (m) {
                            const int g = a0 + c * 64 + __builtin_ctzll(m);
                            m &= m - 1;
                            const double ra = (double)st.radius[g];
                            // bias_agent_towards_res_center (sims.py:544-552)
                            const double x1 = st.x[g] + ra, y1 = st.y[g] + ra;
                            const double th = st.theta[g];
                            const double cl = py_mod_2pi(atan2(pcy - y1, pcx - x1) + th);
                            st.theta[g] = th + (cl - kPi) * 0.02;
                            if (destroy) {
                                notify(st, pool_success, P.tau, g, -1, -1);
                                continue;
                            }
                            const int mode = get_mode(st.overriding_mode[g], st.w[g], P.T_w);
                            if (pooled(mode, pool_success[g], W.pooling_time)) {
                                notify(st, pool_success, P.tau, g, 1, pid);
                                if (W.teleport) {
                                    st.x[g] = (s_px[p] + pr) - ra;
                                    st.y[g] = (s_py[p] + pr) - ra;
                                }
                            }
                            if (mode == ABM_MODE_EXPLOIT) {
                                // Rescource.deplete (rescource.py:118-133)
                                double units = W.consumption;
                                const double q = s_pq[p];
                                if (units > q) units = q;
                                double left = s_pl[p], dep;
                                if (left >= units) { left = left - units; dep = units; }
                                else { dep = left; left = 0.0; }
                                s_pl[p] = left;
                                ps.left[s] = left;
                                destroy = !(left > 0.0);
                                const double cr = st.collected_r[g];
                                st.collected_r_before[g] = cr;
                                st.collected_r[g] = cr + dep;
                                if (destroy) {
                                    // every agent of the group hears of it
                                    for (int c2 = 0; c2 < C; ++c2) {
                                        uint64_t m2 = member[p * C + c2];
                                        while (m2) {
                                            const int g2 = a0 + c2 * 64 + __builtin_ctzll(m2);
                                            m2 &= m2 - 1;
                                            notify(st, pool_success, P.tau, g2, -1, -1);
                                        }
                                    }
                                }
                            }
                        }
                    }
                    if (destroy) {
                        kill_rank[s] = n_kill;
                        pend_id[po + n_kill] = pid;
                        ++n_kill;
                    }
                }
                kind = n_kill ? kPendKill : kPendNone;
                n_pending = n_kill;
            }
        } else {
            for (int base = tid & ~63; base < C * 64; base += nthr) {
                const int j = base + (tid & 63);
                bool ex = false, pl = false;
                if (j < N && ((s_on[j >> 6] >> (j & 63)) & 1ull)) {
                    int p = 0;
                    while (!((member[p * C + (j >> 6)] >> (j & 63)) & 1ull)) ++p;
                    const int g = a0 + j;
                    const double pr = s_pr[p];
                    const double pcx = s_px[p] + pr, pcy = s_py[p] + pr;   // resc.center
                    const double ra = (double)st.radius[g];
                    // bias_agent_towards_res_center (sims.py:544-552)
                    const double x1 = st.x[g] + ra, y1 = st.y[g] + ra;
                    const double th = st.theta[g];
                    const double cl = py_mod_2pi(atan2(pcy - y1, pcx - x1) + th);
                    st.theta[g] = th + (cl - kPi) * 0.02;
                    const int mode = get_mode(st.overriding_mode[g], st.w[g], P.T_w);
                    ex = mode == ABM_MODE_EXPLOIT;
                    pl = pooled(mode, pool_success[g], W.pooling_time);
                }
                const uint64_t me = __ballot(ex), mp = __ballot(pl);
                if ((tid & 63) == 0 && base < C * 64) { s_expl[base >> 6] = me; s_pool[base >> 6] = mp; }
            }
            __syncthreads();
            for (int p = tid; p < np_; p += nthr) {
                // Rescource.deplete (rescource.py:118-133) by the exploiting
                // agents of patch p in id order
                const int64_t s = po + p;
                double units = W.consumption;
                const double q = s_pq[p];
                if (units > q) units = q;
                double left = s_pl[p];
                bool destroy = false;
                for (int c = 0; c < C && !destroy; ++c) {
                    uint64_t m = member[p * C + c] & s_expl[c];
                    while (m && !destroy) {
                        const int j = c * 64 + __builtin_ctzll(m);
                        m &= m - 1;
                        const int g = a0 + j;
                        double dep;
                        if (left >= units) { left = left - units; dep = units; }
                        else { dep = left; left = 0.0; }
                        destroy = !(left > 0.0);
                        const double cr = st.collected_r[g];
                        st.collected_r_before[g] = cr;
                        st.collected_r[g] = cr + dep;
                        if (destroy) s_destroyer[p] = j;
                    }
                }
                s_pl[p] = left;
                ps.left[s] = left;
            }
            __syncthreads();
            for (int j = tid; j < N; j += nthr) {
                if (!((s_on[j >> 6] >> (j & 63)) & 1ull)) continue;
                int p = 0;
                while (!((member[p * C + (j >> 6)] >> (j & 63)) & 1ull)) ++p;
                const int g = a0 + j;
                const int d = s_destroyer[p];
                if (j <= d && ((s_pool[j >> 6] >> (j & 63)) & 1ull)) {
                    notify(st, pool_success, P.tau, g, 1, s_pid[p]);
                    if (W.teleport) {
                        const double pr = s_pr[p], ra = (double)st.radius[g];
                        st.x[g] = (s_px[p] + pr) - ra;
                        st.y[g] = (s_py[p] + pr) - ra;
                    }
                }
                if (d < N) notify(st, pool_success, P.tau, g, -1, -1);   // the destroyer's notify-all
                if (j > d) notify(st, pool_success, P.tau, g, -1, -1);   // its own turn after it
            }
            if (tid == 0) {
                int n_kill = 0;
                for (int p = 0; p < np_; ++p) {
                    if (s_destroyer[p] >= N) continue;
                    kill_rank[po + p] = n_kill;
                    pend_id[po + n_kill] = s_pid[p];
                    ++n_kill;
                }
                kind = n_kill ? kPendKill : kPendNone;
                n_pending = n_kill;
            }
        }
        // ---- agents on no patch (sims.py:847-855)
        for (int j = tid; j < N; j += nthr) {
            if ((s_on[j >> 6] >> (j & 63)) & 1ull) continue;
            const int g = a0 + j;
            const int mode = get_mode(st.overriding_mode[g], st.w[g], P.T_w);
            if (pooled(mode, pool_success[g], W.pooling_time) || mode == ABM_MODE_EXPLOIT)
                notify(st, pool_success, P.tau, g, -1, -1);
        }
    } else if (phase == ABM_WORLD_FILL) {
        kind = kPendFill;
        n_pending = 0;
    }
    if (mt) {
        // ---- placements from the replica's generator on the device: its
        // state in LDS for thread 0's draws, written back after
        if (tid == 0) { s_kind = kind; s_pending = n_pending; }
        __syncthreads();
        kind = s_kind;
        n_pending = s_pending;
        const bool update_follows = kind != kPendFill;
        if (kind == kPendNone) {
            if (tid == 0) {
                pend[2 * b] = kPendNone; pend[2 * b + 1] = 0;
                w_used[b] = 0;
                status[b] = kWorldReady;
            }
            return;
        }
        for (int i = tid; i < kMtN; i += nthr) mkey[i] = mt[b].key[i];
        if (tid == 0) s_pos = mt[b].pos;
        __syncthreads();
        if (tid == 0) {
            MtCursor cur{mkey, s_pos, 0, false};
            const int rc = place_patches(W, ps, b, kind, n_pending, pend_id, kill_rank, cur, fresh);
            w_used[b] = cur.used;
            if (rc == ABM_WORLD_DONE) {
                pend[2 * b] = kPendNone; pend[2 * b + 1] = 0;
                status[b] = update_follows ? kWorldReady : ABM_WORLD_DONE;
                s_pos = cur.pos;
            } else {
                // too many retries: the reference raises; nothing placed, the
                // generator stays where it was
                pend[2 * b] = kind; pend[2 * b + 1] = n_pending;
                status[b] = rc;
                s_pos = -1;
            }
        }
        __syncthreads();
        if (s_pos >= 0) {
            for (int i = tid; i < kMtN; i += nthr) mt[b].key[i] = mkey[i];
            if (tid == 0) mt[b].pos = s_pos;
        }
        return;
    }
    if (tid != 0) return;
    // ---- placements: kept words for the random walk of the update that follows
    const bool update_follows = kind != kPendFill;
    const int reserve = update_follows ? 2 * N : 0;
    const int limit = w_cnt ? max(0, w_cnt[b] - reserve) : 0;
    WordCursor cur{words ? words + w_off[b] : nullptr, 0, words ? limit : 0, false};
    int rc = ABM_WORLD_DONE;
    if (kind != kPendNone) rc = place_patches(W, ps, b, kind, n_pending, pend_id, kill_rank, cur, fresh);
    if (rc == ABM_WORLD_DONE) {
        pend[2 * b] = kPendNone; pend[2 * b + 1] = 0;
        w_used[b] = cur.used;
        status[b] = update_follows ? kWorldReady : ABM_WORLD_DONE;
    } else {
        // nothing placed and nothing drawn; the kill ranks stay for the resume
        pend[2 * b] = kind; pend[2 * b + 1] = n_pending;
        w_used[b] = rc == ABM_WORLD_TIMEOUT ? cur.used : 0;
        status[b] = rc;
    }
}

// Agent.projection_field of one focal agent: obstacles [0, n_social) are the
// social cues, [n_social, n) the non-social occluders (when n_nonsocial >= 0).
__global__ void __launch_bounds__(256)
forage_projection_kernel(int R, double step, double fov0, double fov1, int exclusion,
                         double vision_range, int keep_distance, double fx, double fy,
                         double fth, double fr, const double* __restrict__ ox,
                         const double* __restrict__ oy, int n_social, int n_nonsocial,
                         double* __restrict__ out, int maxN) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) double smem[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int N = n_social + (n_nonsocial > 0 ? n_nonsocial : 0);
    FGSmem S = fg_carve(smem, maxN, R);
    for (int i = tid; i < N; i += nthr) {
        S.x[i] = ox[i]; S.y[i] = oy[i];
        S.cls[i] = i < n_social ? kSocial : kNonExpl;
    }
    if (tid == 0) S.lastd[0] = 0.0;
    __syncthreads();
    const Focal f = make_focal(fx, fy, fth, fr);
    // the distance left in the reference's loop variable: last non-coincident obstacle
    if (tid == 0 && keep_distance) {
        for (int i = N - 1; i >= 0; --i) {
            const PairFG p = pair_forage_exact(f, S.x[i], S.y[i], R, step, fov0, fov1);
            if (!p.skipped) { S.lastd[0] = p.dist; break; }
        }
    }
    project_focal(S, N, R, step, fov0, fov1, exclusion != 0, n_nonsocial >= 0, f, keep_distance != 0);
    const double val = 1.0 - S.lastd[0] / vision_range;
    for (int r = tid; r < R; r += nthr) {
        const int b = soc_bit(S.V, r, R, step, fov0, fov1);
        out[r] = b == 0 ? 0.0 : (b == 2 ? val : 1.0);
    }
}

}  // namespace
}  // namespace abm

using namespace abm;

namespace {

constexpr int kForageMaxAgents = 65536;

// Device bytes of the sweep buffers for replicas of up to M agents: two
// tentative states (nine double and three int arrays each), the uniforms,
// the candidates and the control words.
// generator keys the prep may twist to while making 2M words
int sweep_key_slots(int M) { return (2 * M + kMtN) / kMtN + 1; }

size_t sweep_bytes(int M) {
    return 2 * (size_t)M * (9 * sizeof(double) + 3 * sizeof(int32_t)) + (size_t)M * sizeof(double) +
           (size_t)M * sizeof(int32_t) + sizeof(uint32_t) * kMtN * (size_t)sweep_key_slots(M) +
           kSwWords * sizeof(int32_t) + 256;
}

struct SweepMem {
    SweepX X[2];
    double* uni;
    int32_t* cand;
    uint32_t* keys;
    int32_t* ctrl;
};

SweepMem sweep_carve(void* base, int M) {
    SweepMem m;
    char* p = reinterpret_cast<char*>(base);
    for (int q = 0; q < 2; ++q) {
        double** d[] = {&m.X[q].x, &m.X[q].y, &m.X[q].th, &m.X[q].v, &m.X[q].w, &m.X[q].u, &m.X[q].ip,
                        &m.X[q].dth, &m.X[q].dvel};
        for (double** a : d) { *a = reinterpret_cast<double*>(p); p += sizeof(double) * (size_t)M; }
    }
    m.uni = reinterpret_cast<double*>(p); p += sizeof(double) * (size_t)M;
    for (int q = 0; q < 2; ++q) {
        int32_t** i[] = {&m.X[q].over, &m.X[q].mode, &m.X[q].kind};
        for (int32_t** a : i) { *a = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * (size_t)M; }
    }
    m.cand = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * (size_t)M;
    m.keys = reinterpret_cast<uint32_t*>(p); p += sizeof(uint32_t) * kMtN * (size_t)sweep_key_slots(M);
    m.ctrl = reinterpret_cast<int32_t*>(p);
    return m;
}

// Which replicas the step kernel updates and which are swept: sweeps for
// those beyond one workgroup's LDS, or (ABM_FORAGE_PATH_SWEEP) for every
// replica without visual exclusion.  The step kernel's blocks take their
// replicas largest first (the id-order chain grows with the agents; rays add
// to each agent's step).
int plan_paths(abm_forage_batch* bt, int path) {
    const int n = bt->n_replicas;
    std::vector<int32_t> small, big;
    for (int b = 0; b < n; ++b) {
        const bool sweepable = !bt->exclusion[b] && bt->rep_r[b] <= 32 * kSweepMaxWords;
        if (bt->too_big[b] || (path == ABM_FORAGE_PATH_SWEEP && sweepable)) big.push_back(b);
        else small.push_back(b);
    }
    auto cost = [&](int b) { return (double)bt->rep_n[b] * (64.0 + bt->rep_r[b] / 32.0); };
    std::stable_sort(small.begin(), small.end(), [&](int a, int c) { return cost(a) > cost(c); });
    bt->order_host = small;
    bt->order_host.insert(bt->order_host.end(), big.begin(), big.end());
    bt->n_small = (int32_t)small.size();
    bt->small_max_agents = 0;
    bt->small_max_rays = 0;
    for (int b : small) {
        bt->small_max_agents = std::max(bt->small_max_agents, bt->rep_n[b]);
        bt->small_max_rays = std::max(bt->small_max_rays, bt->rep_r[b]);
    }
    bt->sweep = big;
    bt->path = path;
    return check_hip(hipMemcpy(bt->order_dev, bt->order_host.data(), sizeof(int32_t) * n,
                               hipMemcpyHostToDevice), "upload forage order");
}

int check_state(const abm_forage_state* st, const char* who) {
    const void* req[] = {st->x, st->y, st->theta, st->v, st->w, st->u, st->I_priv, st->collected_r,
                         st->collected_r_before, st->novelty, st->mode, st->overriding_mode,
                         st->env_status, st->exploited_patch_id, st->radius};
    for (const void* p : req)
        if (!p) return set_error(ABM_ERR_INVALID, "%s: every state array is required", who);
    return ABM_OK;
}

// Agent.update of replica rb by sweeps (see sweep_field_kernel): the
// start, then sweeps in rounds of a few (a sweep after convergence returns
// at once) until the control words report a fixed point, then the commit.
// Reads the control words after each round, so the call synchronises the
// stream.
int sweep_replica(abm_forage_batch* b, int rb, const abm_forage_state* st, const double* uniforms,
                  const uint32_t* words, const int64_t* offset, const int32_t* count, int32_t* used,
                  int32_t* status, uint32_t* field_bits, hipStream_t s, abm_mt19937_state* mt) {
    const int N = b->rep_n[rb];
    if (N > b->sweep_max_agents || !b->sweep_mem || !b->sweep_ctrl_host)
        return set_error(ABM_ERR_UNSUPPORTED, "forage replica %d: no sweep buffers for %d agents", rb, N);
    const SweepMem m = sweep_carve(b->sweep_mem, b->sweep_max_agents);
    const FGRep* reps = (const FGRep*)b->reps_dev;
    hipLaunchKernelGGL(sweep_prep_kernel, dim3(1), dim3(1024), 0, s, reps, rb, *st, uniforms, words, offset,
                       count, used, status, mt, m.X[0], m.uni, m.keys, m.cand, m.ctrl);
    if (int rc = check_hip(hipGetLastError(), "sweep_prep_kernel launch")) return rc;
    int par = 0, sweeps = 0, round = 3;
    for (;;) {
        for (int i = 0; i < round; ++i) {
            hipLaunchKernelGGL(sweep_field_kernel, dim3((N + 3) / 4), dim3(256), 0, s, reps, rb, *st,
                               m.X[par], m.X[par ^ 1], m.cand, m.ctrl, field_bits);
            hipLaunchKernelGGL(sweep_motion_kernel, dim3(1), dim3(1024), 0, s, reps, rb, *st, m.X[par],
                               m.X[par ^ 1], m.uni, m.cand, m.ctrl);
            par ^= 1;
        }
        if (int rc = check_hip(hipGetLastError(), "sweep kernels launch")) return rc;
        sweeps += round;
        if (int rc = check_hip(hipMemcpyAsync(b->sweep_ctrl_host, m.ctrl, sizeof(int32_t) * kSwWords,
                                              hipMemcpyDeviceToHost, s), "read sweep control"))
            return rc;
        if (int rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(sweeps)")) return rc;
        if (b->sweep_ctrl_host[kSwSkip]) break;
        if (b->sweep_ctrl_host[kSwConverged]) {
            b->sweep_steps += 1;
            b->sweeps_done += b->sweep_ctrl_host[kSwIters];
            break;
        }
        if (sweeps > N + 2)   // agent k is final after k + 1 sweeps: cannot happen
            return set_error(ABM_ERR_HIP, "forage replica %d: no fixed point after %d sweeps", rb, sweeps);
        round = 2;
    }
    hipLaunchKernelGGL(sweep_commit_kernel, dim3(1), dim3(1024), 0, s, reps, rb, *st, m.X[par], used, status,
                       (words || mt) ? 1 : 0, mt, m.keys, m.ctrl);
    return check_hip(hipGetLastError(), "sweep_commit_kernel launch");
}

// Agent.update for every replica (all, or those a status array marks ready).
int launch_forage_step(abm_forage_batch* b, const abm_forage_state* st, const double* uniforms,
                       const uint32_t* words, const int64_t* offset, const int32_t* count,
                       int32_t* used, int32_t* status, uint32_t* field_bits, void* stream,
                       abm_mt19937_state* mt = nullptr) {
    // the step kernel overlaps one agent's move (wave 0) with the next
    // agent's records (wave 1) and the classes of the one after (the other
    // waves): eight waves when some replica has more agents than two waves
    // classify in one pass, else four (two workgroups then share a CU)
    if (b->n_small > 0) {
        const int nthr = b->small_max_agents > 128 ? 512 : 256;
        const size_t lds = fg_lds_bytes(b->small_max_agents, b->small_max_rays);
        const bool ex = std::any_of(b->exclusion.begin(), b->exclusion.end(), [](uint8_t e) { return e != 0; });
        auto kern = field_bits ? (ex ? forage_step_kernel<true, true> : forage_step_kernel<true, false>)
                               : (ex ? forage_step_kernel<false, true> : forage_step_kernel<false, false>);
        if (lds > 65536) {
            if (int rc = check_hip(hipFuncSetAttribute((const void*)kern,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                                   "hipFuncSetAttribute"))
                return rc;
        }
        hipLaunchKernelGGL(kern, dim3(b->n_small), dim3(nthr), lds, (hipStream_t)stream,
                           (const FGRep*)b->reps_dev, *st, uniforms, words, offset, count, used, status,
                           field_bits, (const int32_t*)b->order_dev, mt, b->small_max_agents,
                           b->small_max_rays);
        if (int rc = check_hip(hipGetLastError(), "forage_step_kernel launch")) return rc;
    }
    for (int rb : b->sweep)
        if (int rc = sweep_replica(b, rb, st, uniforms, words, offset, count, used, status, field_bits,
                                   (hipStream_t)stream, mt))
            return rc;
    return ABM_OK;
}

}  // namespace

extern "C" {

int abm_forage_batch_create(const abm_forage_replica* reps, int32_t n, abm_forage_batch** out) {
    if (!out) return set_error(ABM_ERR_INVALID, "abm_forage_batch_create: out is NULL");
    *out = nullptr;
    if (!reps || n <= 0) return set_error(ABM_ERR_INVALID, "abm_forage_batch_create: no replicas");
    std::vector<FGRep> host(n);
    int64_t agents = 0, words = 0;
    int max_n = 0, max_r = 0;
    for (int b = 0; b < n; ++b) {
        const abm_forage_replica& r = reps[b];
        if (r.n_agents < 1 || r.n_agents > kForageMaxAgents)
            return set_error(ABM_ERR_INVALID, "forage replica %d: n_agents=%d out of range [1, %d]", b,
                             r.n_agents, kForageMaxAgents);
        if (r.n_rays < 2 || r.n_rays > (1 << 16))
            return set_error(ABM_ERR_INVALID, "forage replica %d: n_rays=%d out of range", b, r.n_rays);
        if (r.tau < 1 || r.tau > 4096)
            return set_error(ABM_ERR_INVALID, "forage replica %d: tau=%d out of range", b, r.tau);
        FGRep& d = host[b];
        std::memset(&d, 0, sizeof(d));
        d.agent_offset = (int32_t)agents; d.n_agents = r.n_agents; d.n_rays = r.n_rays;
        d.words = (r.n_rays + 31) / 32; d.visual_exclusion = r.visual_exclusion ? 1 : 0;
        d.patchwise = r.patchwise_exclusion ? 1 : 0; d.tau = r.tau; d.field_offset = words;
        d.fov0 = r.fov0; d.fov1 = r.fov1; d.width = r.width; d.height = r.height;
        d.pad = r.window_pad; d.vision_range = r.vision_range; d.step = linspace_step(r.n_rays);
        d.T_w = r.T_w; d.Eps_w = r.Eps_w; d.g_w = r.g_w; d.B_w = r.B_w; d.w_max = r.w_max;
        d.T_u = r.T_u; d.Eps_u = r.Eps_u; d.g_u = r.g_u; d.B_u = r.B_u; d.u_max = r.u_max;
        d.S_wu = r.S_wu; d.S_uw = r.S_uw; d.F_N = r.F_N; d.F_R = r.F_R;
        d.exp_vel_max = r.exp_vel_max; d.exp_theta_min = r.exp_theta_min;
        d.exp_theta_max = r.exp_theta_max; d.reloc_theta_max = r.reloc_theta_max;
        d.exp_stop_ratio = r.exp_stop_ratio;
        // soc_v_field[k] survives the FOV mask unless linspace_k < fov0 or
        // > fov1 (agent.py:593-595); linspace is monotonic, so the kept
        // indices form one range
        int lo = r.n_rays, hi = -1, kept = 0;
        for (int k = 0; k < r.n_rays; ++k) {
            const double ph = linspace_at(k, r.n_rays, d.step);
            if (ph < r.fov0 || ph > r.fov1) continue;
            if (lo == r.n_rays) lo = k;
            hi = k;
            ++kept;
        }
        if (kept != (hi >= lo ? hi - lo + 1 : 0))
            return set_error(ABM_ERR_INVALID, "forage replica %d: the FOV keeps no single ray range", b);
        d.keep_lo = lo; d.keep_hi = hi;
        agents += r.n_agents;
        words += (int64_t)r.n_agents * d.words;
        max_n = std::max(max_n, r.n_agents);
        max_r = std::max(max_r, r.n_rays);
    }
    abm_forage_batch* bt = new abm_forage_batch();
    bt->n_replicas = n; bt->n_agents_total = (int32_t)agents; bt->field_words = words;
    bt->max_agents = max_n; bt->max_rays = max_r;
    int rc = check_hip(hipGetDevice(&bt->device), "hipGetDevice");
    int dev_max = 0;
    if (!rc) rc = check_hip(hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                                  bt->device), "hipDeviceGetAttribute");
    // a replica beyond one workgroup's LDS is updated by sweeps, which do
    // not implement visual exclusion and hold at most 8192 rays
    bt->too_big.assign(n, 0);
    bt->exclusion.assign(n, 0);
    for (int b = 0; !rc && b < n; ++b) {
        const size_t need = fg_lds_bytes(host[b].n_agents, host[b].n_rays) + sizeof(FGRep);
        bt->too_big[b] = need > (size_t)dev_max;
        bt->exclusion[b] = host[b].visual_exclusion != 0;
        if (bt->too_big[b] && (host[b].visual_exclusion || host[b].n_rays > 32 * kSweepMaxWords))
            rc = set_error(ABM_ERR_UNSUPPORTED, "forage replica %d: %d agents x %d rays need %zu B of LDS "
                           "(one workgroup has %d B); replicas that large are updated by sweeps, which "
                           "support neither visual exclusion nor more than %d rays", b, host[b].n_agents,
                           host[b].n_rays, need, dev_max, 32 * kSweepMaxWords);
        if (!host[b].visual_exclusion && host[b].n_rays <= 32 * kSweepMaxWords)
            bt->sweep_max_agents = std::max(bt->sweep_max_agents, host[b].n_agents);
    }
    bt->order_host.resize(n);
    bt->rep_n.resize(n);
    bt->rep_r.resize(n);
    for (int b = 0; b < n; ++b) { bt->rep_n[b] = host[b].n_agents; bt->rep_r[b] = host[b].n_rays; }
    if (!rc) rc = check_hip(hipMalloc(&bt->reps_dev, sizeof(FGRep) * n), "hipMalloc(forage replicas)");
    if (!rc) rc = check_hip(hipMemcpy(bt->reps_dev, host.data(), sizeof(FGRep) * n, hipMemcpyHostToDevice),
                           "upload forage replicas");
    if (!rc) rc = check_hip(hipMalloc(&bt->order_dev, sizeof(int32_t) * n), "hipMalloc(forage order)");
    if (!rc && bt->sweep_max_agents > 0)
        rc = check_hip(hipMalloc(&bt->sweep_mem, sweep_bytes(bt->sweep_max_agents)), "hipMalloc(sweeps)");
    if (!rc && bt->sweep_max_agents > 0)
        rc = check_hip(hipHostMalloc((void**)&bt->sweep_ctrl_host, sizeof(int32_t) * kSwWords),
                       "hipHostMalloc(sweep control)");
    if (!rc) rc = plan_paths(bt, ABM_FORAGE_PATH_AUTO);
    if (rc) { abm_forage_batch_destroy(bt); return rc; }
    bt->agent_offset_host = new int32_t[n + 1];
    bt->field_offset_host = new int64_t[n + 1];
    for (int b = 0; b < n; ++b) {
        bt->agent_offset_host[b] = host[b].agent_offset;
        bt->field_offset_host[b] = host[b].field_offset;
    }
    bt->agent_offset_host[n] = (int32_t)agents;
    bt->field_offset_host[n] = words;
    *out = bt;
    return ABM_OK;
}

int abm_forage_batch_destroy(abm_forage_batch* b) {
    if (!b) return ABM_OK;
    if (b->reps_dev) (void)hipFree(b->reps_dev);
    if (b->order_dev) (void)hipFree(b->order_dev);
    if (b->sweep_mem) (void)hipFree(b->sweep_mem);
    if (b->sweep_ctrl_host) (void)hipHostFree(b->sweep_ctrl_host);
    delete[] b->agent_offset_host;
    delete[] b->field_offset_host;
    delete b;
    return ABM_OK;
}

int abm_forage_batch_get_info(const abm_forage_batch* b, abm_batch_info* out) {
    if (!b || !out) return set_error(ABM_ERR_INVALID, "abm_forage_batch_get_info: NULL argument");
    out->n_replicas = b->n_replicas; out->n_agents_total = b->n_agents_total;
    out->field_words = b->field_words; out->max_agents = b->max_agents; out->max_rays = b->max_rays;
    return ABM_OK;
}

int abm_forage_batch_get_offsets(const abm_forage_batch* b, int32_t* agent_offset, int64_t* field_offset) {
    if (!b) return set_error(ABM_ERR_INVALID, "abm_forage_batch_get_offsets: NULL batch");
    if (agent_offset) std::copy(b->agent_offset_host, b->agent_offset_host + b->n_replicas + 1, agent_offset);
    if (field_offset) std::copy(b->field_offset_host, b->field_offset_host + b->n_replicas + 1, field_offset);
    return ABM_OK;
}


int abm_forage_sweep_counts(const abm_forage_batch* b, int64_t* updates, int64_t* sweeps) {
    if (!b || !updates || !sweeps) return set_error(ABM_ERR_INVALID, "abm_forage_sweep_counts: NULL argument");
    *updates = b->sweep_steps;
    *sweeps = b->sweeps_done;
    return ABM_OK;
}

int abm_forage_set_path(abm_forage_batch* b, int32_t path) {
    if (!b) return set_error(ABM_ERR_INVALID, "abm_forage_set_path: NULL batch");
    if (path != ABM_FORAGE_PATH_AUTO && path != ABM_FORAGE_PATH_SWEEP)
        return set_error(ABM_ERR_INVALID, "abm_forage_set_path: unknown path %d", path);
    return plan_paths(b, path);
}

int abm_forage_step_f64(abm_forage_batch* b, const abm_forage_state* st, const double* uniforms,
                        const int64_t* uniform_offset, const int32_t* uniform_count,
                        int32_t* uniforms_used, uint32_t* field_bits, void* stream) {
    if (!b || !st) return set_error(ABM_ERR_INVALID, "abm_forage_step_f64: NULL batch or state");
    if (int rc = check_state(st, "abm_forage_step_f64")) return rc;
    if (!uniforms || !uniform_offset || !uniform_count)
        return set_error(ABM_ERR_INVALID, "abm_forage_step_f64: the uniform buffer is required");
    return launch_forage_step(b, st, uniforms, nullptr, uniform_offset, uniform_count, uniforms_used,
                              nullptr, field_bits, stream);
}

int abm_forage_world_create(abm_forage_batch* b, const abm_forage_world_params* params,
                            abm_forage_world** out) {
    if (!out) return set_error(ABM_ERR_INVALID, "abm_forage_world_create: out is NULL");
    *out = nullptr;
    if (!b || !params) return set_error(ABM_ERR_INVALID, "abm_forage_world_create: NULL batch or params");
    const int n = b->n_replicas;
    std::vector<FGRep> reps(n);
    if (int rc = check_hip(hipMemcpy(reps.data(), b->reps_dev, sizeof(FGRep) * n, hipMemcpyDeviceToHost),
                           "download forage replicas"))
        return rc;
    std::vector<WRep> host(n);
    int64_t total = 0;
    size_t lds = 8, collide_lds = 0;
    const int64_t kRange = 0xFFFFFFFEll;   // randint ranges up to 2^32 - 2 values
    for (int i = 0; i < n; ++i) {
        const abm_forage_world_params& q = params[i];
        WRep& w = host[i];
        std::memset(&w, 0, sizeof(w));
        if (q.n_patches < 0 || q.n_patches > kMaxPatches)
            return set_error(ABM_ERR_INVALID, "forage world %d: n_patches=%d out of range [0, %d]", i,
                             q.n_patches, kMaxPatches);
        if (q.patch_radius < 0)
            return set_error(ABM_ERR_INVALID, "forage world %d: negative patch radius", i);
        w.n_patches = q.n_patches; w.patch_radius = q.patch_radius; w.pooling_time = q.pooling_time;
        w.regenerate = q.regenerate ? 1 : 0; w.teleport = q.teleport_exploit ? 1 : 0;
        w.collide = q.collide_agents ? 1 : 0; w.ghost = q.ghost_mode ? 1 : 0;
        if (w.collide && reps[i].n_agents > kCollideMaxAgents)
            return set_error(ABM_ERR_UNSUPPORTED, "forage world %d: agent-agent collisions for up to %d "
                             "agents per replica (got %d)", i, kCollideMaxAgents, reps[i].n_agents);
        if (w.collide)
            collide_lds = std::max(collide_lds, collide_lds_bytes(reps[i].n_agents, reps[i].n_rays));
        w.patch_offset = total;
        // the bounds numpy sees: int() of the Python expressions (sims.py:349-356)
        const double r = (double)q.patch_radius, pad = reps[i].pad;
        const double W = reps[i].width, H = reps[i].height;
        if (q.allow_border_overlap) {
            w.x_lo = (int64_t)trunc(pad - r); w.x_hi = (int64_t)trunc(W + pad - r);
            w.y_lo = (int64_t)trunc(pad - r); w.y_hi = (int64_t)trunc(H + pad - r);
        } else {
            w.x_lo = (int64_t)trunc(pad); w.x_hi = (int64_t)trunc(W + pad - 2.0 * r);
            w.y_lo = (int64_t)trunc(pad); w.y_hi = (int64_t)trunc(H + pad - 2.0 * r);
        }
        w.units_lo = q.min_units; w.units_hi = q.max_units;
        const int64_t lo[] = {w.x_lo, w.y_lo, w.units_lo}, hi[] = {w.x_hi, w.y_hi, w.units_hi};
        for (int k = 0; k < 3; ++k)
            if (q.n_patches > 0 && (hi[k] <= lo[k] || hi[k] - 1 - lo[k] > kRange))
                return set_error(ABM_ERR_INVALID, "forage world %d: randint(%lld, %lld) has no values "
                                 "(numpy raises low >= high) or too many", i, (long long)lo[k],
                                 (long long)hi[k]);
        w.consumption = q.consumption; w.q_lo = q.min_quality; w.q_hi = q.max_quality;
        if (!std::isfinite(q.max_quality - q.min_quality))
            return set_error(ABM_ERR_INVALID, "forage world %d: quality range is not finite", i);
        total += q.n_patches;
        lds = std::max(lds, sizeof(uint64_t) * (size_t)q.n_patches * (size_t)((reps[i].n_agents + 63) / 64));
    }
    abm_forage_world* wd = new abm_forage_world();
    wd->batch = b; wd->n_replicas = n; wd->patch_total = total; wd->member_lds = lds;
    wd->collide_lds = collide_lds;
    if (collide_lds > (size_t)163840) {
        delete wd;
        return set_error(ABM_ERR_UNSUPPORTED, "forage world: the collision kernel needs %zu B of LDS",
                         collide_lds);
    }
    const size_t slots = (size_t)std::max<int64_t>(total, 1);
    int rc = check_hip(hipMalloc(&wd->wreps_dev, sizeof(WRep) * n), "hipMalloc(world replicas)");
    if (!rc) rc = check_hip(hipMalloc(&wd->kill_rank, sizeof(int32_t) * slots), "hipMalloc(kill ranks)");
    if (!rc) rc = check_hip(hipMalloc(&wd->pend, sizeof(int32_t) * 2 * n), "hipMalloc(pending)");
    if (!rc) rc = check_hip(hipMalloc(&wd->pend_id, sizeof(int32_t) * slots), "hipMalloc(pending ids)");
    if (!rc) rc = check_hip(hipMemcpy(wd->wreps_dev, host.data(), sizeof(WRep) * n, hipMemcpyHostToDevice),
                           "upload world replicas");
    if (!rc) rc = check_hip(hipMemset(wd->kill_rank, 0xff, sizeof(int32_t) * slots), "clear kill ranks");
    if (!rc) rc = check_hip(hipMemset(wd->pend, 0, sizeof(int32_t) * 2 * n), "clear pending");
    if (!rc) rc = check_hip(hipDeviceSynchronize(), "world create");
    if (rc) { abm_forage_world_destroy(wd); return rc; }
    wd->patch_offset_host = new int64_t[n + 1];
    for (int i = 0; i < n; ++i) wd->patch_offset_host[i] = host[i].patch_offset;
    wd->patch_offset_host[n] = total;
    *out = wd;
    return ABM_OK;
}

int abm_forage_world_destroy(abm_forage_world* w) {
    if (!w) return ABM_OK;
    if (w->wreps_dev) (void)hipFree(w->wreps_dev);
    if (w->kill_rank) (void)hipFree(w->kill_rank);
    if (w->pend) (void)hipFree(w->pend);
    if (w->pend_id) (void)hipFree(w->pend_id);
    delete[] w->patch_offset_host;
    delete w;
    return ABM_OK;
}

int abm_forage_world_get_offsets(const abm_forage_world* w, int64_t* patch_offset) {
    if (!w || !patch_offset) return set_error(ABM_ERR_INVALID, "abm_forage_world_get_offsets: NULL argument");
    std::copy(w->patch_offset_host, w->patch_offset_host + w->n_replicas + 1, patch_offset);
    return ABM_OK;
}

}  // extern "C"

namespace {

// One phase of the resource loop; the placements and the walk draw from a
// word buffer (words) or from device generators (mt).
int world_run(abm_forage_world* w, int32_t phase, const abm_forage_state* st, int32_t* pool_success,
              const abm_patch_state* ps, const uint32_t* words, const int64_t* word_offset,
              const int32_t* word_count, int32_t* words_used, int32_t* status, uint32_t* field_bits,
              void* stream, abm_mt19937_state* mt, const char* who) {
    if (!w || !st || !ps) return set_error(ABM_ERR_INVALID, "%s: NULL world, state or patches", who);
    if (int rc = check_state(st, who)) return rc;
    if (!pool_success || !words_used || !status)
        return set_error(ABM_ERR_INVALID, "%s: pool_success, words_used and status are required", who);
    if (w->patch_total > 0 && (!ps->count || !ps->id || !ps->x || !ps->y || !ps->radius || !ps->units ||
                               !ps->left || !ps->quality))
        return set_error(ABM_ERR_INVALID, "%s: every patch array is required", who);
    if (!ps->count) return set_error(ABM_ERR_INVALID, "%s: patch count array is required", who);
    abm_forage_batch* b = w->batch;
    const size_t lds = w->member_lds;
    if (lds > 65536) return set_error(ABM_ERR_UNSUPPORTED, "%s: %zu B of membership masks", who, lds);
    if (phase == ABM_WORLD_STEP && w->collide_lds > 0) {
        // agent-agent collisions first (sims.py:736-787)
        if (w->collide_lds > 65536) {
            if (int rc = check_hip(hipFuncSetAttribute((const void*)forage_collide_kernel,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)w->collide_lds), "hipFuncSetAttribute"))
                return rc;
        }
        hipLaunchKernelGGL(forage_collide_kernel, dim3(b->n_replicas), dim3(256), w->collide_lds,
                           (hipStream_t)stream, (const FGRep*)b->reps_dev, (const WRep*)w->wreps_dev, *st,
                           pool_success, w->pend, status);
        if (int rc = check_hip(hipGetLastError(), "forage_collide_kernel launch")) return rc;
    }
    // a replica above 1024 agents takes a block of 1024 threads for the
    // per-agent parts; smaller ones share a CU in blocks of 256
    const int threads = b->max_agents > 1024 ? 1024 : 256;
    hipLaunchKernelGGL(forage_world_kernel, dim3(b->n_replicas), dim3(threads), lds, (hipStream_t)stream,
                       (const FGRep*)b->reps_dev, (const WRep*)w->wreps_dev, (int)phase, *st, pool_success,
                       *ps, w->kill_rank, w->pend, w->pend_id, words, word_offset, word_count, words_used,
                       status, mt);
    if (int rc = check_hip(hipGetLastError(), "forage_world_kernel launch")) return rc;
    if (phase == ABM_WORLD_FILL) return ABM_OK;
    // the update draws its walk after the placements
    return launch_forage_step(b, st, nullptr, words, word_offset, word_count, words_used, status,
                              field_bits, stream, mt);
}

}  // namespace

extern "C" {

int abm_forage_world_run(abm_forage_world* w, int32_t phase, const abm_forage_state* st,
                         int32_t* pool_success, const abm_patch_state* ps, const uint32_t* words,
                         const int64_t* word_offset, const int32_t* word_count, int32_t* words_used,
                         int32_t* status, uint32_t* field_bits, void* stream) {
    if (phase != ABM_WORLD_STEP && phase != ABM_WORLD_RESUME && phase != ABM_WORLD_FILL)
        return set_error(ABM_ERR_INVALID, "abm_forage_world_run: unknown phase %d", phase);
    if (!words || !word_offset || !word_count)
        return set_error(ABM_ERR_INVALID, "abm_forage_world_run: the word buffer is required");
    return world_run(w, phase, st, pool_success, ps, words, word_offset, word_count, words_used, status,
                     field_bits, stream, nullptr, "abm_forage_world_run");
}

int abm_forage_world_run_dev(abm_forage_world* w, int32_t phase, const abm_forage_state* st,
                             int32_t* pool_success, const abm_patch_state* ps, abm_mt19937_state* mt,
                             int32_t* words_used, int32_t* status, uint32_t* field_bits, void* stream) {
    if (phase != ABM_WORLD_STEP && phase != ABM_WORLD_FILL)
        return set_error(ABM_ERR_INVALID, "abm_forage_world_run_dev: phase %d (STEP or FILL)", phase);
    if (!mt) return set_error(ABM_ERR_INVALID, "abm_forage_world_run_dev: NULL generator states");
    return world_run(w, phase, st, pool_success, ps, nullptr, nullptr, nullptr, words_used, status,
                     field_bits, stream, mt, "abm_forage_world_run_dev");
}

int abm_forage_projection(int32_t R, double fov0, double fov1, int32_t visual_exclusion,
                          double vision_range, int32_t keep_distance, double fx, double fy,
                          double ftheta, double fradius, const double* ox, const double* oy,
                          int32_t n_social, int32_t n_nonsocial, double* out_field, void* stream) {
    if (R < 2 || R > (1 << 16)) return set_error(ABM_ERR_INVALID, "abm_forage_projection: n_rays=%d", R);
    if (n_social < 0) return set_error(ABM_ERR_INVALID, "abm_forage_projection: n_social < 0");
    const int n = n_social + (n_nonsocial > 0 ? n_nonsocial : 0);
    if (n > 4096) return set_error(ABM_ERR_INVALID, "abm_forage_projection: more than 4096 obstacles");
    if (!out_field || (n > 0 && (!ox || !oy)))
        return set_error(ABM_ERR_INVALID, "abm_forage_projection: NULL array");
    const int maxN = n > 0 ? n : 1;
    const size_t lds = fg_lds_bytes(maxN, R);
    if (lds > 65536) {
        if (int rc = check_hip(hipFuncSetAttribute((const void*)forage_projection_kernel,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                               "hipFuncSetAttribute"))
            return rc;
    }
    hipLaunchKernelGGL(forage_projection_kernel, dim3(1), dim3(256), lds, (hipStream_t)stream, R,
                       linspace_step(R), fov0, fov1, visual_exclusion, vision_range, keep_distance,
                       fx, fy, ftheta, fradius, ox, oy, n_social, n_nonsocial, out_field, maxN);
    return check_hip(hipGetLastError(), "forage_projection_kernel launch");
}

}  // extern "C"
