// vf_group.hip -- fused float32 VSWRM step with G replicas per wave.
//
// Same step as vf_f32.hip (reference order, vf_sims.py:302; float32 pair
// decision with float64 re-decision inside the guard bands; bit-packed raster;
// float32 sums and motion), laid out so that one wave advances G replicas in
// lockstep: agent k of each of the G replicas is the focal agent of the same
// loop iteration.  Lanes are split into G groups of L = 64 / G lanes, one
// group per replica, and every stage works within a group:
//   pairs   -- lane li of group g decides pairs (k, j) for j = li, li + L, ...
//   sums    -- lane li owns rays [li * RPL, (li + 1) * RPL) of its replica's
//              field; the four sums come from nibble tables: for every 4-ray
//              position q and every 4-bit pattern v the table holds the sum of
//              (trapz weight * cos, * sin) over the set rays, and a second table
//              the sum of (cos, sin), so 4 rays cost two LDS reads and four adds;
//   reduce  -- DPP sums within the group (a DPP row when L = 16);
//   motion  -- computed in every lane of the group; the group's last lane
//              writes the pose, its second-to-last lane computes the heading
//              of the next focal agent with the same sincos instruction.
// The waves of a workgroup run independent groups and share only the nibble
// tables, so the loop has no workgroup barrier.  A group holds replicas that
// agree in every constant except N and the ALP0 / BET0 / V0 defaults; all
// groups of one launch share R (one table).
#include "vf_group_common.hpp"

namespace abm {

namespace {
using namespace f32;

struct GroupWaveSmem {
    float4* pose;      // [G][maxN] live (x, y, radius, theta)
    float4* obj;       // [G][maxN] poses the pairs read (= pose, or a pre-step copy)
    float4* dyn;       // [G][maxN] (v, ALP0, BET0, V0)
    uint32_t* F;       // [G][W + 1] fields (one spare word for the two-word raster)
    int32_t* rlist;    // [G * maxN] pairs for the float64 decision
    int32_t* rfast;    // [G * maxN] their float32 decision (verify mode)
};

__host__ __device__ inline size_t group_wave_bytes(int G, int maxN, int W, bool sync_order) {
    size_t b = sizeof(float4) * (size_t)G * maxN * (sync_order ? 2 : 1) +
               sizeof(float4) * (size_t)G * maxN + sizeof(uint32_t) * (size_t)G * (W + 1) +
               sizeof(int32_t) * (size_t)G * maxN * 2;
    return (b + 15) & ~(size_t)15;
}

__device__ inline GroupWaveSmem group_wave_carve(char* p, int G, int maxN, int W, bool sync_order) {
    GroupWaveSmem s;
    s.pose = reinterpret_cast<float4*>(p); p += sizeof(float4) * (size_t)G * maxN;
    if (sync_order) { s.obj = reinterpret_cast<float4*>(p); p += sizeof(float4) * (size_t)G * maxN; }
    else s.obj = s.pose;
    s.dyn = reinterpret_cast<float4*>(p); p += sizeof(float4) * (size_t)G * maxN;
    s.F = reinterpret_cast<uint32_t*>(p); p += sizeof(uint32_t) * (size_t)G * (W + 1);
    s.rlist = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * (size_t)G * maxN;
    s.rfast = reinterpret_cast<int32_t*>(p);
    return s;
}

// Raster of one float32-decided pair into F, predicated on `draw`.  Outside
// the guard bands the half-width is below R/4, so at most one of start < 0 and
// end >= R holds and the wrap rules of vf_supcalc.py:122-129 reduce to two
// pieces without clamps: [max(start, 0), min(end, R)) and the wrapped piece
// [R + start, R) or [0, end - (R - 1)).  Only lanes with bits to set take
// part in an LDS OR (lanes of the same word serialise on its address, and
// idle lanes would add to that); second words, wrapped pieces and pieces of
// 32 rays or more take wave-uniform branches.
__device__ inline void raster_pred(uint32_t* F, int start, int end, int R, bool draw) {
    const int p0 = start > 0 ? start : 0;
    int n0 = (end < R ? end : R) - p0;
    n0 = draw ? n0 : 0;
    const bool long0 = n0 > 31;
    uint32_t lo, hi;
    span_masks(p0 & 31, long0 ? 0 : n0, &lo, &hi);
    if (lo != 0u) atomicOr(&F[p0 >> 5], lo);
    if (hi != 0u) atomicOr(&F[(p0 >> 5) + 1], hi);
    // the wrapped piece, taken only when some lane has one
    const bool neg = start < 0, over = end >= R;
    if (__ballot(draw & (neg | over))) {
        const int q0 = neg ? R + start : 0;
        int n1 = neg ? -start : (over ? end - (R - 1) : 0);
        n1 = draw ? n1 : 0;
        const bool long1 = n1 > 31;
        uint32_t lo1, hi1;
        span_masks(q0 & 31, long1 ? 0 : n1, &lo1, &hi1);
        if (lo1 != 0u) atomicOr(&F[q0 >> 5], lo1);
        if (hi1 != 0u) atomicOr(&F[(q0 >> 5) + 1], hi1);
        if (long1) or_range(F, q0, q0 + n1);
    }
    if (__ballot(long0))
        if (long0) or_range(F, p0, p0 + n0);
}

// PLAIN: a step with no fields, diagnostics or counters out (the flags and
// pointers that ask for them are ignored), so their tests and addresses leave
// the hot loop.
template <int G, int RPL, bool KH, bool PLAIN>
__global__ void __launch_bounds__(64 * (G == 1 ? 8 : 16 / G))
__attribute__((amdgpu_waves_per_eu(G == 1 ? 4 : 1)))
vf_group_kernel(const VFRep* __restrict__ reps, const int32_t* __restrict__ groups, int n_groups,
                const float2* __restrict__ nib, abm_vf_state_f32 st, uint32_t* __restrict__ bits_out,
                float* __restrict__ diag, uint32_t flags, unsigned long long* __restrict__ counters,
                int maxN, int W) {
    constexpr int L = 64 / G;
    constexpr int NQ = RPL / 4;                 // nibbles per lane
    constexpr int TAB = 4 * RPL * L;            // float2 entries per table (blob, then spike)
    extern __shared__ __align__(16) double smem_d[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
    const int nwg = blockDim.x >> 6;
    const int gl = lane / L, li = lane % L;
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;
    const bool want_diag = !PLAIN && (flags & ABM_WANT_DIAG) != 0 && diag != nullptr;
    const bool count = !PLAIN && (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) != 0 && counters;
    const bool verify = !PLAIN && (flags & ABM_VERIFY_FAST_PATH) != 0 && counters != nullptr;

    float2* tabB = reinterpret_cast<float2*>(smem_d);
    float2* tabS = tabB + TAB;
    {
        const float4* src = reinterpret_cast<const float4*>(nib);
        float4* dst = reinterpret_cast<float4*>(tabB);
        for (int e = tid; e < TAB; e += blockDim.x) dst[e] = src[e];   // 2 * TAB float2
    }
    char* wbase = reinterpret_cast<char*>(tabS + TAB) + (size_t)wave * group_wave_bytes(G, maxN, W, sync_order);
    const GroupWaveSmem S = group_wave_carve(wbase, G, maxN, W, sync_order);
    uint32_t* Fg = S.F + gl * (W + 1);
    float4* poseg = S.pose + gl * maxN;
    float4* objg = S.obj + gl * maxN;
    float4* dyng = S.dyn + gl * maxN;
    const float2* tb = tabB + li;
    const float2* ts = tabS + li;
    for (int w = lane; w < G * (W + 1); w += 64) S.F[w] = 0u;
    __syncthreads();

    unsigned long long n_pairs = 0;   // wave total
    uint32_t n_recheck = 0, n_disagree = 0;   // this lane's
    for (int gi = blockIdx.x * nwg + wave; gi < n_groups; gi += gridDim.x * nwg) {
        const int* grp = groups + (size_t)gi * G;
        const VFRep* rp0 = reps + grp[0];
        // this lane's replica
        const int rid = grp[gl];
        // The group's replicas share these constants.  Loaded through the lane's
        // own replica they sit in VGPRs (no SGPR pressure, no constant-bus
        // limit); the ones that steer branches are taken wave-uniform.
        RepConst C = rep_const(reps + (rid >= 0 ? rid : grp[0]));
        C.R = rp0->n_rays;
        C.W = rp0->words;
        C.torus = rp0->boundary == ABM_BOUNDARY_INFINITE;
        C.limit = rp0->limit_movement;
        C.phi_ok = rp0->phi_ok;
        const int myN = rid >= 0 ? reps[rid].n_agents : 0;
        const int a0 = rid >= 0 ? reps[rid].agent_offset : 0;
        const int64_t fo = rid >= 0 ? reps[rid].field_offset : 0;
        int nmax = 0;
        for (int g = 0; g < G; ++g) {
            const int r = grp[g];
            if (r < 0) continue;
            const VFRep* rg = reps + r;
            const int n = rg->n_agents, b0 = rg->agent_offset;
            const float dalp0 = (float)rg->alp0, dbet0 = (float)rg->bet0, dv0 = (float)rg->v0;
            nmax = n > nmax ? n : nmax;
            for (int i = lane; i < n; i += 64) {
                const int gg = b0 + i;
                const float4 p = make_float4(st.x[gg], st.y[gg], (float)st.radius[gg], st.theta[gg]);
                S.pose[g * maxN + i] = p;
                if (sync_order) S.obj[g * maxN + i] = p;
                S.dyn[g * maxN + i] = make_float4(st.v[gg], st.alp0 ? st.alp0[gg] : dalp0,
                                                  st.bet0 ? st.bet0[gg] : dbet0,
                                                  st.v0 ? st.v0[gg] : dv0);
            }
        }
        wave_sync();
        const int passes = (nmax + L - 1) / L;
        // focal agent 0 of this lane's replica
        float4 fp = poseg[0];
        float4 fd = dyng[0];
        float sth, cth;
        sincos_f32(fp.w, &sth, &cth);

        // the focal agents in four priority stages (no per-agent test)
        const ProgressStages stages = progress_stages(nmax);
        int k = 0;
        for (int stage = 0; stage < 4; ++stage) {
            const int kend = stage_priority(stage, stages, nmax);
            for (; k < kend; ++k) {
                const bool kact = k < myN;
                const float fx = fp.x, fy = fp.y, fr = fp.z, fth = fp.w;
                const float u1x = cth, u1y = -sth;

                // ------------------------------------------------ pairs
                // every lane evaluates a pair (inactive lanes a valid dummy) and
                // only the raster and the counters are predicated
                int nre = 0;   // wave-uniform
                for (int t = 0; t < passes; ++t) {
                    const int j = li + t * L;
                    const bool act = kact && j < myN && j != k;
                    const float4 o = objg[act ? j : 0];
                    PairF32 p = fast_pair(u1x, u1y, fx, fy, fr, o.x, o.y, o.z, C);
                    const bool live = act && !(o.x == fx && o.y == fy);     // coincident (:57)
                    if (count) n_pairs += __popcll(__ballot(live));         // wave total
                    // pairs outside the float32 decision (rare), every pair in verify mode
                    if (__ballot(live && !p.ok) || verify) {
                        if (live && p.tie && tie_in_place(p, fx, fy, fth, fr, o.x, o.y, o.z, reps + rid, C))
                            if (count) ++n_recheck;   // a float64 decision too
                        const bool flag = live && (!p.ok || verify);
                        const uint64_t fm = __ballot(flag);
                        if (flag) {
                            const int slot = nre + (int)__popcll(fm & ((1ull << lane) - 1ull));
                            S.rlist[slot] = entry_code(j, gl, p.ok);
                            S.rfast[slot] = fast_code(p.start, p.end, p.drawn, C.R);
                        }
                        nre += (int)__popcll(fm);
                    }
                    raster_pred(Fg, p.start, p.end, C.R, live && p.ok && p.drawn);
                }
                wave_sync();
                if (nre > 0) {
                    for (int i = lane; i < nre; i += 64) {
                        const int e = S.rlist[i];
                        const int fc = S.rfast[i];
                        const int j = e & 0xfffff;
                        const int g = (e >> 20) & 0xff;
                        const float4 f = S.pose[g * maxN + k];
                        const float4 o = S.obj[g * maxN + j];
                        const VFRep* rpg = reps + grp[g];
                        int dis;
                        const int r = resolve_queued(e, fc, f.x, f.y, f.w, f.z, o.x, o.y, o.z, rpg, verify, &dis);
                        n_disagree += dis;
                        if (e & kEntryForced) {
                            if (count) ++n_recheck;
                            if (r >= 0)
                                raster(S.F + g * (W + 1), (r & 0x7fff) - 2 * C.R, ((r >> 15) & 0x7fff) - 2 * C.R,
                                       C.R);
                        }
                    }
                    wave_sync();
                }

                // ------------------------------------------------ sums
                const int kn = k + 1 < nmax ? k + 1 : k;      // next focal agent
                const float4 fpn = poseg[kn];
                const float4 fdn = dyng[kn];
                uint64_t v, sp;
                span_bits<G, RPL>(Fg, li, C.R, &v, &sp);
                if (!PLAIN && bits_out && kact) {
                    uint32_t* dst = bits_out + fo + (int64_t)k * C.W;
                    for (int w = li; w < C.W; w += L) dst[w] = Fg[w];
                }
                wave_sync();
                if (W + 1 <= L) {   // wave-uniform
                    if (li < W + 1) Fg[li] = 0u;
                } else {
                    for (int w = li; w < W + 1; w += L) Fg[w] = 0u;
                }
                float bc, bs, ec, es;
                float Sv, Sp, Bc, Bs, Ec, Es;   // group totals in diag mode
                if (KH) {
                    // Kahan-compensated ray sums over the lane's nibble entries,
                    // then the group's sums of the compensated values in float64
                    float ks[4], kc[4];
                    span_sums_kahan<RPL, L>(tb, ts, v, sp, ks, kc);
                    const double dBc = group_sum_f64<G>((double)ks[0] - (double)kc[0]);
                    const double dBs = group_sum_f64<G>((double)ks[1] - (double)kc[1]);
                    const double dEc = group_sum_f64<G>((double)ks[2] - (double)kc[2]);
                    const double dEs = group_sum_f64<G>((double)ks[3] - (double)kc[3]);
                    Sv = (float)((double)fd.y * -dBc + (double)(fd.y * C.alp1) * dEc);
                    Sp = (float)((double)fd.z * -dBs + (double)(fd.z * C.bet1) * dEs);
                    Bc = (float)dBc; Bs = (float)dBs; Ec = (float)dEc; Es = (float)dEs;
                } else {
                span_sums<RPL, L>(tb, ts, v, sp, &bc, &bs, &ec, &es);
                Bc = bc; Bs = bs; Ec = ec; Es = es;
                if (!want_diag) {
                    Sv = fd.y * (-bc) + (fd.y * C.alp1) * ec;
                    Sp = fd.z * (-bs) + (fd.z * C.bet1) * es;
                    group_sum2<G>(Sv, Sp);
                } else {
                    group_sum2<G>(Bc, Bs);
                    group_sum2<G>(Ec, Es);
                    Sv = fd.y * -Bc + (fd.y * C.alp1) * Ec;
                    Sp = fd.z * -Bs + (fd.z * C.bet1) * Es;
                }
                }

                // ------------------------------------------------ motion
                float vel, o, dv, dpsi;
                steer(C, fth, fd.x, fd.w, Sv, Sp, &o, &vel, &dv, &dpsi);
                const float ang = li == L - 2 ? fpn.w : o;
                float s_l, c_l;
                sincos_f32(ang, &s_l, &c_l);
                float nx, ny;
                euler_boundary(C, fx, fy, fr, vel, s_l, c_l, &o, &nx, &ny);
                if (li == L - 1 && kact) {
                    poseg[k] = make_float4(nx, ny, fr, o);
                    dyng[k].x = vel;
                    if (want_diag) {
                        float* dd = diag + 6 * (int64_t)(a0 + k);
                        dd[0] = dv; dd[1] = dpsi;
                        dd[2] = C.phi_ok ? fd.y * -Bc : 0.f;
                        dd[3] = C.phi_ok ? (fd.y * C.alp1) * Ec : 0.f;
                        dd[4] = C.phi_ok ? fd.z * -Bs : 0.f;
                        dd[5] = C.phi_ok ? (fd.z * C.bet1) * Es : 0.f;
                    }
                }
                sth = group_bcast_second_last<G>(s_l, gl);
                cth = group_bcast_second_last<G>(c_l, gl);
                fp = fpn;
                fd = fdn;
                wave_sync();
            }
        }
        for (int i = li; i < myN; i += L) {
            const int g = a0 + i;
            const float4 p = poseg[i];
            st.x[g] = p.x;
            st.y[g] = p.y;
            st.theta[g] = p.w;
            st.v[g] = dyng[i].x;
        }
        wave_sync();
    }
    if (count) {
        if (n_pairs && lane == 0) atomicAdd(&counters[0], n_pairs);   // n_pairs is a wave total
        if (n_recheck) atomicAdd(&counters[1], (unsigned long long)n_recheck);
        if (n_disagree) atomicAdd(&counters[2], (unsigned long long)n_disagree);
    }
}

template <int G, int RPL, bool KH, bool PLAIN>
int launch_group_shape(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                       uint32_t flags, hipStream_t s) {
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;
    const GroupPlan& gp = b->gplan[G == 1 ? 0 : (G == 2 ? 1 : 2)];
    const int W = (b->max_rays + 31) / 32;
    const size_t tab = sizeof(float2) * 2 * (size_t)gp.tab_entries;
    const size_t per_wave = group_wave_bytes(G, b->max_agents, W, sync_order);
    int nwg = G == 1 ? 8 : 16 / G;
    // enough workgroups to spread the groups over the CUs, within the LDS limit
    while (nwg > 1 && ((long)(gp.n_groups + nwg - 1) / nwg < b->n_cu ||
                       tab + per_wave * nwg > (size_t)b->dev_max_lds))
        nwg /= 2;
    const size_t lds = tab + per_wave * nwg;
    if (lds > (size_t)b->dev_max_lds)
        return set_error(ABM_ERR_UNSUPPORTED, "abm_vf_step_f32 (grouped): %zu B of LDS needed (device max %d)",
                         lds, b->dev_max_lds);
    auto kern = vf_group_kernel<G, RPL, KH, PLAIN>;
    if (lds > 65536) {
        if (int rc = check_hip(hipFuncSetAttribute((const void*)kern,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                               "hipFuncSetAttribute"))
            return rc;
    }
    int occ = 0;
    if (int rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64 * nwg, lds),
                           "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return rc;
    if (occ < 1) occ = 1;
    long grid = (gp.n_groups + nwg - 1) / nwg;
    if (grid > (long)occ * b->n_cu) grid = (long)occ * b->n_cu;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * nwg), lds, s, b->reps_dev, gp.groups,
                       gp.n_groups, gp.nib, *st, bits, diag, flags, b->counters, b->max_agents, W);
    b->last_kernels |= ABM_KERNEL_GROUPED;
    return check_hip(hipGetLastError(), "vf_group_kernel launch");
}

}  // namespace

// Rays per lane for G replicas per wave: a power of two, at least 4.
int group_rpl(int G, int R) {
    const int L = 64 / G;
    int rpl = 4;
    while (rpl * L < R) rpl *= 2;
    return rpl;
}

int launch_vf_group_f32(abm_vf_batch* b, int G, const abm_vf_state_f32* st, uint32_t* bits,
                        float* diag, uint32_t flags, hipStream_t s) {
    const int rpl = group_rpl(G, b->max_rays);
    const bool kahan = (flags & ABM_KAHAN_SUMS) != 0;
    const bool plain = bits == nullptr &&
                       (flags & (ABM_WANT_DIAG | ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) == 0;
#define ABM_GLAUNCH(GV, RV) \
    if (G == GV && rpl == RV) \
        return kahan ? launch_group_shape<GV, RV, true, false>(b, st, bits, diag, flags, s) \
                     : (plain ? launch_group_shape<GV, RV, false, true>(b, st, bits, diag, flags, s) \
                              : launch_group_shape<GV, RV, false, false>(b, st, bits, diag, flags, s));
    ABM_GLAUNCH(1, 4) ABM_GLAUNCH(1, 8) ABM_GLAUNCH(1, 16)
    ABM_GLAUNCH(2, 4) ABM_GLAUNCH(2, 8) ABM_GLAUNCH(2, 16) ABM_GLAUNCH(2, 32)
    ABM_GLAUNCH(4, 4) ABM_GLAUNCH(4, 8) ABM_GLAUNCH(4, 16) ABM_GLAUNCH(4, 32) ABM_GLAUNCH(4, 64)
#undef ABM_GLAUNCH
    return set_error(ABM_ERR_UNSUPPORTED, "abm_vf_step_f32 (grouped): no shape for %d replicas per wave "
                     "and %d rays", G, b->max_rays);
}

}  // namespace abm
