// vf_f32.hip -- fused float32 VSWRM step for batches of replicas.
//
// Persistent workgroups of NW waves; each takes replicas (largest first) and
// runs the reference-order step (agent k sees agents <k moved,
// vf_sims.py:302) with the replica's poses and its float32 ray table in LDS.
//
// Per focal agent k:
//   pairs   -- lanes over the other agents j.  The centre differences, the
//              periodic image, atan2 for the closed angle, atan for the
//              half-width and the bin arithmetic are float32 (fast_pair in
//              vf_f32_common.hpp, within 1e-6 rad of the exact angle).  When the
//              float32 angle lies within a guard band of a bin tie or of the
//              +-pi wrap, the half-width within its band of an integer, or a
//              difference within its band of the half-arena image switch, the
//              pair is re-decided with the exact float64 code of vf_exact.hpp
//              (vf_supcalc.py:55-119), so the raster equals the float64 one.
//   raster  -- each drawn interval (wrap rules of vf_supcalc.py:122-129) is
//              OR-ed into bit-packed uint32 words in LDS.
//   sums    -- lanes over contiguous ray spans: V bits, edges by shift-XOR
//              with the dPhi alignment rule (vf_supcalc.py:266-274), and the
//              four sums over float32 tables (trapz weight * cos/sin, cos, sin);
//              float32 per-lane partials, float64 cross-lane tree.
//   motion  -- float32 VSWRM update, Euler step, boundary (vf_agent.py:188-204,
//              263-326; agent.py:347-395, 605-610).
#include <cmath>

#include "vf_f32_common.hpp"

namespace abm {

namespace {
using namespace f32;

// Replicas with more agents than this take the wide (synchronous) kernel.
constexpr int kWideThreshold = 2048;

// LDS layout: per-agent records as float4 so one 16-byte read serves a pair.
struct Smem {
    RayTab32* tab;     // transposed ray table: entry i*NT + lane holds ray lane*RPL + i
    float4* pose;      // live (x, y, radius, theta) of every agent
    float4* obj;       // positions the pairs read: = pose, or a pre-step copy (synchronous)
    float4* dyn;       // (v, ALP0, BET0, V0)
    int32_t* rlist;    // pairs queued for the float64 decision
    int32_t* rfast;    // their float32 decision (verify mode)
    uint32_t* F;       // the focal agent's bit-packed field
    double* red;       // cross-wave partial sums
    int32_t* nre;
};

__device__ inline Smem carve(void* base, int maxN, int tab_entries, int nw, bool sync_order) {
    Smem s;
    char* p = reinterpret_cast<char*>(base);
    s.tab = reinterpret_cast<RayTab32*>(p); p += sizeof(RayTab32) * tab_entries;
    s.pose = reinterpret_cast<float4*>(p); p += sizeof(float4) * maxN;
    s.dyn = reinterpret_cast<float4*>(p); p += sizeof(float4) * maxN;
    if (sync_order) { s.obj = reinterpret_cast<float4*>(p); p += sizeof(float4) * maxN; }
    else s.obj = s.pose;
    s.red = reinterpret_cast<double*>(p); p += sizeof(double) * (2 * nw + 2);
    s.rlist = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * maxN;
    s.rfast = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * maxN;
    s.nre = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * 4;
    s.F = reinterpret_cast<uint32_t*>(p);
    return s;
}

template <int NW, int RPL, bool PLAIN>   // PLAIN: as the push kernel's
__global__ void __launch_bounds__(64 * NW)
vf_step_f32_kernel(const VFRep* __restrict__ reps, const int32_t* __restrict__ order, int n_rep,
                   const RayTab32* __restrict__ tab32, abm_vf_state_f32 st,
                   uint32_t* __restrict__ bits_out, float* __restrict__ diag, uint32_t flags,
                   unsigned long long* __restrict__ counters, int maxN) {
    extern __shared__ __align__(16) double smem_d[];
    constexpr int NT = 64 * NW;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;
    const bool want_diag = !PLAIN && (flags & ABM_WANT_DIAG) != 0 && diag != nullptr;
    const bool count = !PLAIN && (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) != 0 && counters;
    const bool verify = !PLAIN && (flags & ABM_VERIFY_FAST_PATH) != 0 && counters != nullptr;
    Smem S = carve(smem_d, maxN, RPL * NT, NW, sync_order);
    int loaded_tab = -1;
    uint32_t n_pairs = 0, n_recheck = 0, n_disagree = 0;   // this lane's

    for (int idx = blockIdx.x; idx < n_rep; idx += gridDim.x) {
        const VFRep* rp = reps + order[idx];
        const RepConst C = rep_const(rp);
        const int tab_off = rp->table_offset;
        if (tab_off != loaded_tab) {
            for (int e = tid; e < RPL * NT; e += NT) {
                const int r = (e % NT) * RPL + e / NT;
                RayTab32 z;
                z.wcos = z.wsin = z.cos = z.sin = 0.f;
                S.tab[e] = r < C.R ? tab32[tab_off + r] : z;
            }
            loaded_tab = tab_off;
        }
        const float a0d = (float)rp->alp0, b0d = (float)rp->bet0, v0d = (float)rp->v0;
        for (int i = tid; i < C.N; i += NT) {
            const int g = C.a0 + i;
            const float4 p = make_float4(st.x[g], st.y[g], (float)st.radius[g], st.theta[g]);
            S.pose[i] = p;
            if (sync_order) S.obj[i] = p;
            S.dyn[i] = make_float4(st.v[g], st.alp0 ? st.alp0[g] : a0d, st.bet0 ? st.bet0[g] : b0d,
                                   st.v0 ? st.v0[g] : v0d);
        }
        const int rs = tid * RPL;            // this lane's ray span [rs, rs + rn)
        const int rn = rs >= C.R ? 0 : (C.R - rs < RPL ? C.R - rs : RPL);
        if (tid == 0) S.nre[0] = 0;
        __syncthreads();
        // heading of agent 0 (later headings come with the previous agent's motion)
        float sth, cth;
        sincos_f32(S.pose[0].w, &sth, &cth);

        for (int k = 0; k < C.N; ++k) {
            for (int w = tid; w < C.W; w += NT) S.F[w] = 0u;
            const float4 fp = S.pose[k];
            const float fx = fp.x, fy = fp.y, fr = fp.z, fth = fp.w;
            const float u1x = cth, u1y = -sth;           // unit(v1) = (cos th, -sin th)
            __syncthreads();

            // ------------------------------------------------ pairs, float32
            for (int j = tid; j < C.N; j += NT) {
                if (j == k) continue;
                const float4 o = S.obj[j];
                if (o.x == fx && o.y == fy) continue;          // coincident (:57)
                PairF32 p = fast_pair(u1x, u1y, fx, fy, fr, o.x, o.y, o.z, C);
                if (count) ++n_pairs;
                if (p.tie && tie_in_place(p, fx, fy, fth, fr, o.x, o.y, o.z, rp, C))   // rare
                    if (count) ++n_recheck;   // a float64 decision too
                if (!p.ok || verify) {
                    const int slot = atomicAdd(S.nre, 1);
                    S.rlist[slot] = entry_code(j, 0, p.ok);
                    S.rfast[slot] = fast_code(p.start, p.end, p.drawn, C.R);
                    if (!p.ok) continue;
                }
                if (p.drawn) raster(S.F, p.start, p.end, C.R);
            }
            __syncthreads();
            // ------------------------------------- pairs re-decided in float64
            const int nre = S.nre[0];
            if (nre > 0) {
                for (int i = tid; i < nre; i += NT) {
                    const int e = S.rlist[i];
                    const int fc = S.rfast[i];
                    const int j = e & 0xfffff;
                    const float4 o = S.obj[j];
                    int dis;
                    const int r = resolve_queued(e, fc, fx, fy, fth, fr, o.x, o.y, o.z, rp, verify, &dis);
                    n_disagree += dis;
                    if (e & kEntryForced) {
                        if (count) ++n_recheck;
                        if (r >= 0) raster(S.F, (r & 0x7fff) - 2 * C.R, ((r >> 15) & 0x7fff) - 2 * C.R, C.R);
                    }
                }
                __syncthreads();
                if (tid == 0) S.nre[0] = 0;
            }

            // ------------------------------------------------ sums
            const bool case1 = (S.F[0] & 1u) == 1u && field_bit(S.F, C.R - 1) == 0u;
            float bc = 0.f, bs = 0.f, ec = 0.f, es = 0.f;
            {
                // rays at or beyond R have zero table entries and zero bits
                const int w0 = rs >> 5;
                const uint32_t lo = w0 < C.W ? S.F[w0] : 0u;
                const uint32_t hi = (RPL > 1 && w0 + 1 < C.W) ? S.F[w0 + 1] : 0u;
                const uint64_t v64 = (((uint64_t)hi << 32) | lo) >> (rs & 31);
                const uint32_t m = low_mask(rn);
                const uint32_t cur = (uint32_t)v64 & m;
                uint32_t spike = 0u;
                if (rn > 0) {
                    if (case1) {
                        const uint32_t prev = field_bit(S.F, rs == 0 ? C.R - 1 : rs - 1);
                        spike = (cur ^ ((cur << 1) | prev)) & m;
                    } else {
                        const int nx = rs + rn;
                        const uint32_t next = field_bit(S.F, nx >= C.R ? 0 : nx);
                        spike = (cur ^ ((cur >> 1) | (next << (rn - 1)))) & m;
                    }
                }
#pragma unroll
                for (int i = 0; i < RPL; ++i) {
                    const RayTab32 T = S.tab[i * NT + tid];
                    const float vb = (float)((cur >> i) & 1u);   // acc + (bit ? t : 0) as fma(bit, t, acc)
                    const float sb = (float)((spike >> i) & 1u);
                    bc = __builtin_fmaf(vb, T.wcos, bc);
                    bs = __builtin_fmaf(vb, T.wsin, bs);
                    ec = __builtin_fmaf(sb, T.cos, ec);
                    es = __builtin_fmaf(sb, T.sin, es);
                }
            }
            const float4 dk = S.dyn[k];
            const float alp0 = dk.y, bet0 = dk.z, v0 = dk.w;
            // blob terms: trapz(cos(PHI) * (-V)) = -sum(w cos V)
            double Sv, Sp, Bc = 0.0, Bs = 0.0, Ec = 0.0, Es = 0.0;
            if (!want_diag) {
                Sv = wave_total(alp0 * (-bc) + (alp0 * C.alp1) * ec);
                Sp = wave_total(bet0 * (-bs) + (bet0 * C.bet1) * es);
            } else {
                Bc = wave_total(bc); Bs = wave_total(bs);
                Ec = wave_total(ec); Es = wave_total(es);
                Sv = (double)alp0 * -Bc + (double)(alp0 * C.alp1) * Ec;
                Sp = (double)bet0 * -Bs + (double)(bet0 * C.bet1) * Es;
            }
            if (NW > 1) {
                if (lane == 0) {
                    S.red[2 * (tid >> 6)] = Sv; S.red[2 * (tid >> 6) + 1] = Sp;
                }
                __syncthreads();
                Sv = Sp = 0.0;
                for (int w = 0; w < NW; ++w) { Sv += S.red[2 * w]; Sp += S.red[2 * w + 1]; }
                if (want_diag) {   // components of the other waves
                    __syncthreads();
                    double comp[4] = {Bc, Bs, Ec, Es};
                    for (int c = 0; c < 4; ++c) {
                        if (lane == 0) S.red[tid >> 6] = comp[c];
                        __syncthreads();
                        double t = 0.0;
                        for (int w = 0; w < NW; ++w) t += S.red[w];
                        comp[c] = t;
                        __syncthreads();
                    }
                    Bc = comp[0]; Bs = comp[1]; Ec = comp[2]; Es = comp[3];
                }
            }
            if (!PLAIN && bits_out) {
                uint32_t* dst = bits_out + C.field_offset + (int64_t)k * C.W;
                for (int w = tid; w < C.W; w += NT) dst[w] = S.F[w];
            }

            // ------------------------------------------------ motion
            float vel, o, dv, dpsi;
            steer(C, fth, dk.x, v0, (float)Sv, (float)Sp, &o, &vel, &dv, &dpsi);
            // one sincos serves this agent's motion (lane 0) and the next
            // agent's heading (lane 1)
            const float ang = (lane == 1 && k + 1 < C.N) ? S.pose[k + 1].w : o;
            float s_l, c_l;
            sincos_f32(ang, &s_l, &c_l);
            const float so = lane_f32(s_l, 0), co = lane_f32(c_l, 0);
            sth = lane_f32(s_l, 1);
            cth = lane_f32(c_l, 1);
            float nx, ny;
            euler_boundary(C, fx, fy, fr, vel, so, co, &o, &nx, &ny);
            if (tid == 0) {
                S.pose[k] = make_float4(nx, ny, fr, o);
                S.dyn[k].x = vel;
                if (want_diag) {
                    float* dd = diag + 6 * (int64_t)(C.a0 + k);
                    dd[0] = dv; dd[1] = dpsi;
                    dd[2] = C.phi_ok ? (float)((double)alp0 * -Bc) : 0.f;
                    dd[3] = C.phi_ok ? (float)((double)(alp0 * C.alp1) * Ec) : 0.f;
                    dd[4] = C.phi_ok ? (float)((double)bet0 * -Bs) : 0.f;
                    dd[5] = C.phi_ok ? (float)((double)(bet0 * C.bet1) * Es) : 0.f;
                }
            }
            __syncthreads();
        }
        for (int i = tid; i < C.N; i += NT) {
            const int g = C.a0 + i;
            const float4 p = S.pose[i];
            st.x[g] = p.x;
            st.y[g] = p.y;
            st.theta[g] = p.w;
            st.v[g] = S.dyn[i].x;
        }
        __syncthreads();
    }
    if (count) {
        if (n_pairs) atomicAdd(&counters[0], (unsigned long long)n_pairs);
        if (n_recheck) atomicAdd(&counters[1], (unsigned long long)n_recheck);
        if (n_disagree) atomicAdd(&counters[2], (unsigned long long)n_disagree);
    }
}

// ---------------------------------------------------------------------------
// Wide variant for single large flocks, synchronous order only: one wave per
// focal agent, four waves per workgroup.  The waves of a workgroup share the
// ray table and 256-agent tiles of the (pre-step) poses in LDS.  A cheap
// float32 distance pass first drops every pair beyond the exact visibility
// cut-off -- the raster of a pair is empty when floor(proj_size / 2) = 0,
// i.e. when d > size * cot(2 pi / R) -- with a relative guard of 2e-4, and
// compacts the others; only those take the float32 decision.  New poses go
// to a scratch buffer and are copied back by a second kernel.
constexpr int kWideWaves = 8;    // focal agents per block, sharing the ray table
constexpr int kWideTile = 512;  // objects per LDS tile of the full scan (one per thread)
constexpr int kWideQueue = 128;   // grid path: candidates queued per wave (tile + cand memory)
static_assert(sizeof(float4) * kWideTile + sizeof(int32_t) * kWideTile * kWideWaves >=
              sizeof(float4) * kWideQueue * kWideWaves, "grid queues reuse the tile memory");

// Grid cell of a centre: the coordinate taken modulo the arena, clamped
// into the grid (rounding at the far edge).
__device__ inline int grid_cell(const WideGrid& G, float cx, float cy) {
    float u = cx - G.x0, v = cy - G.y0;
    u = u - G.W * floorf(u / G.W);
    v = v - G.H * floorf(v / G.H);
    const int ix = min(G.nx - 1, max(0, (int)(u / G.cw)));
    const int iy = min(G.ny - 1, max(0, (int)(v / G.ch)));
    return iy * G.nx + ix;
}

// Decision and raster of one pair of the grid path (no counters, no verify).
__device__ __forceinline__ void wide_decide(const float4 o, float u1x, float u1y, float fx, float fy,
                                            float fth, float fr, const VFRep* rp, const RepConst& C,
                                            uint32_t* F) {
    PairF32 p = fast_pair(u1x, u1y, fx, fy, fr, o.x, o.y, o.z, C);
    if (p.tie) tie_in_place(p, fx, fy, fth, fr, o.x, o.y, o.z, rp, C);   // rare
    if (!p.ok) {
        int es = 0, ee = 0, ed = 0;
        exact_pair(fx, fy, fth, fr, o.x, o.y, o.z, rp, &es, &ee, &ed);
        p.start = es; p.end = ee; p.drawn = ed != 0;
    }
    if (p.drawn) raster(F, p.start, p.end, C.R);
}

// LDS written by some lanes of a wave and read by others: in-order per
// wave once the writes have completed.
__device__ inline void wave_lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}


struct WideSmem {
    RayTab32* tab;      // transposed for 64 lanes
    float4* tile;       // (x, y, radius, -) of the current object tile
    uint32_t* F;        // kWideWaves fields of maxW words
    int32_t* cand;      // kWideWaves lists of tile-local candidates
};

// START: the chain step's first pass instead (see vf_step_f32_chain_kernel):
// agent k's pairs with the agents after it in id order only, its field written
// to start_fields[k W ..] and its heading vector to start_u[k]; no move.
template <int RPL, bool START, bool PLAIN>   // PLAIN: as the push kernel's
__global__ void __launch_bounds__(64 * kWideWaves)
vf_step_f32_wide_kernel(const VFRep* __restrict__ rp, const RayTab32* __restrict__ tab32,
                        abm_vf_state_f32 st, float4* __restrict__ out_pose,
                        uint32_t* __restrict__ bits_out, float* __restrict__ diag, uint32_t flags,
                        unsigned long long* __restrict__ counters, float cull_c,
                        const WideGrid* __restrict__ grid, const float4* __restrict__ sorted,
                        const int32_t* __restrict__ cell_start, uint32_t* __restrict__ start_fields,
                        float2* __restrict__ start_u) {
    extern __shared__ __align__(16) double smem_d[];
    const RepConst C = rep_const(rp);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool want_diag = !PLAIN && (flags & ABM_WANT_DIAG) != 0 && diag != nullptr;
    const bool count = !PLAIN && (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) != 0 && counters;
    const bool verify = !PLAIN && (flags & ABM_VERIFY_FAST_PATH) != 0 && counters != nullptr;
    // Grid path (counters and verify mode keep the full scan): focal agents
    // in cell order, so a block's waves read the same neighbourhoods
    const WideGrid G = *grid;
    const bool use_grid = G.ok && !count && !verify;
    const int slot = blockIdx.x * kWideWaves + wave;
    const bool active = slot < C.N;
    const int k = (use_grid && active) ? __float_as_int(sorted[slot].w) : slot;
    const int maxW = (RPL * 64 + 31) >> 5;
    WideSmem S;
    {
        char* p = reinterpret_cast<char*>(smem_d);
        S.tab = reinterpret_cast<RayTab32*>(p); p += sizeof(RayTab32) * RPL * 64;
        S.tile = reinterpret_cast<float4*>(p); p += sizeof(float4) * kWideTile;
        S.cand = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * kWideTile * kWideWaves;
        S.F = reinterpret_cast<uint32_t*>(p);
    }
    uint32_t* F = S.F + wave * maxW;
    int32_t* cand = S.cand + wave * kWideTile;
    const int tab_off = rp->table_offset;
    for (int e = tid; e < (START ? 0 : RPL * 64); e += 64 * kWideWaves) {
        const int r = (e % 64) * RPL + e / 64;
        RayTab32 z = tab32[tab_off + (r < C.R ? r : 0)];
        if (r >= C.R) z.wcos = z.wsin = z.cos = z.sin = 0.f;
        S.tab[e] = z;
    }
    for (int w = lane; w < maxW; w += 64) F[w] = 0u;
    const int g = C.a0 + (active ? k : 0);
    const float fx = st.x[g], fy = st.y[g], fth = st.theta[g], fr = (float)st.radius[g];
    float sth, cth;
    sincos_f32(fth, &sth, &cth);
    const float u1x = cth, u1y = -sth;
    unsigned long long n_pairs = 0, n_recheck = 0, n_disagree = 0;
    // Grid path: each wave takes its focal agent's 3 x 3 cells from the
    // sorted agents (the next 64 loaded while the current 64 are tested),
    // queues the pairs inside the cut-off and decides them 64 at a time.  No
    // block barrier inside: the waves share only the ray table.
    if (use_grid) {
        __syncthreads();   // the ray table
        if (active) {
            float4* cq = reinterpret_cast<float4*>(S.tile) + wave * kWideQueue;
            int nq = 0;
            const int fc = grid_cell(G, fx + fr, fy + fr);
            const int fiy = fc / G.nx, fix = fc - fiy * G.nx;
            const int xl = fix == 0 ? G.nx - 1 : fix - 1, xr = fix == G.nx - 1 ? 0 : fix + 1;
            const int yl = fiy == 0 ? G.ny - 1 : fiy - 1, yr = fiy == G.ny - 1 ? 0 : fiy + 1;
            for (int t = 0; t < 9; ++t) {   // the 3 x 3 cells, wrapped
                const int ty = t >= 6 ? 2 : (t >= 3 ? 1 : 0), tx = t - 3 * ty;
                const int cy = ty == 0 ? yl : (ty == 1 ? fiy : yr);
                const int cx = tx == 0 ? xl : (tx == 1 ? fix : xr);
                const int c = cy * G.nx + cx;
                const int end = cell_start[c + 1];
                int pos = cell_start[c];
                // loads always from a valid index; the next 64 are in flight
                // while these are tested
                float4 nxt = sorted[pos + lane < end ? pos + lane : 0];
                while (pos < end) {
                    const bool valid = pos + lane < end;
                    const float4 o = nxt;
                    pos += 64;
                    nxt = sorted[pos + lane < end ? pos + lane : 0];
                    bool cnd = false;
                    const int j = __float_as_int(o.w);
                    if (valid && (START ? j > k : j != k) && !(o.x == fx && o.y == fy)) {
                        float dx = (o.x - fx) + (o.z - fr), dy = (o.y - fy) + (o.z - fr);
                        const float ex = fabsf(dx) - C.hw, ey = fabsf(dy) - C.hh;
                        const bool near_wrap = fabsf(ex) < C.tguard || fabsf(ey) < C.tguard;
                        if (ex > 0.f) dx = dx > 0.f ? dx - C.fW : dx + C.fW;
                        if (ey > 0.f) dy = dy > 0.f ? dy - C.fH : dy + C.fH;
                        const float lim = o.z * cull_c;
                        cnd = near_wrap || fmaf(dy, dy, dx * dx) <= lim * lim;
                    }
                    const uint64_t m = __ballot(cnd);
                    if (cnd) cq[nq + __popcll(m & ((1ull << lane) - 1ull))] = o;
                    nq += __popcll(m);
                    if (nq >= 64) {
                        wave_lds_sync();
                        wide_decide(cq[lane], u1x, u1y, fx, fy, fth, fr, rp, C, F);
                        const int rest = nq - 64;
                        const float4 tq = cq[64 + lane];   // lanes >= rest read a stale entry, unused
                        wave_lds_sync();
                        if (lane < rest) cq[lane] = tq;
                        nq = rest;
                        wave_lds_sync();
                    }
                }
            }
            wave_lds_sync();
            if (lane < nq) wide_decide(cq[lane], u1x, u1y, fx, fy, fth, fr, rp, C, F);
            wave_lds_sync();
        }
    }

    // (START: the tiles before the block's first agent hold no pair)
    const int t_first = START ? (blockIdx.x * kWideWaves / kWideTile) * kWideTile : 0;
    for (int t0 = t_first; t0 < (use_grid ? 0 : C.N); t0 += kWideTile) {
        __syncthreads();
        if (t0 + tid < C.N) {
            const int gj = C.a0 + t0 + tid;
            S.tile[tid] = make_float4(st.x[gj], st.y[gj], (float)st.radius[gj], 0.f);
        }
        __syncthreads();
        // cheap pass: candidates inside the cut-off (or inside a torus guard band)
        int ncand = 0;
        const int tn = C.N - t0 < kWideTile ? C.N - t0 : kWideTile;
        for (int sub = 0; sub < kWideTile; sub += 64) {
            const int jt = sub + lane;
            bool c = false, culled = false;
            if (active && jt < tn && (START ? t0 + jt > k : t0 + jt != k)) {
                const float4 o = S.tile[jt];
                if (!(o.x == fx && o.y == fy)) {
                    if (count) ++n_pairs;
                    float dx = (o.x - fx) + (o.z - fr), dy = (o.y - fy) + (o.z - fr);
                    bool near_wrap = false;
                    if (C.torus) {
                        const float ex = fabsf(dx) - C.hw, ey = fabsf(dy) - C.hh;
                        near_wrap = fabsf(ex) < C.tguard || fabsf(ey) < C.tguard;
                        if (ex > 0.f) dx = dx > 0.f ? dx - C.fW : dx + C.fW;
                        if (ey > 0.f) dy = dy > 0.f ? dy - C.fH : dy + C.fH;
                    }
                    const float lim = o.z * cull_c;
                    culled = !near_wrap && fmaf(dy, dy, dx * dx) > lim * lim;
                    c = !culled || verify;
                }
            }
            const uint64_t m = __ballot(c);
            if (c) {
                const int pos = ncand + __popcll(m & ((1ull << lane) - 1ull));
                cand[pos] = jt | (culled ? (1 << 30) : 0);
            }
            ncand += __popcll(m);
        }
        __syncthreads();
        // float32 decisions of the candidates (float64 when inside a guard band)
        for (int c0 = 0; c0 < ncand; c0 += 64) {
            const int ci = c0 + lane;
            if (ci >= ncand) continue;
            const int e = cand[ci];
            const int jt = e & ((1 << 30) - 1);
            const bool culled = (e >> 30) & 1;
            const float4 o = S.tile[jt];
            PairF32 p = fast_pair(u1x, u1y, fx, fy, fr, o.x, o.y, o.z, C);
            if (p.tie && tie_in_place(p, fx, fy, fth, fr, o.x, o.y, o.z, rp, C))   // rare
                if (count) ++n_recheck;   // a float64 decision too
            if (!p.ok || verify) {
                int es = 0, ee = 0, ed = 0;
                exact_pair(fx, fy, fth, fr, o.x, o.y, o.z, rp, &es, &ee, &ed);
                if (verify) {
                    // a culled pair must have an empty raster; a float32 decision must match
                    const bool empty = !ed || es == ee;
                    if ((culled && !empty) ||
                        (!culled && p.ok && (es != p.start || ee != p.end || ed != (int)p.drawn)))
                        ++n_disagree;
                }
                if (!p.ok) {
                    if (count) ++n_recheck;
                    p.start = es; p.end = ee; p.drawn = ed != 0;
                }
            }
            if (culled) continue;
            if (p.drawn) raster(F, p.start, p.end, C.R);
        }
    }
    if (!use_grid) __syncthreads();
    if (START) {
        if (active) {
            uint32_t* dst = start_fields + (size_t)k * C.W;
            for (int w = lane; w < C.W; w += 64) dst[w] = F[w];
            if (lane == 0) start_u[k] = make_float2(u1x, u1y);
        }
        if (count) {
            if (n_pairs) atomicAdd(&counters[0], n_pairs);
            if (n_recheck) atomicAdd(&counters[1], n_recheck);
            if (n_disagree) atomicAdd(&counters[2], n_disagree);
        }
        return;
    }
    if (!active) {
        if (count) {
            if (n_pairs) atomicAdd(&counters[0], n_pairs);
        }
        return;
    }
    // sums over this wave's field (lanes own RPL consecutive rays)
    const int rs = lane * RPL;
    const int rn = rs >= C.R ? 0 : (C.R - rs < RPL ? C.R - rs : RPL);
    const bool case1 = (F[0] & 1u) == 1u && field_bit(F, C.R - 1) == 0u;
    float bc = 0.f, bs = 0.f, ec = 0.f, es = 0.f;
    {
        const int w0 = rs >> 5;
        const uint32_t lo = w0 < C.W ? F[w0] : 0u;
        const uint32_t hi = (RPL > 1 && w0 + 1 < C.W) ? F[w0 + 1] : 0u;
        const uint64_t v64 = (((uint64_t)hi << 32) | lo) >> (rs & 31);
        const uint32_t m = low_mask(rn);
        const uint32_t cur = (uint32_t)v64 & m;
        uint32_t spike = 0u;
        if (rn > 0) {
            if (case1) {
                const uint32_t prev = field_bit(F, rs == 0 ? C.R - 1 : rs - 1);
                spike = (cur ^ ((cur << 1) | prev)) & m;
            } else {
                const int nx = rs + rn;
                const uint32_t next = field_bit(F, nx >= C.R ? 0 : nx);
                spike = (cur ^ ((cur >> 1) | (next << (rn - 1)))) & m;
            }
        }
#pragma unroll 8
        for (int i = 0; i < RPL; ++i) {
            const RayTab32 T = S.tab[i * 64 + lane];
            const float vb = (float)((cur >> i) & 1u);   // acc + (bit ? t : 0) as fma(bit, t, acc)
            const float sb = (float)((spike >> i) & 1u);
            bc = __builtin_fmaf(vb, T.wcos, bc);
            bs = __builtin_fmaf(vb, T.wsin, bs);
            ec = __builtin_fmaf(sb, T.cos, ec);
            es = __builtin_fmaf(sb, T.sin, es);
        }
    }
    const float alp0 = st.alp0 ? st.alp0[g] : (float)rp->alp0;
    const float bet0 = st.bet0 ? st.bet0[g] : (float)rp->bet0;
    const float v0 = st.v0 ? st.v0[g] : (float)rp->v0;
    const double Bc = wave_total(bc), Bs = wave_total(bs), Ec = wave_total(ec), Es = wave_total(es);
    const double Sv = (double)alp0 * -Bc + (double)(alp0 * C.alp1) * Ec;
    const double Sp = (double)bet0 * -Bs + (double)(bet0 * C.bet1) * Es;
    if (!PLAIN && bits_out) {
        uint32_t* dst = bits_out + C.field_offset + (int64_t)k * C.W;
        for (int w = lane; w < C.W; w += 64) dst[w] = F[w];
    }
    float vel, o, dv, dpsi;
    steer(C, fth, st.v[g], v0, (float)Sv, (float)Sp, &o, &vel, &dv, &dpsi);
    float so, co;
    sincos_f32(o, &so, &co);
    float nx, ny;
    euler_boundary(C, fx, fy, fr, vel, so, co, &o, &nx, &ny);
    if (lane == 0) {
        out_pose[g] = make_float4(nx, ny, o, vel);
        if (want_diag) {
            float* dd = diag + 6 * (int64_t)g;
            dd[0] = dv; dd[1] = dpsi;
            dd[2] = C.phi_ok ? (float)((double)alp0 * -Bc) : 0.f;
            dd[3] = C.phi_ok ? (float)((double)(alp0 * C.alp1) * Ec) : 0.f;
            dd[4] = C.phi_ok ? (float)((double)bet0 * -Bs) : 0.f;
            dd[5] = C.phi_ok ? (float)((double)(bet0 * C.bet1) * Es) : 0.f;
        }
    }
    if (count) {
        if (n_pairs) atomicAdd(&counters[0], n_pairs);
        if (n_recheck) atomicAdd(&counters[1], n_recheck);
        if (n_disagree) atomicAdd(&counters[2], n_disagree);
    }
}

// One block: the replica's grid (cells no narrower than the largest
// visibility cut-off r_max * cull_c, plus 0.1 % and 1 px) and its agents
// sorted by cell (counting sort in LDS).  The order inside a cell is free:
// the field is an OR of rasters.
__global__ void __launch_bounds__(1024)
wide_grid_kernel(const VFRep* __restrict__ rp, abm_vf_state_f32 st, float cull_c,
                 float4* __restrict__ sorted, int32_t* __restrict__ cell_start,
                 WideGrid* __restrict__ grid) {
    __shared__ int cnt[kWideGridMaxCells + 1];
    __shared__ int part[1024];
    __shared__ int rmax;
    __shared__ WideGrid G;
    const RepConst C = rep_const(rp);
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (tid == 0) rmax = 0;
    __syncthreads();
    // eight agents per thread per pass, their loads issued together
    constexpr int kPer = 8;
    int rm = 0;
    for (int base = 0; base < C.N; base += nthr * kPer) {
        int rv[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = base + q * nthr + tid;
            rv[q] = i < C.N ? (int)st.radius[C.a0 + i] : 0;
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) rm = max(rm, rv[q]);
    }
    atomicMax(&rmax, rm);
    __syncthreads();
    if (tid == 0) {
        WideGrid g;
        g.pad_ = 0;
        g.x0 = C.pad; g.y0 = C.pad; g.W = C.fW; g.H = C.fH;
        const float reach = (float)rmax * cull_c * 1.001f + 1.0f;
        g.nx = min(kWideGridMaxSide, (int)floorf(C.fW / reach));
        g.ny = min(kWideGridMaxSide, (int)floorf(C.fH / reach));
        g.ok = C.torus && g.nx >= 3 && g.ny >= 3 && C.hw - C.tguard > reach && C.hh - C.tguard > reach;
        g.cw = g.nx > 0 ? C.fW / (float)g.nx : 1.f;
        g.ch = g.ny > 0 ? C.fH / (float)g.ny : 1.f;
        G = g;
        *grid = g;
    }
    __syncthreads();
    if (!G.ok) return;
    const int ncell = G.nx * G.ny;
    for (int c = tid; c <= ncell; c += nthr) cnt[c] = 0;
    __syncthreads();
    for (int base = 0; base < C.N; base += nthr * kPer) {
        float px[kPer], py[kPer], pr[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = base + q * nthr + tid;
            const int g = C.a0 + (i < C.N ? i : 0);
            px[q] = st.x[g]; py[q] = st.y[g]; pr[q] = (float)st.radius[g];
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if (base + q * nthr + tid < C.N) atomicAdd(&cnt[grid_cell(G, px[q] + pr[q], py[q] + pr[q])], 1);
    }
    __syncthreads();
    // exclusive scan of the counts: per-thread runs of 4 cells, then a
    // Hillis-Steele scan of the run totals
    const int per = (ncell + nthr - 1) / nthr;
    int run = 0;
    for (int q = 0; q < per; ++q) {
        const int c = tid * per + q;
        if (c < ncell) run += cnt[c];
    }
    part[tid] = run;
    __syncthreads();
    for (int o = 1; o < nthr; o <<= 1) {
        const int v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int acc = part[tid] - run;
    for (int q = 0; q < per; ++q) {
        const int c = tid * per + q;
        if (c < ncell) {
            const int n = cnt[c];
            cnt[c] = acc;          // now the cell's write cursor
            cell_start[c] = acc;
            acc += n;
        }
    }
    if (tid == 0) cell_start[ncell] = C.N;
    __syncthreads();
    for (int base = 0; base < C.N; base += nthr * kPer) {
        float px[kPer], py[kPer], pr[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = base + q * nthr + tid;
            const int g = C.a0 + (i < C.N ? i : 0);
            px[q] = st.x[g]; py[q] = st.y[g]; pr[q] = (float)st.radius[g];
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = base + q * nthr + tid;
            if (i < C.N) {
                const int at = atomicAdd(&cnt[grid_cell(G, px[q] + pr[q], py[q] + pr[q])], 1);
                sorted[at] = make_float4(px[q], py[q], pr[q], __int_as_float(i));
            }
        }
    }
}

__global__ void copy_back_kernel(const float4* __restrict__ out_pose, abm_vf_state_f32 st, int a0, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = out_pose[a0 + i];
    st.x[a0 + i] = p.x;
    st.y[a0 + i] = p.y;
    st.theta[a0 + i] = p.z;
    st.v[a0 + i] = p.w;
}

size_t lds_bytes_wide(int rpl) {
    const int maxW = (rpl * 64 + 31) >> 5;
    return sizeof(RayTab32) * (size_t)rpl * 64 + sizeof(float4) * kWideTile +
           sizeof(int32_t) * kWideTile * kWideWaves + sizeof(uint32_t) * (size_t)maxW * kWideWaves;
}

template <int RPL>
int launch_wide(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                uint32_t flags, hipStream_t s) {
    const size_t lds = lds_bytes_wide(RPL);
    const bool plain = bits == nullptr && (flags & (ABM_WANT_DIAG | ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) == 0;
    auto kern = plain ? vf_step_f32_wide_kernel<RPL, false, true> : vf_step_f32_wide_kernel<RPL, false, false>;
    if (lds > 65536) {
        if (int rc = check_hip(hipFuncSetAttribute((const void*)kern,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                               "hipFuncSetAttribute"))
            return rc;
    }
    for (int rep = 0; rep < b->n_replicas; ++rep) {
        const VFRep& h = b->reps_host[rep];
        // visibility cut-off: floor(proj_size/2) >= 1 needs d <= size * cot(2 pi / R)
        const float cull_c = (float)((1.0 / std::tan(kTwoPi / h.n_rays)) * (1.0 + 2e-4));
        const int blocks = (h.n_agents + kWideWaves - 1) / kWideWaves;
        WideGrid* grid = (WideGrid*)b->grid_param;
        hipLaunchKernelGGL(wide_grid_kernel, dim3(1), dim3(1024), 0, s, b->reps_dev + rep, *st, cull_c,
                           b->grid_sorted, b->grid_start, grid);
        if (int rc = check_hip(hipGetLastError(), "wide_grid_kernel launch")) return rc;
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * kWideWaves), lds, s, b->reps_dev + rep,
                           b->tab32, *st, b->wide_out, bits, diag, flags, b->counters, cull_c,
                           (const WideGrid*)grid, (const float4*)b->grid_sorted,
                           (const int32_t*)b->grid_start, nullptr, nullptr);
        if (int rc = check_hip(hipGetLastError(), "vf_step_f32_wide_kernel launch")) return rc;
        hipLaunchKernelGGL(copy_back_kernel, dim3((h.n_agents + 255) / 256), dim3(256), 0, s,
                           (const float4*)b->wide_out, *st, h.agent_offset, h.n_agents);
        if (int rc = check_hip(hipGetLastError(), "copy_back_kernel launch")) return rc;
    }
    b->last_kernels |= ABM_KERNEL_WIDE;
    return ABM_OK;
}

// ---------------------------------------------------------------------------
// Reference (id) order for replicas beyond kWideThreshold agents: one
// workgroup of 1024 threads per replica, agents in id order as in the
// per-replica kernel, with every agent's pose (x, y, radius, theta) in LDS,
// 16 bytes each (up to about 8,900 agents in 160 KB).  Velocities and the
// ray table stay in global memory (each thread keeps its rays' table entries
// in registers for the whole step).  Per focal agent k: a cheap float32 pass
// over all others drops the pairs beyond the exact visibility cut-off (as
// the wide kernel) and queues the rest in LDS (16-bit indices); the queued
// pairs take the float32 decision (float64 inside the guard bands) and are
// ORed into the field; each thread sums its rays, the waves' sums meet in
// LDS, wave 0 moves agent k.
constexpr int kBigThreads = 1024;
constexpr int kBigWaves = kBigThreads / 64;

// Queue segment of one wave: wave w owns the objects j with j mod 1024 in
// [64 w, 64 w + 64), at most 64 per 1024 agents.
__host__ __device__ inline int big_segment(int maxN) { return 64 * ((maxN + kBigThreads - 1) / kBigThreads); }

// The queues of 16-bit indices, padded to 16 bytes (what follows is int32).
__host__ __device__ inline size_t big_queue_bytes(int maxN) {
    return ((size_t)kBigWaves * big_segment(maxN) * sizeof(uint16_t) + 15) & ~(size_t)15;
}

size_t lds_bytes_big(int maxN, int maxR) {
    const int maxW = (maxR + 31) >> 5;
    size_t b = sizeof(float4) * (size_t)maxN + big_queue_bytes(maxN) +
               sizeof(double) * (4 * kBigWaves + 8) + sizeof(uint32_t) * (size_t)(maxW + 4);
    return (b + 15) & ~(size_t)15;
}

template <int RPL, bool PLAIN>   // PLAIN: as the push kernel's
__global__ void __launch_bounds__(kBigThreads)
vf_step_f32_big_kernel(const VFRep* __restrict__ reps, const int32_t* __restrict__ order,
                       const RayTab32* __restrict__ tab32, abm_vf_state_f32 st,
                       uint32_t* __restrict__ bits_out, float* __restrict__ diag, uint32_t flags,
                       unsigned long long* __restrict__ counters, float cull_c, int maxN, int maxR) {
    extern __shared__ __align__(16) double smem_d[];
    const VFRep* rp = reps + order[blockIdx.x];
    const RepConst C = rep_const(rp);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool want_diag = !PLAIN && (flags & ABM_WANT_DIAG) != 0 && diag != nullptr;
    const bool count = !PLAIN && (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) != 0 && counters;
    const bool verify = !PLAIN && (flags & ABM_VERIFY_FAST_PATH) != 0 && counters != nullptr;
    char* p = reinterpret_cast<char*>(smem_d);
    double* red = reinterpret_cast<double*>(p); p += sizeof(double) * (4 * kBigWaves + 8);
    float4* pose = reinterpret_cast<float4*>(p); p += sizeof(float4) * (size_t)maxN;
    uint16_t* queue = reinterpret_cast<uint16_t*>(p); p += big_queue_bytes(maxN);
    p += sizeof(int32_t) * 4;
    uint32_t* F = reinterpret_cast<uint32_t*>(p);
    // each wave queues the candidates among its own objects (64 of every
    // 1024) in its own segment, so no counter is shared between waves
    uint16_t* wq = queue + wave * big_segment(maxN);
    for (int i = tid; i < C.N; i += kBigThreads) {
        const int g = C.a0 + i;
        pose[i] = make_float4(st.x[g], st.y[g], (float)st.radius[g], st.theta[g]);
    }
    // this thread's rays [rs, rs + rn) and their table entries
    const int rs = tid * RPL;
    const int rn = rs >= C.R ? 0 : (C.R - rs < RPL ? C.R - rs : RPL);
    RayTab32 T[RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
        const int r = rs + i;
        if (r < C.R) T[i] = tab32[rp->table_offset + r];
        else T[i].wcos = T[i].wsin = T[i].cos = T[i].sin = 0.f;
    }
    const float a0d = (float)rp->alp0, b0d = (float)rp->bet0, v0d = (float)rp->v0;
    unsigned long long n_pairs = 0, n_recheck = 0, n_disagree = 0;
    // wave 0 moves the agents: agent k's speed and parameters are loaded one
    // agent ahead, so no global load waits on the id-order chain
    float vel_next = 0.f, a0_next = a0d, b0_next = b0d, v0_next = v0d;
    if (wave == 0) {
        vel_next = st.v[C.a0];
        if (st.alp0) a0_next = st.alp0[C.a0];
        if (st.bet0) b0_next = st.bet0[C.a0];
        if (st.v0) v0_next = st.v0[C.a0];
    }
    __syncthreads();
    for (int k = 0; k < C.N; ++k) {
        for (int w = tid; w < C.W; w += kBigThreads) F[w] = 0u;
        const float vel_k = vel_next, alp0_k = a0_next, bet0_k = b0_next, v0_k = v0_next;
        if (wave == 0 && k + 1 < C.N) {
            const int gn = C.a0 + k + 1;
            vel_next = st.v[gn];
            if (st.alp0) a0_next = st.alp0[gn];
            if (st.bet0) b0_next = st.bet0[gn];
            if (st.v0) v0_next = st.v0[gn];
        }
        const float4 fp = pose[k];
        const float fx = fp.x, fy = fp.y, fr = fp.z, fth = fp.w;
        float sth, cth;
        sincos_f32(fth, &sth, &cth);
        const float u1x = cth, u1y = -sth;
        __syncthreads();
        // ---- cheap pass: the pairs inside the cut-off (or a torus guard band)
        int nq = 0;   // wave-uniform
        for (int base = tid - lane; base < C.N; base += kBigThreads) {
            const int j = base + lane;
            bool c = false;
            if (j < C.N && j != k) {
                const float4 o = pose[j];
                if (!(o.x == fx && o.y == fy)) {
                    float dx = (o.x - fx) + (o.z - fr), dy = (o.y - fy) + (o.z - fr);
                    bool near_wrap = false;
                    if (C.torus) {
                        const float ex = fabsf(dx) - C.hw, ey = fabsf(dy) - C.hh;
                        near_wrap = fabsf(ex) < C.tguard || fabsf(ey) < C.tguard;
                        if (ex > 0.f) dx = dx > 0.f ? dx - C.fW : dx + C.fW;
                        if (ey > 0.f) dy = dy > 0.f ? dy - C.fH : dy + C.fH;
                    }
                    const float lim = o.z * cull_c;
                    const bool culled = !near_wrap && fmaf(dy, dy, dx * dx) > lim * lim;
                    if (culled && count) ++n_pairs;   // decided here: out of sight
                    if (culled && verify) {
                        // a culled pair must have an empty raster
                        int es = 0, ee = 0, ed = 0;
                        exact_pair(fx, fy, fth, fr, o.x, o.y, o.z, rp, &es, &ee, &ed);
                        if (ed && es != ee) ++n_disagree;
                    }
                    c = !culled;
                }
            }
            const uint64_t m = __ballot(c);
            if (c) wq[nq + (int)__popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)j;
            nq += (int)__popcll(m);
        }
        wave_lds_sync();
        // ---- this wave's queued pairs: float32 decision, float64 inside the guard bands
        for (int q = lane; q < nq; q += 64) {
            const float4 o = pose[wq[q]];
            PairF32 pr = fast_pair(u1x, u1y, fx, fy, fr, o.x, o.y, o.z, C);
            if (count) ++n_pairs;
            if (pr.tie && tie_in_place(pr, fx, fy, fth, fr, o.x, o.y, o.z, rp, C))   // rare
                if (count) ++n_recheck;
            if (!pr.ok || verify) {
                int es = 0, ee = 0, ed = 0;
                exact_pair(fx, fy, fth, fr, o.x, o.y, o.z, rp, &es, &ee, &ed);
                if (verify && pr.ok && (es != pr.start || ee != pr.end || ed != (int)pr.drawn))
                    ++n_disagree;
                if (!pr.ok) {
                    if (count) ++n_recheck;
                    pr.start = es; pr.end = ee; pr.drawn = ed != 0;
                }
            }
            if (pr.drawn) raster(F, pr.start, pr.end, C.R);
        }
        __syncthreads();
        // ---- sums over this thread's rays
        const bool case1 = (F[0] & 1u) == 1u && field_bit(F, C.R - 1) == 0u;
        float bc = 0.f, bs = 0.f, ec = 0.f, es = 0.f;
        if (rn > 0) {
            const int w0 = rs >> 5;
            const uint32_t lo = F[w0];
            const uint32_t hi = (RPL > 1 && w0 + 1 < C.W) ? F[w0 + 1] : 0u;
            const uint64_t v64 = (((uint64_t)hi << 32) | lo) >> (rs & 31);
            const uint32_t m = low_mask(rn);
            const uint32_t cur = (uint32_t)v64 & m;
            uint32_t spike;
            if (case1) {
                const uint32_t prev = field_bit(F, rs == 0 ? C.R - 1 : rs - 1);
                spike = (cur ^ ((cur << 1) | prev)) & m;
            } else {
                const int nx = rs + rn;
                const uint32_t next = field_bit(F, nx >= C.R ? 0 : nx);
                spike = (cur ^ ((cur >> 1) | (next << (rn - 1)))) & m;
            }
#pragma unroll
            for (int i = 0; i < RPL; ++i) {
                const float vb = (float)((cur >> i) & 1u);   // acc + (bit ? t : 0) as fma(bit, t, acc)
                const float sb = (float)((spike >> i) & 1u);
                bc = __builtin_fmaf(vb, T[i].wcos, bc);
                bs = __builtin_fmaf(vb, T[i].wsin, bs);
                ec = __builtin_fmaf(sb, T[i].cos, ec);
                es = __builtin_fmaf(sb, T[i].sin, es);
            }
        }
        const int gk = C.a0 + k;
        const float alp0 = alp0_k, bet0 = bet0_k, v0 = v0_k;
        // blob terms: trapz(cos(PHI) * (-V)) = -sum(w cos V); components by
        // wave in LDS, summed in wave order by every thread
        {
            const double Bc = wave_total(bc), Bs = wave_total(bs), Ec = wave_total(ec), Es = wave_total(es);
            if (lane == 0) {
                red[4 * wave] = Bc; red[4 * wave + 1] = Bs; red[4 * wave + 2] = Ec; red[4 * wave + 3] = Es;
            }
        }
        __syncthreads();
        if (wave == 0) {
            double Bc = 0.0, Bs = 0.0, Ec = 0.0, Es = 0.0;
            for (int w = 0; w < kBigWaves; ++w) {
                Bc += red[4 * w]; Bs += red[4 * w + 1]; Ec += red[4 * w + 2]; Es += red[4 * w + 3];
            }
            const double Sv = (double)alp0 * -Bc + (double)(alp0 * C.alp1) * Ec;
            const double Sp = (double)bet0 * -Bs + (double)(bet0 * C.bet1) * Es;
            float vel, o, dv, dpsi;
            steer(C, fth, vel_k, v0, (float)Sv, (float)Sp, &o, &vel, &dv, &dpsi);
            float so, co;
            sincos_f32(o, &so, &co);
            float nx, ny;
            euler_boundary(C, fx, fy, fr, vel, so, co, &o, &nx, &ny);
            if (lane == 0) {
                pose[k] = make_float4(nx, ny, fr, o);
                st.v[gk] = vel;
                if (want_diag) {
                    float* dd = diag + 6 * (int64_t)gk;
                    dd[0] = dv; dd[1] = dpsi;
                    dd[2] = C.phi_ok ? (float)((double)alp0 * -Bc) : 0.f;
                    dd[3] = C.phi_ok ? (float)((double)(alp0 * C.alp1) * Ec) : 0.f;
                    dd[4] = C.phi_ok ? (float)((double)bet0 * -Bs) : 0.f;
                    dd[5] = C.phi_ok ? (float)((double)(bet0 * C.bet1) * Es) : 0.f;
                }
            }
        }
        if (!PLAIN && bits_out) {
            uint32_t* dst = bits_out + C.field_offset + (int64_t)k * C.W;
            for (int w = tid; w < C.W; w += kBigThreads) dst[w] = F[w];
        }
        __syncthreads();
    }
    for (int i = tid; i < C.N; i += kBigThreads) {
        const int g = C.a0 + i;
        const float4 q = pose[i];
        st.x[g] = q.x;
        st.y[g] = q.y;
        st.theta[g] = q.w;
    }
    if (count) {
        if (n_pairs) atomicAdd(&counters[0], n_pairs);
        if (n_recheck) atomicAdd(&counters[1], n_recheck);
        if (n_disagree) atomicAdd(&counters[2], n_disagree);
    }
}

// The 16-wave kernel over the replicas order[0 .. n_rep), the largest with
// maxN agents and maxR rays.
template <int RPL>
int launch_big(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
               uint32_t flags, hipStream_t s, const int32_t* order, int n_rep, int maxN, int maxR) {
    const size_t lds = lds_bytes_big(maxN, maxR);
    if (lds > (size_t)b->dev_max_lds)
        return set_error(ABM_ERR_UNSUPPORTED,
                         "abm_vf_step_f32: %d agents in one replica need %zu B of LDS in the reference "
                         "(sequential) order on the 16-wave kernel (device max %d); the chain step (the default "
                         "for single flocks, abm_vf_set_launch(64)) takes it", maxN,
                         lds, b->dev_max_lds);
    const bool plain = bits == nullptr && (flags & (ABM_WANT_DIAG | ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) == 0;
    auto kern = plain ? vf_step_f32_big_kernel<RPL, true> : vf_step_f32_big_kernel<RPL, false>;
    if (lds > 65536) {
        if (int rc = check_hip(hipFuncSetAttribute((const void*)kern,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                               "hipFuncSetAttribute"))
            return rc;
    }
    // the cut-off of the largest R: no smaller for a replica of fewer rays
    // (cot(2 pi / R) grows with R), so it drops no drawable pair there either
    const float cull_c = (float)((1.0 / std::tan(kTwoPi / maxR)) * (1.0 + 2e-4));
    hipLaunchKernelGGL(kern, dim3(n_rep), dim3(kBigThreads), lds, s, b->reps_dev, order, b->tab32, *st,
                       bits, diag, flags, b->counters, cull_c, maxN, maxR);
    b->last_kernels |= ABM_KERNEL_BIG;
    return check_hip(hipGetLastError(), "vf_step_f32_big_kernel launch");
}

int launch_big_any(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                   uint32_t flags, hipStream_t s, const int32_t* order, int n_rep, int maxN, int maxR) {
    if (maxR <= kBigThreads) return launch_big<1>(b, st, bits, diag, flags, s, order, n_rep, maxN, maxR);
    return launch_big<2>(b, st, bits, diag, flags, s, order, n_rep, maxN, maxR);
}

// ---------------------------------------------------------------------------
// Reference (id) order with the pairs off the id-order chain (one workgroup
// of 512 threads per replica, every agent's field in LDS).  Agent m's field
// is the OR of its pairs with every other agent j, seen at j's new pose for
// j < m and at its start pose for j > m, and m itself at its start pose.
// Step k of the chain (one barrier each):
//   * wave 0 sums agent k's field, moves agent k and decides the pair
//     (k + 1, k) at k's new pose;
//   * waves 1..7 decide the pairs (m, k - 1) for m > k with agent k - 1's
//     new pose.
// The start-pose pairs (m, j > m) are decided before the chain by every
// wave.  Pairs beyond the visibility cut-off are dropped by a cheap float32
// test (as in the 16-wave kernel) and the rest queued, so that the exact
// decisions run on full waves.
// So only one pair, agent k's sums and its move stay on the chain.  Wave 0
// sums and moves exactly as the one-wave per-replica kernel does (same ray
// spans, table order and reductions), and the decisions are exact, so the
// two kernels give identical results bit for bit.  Wave 0 keeps its rays'
// table entries in registers (up to 128), which two waves per SIMD allow;
// from LDS, the loads of the sums waited on the other waves' LDS traffic.
constexpr int kPushThreads = 64 * kPushWaves;

// The one-wave sums of a field Fk, as the one-wave per-replica kernel takes
// them: lane l sums its rays [l RPL, (l + 1) RPL) (table entries T) in float32,
// the wave totals are float64.  Without diagnostics the two integrals are
// totalled directly (Sv, Sp); with them the four parts (Bc, Bs, Ec, Es) and
// the integrals from those.  Shared by the push kernel and the chain step's
// segment kernel, which thereby give the same results bit for bit.
template <int RPL>
__device__ __forceinline__ void one_wave_sums(const uint32_t* Fk, const RayTab32* T, int lane, const RepConst& C,
                                              bool want_diag, float alp0, float bet0, double& Sv, double& Sp,
                                              double& Bc, double& Bs, double& Ec, double& Es) {
    const int rs = lane * RPL;
    const int rn = rs >= C.R ? 0 : (C.R - rs < RPL ? C.R - rs : RPL);
    const bool case1 = (Fk[0] & 1u) == 1u && field_bit(Fk, C.R - 1) == 0u;
    float bc = 0.f, bs = 0.f, ec = 0.f, es = 0.f;
    {
        const int w0 = rs >> 5;
        const uint32_t lo = w0 < C.W ? Fk[w0] : 0u;
        const uint32_t hi = (RPL > 1 && w0 + 1 < C.W) ? Fk[w0 + 1] : 0u;
        const uint64_t v64 = (((uint64_t)hi << 32) | lo) >> (rs & 31);
        const uint32_t m = low_mask(rn);
        const uint32_t cur = (uint32_t)v64 & m;
        uint32_t spike = 0u;
        if (rn > 0) {
            if (case1) {
                const uint32_t prev = field_bit(Fk, rs == 0 ? C.R - 1 : rs - 1);
                spike = (cur ^ ((cur << 1) | prev)) & m;
            } else {
                const int nx = rs + rn;
                const uint32_t next = field_bit(Fk, nx >= C.R ? 0 : nx);
                spike = (cur ^ ((cur >> 1) | (next << (rn - 1)))) & m;
            }
        }
        // acc + (bit ? t : 0) as fma(bit, t, acc): the same value (a 0 bit adds
        // a zero to an accumulator that is never -0), without a lane mask
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const float vb = (float)((cur >> i) & 1u);
            const float sb = (float)((spike >> i) & 1u);
            bc = __builtin_fmaf(vb, T[i].wcos, bc);
            bs = __builtin_fmaf(vb, T[i].wsin, bs);
            ec = __builtin_fmaf(sb, T[i].cos, ec);
            es = __builtin_fmaf(sb, T[i].sin, es);
        }
    }
    Bc = Bs = Ec = Es = 0.0;
    if (!want_diag) {
        Sv = wave_total(alp0 * (-bc) + (alp0 * C.alp1) * ec);
        Sp = wave_total(bet0 * (-bs) + (bet0 * C.bet1) * es);
    } else {
        Bc = wave_total(bc); Bs = wave_total(bs);
        Ec = wave_total(ec); Es = wave_total(es);
        Sv = (double)alp0 * -Bc + (double)(alp0 * C.alp1) * Ec;
        Sp = (double)bet0 * -Bs + (double)(bet0 * C.bet1) * Es;
    }
}

// One pair of the push kernel: focal f (start pose, heading vector u), object
// o; the float32 decision with the float64 one inside the guard bands (and
// every pair in verify mode), rastered into the focal's field Fm.
// The decision alone: whether the pair draws, and its span in pr.start/end.
__device__ inline bool decide_pair(PairF32& pr, float2 u, float4 f, float4 o, const VFRep* rp,
                                   const RepConst& C, bool count, bool verify, unsigned long long& n_pairs,
                                   unsigned long long& n_recheck, unsigned long long& n_disagree) {
    if (o.x == f.x && o.y == f.y) return false;   // coincident (vf_supcalc.py:57)
    pr = fast_pair(u.x, u.y, f.x, f.y, f.z, o.x, o.y, o.z, C);
    if (count) ++n_pairs;
    if (pr.tie && tie_in_place(pr, f.x, f.y, f.w, f.z, o.x, o.y, o.z, rp, C))   // rare
        if (count) ++n_recheck;
    if (!pr.ok || verify) {
        int es = 0, ee = 0, ed = 0;
        exact_pair(f.x, f.y, f.w, f.z, o.x, o.y, o.z, rp, &es, &ee, &ed);
        if (verify && pr.ok && (es != pr.start || ee != pr.end || ed != (int)pr.drawn)) ++n_disagree;
        if (!pr.ok) {
            if (count) ++n_recheck;
            pr.start = es; pr.end = ee; pr.drawn = ed != 0;
        }
    }
    return pr.drawn;
}

__device__ inline void push_pair(uint32_t* Fm, float2 u, float4 f, float4 o, const VFRep* rp,
                                 const RepConst& C, bool count, bool verify, unsigned long long& n_pairs,
                                 unsigned long long& n_recheck, unsigned long long& n_disagree) {
    PairF32 pr;
    if (decide_pair(pr, u, f, o, rp, C, count, verify, n_pairs, n_recheck, n_disagree))
        raster(Fm, pr.start, pr.end, C.R);
}

// Whether a pair lies beyond the visibility cut-off d > size cot(2 pi / R)
// (floor(proj_size / 2) = 0: nothing drawn), in float32 with the periodic
// image's guard band, as the 16-wave kernel's cheap pass.  A culled pair is
// counted as decided; in verify mode its float64 raster must be empty.
__device__ inline bool push_culled(float4 f, float4 o, const RepConst& C, float cull_c) {
    float dx = (o.x - f.x) + (o.z - f.z), dy = (o.y - f.y) + (o.z - f.z);
    bool near_wrap = false;
    if (C.torus) {
        const float ex = fabsf(dx) - C.hw, ey = fabsf(dy) - C.hh;
        near_wrap = fabsf(ex) < C.tguard || fabsf(ey) < C.tguard;
        if (ex > 0.f) dx = dx > 0.f ? dx - C.fW : dx + C.fW;
        if (ey > 0.f) dy = dy > 0.f ? dy - C.fH : dy + C.fH;
    }
    const float lim = o.z * cull_c;
    return !near_wrap && fmaf(dy, dy, dx * dx) > lim * lim;
}

__device__ inline void push_culled_check(float4 f, float4 o, const VFRep* rp, bool count, bool verify,
                                         unsigned long long& n_pairs, unsigned long long& n_disagree) {
    if (count) ++n_pairs;
    if (verify) {
        int es = 0, ee = 0, ed = 0;
        exact_pair(f.x, f.y, f.w, f.z, o.x, o.y, o.z, rp, &es, &ee, &ed);
        if (ed && es != ee) ++n_disagree;
    }
}

// PLAIN: a step with no fields, diagnostics or counters out (the flags and
// pointers that ask for them are ignored), so their tests leave the kernel.
template <int RPL, bool PLAIN>
__global__ void __launch_bounds__(kPushThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
vf_step_f32_push_kernel(const VFRep* __restrict__ reps, const int32_t* __restrict__ order,
                        const RayTab32* __restrict__ tab32, abm_vf_state_f32 st,
                        uint32_t* __restrict__ bits_out, float* __restrict__ diag, uint32_t flags,
                        unsigned long long* __restrict__ counters) {
    extern __shared__ __align__(16) double smem_d[];
    const VFRep* rp = reps + order[blockIdx.x];
    const RepConst C = rep_const(rp);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int W1 = C.W + 1;
    const bool want_diag = !PLAIN && (flags & ABM_WANT_DIAG) != 0 && diag != nullptr;
    const bool count = !PLAIN && (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) != 0 && counters;
    const bool verify = !PLAIN && (flags & ABM_VERIFY_FAST_PATH) != 0 && counters != nullptr;
    // laid out for this replica (lds_bytes_push(N, R) bytes)
    char* p = reinterpret_cast<char*>(smem_d);
    float4* P0 = reinterpret_cast<float4*>(p); p += sizeof(float4) * (size_t)C.N;    // start poses
    float4* P1 = reinterpret_cast<float4*>(p); p += sizeof(float4) * (size_t)C.N;    // new poses
    float4* dyn = reinterpret_cast<float4*>(p); p += sizeof(float4) * (size_t)C.N;   // v, ALP0, BET0, V0
    float2* U = reinterpret_cast<float2*>(p); p += sizeof(float2) * (size_t)C.N;     // (cos, -sin) of start headings
    uint32_t* queue = reinterpret_cast<uint32_t*>(p); p += sizeof(uint32_t) * kPushQueue * kPushWaves;
    uint32_t* F = reinterpret_cast<uint32_t*>(p);                                     // [N][W + 1] fields
    const float cull_c = (float)rp->cull_c;
    uint32_t* wq = queue + wave * kPushQueue;   // this wave's ring of queued pairs
    const float a0d = (float)rp->alp0, b0d = (float)rp->bet0, v0d = (float)rp->v0;
    for (int i = tid; i < C.N; i += kPushThreads) {
        const int g = C.a0 + i;
        const float4 q = make_float4(st.x[g], st.y[g], (float)st.radius[g], st.theta[g]);
        P0[i] = q;
        dyn[i] = make_float4(st.v[g], st.alp0 ? st.alp0[g] : a0d, st.bet0 ? st.bet0[g] : b0d,
                             st.v0 ? st.v0[g] : v0d);
        float sth, cth;
        sincos_f32(q.w, &sth, &cth);
        U[i] = make_float2(cth, -sth);
    }
    // wave 0's lane l owns rays [l RPL, (l + 1) RPL) and their table entries
    RayTab32 T[RPL];
    if (wave == 0) {
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int r = lane * RPL + i;
            if (r < C.R) T[i] = tab32[rp->table_offset + r];
            else T[i].wcos = T[i].wsin = T[i].cos = T[i].sin = 0.f;
        }
    }
    for (int w = tid; w < C.N * W1; w += kPushThreads) F[w] = 0u;
    unsigned long long n_pairs = 0, n_recheck = 0, n_disagree = 0;
    __syncthreads();
    // ---- the start-pose pairs (m, j > m), wave w taking the rows m = w mod 8;
    // queue entries m | j << 16
    {
        int head = 0, tail = 0;   // wave-uniform ring positions
        for (int m = wave; m < C.N; m += kPushWaves) {
            const float4 f = P0[m];
            for (int jb = m + 1; jb < C.N; jb += 64) {
                const int j = jb + lane;
                bool keep = false;
                if (j < C.N) {
                    const float4 ob = P0[j];
                    if (!(ob.x == f.x && ob.y == f.y)) {   // coincident pairs are skipped (vf_supcalc.py:57)
                        keep = !push_culled(f, ob, C, cull_c);
                        if (!keep) push_culled_check(f, ob, rp, count, verify, n_pairs, n_disagree);
                    }
                }
                const uint64_t mk = __ballot(keep);
                if (keep)
                    wq[(tail + (int)__popcll(mk & ((1ull << lane) - 1ull))) & (kPushQueue - 1)] =
                        (uint32_t)m | ((uint32_t)j << 16);
                tail += (int)__popcll(mk);
                wave_lds_sync();
                if (tail - head >= 64) {
                    const uint32_t e = wq[(head + lane) & (kPushQueue - 1)];
                    const int mm = (int)(e & 0xffffu), jj = (int)(e >> 16);
                    push_pair(F + (size_t)mm * W1, U[mm], P0[mm], P0[jj], rp, C, count, verify, n_pairs,
                              n_recheck, n_disagree);
                    head += 64;
                    wave_lds_sync();
                }
            }
        }
        if (lane < tail - head) {
            const uint32_t e = wq[(head + lane) & (kPushQueue - 1)];
            const int mm = (int)(e & 0xffffu), jj = (int)(e >> 16);
            push_pair(F + (size_t)mm * W1, U[mm], P0[mm], P0[jj], rp, C, count, verify, n_pairs, n_recheck,
                      n_disagree);
        }
    }
    __syncthreads();
    if (wave == 0) __builtin_amdgcn_s_setprio(3);   // the chain's wave issues first
    for (int k = 0; k < C.N; ++k) {
        if (wave == 0) {
            // ---- agent k's sums, as the one-wave per-replica kernel
            const uint32_t* Fk = F + (size_t)k * W1;
            const float4 fp = P0[k];
            const float4 dk = dyn[k];
            const float fx = fp.x, fy = fp.y, fr = fp.z, fth = fp.w;
            const float alp0 = dk.y, bet0 = dk.z, v0 = dk.w;
            double Sv, Sp, Bc, Bs, Ec, Es;
            one_wave_sums<RPL>(Fk, T, lane, C, want_diag, alp0, bet0, Sv, Sp, Bc, Bs, Ec, Es);
            if (!PLAIN && bits_out) {
                uint32_t* dst = bits_out + C.field_offset + (int64_t)k * C.W;
                for (int w = lane; w < C.W; w += 64) dst[w] = Fk[w];
            }
            // ---- agent k's move, then the pair (k + 1, k) at its new pose
            float vel, o, dv, dpsi;
            steer(C, fth, dk.x, v0, (float)Sv, (float)Sp, &o, &vel, &dv, &dpsi);
            float so, co;
            sincos_f32(o, &so, &co);
            float nx, ny;
            euler_boundary(C, fx, fy, fr, vel, so, co, &o, &nx, &ny);
            const float4 moved = make_float4(nx, ny, fr, o);
            if (lane == 0) {
                const int gk = C.a0 + k;
                P1[k] = moved;
                st.v[gk] = vel;
                if (want_diag) {
                    float* dd = diag + 6 * (int64_t)gk;
                    dd[0] = dv; dd[1] = dpsi;
                    dd[2] = C.phi_ok ? (float)((double)alp0 * -Bc) : 0.f;
                    dd[3] = C.phi_ok ? (float)((double)(alp0 * C.alp1) * Ec) : 0.f;
                    dd[4] = C.phi_ok ? (float)((double)bet0 * -Bs) : 0.f;
                    dd[5] = C.phi_ok ? (float)((double)(bet0 * C.bet1) * Es) : 0.f;
                }
                if (k + 1 < C.N) {
                    const float4 f1 = P0[k + 1];
                    if (push_culled(f1, moved, C, cull_c))
                        push_culled_check(f1, moved, rp, count, verify, n_pairs, n_disagree);
                    else
                        push_pair(F + (size_t)(k + 1) * W1, U[k + 1], f1, moved, rp, C, count, verify,
                                  n_pairs, n_recheck, n_disagree);
                }
            }
        } else {
            // ---- agent k - 1's new pose into the fields of agents k + 1 ..:
            // each wave tests its focal agents against the cut-off 64 at a
            // time and queues the rest in a ring of 128 in LDS, deciding 64
            // at a time
            if (k >= 1) {
                const float4 o = P1[k - 1];
                int head = 0, tail = 0;   // wave-uniform ring positions
                for (int mb = k + 1 + (wave - 1) * 64; mb < C.N; mb += kPushThreads - 64) {
                    const int m = mb + lane;
                    bool keep = false;
                    if (m < C.N) {
                        const float4 f = P0[m];
                        if (!(o.x == f.x && o.y == f.y)) {   // coincident pairs are skipped (vf_supcalc.py:57)
                            keep = !push_culled(f, o, C, cull_c);
                            if (!keep) push_culled_check(f, o, rp, count, verify, n_pairs, n_disagree);
                        }
                    }
                    const uint64_t mk = __ballot(keep);
                    if (keep) wq[(tail + (int)__popcll(mk & ((1ull << lane) - 1ull))) & (kPushQueue - 1)] = (uint32_t)m;
                    tail += (int)__popcll(mk);
                    wave_lds_sync();
                    if (tail - head >= 64) {
                        const int mm = (int)wq[(head + lane) & (kPushQueue - 1)];
                        push_pair(F + (size_t)mm * W1, U[mm], P0[mm], o, rp, C, count, verify, n_pairs,
                                  n_recheck, n_disagree);
                        head += 64;
                        wave_lds_sync();
                    }
                }
                if (lane < tail - head) {
                    const int mm = (int)wq[(head + lane) & (kPushQueue - 1)];
                    push_pair(F + (size_t)mm * W1, U[mm], P0[mm], o, rp, C, count, verify, n_pairs, n_recheck,
                              n_disagree);
                }
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < C.N; i += kPushThreads) {
        const int g = C.a0 + i;
        const float4 q = P1[i];
        st.x[g] = q.x;
        st.y[g] = q.y;
        st.theta[g] = q.w;
    }
    if (count) {
        if (n_pairs) atomicAdd(&counters[0], n_pairs);
        if (n_recheck) atomicAdd(&counters[1], n_recheck);
        if (n_disagree) atomicAdd(&counters[2], n_disagree);
    }
}

// Whether the push kernel holds every replica (all) or the large size class.
bool push_fits_in(const abm_vf_batch* b, bool all) {
    const size_t lds = all ? b->push_lds_all : b->push_lds_big;
    return lds > 0 && b->max_rays <= 64 * 32 && lds <= (size_t)b->dev_max_lds &&
           (all ? b->max_agents : b->big_max_agents) <= kWideThreshold;
}

template <int RPL>
int launch_push(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                uint32_t flags, hipStream_t s, const int32_t* order, int n_rep, size_t lds) {
    const bool plain = bits == nullptr && (flags & (ABM_WANT_DIAG | ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) == 0;
    auto kern = plain ? vf_step_f32_push_kernel<RPL, true> : vf_step_f32_push_kernel<RPL, false>;
    if (lds > 65536) {
        if (int rc = check_hip(hipFuncSetAttribute((const void*)kern,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                               "hipFuncSetAttribute"))
            return rc;
    }
    hipLaunchKernelGGL(kern, dim3(n_rep), dim3(kPushThreads), lds, s, b->reps_dev, order, b->tab32, *st,
                       bits, diag, flags, b->counters);
    b->last_kernels |= ABM_KERNEL_PUSH;
    return check_hip(hipGetLastError(), "vf_step_f32_push_kernel launch");
}

// The push kernel over order[0 .. n_rep): every replica (all) or the large
// size class.  Rays per lane of wave 0 follow the one-wave per-replica
// kernel's rule for the largest R of the launch.
int launch_push_any(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                    uint32_t flags, hipStream_t s, const int32_t* order, int n_rep, bool all) {
    const size_t lds = all ? b->push_lds_all : b->push_lds_big;
    if (!push_fits_in(b, all))
        return set_error(ABM_ERR_UNSUPPORTED,
                         "abm_vf_step_f32: the push kernel holds no replica of this batch (%zu B of LDS "
                         "for its fields, device max %d; up to %d agents)", lds, b->dev_max_lds,
                         kWideThreshold);
    const int rpl = push_rpl(all ? b->max_rays : b->big_max_rays);
    if (rpl == 4) return launch_push<4>(b, st, bits, diag, flags, s, order, n_rep, lds);
    if (rpl == 8) return launch_push<8>(b, st, bits, diag, flags, s, order, n_rep, lds);
    if (rpl == 16) return launch_push<16>(b, st, bits, diag, flags, s, order, n_rep, lds);
    return launch_push<32>(b, st, bits, diag, flags, s, order, n_rep, lds);
}

// ---------------------------------------------------------------------------
// Reference (id) order for flocks above kWideThreshold agents, the chain
// step.  Agent m's field is the OR of its pairs with every other agent j at
// j's new pose for j < m and its start pose for j > m (m itself at its start
// pose).  The pairs are decided as soon as both poses exist, mostly off the
// id-order chain and over the whole GPU:
//   1. wide_grid_kernel: the grid of the start poses (torus arenas whose
//      cells can be at least the visibility cut-off wide).
//   2. vf_step_f32_wide_kernel<RPL, true, *>: each agent's field with the agents
//      after it, all at their start poses, into a global buffer of fields,
//      and its start heading vector (one wave per agent).
//   3. For each segment of kChainSeg agents in id order:
//      a. vf_step_f32_segment_kernel (one workgroup of 512 threads) moves the
//         segment's agents in id order, the segment's fields in LDS (their
//         pairs with every earlier agent already ORed in), summing and moving
//         as the push kernel does, so the two give the same fields and poses
//         bit for bit.
//      b. vf_step_f32_chain_push_kernel (the whole GPU) decides the segment's
//         new poses in the global fields of every later agent: one wave per
//         (moved agent, grid cell of its 3 x 3 neighbourhood), or per range
//         of later agents without a grid (and with counters or verify mode).
// Kernel boundaries order the global fields between the launches.
constexpr int kChainThreads = 512;
constexpr int kChainWaves = kChainThreads / 64;
constexpr int kChainCullWaves = 3;    // segment kernel: three waves test, three decide
static_assert(kChainWaves == 2 + 2 * kChainCullWaves, "segment kernel wave roles");
constexpr int kChainSeg = 256;        // agents per segment
constexpr int kChainQueue = 128;      // queued agents per wave of the chain push kernel
constexpr int kChainPushWaves = 4;    // waves per workgroup of the push kernel
constexpr int kChainScanRange = 512;  // later agents per wave of the push kernel without a grid

// Barrier of the segment kernel: this wave's LDS operations complete, then
// s_barrier (the kernel shares no global memory between its waves).
__device__ inline void chain_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// The agents [s0, s0 + n) of one replica, n <= kChainSeg, one barrier per
// agent k:
//   * wave 0 takes agent k's sums (made by wave 1 during agent k - 1), moves
//     agent k and decides the pairs (k + 1, k) and (k + 2, k) at k's new pose
//     and (k + 2, k - 1) at k - 1's (one lane each);
//   * wave 1 sums agent k + 1's field, which then lacks only the pair
//     (k + 1, k), as the push kernel's wave 0 sums (one_wave_sums);
//   * three waves test agent k - 1's new pose against the segment's agents
//     from k + 3 on (the cheap float32 cut-off test) and list the pairs that
//     remain; three more decide the pairs listed in the previous step (agent
//     k - 2's new pose), 64 per wave at a time.  A pair (m, j) with
//     m >= j + 4 is so decided two steps after j moved, before wave 1 sums
//     agent m's field (in step m - 1), and each wave makes one pass a step.
//     The busiest of them (the first lister and the first decider) run as
//     waves 3 and 2, which share their SIMDs (wave w on SIMD w mod 4) with
//     the lightly loaded waves 7 and 6, not with waves 0 and 1.
// When the pair (k + 1, k) draws, wave 1 sums agent k + 1's field again,
// behind an extra barrier, before wave 0 uses it.  The decisions are exact
// and the sums and moves are the push kernel's, so the two give the same
// fields and poses bit for bit.
template <int RPL, bool PLAIN>   // PLAIN: as the push kernel's
__global__ void __launch_bounds__(kChainThreads)
vf_step_f32_segment_kernel(const VFRep* __restrict__ rp, const RayTab32* __restrict__ tab32,
                           abm_vf_state_f32 st, const uint32_t* __restrict__ fields,
                           const float2* __restrict__ ustart, uint32_t* __restrict__ bits_out,
                           float* __restrict__ diag, uint32_t flags, unsigned long long* __restrict__ counters,
                           float cull_c, int s0, int n) {
    __shared__ uint32_t rows[kChainSeg][64];
    __shared__ float4 srec[kChainSeg];     // start poses (x, y, radius, theta)
    __shared__ float2 su[kChainSeg];       // start heading vectors
    __shared__ float4 sdyn[kChainSeg];     // speed, ALP0, BET0, V0
    __shared__ double red[2][6];           // Sv, Sp, Bc, Bs, Ec, Es of agents k and k + 1
    __shared__ float4 newpose[4];          // new pose of local agent kk at kk & 3
    __shared__ int dirty[2];               // the pair (k, k - 1) drew after k's sums were taken
    __shared__ uint16_t cq[2][kChainCullWaves][kChainSeg];   // listed rows, by step parity and cull wave
    __shared__ int cn[2][kChainCullWaves];
    const RepConst C = rep_const(rp);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool want_diag = !PLAIN && (flags & ABM_WANT_DIAG) != 0 && diag != nullptr;
    const bool count = !PLAIN && (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) != 0 && counters;
    const bool verify = !PLAIN && (flags & ABM_VERIFY_FAST_PATH) != 0 && counters != nullptr;
    const float a0d = (float)rp->alp0, b0d = (float)rp->bet0, v0d = (float)rp->v0;
    for (int i = tid; i < n; i += kChainThreads) {
        const int g = C.a0 + s0 + i;
        srec[i] = make_float4(st.x[g], st.y[g], (float)st.radius[g], st.theta[g]);
        su[i] = ustart[s0 + i];
        sdyn[i] = make_float4(st.v[g], st.alp0 ? st.alp0[g] : a0d, st.bet0 ? st.bet0[g] : b0d,
                              st.v0 ? st.v0[g] : v0d);
    }
    for (int w = tid; w < n * 64; w += kChainThreads) {
        const int i = w >> 6, c = w & 63;
        rows[i][c] = c < C.W ? fields[(size_t)(s0 + i) * C.W + c] : 0u;
    }
    if (tid < 2) dirty[tid] = 0;
    // wave 1's lane l owns rays [l RPL, (l + 1) RPL) and their table entries
    RayTab32 T[RPL];
    if (wave == 1) {
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int r = lane * RPL + i;
            if (r < C.R) T[i] = tab32[rp->table_offset + r];
            else T[i].wcos = T[i].wsin = T[i].cos = T[i].sin = 0.f;
        }
    }
    // wave 1: the sums of local agent kk's field into red[buf]
    auto sums = [&](int kk, int buf) {
        const float4 d = sdyn[kk];
        double Sv, Sp, Bc, Bs, Ec, Es;
        one_wave_sums<RPL>(rows[kk], T, lane, C, want_diag, d.y, d.z, Sv, Sp, Bc, Bs, Ec, Es);
        if (lane == 0) {
            double* r = red[buf];
            r[0] = Sv; r[1] = Sp; r[2] = Bc; r[3] = Bs; r[4] = Ec; r[5] = Es;
        }
    };
    unsigned long long n_pairs = 0, n_recheck = 0, n_disagree = 0;
    __syncthreads();
    if (wave == 1) sums(0, 0);   // agent s0's field is complete
    __syncthreads();
    if (wave == 0) __builtin_amdgcn_s_setprio(3);   // the chain's wave issues first,
    if (wave == 1) __builtin_amdgcn_s_setprio(2);   // then the summing wave
    for (int kk = 0; kk < n; ++kk) {
        const int k = s0 + kk;
        const int buf = kk & 1;
        if (dirty[buf]) {   // uniform: agent k's field changed after its sums
            if (wave == 1) sums(kk, buf);
            chain_barrier();
        }
        if (wave == 0) {
            const int gk = C.a0 + k;
            const double* r = red[buf];
            const double Sv = r[0], Sp = r[1], Bc = r[2], Bs = r[3], Ec = r[4], Es = r[5];
            const float4 pk = srec[kk];
            const float4 dk = sdyn[kk];
            const float fx = pk.x, fy = pk.y, fr = pk.z, fth = pk.w;
            const float alp0 = dk.y, bet0 = dk.z, v0 = dk.w;
            float vel, o, dv, dpsi;
            steer(C, fth, dk.x, v0, (float)Sv, (float)Sp, &o, &vel, &dv, &dpsi);
            float so, co;
            sincos_f32(o, &so, &co);
            float nx, ny;
            euler_boundary(C, fx, fy, fr, vel, so, co, &o, &nx, &ny);
            const float4 moved = make_float4(nx, ny, fr, o);
            if (lane == 0) {
                newpose[kk & 3] = moved;
                st.x[gk] = nx;
                st.y[gk] = ny;
                st.theta[gk] = o;
                st.v[gk] = vel;
                if (want_diag) {
                    float* dd = diag + 6 * (int64_t)gk;
                    dd[0] = dv; dd[1] = dpsi;
                    dd[2] = C.phi_ok ? (float)((double)alp0 * -Bc) : 0.f;
                    dd[3] = C.phi_ok ? (float)((double)(alp0 * C.alp1) * Ec) : 0.f;
                    dd[4] = C.phi_ok ? (float)((double)bet0 * -Bs) : 0.f;
                    dd[5] = C.phi_ok ? (float)((double)(bet0 * C.bet1) * Es) : 0.f;
                }
            }
            // lane d (1, 2): the pair (k + d, k); lane 3: the pair (k + 2, k - 1)
            // at k - 1's new pose (agent s0 - 1's pairs came with the fields)
            bool drew = false;
            const int m = kk + (lane == 3 ? 2 : lane);
            if (lane >= 1 && lane <= 3 && m < n && (lane < 3 || kk >= 1)) {
                const float4 f = srec[m];
                const float4 ob = lane == 3 ? newpose[(kk - 1) & 3] : moved;
                if (push_culled(f, ob, C, cull_c)) {
                    push_culled_check(f, ob, rp, count, verify, n_pairs, n_disagree);
                } else {
                    PairF32 pr;
                    if (decide_pair(pr, su[m], f, ob, rp, C, count, verify, n_pairs, n_recheck, n_disagree)) {
                        raster(rows[m], pr.start, pr.end, C.R);
                        drew = true;
                    }
                }
            }
            const uint64_t dm = __ballot(drew && lane == 1);
            if (lane == 0) dirty[buf ^ 1] = dm != 0 ? 1 : 0;
        } else if (wave == 1) {
            if (kk + 1 < n) sums(kk + 1, buf ^ 1);   // lacks only the pair (k + 1, k)
        } else if (wave >= 3 && wave <= 5) {
            if (!PLAIN && bits_out && wave == 3 && lane < C.W) bits_out[C.field_offset + (int64_t)k * C.W + lane] = rows[kk][lane];
            // agent k - 1's new pose against the segment's agents from k + 3
            // on: the cut-off test, the rest listed for the next step
            const int c = wave - 3;
            int cnt = 0;   // wave-uniform
            if (kk >= 1) {
                const float4 o = newpose[(kk - 1) & 3];
                for (int mb = kk + 3 + c * 64; mb < n; mb += kChainCullWaves * 64) {
                    const int m = mb + lane;
                    bool keep = false;
                    if (m < n) {
                        const float4 f = srec[m];
                        if (!(o.x == f.x && o.y == f.y)) {   // coincident pairs are skipped (vf_supcalc.py:57)
                            keep = !push_culled(f, o, C, cull_c);
                            if (!keep) push_culled_check(f, o, rp, count, verify, n_pairs, n_disagree);
                        }
                    }
                    const uint64_t mk = __ballot(keep);
                    if (keep) cq[buf][c][cnt + (int)__popcll(mk & ((1ull << lane) - 1ull))] = (uint16_t)m;
                    cnt += (int)__popcll(mk);
                }
            }
            if (lane == 0) cn[buf][c] = cnt;
        } else if (kk >= 2) {
            // the pairs listed in the previous step, at agent k - 2's new pose
            const int pb = buf ^ 1;
            const int c0 = cn[pb][0], c1 = cn[pb][1], c2 = cn[pb][2];
            const float4 o = newpose[(kk - 2) & 3];
            const int d = wave == 2 ? 0 : wave - 5;   // waves 2, 6, 7
            for (int e = d * 64 + lane; e < c0 + c1 + c2; e += kChainCullWaves * 64) {
                const int mm = e < c0 ? cq[pb][0][e] : (e < c0 + c1 ? cq[pb][1][e - c0] : cq[pb][2][e - c0 - c1]);
                push_pair(rows[mm], su[mm], srec[mm], o, rp, C, count, verify, n_pairs, n_recheck, n_disagree);
            }
        }
        chain_barrier();
    }
    if (count) {
        if (n_pairs) atomicAdd(&counters[0], n_pairs);
        if (n_recheck) atomicAdd(&counters[1], n_recheck);
        if (n_disagree) atomicAdd(&counters[2], n_disagree);
    }
}

// The new poses of the agents [s0, s0 + n) (just moved) into the global
// fields of the agents from s1 = s0 + n on (at their start poses).  Wave
// (j, c): moved agent s0 + j and chunk c of the later agents -- grid cell c
// of the 3 x 3 around its new centre, or the range [s1 + c R, s1 + (c + 1) R)
// (R = kChainScanRange) without a grid.
__global__ void __launch_bounds__(64 * kChainPushWaves)
vf_step_f32_chain_push_kernel(const VFRep* __restrict__ rp, abm_vf_state_f32 st, uint32_t* __restrict__ fields,
                              const float2* __restrict__ ustart, uint32_t flags,
                              unsigned long long* __restrict__ counters, float cull_c,
                              const WideGrid* __restrict__ grid, const float4* __restrict__ sorted,
                              const int32_t* __restrict__ cell_start, int s0, int n, int n_chunks) {
    __shared__ uint32_t queue[kChainPushWaves][kChainQueue];
    const RepConst C = rep_const(rp);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool count = (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) != 0 && counters;
    const bool verify = (flags & ABM_VERIFY_FAST_PATH) != 0 && counters != nullptr;
    const WideGrid G = *grid;
    const bool use_grid = G.ok && !count && !verify;
    const int item = blockIdx.x * kChainPushWaves + wave;
    const int j = item / n_chunks, c = item - j * n_chunks;
    const int s1 = s0 + n;
    // whole waves only (no barrier below): past the moved agents, or past the
    // nine cells of the grid
    if (j >= n || (use_grid && c >= 9)) return;
    const int gj = C.a0 + s0 + j;
    const float4 o = make_float4(st.x[gj], st.y[gj], (float)st.radius[gj], st.theta[gj]);
    uint32_t* wq = queue[wave];
    unsigned long long n_pairs = 0, n_recheck = 0, n_disagree = 0;
    int beg, end;
    if (use_grid) {
        const int oc = grid_cell(G, o.x + o.z, o.y + o.z);
        const int oiy = oc / G.nx, oix = oc - oiy * G.nx;
        const int ty = c / 3, tx = c - 3 * ty;
        int cy = oiy + ty - 1, cx = oix + tx - 1;   // the 3 x 3 cells, wrapped
        cy = cy < 0 ? G.ny - 1 : (cy >= G.ny ? 0 : cy);
        cx = cx < 0 ? G.nx - 1 : (cx >= G.nx ? 0 : cx);
        const int cell = cy * G.nx + cx;
        beg = cell_start[cell];
        end = cell_start[cell + 1];
    } else {
        beg = s1 + c * kChainScanRange;
        end = min(C.N, beg + kChainScanRange);
    }
    // focal agent m (record i), at its start pose
    auto focal = [&](int i, int& m, bool with_theta) -> float4 {
        if (use_grid) {
            const float4 s = sorted[i];
            m = __float_as_int(s.w);
            return make_float4(s.x, s.y, s.z, with_theta ? st.theta[C.a0 + m] : 0.f);
        }
        m = i;
        const int g = C.a0 + i;
        return make_float4(st.x[g], st.y[g], (float)st.radius[g], with_theta ? st.theta[g] : 0.f);
    };
    auto decide = [&](int head, int tail) {
        if (head + lane < tail) {
            int m;
            const float4 f = focal((int)wq[(head + lane) & (kChainQueue - 1)], m, true);
            push_pair(fields + (size_t)m * C.W, ustart[m], f, o, rp, C, count, verify, n_pairs, n_recheck,
                      n_disagree);
        }
    };
    int head = 0, tail = 0;   // wave-uniform ring positions
    for (int b = beg; b < end; b += 64) {
        const int i = b + lane;
        bool keep = false;
        if (i < end) {
            int m;
            const float4 f = focal(i, m, verify);
            if (m >= s1 && !(o.x == f.x && o.y == f.y)) {   // coincident pairs are skipped (vf_supcalc.py:57)
                keep = !push_culled(f, o, C, cull_c);
                if (!keep) push_culled_check(f, o, rp, count, verify, n_pairs, n_disagree);
            }
        }
        const uint64_t mk = __ballot(keep);
        if (keep) wq[(tail + (int)__popcll(mk & ((1ull << lane) - 1ull))) & (kChainQueue - 1)] = (uint32_t)i;
        tail += (int)__popcll(mk);
        wave_lds_sync();
        if (tail - head >= 64) {
            decide(head, tail);
            head += 64;
            wave_lds_sync();
        }
    }
    decide(head, tail);
    if (count) {
        if (n_pairs) atomicAdd(&counters[0], n_pairs);
        if (n_recheck) atomicAdd(&counters[1], n_recheck);
        if (n_disagree) atomicAdd(&counters[2], n_disagree);
    }
}

template <int RPLW>
int launch_chain_start(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t flags, hipStream_t s,
                       int rep, float cull_c, int n_agents) {
    const size_t lds = lds_bytes_wide(RPLW);
    const bool plain = (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) == 0;
    auto kern = plain ? vf_step_f32_wide_kernel<RPLW, true, true> : vf_step_f32_wide_kernel<RPLW, true, false>;
    if (lds > 65536) {
        if (int rc = check_hip(hipFuncSetAttribute((const void*)kern,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                               "hipFuncSetAttribute"))
            return rc;
    }
    const int blocks = (n_agents + kWideWaves - 1) / kWideWaves;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * kWideWaves), lds, s, b->reps_dev + rep, b->tab32, *st,
                       (float4*)nullptr, (uint32_t*)nullptr, (float*)nullptr, flags, b->counters, cull_c,
                       (const WideGrid*)b->grid_param, (const float4*)b->grid_sorted,
                       (const int32_t*)b->grid_start, b->chain_fields, (float2*)b->chain_u);
    return check_hip(hipGetLastError(), "vf_step_f32_wide_kernel (start pairs) launch");
}

// Whether the chain step can run the batch (its buffers, R up to 2048), and
// whether it does by default: every replica above kWideThreshold agents (a
// single large flock).
bool chain_ready(const abm_vf_batch* b) {
    return b->chain_fields != nullptr && b->chain_u != nullptr && b->grid_sorted != nullptr &&
           b->grid_start != nullptr && b->grid_param != nullptr && b->n_replicas > 0 && b->max_rays <= 2048;
}
bool chain_fits(const abm_vf_batch* b) { return chain_ready(b) && b->min_agents > kWideThreshold; }

int launch_chain(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag, uint32_t flags,
                 hipStream_t s) {
    const bool counting = (flags & (ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) != 0;
    const bool plain = bits == nullptr && !counting && (flags & ABM_WANT_DIAG) == 0;
    const int rpl = push_rpl(b->max_rays);
    for (int rep = 0; rep < b->n_replicas; ++rep) {
        const VFRep& h = b->reps_host[rep];
        const int N = h.n_agents;
        const float cull_c = (float)h.cull_c;
        hipLaunchKernelGGL(wide_grid_kernel, dim3(1), dim3(1024), 0, s, b->reps_dev + rep, *st, cull_c,
                           b->grid_sorted, b->grid_start, (WideGrid*)b->grid_param);
        if (int rc = check_hip(hipGetLastError(), "wide_grid_kernel launch")) return rc;
        const int need = (h.n_rays + 63) / 64;
        int rc;
        if (need <= 4) rc = launch_chain_start<4>(b, st, flags, s, rep, cull_c, N);
        else if (need <= 8) rc = launch_chain_start<8>(b, st, flags, s, rep, cull_c, N);
        else if (need <= 16) rc = launch_chain_start<16>(b, st, flags, s, rep, cull_c, N);
        else rc = launch_chain_start<32>(b, st, flags, s, rep, cull_c, N);
        if (rc) return rc;
        for (int s0 = 0; s0 < N; s0 += kChainSeg) {
            const int n = std::min(kChainSeg, N - s0);
            // rays per lane of the summing wave: the push kernel's rule for the batch
#define ABM_SEGMENT(RPLV)                                                                                   \
    if (rpl == RPLV)                                                                                        \
        hipLaunchKernelGGL((plain ? vf_step_f32_segment_kernel<RPLV, true>                                  \
                                  : vf_step_f32_segment_kernel<RPLV, false>),                               \
                           dim3(1), dim3(kChainThreads), 0, s,                                              \
                           b->reps_dev + rep, b->tab32, *st, (const uint32_t*)b->chain_fields,             \
                           (const float2*)b->chain_u, bits, diag, flags, b->counters, cull_c, s0, n);
            ABM_SEGMENT(4) ABM_SEGMENT(8) ABM_SEGMENT(16) ABM_SEGMENT(32)
#undef ABM_SEGMENT
            if (int rc2 = check_hip(hipGetLastError(), "vf_step_f32_segment_kernel launch")) return rc2;
            const int s1 = s0 + n;
            if (s1 >= N) break;
            // chunks per moved agent: enough for the nine cells of a grid and
            // for the ranges of the later agents (the device's grid decides
            // which apply; the other waves return at once)
            const int n_ranges = (N - s1 + kChainScanRange - 1) / kChainScanRange;
            const int n_chunks = counting ? n_ranges : std::max(9, n_ranges);
            const int items = n * n_chunks;
            hipLaunchKernelGGL(vf_step_f32_chain_push_kernel, dim3((items + kChainPushWaves - 1) / kChainPushWaves),
                               dim3(64 * kChainPushWaves), 0, s, b->reps_dev + rep, *st, b->chain_fields,
                               (const float2*)b->chain_u, flags, b->counters, cull_c, (const WideGrid*)b->grid_param,
                               (const float4*)b->grid_sorted, (const int32_t*)b->grid_start, s0, n, n_chunks);
            if (int rc2 = check_hip(hipGetLastError(), "vf_step_f32_chain_push_kernel launch")) return rc2;
        }
    }
    b->last_kernels |= ABM_KERNEL_CHAIN;
    return ABM_OK;
}

size_t lds_bytes_f32(int maxN, int maxR, int nw, int rpl, bool sync_order) {
    const int maxW = (maxR + 31) >> 5;
    size_t b = sizeof(RayTab32) * (size_t)rpl * 64 * nw + sizeof(float4) * (size_t)maxN * (sync_order ? 3 : 2) +
               sizeof(double) * (2 * nw + 2) + sizeof(int32_t) * ((size_t)2 * maxN + 4) +
               sizeof(uint32_t) * (maxW + 2);
    return (b + 15) & ~(size_t)15;
}

// The per-replica kernel over the replicas order[0 .. n_rep), the largest
// with maxN agents and maxR rays (split: the small size class of a batch
// whose large class goes to the 16-wave kernel).
template <int NW, int RPL>
int launch_nw(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
              uint32_t flags, hipStream_t s, const int32_t* order, int n_rep, int maxN, int maxR,
              bool split) {
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;
    const size_t lds = lds_bytes_f32(maxN, maxR, NW, RPL, sync_order);
    if (lds > (size_t)b->dev_max_lds)
        return set_error(ABM_ERR_UNSUPPORTED,
                         "abm_vf_step_f32: %d agents x %d rays need %zu B of LDS (device max %d)",
                         maxN, maxR, lds, b->dev_max_lds);
    const bool plain = bits == nullptr && (flags & (ABM_WANT_DIAG | ABM_COUNT_RECHECKS | ABM_VERIFY_FAST_PATH)) == 0;
    auto kern = plain ? vf_step_f32_kernel<NW, RPL, true> : vf_step_f32_kernel<NW, RPL, false>;
    const int slot = NW == 1 ? 0 : (NW == 2 ? 1 : 2);
    int& per_cu = b->occ_f32[split ? 1 : 0][sync_order ? 1 : 0][plain ? 1 : 0][slot];
    if (per_cu == 0) {   // first launch of this shape: set the LDS limit, read the occupancy
        if (lds > 65536) {
            if (int rc = check_hip(hipFuncSetAttribute((const void*)kern,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)lds), "hipFuncSetAttribute"))
                return rc;
        }
        int occ = 0;
        if (int rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64 * NW, lds),
                               "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
            return rc;
        per_cu = occ < 1 ? 1 : occ;
    }
    const int n_cu = b->n_cu;
    long grid = (long)per_cu * n_cu;
    if (grid > n_rep) grid = n_rep;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * NW), lds, s, b->reps_dev, order,
                       n_rep, b->tab32, *st, bits, diag, flags, b->counters, maxN);
    b->last_kernels |= ABM_KERNEL_PER_REPLICA;
    return check_hip(hipGetLastError(), "vf_step_f32_kernel launch");
}

}  // namespace

bool push_fits(const abm_vf_batch* b, bool all) { return push_fits_in(b, all); }

// Replicas per wave of the grouped kernel for this batch, or 0 for the
// per-replica kernel.  Automatic choice (measured on MI355X): batches with at
// least 8 replicas per CU run grouped, one replica per wave (four waves per
// SIMD), or four per wave when that still gives every SIMD two waves; so do
// batches with 4 to 8 replicas per CU of at most 128 agents (one replica per
// wave: 1024 replicas of 50 agents at R = 512 take 0.070 ms grouped against
// 0.095 ms on the per-replica kernel, of 100 agents at R = 1024 0.201 against
// 0.213; of 200 agents the per-replica kernel wins, 0.546 against 0.597);
// smaller batches are latency-bound and take the per-replica kernel with
// several waves per replica.
int selected_group_size(const abm_vf_batch* b) {
    if (!b->group_ok || b->group_g < 0) return 0;
    if (b->group_g > 0) return b->group_g;
    if (b->waves_per_group > 0) return 0;
    if (b->n_replicas < 4 * b->n_cu) return 0;
    if (b->n_replicas < 8 * b->n_cu) return b->max_agents <= 128 ? 1 : 0;
    if (b->gplan[2].n_groups >= 8 * b->n_cu) return 4;
    return 1;
}

// Waves per replica of the per-replica kernel: four while the batch gives each
// CU at most two such workgroups, two while it fits at 16 waves per CU, else
// one (measured on MI355X: 1024 replicas of 100 agents take 0.223 ms on two
// waves and 0.256 ms on four; 512 take 0.216 and 0.219).
static int default_waves(const abm_vf_batch* b, int maxN, int n_rep) {
    if (maxN > 64 && (long)n_rep * 4 <= 8L * b->n_cu) return 4;
    if (maxN > 32 && (long)n_rep * 2 <= 16L * b->n_cu) return 2;
    return 1;
}

// The per-replica kernel with nw waves (0 = automatic) over order[0 .. n_rep).
static int launch_per_replica(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                              uint32_t flags, hipStream_t s, int nw, const int32_t* order, int n_rep,
                              int maxN, int maxR, bool split) {
    if (nw == 0) nw = default_waves(b, maxN, n_rep);
    const int need = (maxR + 64 * nw - 1) / (64 * nw);   // rays per lane
    const int rpl = need <= 4 ? 4 : (need <= 8 ? 8 : (need <= 16 ? 16 : 32));
#define ABM_LAUNCH(NWV, RPLV) \
    if (nw == NWV && rpl == RPLV) \
        return launch_nw<NWV, RPLV>(b, st, bits, diag, flags, s, order, n_rep, maxN, maxR, split);
    ABM_LAUNCH(1, 4) ABM_LAUNCH(1, 8) ABM_LAUNCH(1, 16) ABM_LAUNCH(1, 32)
    ABM_LAUNCH(2, 4) ABM_LAUNCH(2, 8) ABM_LAUNCH(2, 16) ABM_LAUNCH(2, 32)
    ABM_LAUNCH(4, 4) ABM_LAUNCH(4, 8) ABM_LAUNCH(4, 16) ABM_LAUNCH(4, 32)
#undef ABM_LAUNCH
    return set_error(ABM_ERR_INVALID, "abm_vf_step_f32: no launch shape for %d waves", nw);
}

int launch_vf_step_f32(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                       uint32_t flags, hipStream_t s) {
    b->last_kernels = 0;
    if (b->max_rays > 2048)
        return set_error(ABM_ERR_UNSUPPORTED,
                         "abm_vf_step_f32: n_rays up to 2048 (got %d); use abm_vf_step_f64", b->max_rays);
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;
    const bool kahan = (flags & ABM_KAHAN_SUMS) != 0;
    if (kahan && b->max_agents > kWideThreshold)
        return set_error(ABM_ERR_UNSUPPORTED, "abm_vf_step_f32: ABM_KAHAN_SUMS is implemented by the "
                         "grouped kernel only (replicas of at most %d agents)", kGroupMaxAgents);
    if (b->max_agents > kWideThreshold) {
        // reference order: the chain step for single large flocks (unless the
        // 16-wave kernel is forced), else one workgroup per replica with every
        // pose in LDS
        if (!sync_order && b->waves_per_group != 16 && chain_fits(b))
            return launch_chain(b, st, bits, diag, flags, s);
        if (!sync_order)
            return launch_big_any(b, st, bits, diag, flags, s, b->order_dev, b->n_replicas, b->max_agents,
                                  b->max_rays);
        if (!b->wide_out)
            return set_error(ABM_ERR_INVALID, "abm_vf_step_f32: batch has no wide-step buffer");
        const int need = (b->max_rays + 63) / 64;
        if (need <= 4) return launch_wide<4>(b, st, bits, diag, flags, s);
        if (need <= 8) return launch_wide<8>(b, st, bits, diag, flags, s);
        if (need <= 16) return launch_wide<16>(b, st, bits, diag, flags, s);
        return launch_wide<32>(b, st, bits, diag, flags, s);
    }
    // Kahan sums run on the grouped kernel even where the automatic choice
    // (or a forced launch shape) would take the per-replica kernel
    const int G0 = selected_group_size(b);
    const int G = kahan && b->group_ok && G0 == 0 ? 1 : G0;
    if (G > 0) return launch_vf_group_f32(b, G, st, bits, diag, flags, s);
    if (kahan)
        return set_error(ABM_ERR_UNSUPPORTED, "abm_vf_step_f32: ABM_KAHAN_SUMS is implemented by the "
                         "grouped kernel only, which this batch does not take (one R up to %d rays, "
                         "replicas of at most %d agents)", kGroupMaxRays, kGroupMaxAgents);
    const int nw = b->waves_per_group;
    // the 16-wave and push kernels and the chain step move agents in the
    // reference order only; the synchronous order takes the per-replica kernel
    if (sync_order && (nw == 16 || nw == 32 || nw == 64))
        return set_error(ABM_ERR_UNSUPPORTED, "abm_vf_step_f32: the %s runs the reference order only "
                         "(ABM_ORDER_SYNCHRONOUS takes the per-replica kernel, abm_vf_set_launch(0))",
                         nw == 16 ? "16-wave kernel" : (nw == 32 ? "push kernel" : "chain step"));
    if (nw == 64) {   // the chain step for every replica
        if (!chain_ready(b))
            return set_error(ABM_ERR_INVALID, "abm_vf_step_f32: the chain step has no buffers (abm_vf_set_launch(64))");
        return launch_chain(b, st, bits, diag, flags, s);
    }
    if (nw == 32)   // the push kernel for every replica
        return launch_push_any(b, st, bits, diag, flags, s, b->order_dev, b->n_replicas, true);
    if (nw == 16)   // the 16-wave kernel for every replica
        return launch_big_any(b, st, bits, diag, flags, s, b->order_dev, b->n_replicas, b->max_agents,
                              b->max_rays);
    const int nb = b->n_big_class;
    const bool push_ok = !sync_order && nb > 0 && push_fits(b, false);
    // The push kernel holds one replica per CU (eight waves of up to 256
    // VGPRs), so it shortens the step only while the large class's agents,
    // spread over the CUs, do not outlast the chain of its largest replica;
    // beyond that (many replicas of similar size) the per-replica kernel's
    // occupancy wins.
    const bool push = push_ok && b->big_sum_agents <= (int64_t)b->big_max_agents * b->n_cu;
    if (nw == 0 && nb > 0 && !sync_order && (push || (!push_ok && nb < b->n_replicas))) {
        // size classes: the small replicas on the per-replica kernel, then
        // the large ones, whose id-order chains are the step's critical
        // path, on the push kernel (pairs off the chain) where their fields
        // fit in LDS, else on the 16-wave kernel
        if (nb < b->n_replicas) {
            if (int rc = launch_per_replica(b, st, bits, diag, flags, s, 0, b->order_dev + nb,
                                            b->n_replicas - nb, b->small_max_agents, b->small_max_rays, true))
                return rc;
        }
        if (push) return launch_push_any(b, st, bits, diag, flags, s, b->order_dev, nb, false);
        return launch_big_any(b, st, bits, diag, flags, s, b->order_dev, nb, b->big_max_agents,
                              b->big_max_rays);
    }
    return launch_per_replica(b, st, bits, diag, flags, s, nw, b->order_dev, b->n_replicas, b->max_agents,
                              b->max_rays, false);
}

}  // namespace abm
