// vf_f32.hip -- fused float32 VSWRM step for batches of replicas.
//
// Persistent workgroups of NW waves; each takes replicas (largest first) and
// runs the reference-order step (agent k sees agents <k moved,
// vf_sims.py:302) with the replica's poses and its float32 ray table in LDS.
//
// Per focal agent k:
//   pairs   -- lanes over the other agents j.  The centre differences, torus
//              image decisions and the bin arithmetic are float64 adds and
//              compares on the float32 state, hence identical to the float64
//              reference arithmetic; only the transcendentals (atan2 for the
//              closed angle, atan for the half-width, sqrt) are float32.  When
//              the float32 angle lies within a guard band of a bin tie or of the
//              +-pi wrap, or the half-width within its band of an integer, the
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
#include "vf_batch.hpp"
#include "vf_exact.hpp"

namespace abm {

namespace {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kTwoPiF = 6.28318530717958647692f;
constexpr float kHalfPiF = 1.57079632679489661923f;
constexpr float kThreeHalfPiF = 4.71238898038468985769f;

// Guard bands.  Float32 angle error (sincos of the heading, cross/dot,
// atan2f) is below 1e-6 rad and the float64 reference angle is within 5e-8
// rad of the exact one, so 4e-6 rad covers both with margin; the half-width
// guard is relative (atanf, the division and sqrt: < 1e-6 relative).
constexpr double kAngleGuard = 4e-6;
constexpr double kHalfGuardRel = 4e-6;
constexpr double kHalfGuardAbs = 1e-6;

// Cross-lane sum of a double over the wave with DPP moves only (no LDS):
// quad xor 1, xor 2, half-row mirror, row mirror, then row_bcast15 into rows
// 1 and 3 and row_bcast31 into rows 2 and 3; lane 63 ends with the total.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double wave_total_f64(double v) {
    v += dpp_f64<0xB1>(v);         // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);         // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);        // row_half_mirror
    v += dpp_f64<0x140>(v);        // row_mirror
    v += dpp_f64<0x142, 0xa>(v);   // row_bcast15 -> rows 1, 3
    v += dpp_f64<0x143, 0xc>(v);   // row_bcast31 -> rows 2, 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

__device__ inline float lane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ inline uint32_t low_mask(int n) { return n >= 32 ? 0xffffffffu : ((1u << n) - 1u); }

__device__ inline uint32_t field_bit(const uint32_t* F, int r) { return (F[r >> 5] >> (r & 31)) & 1u; }

// OR the rays [a, b) into the bit field.
__device__ inline void or_range(uint32_t* F, int a, int b) {
    while (a < b) {
        const int w = a >> 5;
        const int lo = a & 31;
        const int hi = (b - (w << 5)) < 32 ? (b - (w << 5)) : 32;
        const uint32_t m = low_mask(hi) & ~low_mask(lo);
        atomicOr(&F[w], m);
        a = (w + 1) << 5;
    }
}

__device__ inline float wrap_2pi(float o) {
    if (o < 0.f) o = kTwoPiF + o;
    if (o > kTwoPiF) o = o - kTwoPiF;
    return o;
}

struct Smem {
    RayTab32* tab;
    float *px, *py, *th, *vv, *rad, *ox, *oy;
    int32_t* rlist;     // pairs queued for the float64 decision
    int32_t* rfast;     // their float32 decision (verify mode)
    uint32_t* F;
    double* red;
    int32_t* nre;
};

__device__ inline Smem carve(void* base, int maxN, int maxR, int nw, bool sync_order) {
    Smem s;
    char* p = reinterpret_cast<char*>(base);
    s.tab = reinterpret_cast<RayTab32*>(p); p += sizeof(RayTab32) * maxR;
    s.red = reinterpret_cast<double*>(p); p += sizeof(double) * (4 * nw + 8);
    s.px = reinterpret_cast<float*>(p); p += sizeof(float) * maxN;
    s.py = reinterpret_cast<float*>(p); p += sizeof(float) * maxN;
    s.th = reinterpret_cast<float*>(p); p += sizeof(float) * maxN;
    s.vv = reinterpret_cast<float*>(p); p += sizeof(float) * maxN;
    s.rad = reinterpret_cast<float*>(p); p += sizeof(float) * maxN;
    if (sync_order) {
        s.ox = reinterpret_cast<float*>(p); p += sizeof(float) * maxN;
        s.oy = reinterpret_cast<float*>(p); p += sizeof(float) * maxN;
    } else {
        s.ox = s.px; s.oy = s.py;
    }
    s.rlist = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * maxN;
    s.rfast = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * maxN;
    s.nre = reinterpret_cast<int32_t*>(p); p += sizeof(int32_t) * 4;
    s.F = reinterpret_cast<uint32_t*>(p);
    return s;
}

// Float32 decision of one pair; returns false when the pair must be
// re-decided in float64 (inside a guard band, or degenerate).
__device__ inline bool fast_pair(float u1x, float u1y, double cx, double cy, float ojx, float ojy,
                                 float ojr, bool torus, const VFRep& P, double inv_step,
                                 double half_scale, int R, int* start, int* end, int* drawn) {
    double ocx = (double)ojx + (double)ojr;
    double ocy = (double)ojy + (double)ojr;
    double dx = ocx - cx, dy = ocy - cy;
    if (torus) {
        if (fabs(dx) > P.width / 2.0) {
            if (cx < ocx) ocx -= P.width;
            else if (cx > ocx) ocx += P.width;
        }
        if (fabs(dy) > P.height / 2.0) {
            if (cy < ocy) ocy -= P.height;
            else if (cy > ocy) ocy += P.height;
        }
        dx = ocx - cx;
        dy = ocy - cy;
    }
    const float fdx = (float)dx, fdy = (float)dy;
    const float d = sqrtf(fdx * fdx + fdy * fdy);
    const float a = atan2f(u1x * fdy - u1y * fdx, u1x * fdx + u1y * fdy);
    const float q = atanf(ojr / d);
    const double xb = (-(double)a + kPi) * inv_step;   // bin coordinate of ca = -a
    const double fl = floor(xb);
    const double fb = xb - fl;
    const double half = (double)q * half_scale;        // proj_size / 2
    const double hfl = floor(half);
    const double hfrac = half - hfl;
    const double hg = kHalfGuardRel * half + kHalfGuardAbs;
    const bool ok = (d > 0.0f) && fabs((double)a) < kPi - kAngleGuard &&
                    fabs(fb - 0.5) >= kAngleGuard * inv_step && hfrac >= hg && hfrac <= 1.0 - hg;
    int t = (int)fl + (fb >= 0.5 ? 1 : 0);
    t = t < 0 ? 0 : (t > R - 1 ? R - 1 : t);
    const int h = (int)hfl;
    *start = t - h;
    *end = t + h;
    *drawn = ((P.fov_px0 < *start && *start < P.fov_px1) ||
              (P.fov_px0 < *end && *end < P.fov_px1)) ? 1 : 0;
    return ok;
}

__device__ __noinline__ void exact_pair(float fx, float fy, float fth, float fr, float ojx, float ojy,
                                        float ojr, bool torus, const VFRep& P, int* start, int* end,
                                        int* drawn) {
    const Focal f = make_focal((double)fx, (double)fy, (double)fth, (double)fr);
    const PairVF e = pair_vf_exact(f, (double)ojx, (double)ojy, (double)ojr, torus ? 1 : 0,
                                   P.width, P.height, P.n_rays, P.step, P.fov_px0, P.fov_px1);
    *start = e.start; *end = e.end; *drawn = e.skipped ? 0 : e.drawn;
}

template <int NW, int RPL>
__global__ void __launch_bounds__(64 * NW)
vf_step_f32_kernel(const VFRep* __restrict__ reps, const int32_t* __restrict__ order, int n_rep,
                   const RayTab32* __restrict__ tab32, abm_vf_state_f32 st,
                   uint32_t* __restrict__ bits_out, float* __restrict__ diag, uint32_t flags,
                   unsigned long long* __restrict__ counters, int maxN, int maxR) {
    extern __shared__ __align__(16) double smem_d[];
    constexpr int NT = 64 * NW;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;
    const bool want_diag = (flags & ABM_WANT_DIAG) != 0 && diag != nullptr;
    const bool count = (flags & ABM_COUNT_RECHECKS) != 0 && counters != nullptr;
    const bool verify = (flags & ABM_VERIFY_FAST_PATH) != 0 && counters != nullptr;
    Smem S = carve(smem_d, maxN, maxR, NW, sync_order);
    int loaded_tab = -1;
    unsigned long long n_pairs = 0, n_recheck = 0, n_disagree = 0;

    for (int idx = blockIdx.x; idx < n_rep; idx += gridDim.x) {
        const VFRep P = reps[order[idx]];
        const int N = P.n_agents, R = P.n_rays, W = P.words, a0 = P.agent_offset;
        const bool torus = P.boundary == ABM_BOUNDARY_INFINITE;
        if (P.table_offset != loaded_tab) {
            for (int r = tid; r < R; r += NT) S.tab[r] = tab32[P.table_offset + r];
            loaded_tab = P.table_offset;
        }
        for (int i = tid; i < N; i += NT) {
            S.px[i] = st.x[a0 + i];
            S.py[i] = st.y[a0 + i];
            S.th[i] = st.theta[a0 + i];
            S.vv[i] = st.v[a0 + i];
            S.rad[i] = (float)st.radius[a0 + i];
            if (sync_order) { S.ox[i] = S.px[i]; S.oy[i] = S.py[i]; }
        }
        const double inv_step = 1.0 / P.step;
        const double half_scale = (double)R / kTwoPi;
        const float fW = (float)P.width, fH = (float)P.height, fpad = (float)P.pad;
        const float gam = (float)P.gam, alp1 = (float)P.alp1, bet1 = (float)P.bet1;
        const int rs = tid * RPL;            // this lane's ray span [rs, rs + rn)
        const int rn = rs >= R ? 0 : (R - rs < RPL ? R - rs : RPL);
        if (tid == 0) S.nre[0] = 0;
        __syncthreads();
        // heading of agent 0 (later headings come with the previous agent's motion)
        float sth, cth;
        sincosf(S.th[0], &sth, &cth);

        for (int k = 0; k < N; ++k) {
            for (int w = tid; w < W; w += NT) S.F[w] = 0u;
            const float fx = S.px[k], fy = S.py[k], fth = S.th[k], fr = S.rad[k];
            const float u1x = cth, u1y = -sth;           // unit(v1) = (cos th, -sin th)
            const double cx = (double)fx + (double)fr;
            const double cy = (double)fy + (double)fr;
            __syncthreads();

            // ------------------------------------------------ pairs, float32
            for (int j = tid; j < N; j += NT) {
                if (j == k) continue;
                const float ojx = S.ox[j], ojy = S.oy[j], ojr = S.rad[j];
                if (ojx == fx && ojy == fy) continue;          // coincident (:57)
                int start, end, drawn;
                const bool ok = fast_pair(u1x, u1y, cx, cy, ojx, ojy, ojr, torus, P, inv_step,
                                          half_scale, R, &start, &end, &drawn);
                if (count || verify) ++n_pairs;
                if (!ok || verify) {
                    const int slot = atomicAdd(S.nre, 1);
                    S.rlist[slot] = ok ? j : (j | (1 << 30));
                    S.rfast[slot] = drawn ? ((start + 2 * R) | ((end + 2 * R) << 15)) : -1;
                    if (!ok) continue;
                }
                if (drawn) {
                    int p0, p1, q0, q1;
                    vf_pieces(start, end, R, &p0, &p1, &q0, &q1);
                    or_range(S.F, p0, p1);
                    or_range(S.F, q0, q1);
                }
            }
            __syncthreads();
            // ------------------------------------- pairs re-decided in float64
            const int nre = S.nre[0];
            if (nre > 0) {
                for (int i = tid; i < nre; i += NT) {
                    const int e = S.rlist[i];
                    const int j = e & ((1 << 30) - 1);
                    const bool forced = (e >> 30) & 1;
                    int start, end, drawn;
                    exact_pair(fx, fy, fth, fr, S.ox[j], S.oy[j], S.rad[j], torus, P, &start,
                               &end, &drawn);
                    if (forced) {
                        if (count || verify) ++n_recheck;
                        if (drawn) {
                            int p0, p1, q0, q1;
                            vf_pieces(start, end, R, &p0, &p1, &q0, &q1);
                            or_range(S.F, p0, p1);
                            or_range(S.F, q0, q1);
                        }
                    } else {
                        const int fast = S.rfast[i];
                        const int ex = drawn ? ((start + 2 * R) | ((end + 2 * R) << 15)) : -1;
                        if (fast != ex) ++n_disagree;
                    }
                }
                __syncthreads();
                if (tid == 0) S.nre[0] = 0;
            }

            // ------------------------------------------------ sums
            const uint32_t v0bit = S.F[0] & 1u;
            const uint32_t vlast = field_bit(S.F, R - 1);
            const bool case1 = v0bit == 1u && vlast == 0u;
            float bc = 0.f, bs = 0.f, ec = 0.f, es = 0.f;
            if (rn > 0) {
                const int w0 = rs >> 5;
                const uint32_t lo = S.F[w0];
                const uint32_t hi = (RPL > 1 && w0 + 1 < W) ? S.F[w0 + 1] : 0u;
                const uint64_t v64 = (((uint64_t)hi << 32) | lo) >> (rs & 31);
                const uint32_t m = low_mask(rn);
                const uint32_t cur = (uint32_t)v64 & m;
                uint32_t spike;
                if (case1) {
                    const uint32_t prev = field_bit(S.F, rs == 0 ? R - 1 : rs - 1);
                    spike = (cur ^ ((cur << 1) | prev)) & m;
                } else {
                    const int nx = rs + rn;
                    const uint32_t next = field_bit(S.F, nx >= R ? 0 : nx);
                    spike = (cur ^ ((cur >> 1) | (next << (rn - 1)))) & m;
                }
#pragma unroll
                for (int i = 0; i < RPL; ++i) {
                    if (i < rn) {
                        const RayTab32 T = S.tab[rs + i];
                        const bool vb = (cur >> i) & 1u;
                        const bool sb = (spike >> i) & 1u;
                        bc += vb ? T.wcos : 0.f;
                        bs += vb ? T.wsin : 0.f;
                        ec += sb ? T.cos : 0.f;
                        es += sb ? T.sin : 0.f;
                    }
                }
            }
            const int ag = a0 + k;
            const float alp0 = st.alp0 ? st.alp0[ag] : (float)P.alp0;
            const float bet0 = st.bet0 ? st.bet0[ag] : (float)P.bet0;
            const float v0 = st.v0 ? st.v0[ag] : (float)P.v0;
            // blob terms: trapz(cos(PHI) * (-V)) = -sum(w cos V)
            double Sv, Sp, Bc = 0.0, Bs = 0.0, Ec = 0.0, Es = 0.0;
            if (!want_diag) {
                Sv = wave_total_f64((double)(alp0 * (-bc) + (alp0 * alp1) * ec));
                Sp = wave_total_f64((double)(bet0 * (-bs) + (bet0 * bet1) * es));
            } else {
                Bc = wave_total_f64((double)bc); Bs = wave_total_f64((double)bs);
                Ec = wave_total_f64((double)ec); Es = wave_total_f64((double)es);
                Sv = (double)alp0 * -Bc + (double)(alp0 * alp1) * Ec;
                Sp = (double)bet0 * -Bs + (double)(bet0 * bet1) * Es;
            }
            if (NW > 1) {
                if (lane == 0) {
                    double* rr = S.red + 6 * (tid >> 6);
                    rr[0] = Sv; rr[1] = Sp; rr[2] = Bc; rr[3] = Bs; rr[4] = Ec; rr[5] = Es;
                }
                __syncthreads();
                Sv = Sp = Bc = Bs = Ec = Es = 0.0;
                for (int w = 0; w < NW; ++w) {
                    const double* rr = S.red + 6 * w;
                    Sv += rr[0]; Sp += rr[1]; Bc += rr[2]; Bs += rr[3]; Ec += rr[4]; Es += rr[5];
                }
            }
            if (bits_out) {
                uint32_t* dst = bits_out + P.field_offset + (int64_t)k * W;
                for (int w = tid; w < W; w += NT) dst[w] = S.F[w];
            }

            // ------------------------------------------------ motion
            float vel = S.vv[k];
            float dv = 0.f, dpsi = 0.f;
            if (P.phi_ok) {
                dv = gam * (v0 - vel) + (float)Sv;
                dpsi = (float)Sp;
            }
            float theta = dpsi;
            if (P.limit_movement) {
                const float lim = (float)P.max_th;
                if (fabsf(theta) > lim) theta = theta < 0.f ? -lim : lim;
            }
            float o = wrap_2pi(fth + theta);
            vel = vel + dv;
            if (P.limit_movement) {
                const float lim = (float)P.max_vel;
                if (fabsf(vel) > lim) vel = vel < 0.f ? -lim : lim;
            }
            // one sincos serves this agent's motion (lane 0) and the next
            // agent's heading (lane 1)
            const float ang = (lane == 1 && k + 1 < N) ? S.th[k + 1] : o;
            float s_l, c_l;
            sincosf(ang, &s_l, &c_l);
            const float so = lane_f32(s_l, 0), co = lane_f32(c_l, 0);
            sth = lane_f32(s_l, 1);
            cth = lane_f32(c_l, 1);
            float nx = fx + vel * co;
            float ny = fy - vel * so;
            const float xc = nx + fr, yc = ny + fr;
            if (torus) {
                if (xc < fpad) nx = fpad + fW - fr;
                else if (xc > fpad + fW) nx = fpad + fr;
                if (yc < fpad) ny = fpad + fH - fr;
                else if (yc > fpad + fH) ny = fpad + fr;
            } else {
                if (xc < fpad) {
                    nx = fpad - fr;
                    if (kHalfPiF <= o && o < kPiF) o -= kHalfPiF;
                    else if (kPiF <= o && o <= kThreeHalfPiF) o += kHalfPiF;
                    o = wrap_2pi(o);
                }
                if (xc > fpad + fW) {
                    nx = fpad + fW - fr - 1.f;
                    if (kThreeHalfPiF <= o && o < kTwoPiF) o -= kHalfPiF;
                    else if (0.f <= o && o <= kHalfPiF) o += kHalfPiF;
                    o = wrap_2pi(o);
                }
                if (yc < fpad) {
                    ny = fpad - fr;
                    if (kHalfPiF <= o && o <= kPiF) o += kHalfPiF;
                    else if (0.f <= o && o < kHalfPiF) o -= kHalfPiF;
                    o = wrap_2pi(o);
                }
                if (yc > fpad + fH) {
                    ny = fpad + fH - fr - 1.f;
                    if (kThreeHalfPiF <= o && o <= kTwoPiF) o += kHalfPiF;
                    else if (kPiF <= o && o < kThreeHalfPiF) o -= kHalfPiF;
                    o = wrap_2pi(o);
                }
            }
            if (tid == 0) {
                S.px[k] = nx; S.py[k] = ny; S.th[k] = o; S.vv[k] = vel;
                if (want_diag) {
                    float* dd = diag + 6 * (int64_t)ag;
                    dd[0] = dv; dd[1] = dpsi;
                    dd[2] = P.phi_ok ? (float)((double)alp0 * -Bc) : 0.f;
                    dd[3] = P.phi_ok ? (float)((double)(alp0 * alp1) * Ec) : 0.f;
                    dd[4] = P.phi_ok ? (float)((double)bet0 * -Bs) : 0.f;
                    dd[5] = P.phi_ok ? (float)((double)(bet0 * bet1) * Es) : 0.f;
                }
            }
            __syncthreads();
        }
        for (int i = tid; i < N; i += NT) {
            st.x[a0 + i] = S.px[i];
            st.y[a0 + i] = S.py[i];
            st.theta[a0 + i] = S.th[i];
            st.v[a0 + i] = S.vv[i];
        }
        __syncthreads();
    }
    if (count || verify) {
        if (n_pairs) atomicAdd(&counters[0], n_pairs);
        if (n_recheck) atomicAdd(&counters[1], n_recheck);
        if (n_disagree) atomicAdd(&counters[2], n_disagree);
    }
}

size_t lds_bytes_f32(int maxN, int maxR, int nw, bool sync_order) {
    const int maxW = (maxR + 31) >> 5;
    size_t b = sizeof(RayTab32) * (size_t)maxR + sizeof(double) * (4 * nw + 8) +
               sizeof(float) * (size_t)maxN * (sync_order ? 7 : 5) +
               sizeof(int32_t) * ((size_t)2 * maxN + 4) + sizeof(uint32_t) * (maxW + 2);
    return (b + 15) & ~(size_t)15;
}

template <int NW, int RPL>
int launch_nw(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
              uint32_t flags, hipStream_t s) {
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;
    const size_t lds = lds_bytes_f32(b->max_agents, b->max_rays, NW, sync_order);
    if (lds > (size_t)b->dev_max_lds)
        return set_error(ABM_ERR_UNSUPPORTED,
                         "abm_vf_step_f32: %d agents x %d rays need %zu B of LDS (device max %d)",
                         b->max_agents, b->max_rays, lds, b->dev_max_lds);
    auto kern = vf_step_f32_kernel<NW, RPL>;
    const int slot = NW == 1 ? 0 : (NW == 2 ? 1 : 2);
    int& per_cu = b->occ_f32[sync_order ? 1 : 0][slot];
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
    if (grid > b->n_replicas) grid = b->n_replicas;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * NW), lds, s, b->reps_dev, b->order_dev,
                       b->n_replicas, b->tab32, *st, bits, diag, flags, b->counters,
                       b->max_agents, b->max_rays);
    return check_hip(hipGetLastError(), "vf_step_f32_kernel launch");
}

}  // namespace

int launch_vf_step_f32(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                       uint32_t flags, hipStream_t s) {
    if (b->max_rays > 2048)
        return set_error(ABM_ERR_UNSUPPORTED,
                         "abm_vf_step_f32: n_rays up to 2048 (got %d); use abm_vf_step_f64", b->max_rays);
    int nw = b->waves_per_group;
    if (nw == 0) nw = b->max_agents > 256 ? 4 : (b->max_agents > 96 ? 2 : 1);
    const int need = (b->max_rays + 64 * nw - 1) / (64 * nw);   // rays per lane
    const int rpl = need <= 4 ? 4 : (need <= 8 ? 8 : (need <= 16 ? 16 : 32));
#define ABM_LAUNCH(NWV, RPLV) \
    if (nw == NWV && rpl == RPLV) return launch_nw<NWV, RPLV>(b, st, bits, diag, flags, s);
    ABM_LAUNCH(1, 4) ABM_LAUNCH(1, 8) ABM_LAUNCH(1, 16) ABM_LAUNCH(1, 32)
    ABM_LAUNCH(2, 4) ABM_LAUNCH(2, 8) ABM_LAUNCH(2, 16) ABM_LAUNCH(2, 32)
    ABM_LAUNCH(4, 4) ABM_LAUNCH(4, 8) ABM_LAUNCH(4, 16) ABM_LAUNCH(4, 32)
#undef ABM_LAUNCH
    return set_error(ABM_ERR_INVALID, "abm_vf_step_f32: no launch shape for %d waves", nw);
}

}  // namespace abm
