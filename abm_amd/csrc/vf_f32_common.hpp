// vf_f32_common.hpp -- float32 building blocks shared by the fused VSWRM
// kernels: wave reductions, the float32 pair decision with its guard bands and
// the float64 re-decision, the raster into bit-packed words, sincos, and the
// motion and boundary update (vf_agent.py:188-204, 263-326; agent.py:347-395,
// 605-610).
#pragma once

#include <cmath>

#include "vf_batch.hpp"
#include "vf_exact.hpp"

namespace abm {
namespace f32 {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kTwoPiF = 6.28318530717958647692f;
constexpr float kHalfPiF = 1.57079632679489661923f;
constexpr float kThreeHalfPiF = 4.71238898038468985769f;

// Guard bands.  The float32 angle (sincos of the heading, centre
// differences, cross/dot, atan2_f32) is within 1e-6 rad of the exact one and
// the float64 reference angle within 5e-8 rad, so a 4e-6 rad band around
// every bin tie and around the +-pi wrap covers both with margin.  The
// half-width band is 4e-6 relative + 1e-6 (its float32 error is < 1e-6
// relative).  Pairs inside a band are re-decided in float64.
constexpr float kAngleGuardF = 4e-6f;

// Wave sum: float32 DPP adds inside each 16-lane row (quad xor 1, xor 2,
// half-row mirror, row mirror), then float64 across rows (row_bcast15 into
// rows 1 and 3, row_bcast31 into rows 2 and 3); lane 63 ends with the total,
// returned as a wave-uniform double.
template <int CTRL>
__device__ inline float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_f64_masked(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double wave_total(float x) {
    x += dpp_f32<0xB1>(x);          // quad_perm [1,0,3,2]
    x += dpp_f32<0x4E>(x);          // quad_perm [2,3,0,1]
    x += dpp_f32<0x141>(x);         // row_half_mirror
    x += dpp_f32<0x140>(x);         // row_mirror: every lane holds its row's sum
    double v = (double)x;
    v += dpp_f64_masked<0x142, 0xa>(v);   // row_bcast15 -> rows 1, 3
    v += dpp_f64_masked<0x143, 0xc>(v);   // row_bcast31 -> rows 2, 3
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

__device__ inline void raster(uint32_t* F, int start, int end, int R) {
    int p0, p1, q0, q1;
    vf_pieces(start, end, R, &p0, &p1, &q0, &q1);
    or_range(F, p0, p1);
    or_range(F, q0, q1);
}

__device__ inline float wrap_2pi(float o) {
    if (o < 0.f) o = kTwoPiF + o;
    if (o > kTwoPiF) o = o - kTwoPiF;
    return o;
}

// atan2 for float32 from a degree-7 minimax polynomial of atan(z)/z in z^2 on
// [0, 1] (|error| <= 1.5e-7 with fused multiply-adds) and octant reduction;
// total error below 4e-7 rad.  The guard bands allow for it.
__device__ inline float atan2_f32(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float z = mn * __builtin_amdgcn_rcpf(mx);
    const float s = z * z;
    float p = -0.0046934f;
    p = fmaf(p, s, 0.02425289f);
    p = fmaf(p, s, -0.05948712f);
    p = fmaf(p, s, 0.09914348f);
    p = fmaf(p, s, -0.14019503f);
    p = fmaf(p, s, 0.19969729f);
    p = fmaf(p, s, -0.3333199f);
    p = fmaf(p, s, 0.9999999f);
    p = p * z;
    if (ay > ax) p = kHalfPiF - p;
    if (x < 0.f) p = kPiF - p;
    return y < 0.f ? -p : p;
}

// atan(y / x) for y > 0, x >= 0 (the half-width angle atan(size / d)): the
// same polynomial without the sign handling.
__device__ inline float atan_pos_f32(float y, float x) {
    const float mx = fmaxf(x, y), mn = fminf(x, y);
    const float z = mn * __builtin_amdgcn_rcpf(mx);
    const float s = z * z;
    float p = -0.0046934f;
    p = fmaf(p, s, 0.02425289f);
    p = fmaf(p, s, -0.05948712f);
    p = fmaf(p, s, 0.09914348f);
    p = fmaf(p, s, -0.14019503f);
    p = fmaf(p, s, 0.19969729f);
    p = fmaf(p, s, -0.3333199f);
    p = fmaf(p, s, 0.9999999f);
    p = p * z;
    return y > x ? kHalfPiF - p : p;
}

// atan(z) for 0 <= z < 1 (the half-width angle atan(size / d) with d > size):
// the same polynomial without the octant reduction.
__device__ inline float atan_unit_f32(float z) {
    const float s = z * z;
    float p = -0.0046934f;
    p = fmaf(p, s, 0.02425289f);
    p = fmaf(p, s, -0.05948712f);
    p = fmaf(p, s, 0.09914348f);
    p = fmaf(p, s, -0.14019503f);
    p = fmaf(p, s, 0.19969729f);
    p = fmaf(p, s, -0.3333199f);
    p = fmaf(p, s, 0.9999999f);
    return p * z;
}

// Per-replica constants, all float32 except the few integers.
struct RepConst {
    int N, R, W, a0, torus, fov_px0, fov_px1, phi_ok, limit;
    long field_offset;
    float inv_step;   // (R - 1) / (2 pi): bins per radian of linspace(-pi, pi, R)
    float g_bins;     // bin-tie guard in bins (angle guard + float32 rounding of the coordinate)
    float hs;         // R / (2 pi): proj_size / 2 per radian of atan(size / d)
    float hw, hh;     // W/2, H/2
    float tguard;     // guard on |dx| - W/2 for the torus image decision
    float fW, fH, pad;
    float gam, alp1, bet1, max_th, max_vel;
};

__device__ inline RepConst rep_const(const VFRep* rp) {
    RepConst c;
    c.N = rp->n_agents; c.R = rp->n_rays; c.W = rp->words; c.a0 = rp->agent_offset;
    c.torus = rp->boundary == ABM_BOUNDARY_INFINITE;
    c.fov_px0 = rp->fov_px0; c.fov_px1 = rp->fov_px1; c.phi_ok = rp->phi_ok;
    c.limit = rp->limit_movement; c.field_offset = rp->field_offset;
    c.inv_step = (float)(1.0 / rp->step);
    c.g_bins = (float)(kAngleGuardF / rp->step + 1.5e-7 * rp->n_rays);
    c.hs = (float)((double)rp->n_rays / kTwoPi);
    c.hw = (float)(rp->width / 2.0);
    c.hh = (float)(rp->height / 2.0);
    c.tguard = (float)(1e-6 * (rp->width > rp->height ? rp->width : rp->height) + 1e-4);
    c.fW = (float)rp->width; c.fH = (float)rp->height; c.pad = (float)rp->pad;
    c.gam = (float)rp->gam; c.alp1 = (float)rp->alp1; c.bet1 = (float)rp->bet1;
    c.max_th = (float)rp->max_th; c.max_vel = (float)rp->max_vel;
    return c;
}

// Float32 decision of one pair; returns false when the pair must be
// re-decided in float64 (inside a guard band, or degenerate).  The centre
// difference is exact or rounded once relative to its own size: coordinates
// close together subtract exactly (Sterbenz), distant ones give a large
// difference, and for the periodic image the shift by W is applied to the
// coordinate it leaves within a factor two of W (exact) before the final
// subtraction -- subtracting first would round at the ulp of W, an angle error
// of ~3e-5 px / d for a wrapped pair at distance d.  With sincos (1.1e-7),
// cross/dot, atan2_f32 (2e-7) and the rounding of pi - a (2.4e-7) the float32
// angle stays within 1e-6 rad of the exact one; the float64 reference is within
// 5e-8 rad of it.
// The result: .ok false when the pair must be decided in float64; .tie when
// only the bin-tie band was hit (the tie is between t and t + 1 for dfr > 0,
// else t - 1 and t), for tie_in_place.
struct PairF32 {
    int start, end;   // t - h, t + h
    bool drawn;       // an end inside the FOV (vf_supcalc.py:116-119)
    bool ok;          // outside every guard band: the float32 decision stands
    bool tie;         // only the bin-tie band was hit
    float dfr;        // bin coordinate minus its nearest bin
};

__device__ inline PairF32 fast_pair(float u1x, float u1y, float fx, float fy, float fr, float ojx,
                                   float ojy, float ojr, const RepConst& C) {
    const float dr = ojr - fr;                       // radii are small integers: exact
    float dx = (ojx - fx) + dr;
    float dy = (ojy - fy) + dr;
    bool ok = true;
    if (C.torus) {
        // nearest image: shift by W toward the focal when |dx| > W/2 (same for y)
        const float ex = fabsf(dx) - C.hw, ey = fabsf(dy) - C.hh;
        ok = (fabsf(ex) >= C.tguard) & (fabsf(ey) >= C.tguard);
        // object shifted down by W (object above), or focal shifted down by W
        // (object below); selects, no min / max (those would canonicalise)
        // (x - 0 is x: the unshifted sides are the coordinates themselves, and
        // the focal's shifted coordinates do not depend on the object)
        const bool wx = ex > 0.f, px = dx > 0.f, wy = ey > 0.f, py = dy > 0.f;
        const float oxs = (wx & px) ? ojx - C.fW : ojx, fxs = (wx & !px) ? fx - C.fW : fx;
        const float oys = (wy & py) ? ojy - C.fH : ojy, fys = (wy & !py) ? fy - C.fH : fy;
        dx = (oxs - fxs) + dr;
        dy = (oys - fys) + dr;
    }
    const float d2 = fmaf(dy, dy, dx * dx);
    const float a = atan2_f32(fmaf(u1x, dy, -u1y * dx), fmaf(u1x, dx, u1y * dy));
    // atan(size / d) from size / d = size * rsqrt(d^2) (v_rsq_f32, 1 ulp);
    // a pair with d <= size (overlapping circles) goes to float64
    const float z = ojr * __builtin_amdgcn_rsqf(d2);
    const float q = atan_unit_f32(z);
    const float xb = (kPiF - a) * C.inv_step;       // bin coordinate of ca = -a
    // nearest bin; a tie (argmin takes the first index) lies inside the guard
    const float tr = rintf(xb);
    const float dfr = xb - tr;
    const float half = q * C.hs;                    // proj_size / 2
    const float hfl = floorf(half);
    const float hfrac = half - hfl;
    const float hg = 4e-6f * half + 1e-6f;
    // bitwise & (no short circuit): the compiler keeps the tests branch-free.
    // Outside the guard |a| < pi - guard, so t = rint(xb) lies in [0, R - 1].
    const bool other = ok & (d2 > 1e-30f) & (z < 1.0f) & (fabsf(a) < kPiF - kAngleGuardF) &
                       (hfrac >= hg) & (hfrac <= 1.0f - hg);
    const bool tie_ok = fabsf(dfr) <= 0.5f - C.g_bins;
    const int t = (int)tr;
    const int h = (int)hfl;
    PairF32 p;
    p.start = t - h;
    p.end = t + h;
    p.drawn = ((C.fov_px0 < p.start) & (p.start < C.fov_px1)) | ((C.fov_px0 < p.end) & (p.end < C.fov_px1));
    p.ok = other & tie_ok;
    p.tie = other & !tie_ok;
    p.dfr = dfr;
    return p;
}

// start / end / drawn of a VF pair from its bin and half-width (vf_supcalc.py:116-119).
__device__ inline void vf_interval(int t, int h, const VFRep* rp, int* start, int* end, int* drawn) {
    *start = t - h;
    *end = t + h;
    *drawn = ((rp->fov_px0 < *start && *start < rp->fov_px1) ||
              (rp->fov_px0 < *end && *end < rp->fov_px1)) ? 1 : 0;
}

// Queued-pair entries: j | slot << 20 | kEntryForced (the float64 replica
// decides) and the float32 decision (start + 2R) | (end + 2R) << 15 |
// kFastDrawn (compared in verify mode).
constexpr int kEntryForced = 1 << 30;
constexpr int kFastDrawn = 1 << 30;

__device__ inline int entry_code(int j, int slot, bool ok) {
    return j | (slot << 20) | (ok ? 0 : kEntryForced);
}
__device__ inline int fast_code(int start, int end, int drawn, int R) {
    return (start + 2 * R) | ((end + 2 * R) << 15) | (drawn ? kFastDrawn : 0);
}

// The decision of a queued pair: for a pair flagged only for the bin tie the
// float64 side test (tie_bin), else (or when that is undecided) the float64
// replica of the reference arithmetic.  Sets start / end / drawn.
static __device__ __noinline__ void exact_pair(float fx, float fy, float fth, float fr, float ojx,
                                               float ojy, float ojr, const VFRep* rp, int* start,
                                               int* end, int* drawn);

// A queued pair of the fused kernels' float64 lists, resolved out of line by
// the float64 replica (pairs flagged only for the bin tie are decided in the
// pair pass by tie_in_place).  Forced entries return their decision, start |
// end << 15 (+ 2R each) | drawn << 30, or -1 when not drawn; in verify mode an
// unforced entry's float32 decision is compared and *disagree set on a
// difference.
static __device__ __noinline__ int resolve_queued(int e, int fc, float fx, float fy, float fth, float fr,
                                                  float ojx, float ojy, float ojr, const VFRep* rp,
                                                  bool verify, int* disagree);

// The bin of a pair whose float32 angle fell only inside the bin-tie band
// (why & kTieOnly): which side of the tie between bins b and b + 1 the
// reference's closed angle lies on, from float64 geometry instead of the
// reference's arithmetic.  The closed angle of vf_supcalc.py:142-158 is
// ca = phi' - theta, phi' the direction of the centre difference v with y up
// (v exact in float64 from the float32 state, periodic image as in
// vf_supcalc.py:70-83), so with psi = theta + m_b, m_b the midpoint of the
// linspace values b and b + 1, s = -v_y cos(psi) - v_x sin(psi) =
// |v| sin(ca - m_b).  The reference's own angle is within 2e-8 rad of the
// exact one (acos near 0 is the worst case), so when |s| > 1e-7 |v| its argmin
// picks b + 1 for s > 0 and b for s < 0.  Returns the bin, or -1 when the pair
// lies within 1e-7 rad of the tie and the float64 replica must decide.
// sin and cos in float64 for |x| < 1e5: Cody-Waite reduction by pi/2 with a
// two-part constant (fdlibm's published pio2_1 / pio2_1t split: the first
// part has 33 significant bits, so k * part is exact for |k| < 2^20 and
// x - k * part is exact), then the Taylor series (coefficients 1/n!) of sin
// to r^11 and of cos to r^12 on |r| <= pi/4 (truncation below 1e-11).
// Returns false (outputs unset) for larger arguments or NaN.
__device__ __forceinline__ bool sincos_f64_small(double x, double* s, double* c) {
    if (!(fabs(x) < 1e5)) return false;
    const double k = rint(x * 0.63661977236758138);
    double r = fma(-k, 1.5707963267341256, x);
    r = fma(-k, 6.077100506506192e-11, r);
    const double r2 = r * r;
    double ps = fma(r2, -2.5052108385441720e-08, 2.7557319223985893e-06);
    ps = fma(r2, ps, -1.9841269841269841e-04);
    ps = fma(r2, ps, 8.3333333333333333e-03);
    ps = fma(r2, ps, -1.6666666666666667e-01);
    const double sn = fma(r * r2, ps, r);
    double pc = fma(r2, 2.0876756987868099e-09, -2.7557319223985891e-07);
    pc = fma(r2, pc, 2.4801587301587302e-05);
    pc = fma(r2, pc, -1.3888888888888889e-03);
    pc = fma(r2, pc, 4.1666666666666667e-02);
    pc = fma(r2, pc, -0.5);
    const double cn = fma(r2, pc, 1.0);
    const int q = (int)k;
    const double s0 = (q & 1) ? cn : sn, c0 = (q & 1) ? sn : cn;
    *s = (q & 2) ? -s0 : s0;
    *c = ((q + 1) & 2) ? -c0 : c0;
    return true;
}

__device__ __forceinline__ int tie_bin_inl(float fx, float fy, float fth, float fr, float ojx,
                                           float ojy, float ojr, const VFRep* rp, int b) {
    const double cx = (double)fx + (double)fr, cy = (double)fy + (double)fr;
    double ocx = (double)ojx + (double)ojr, ocy = (double)ojy + (double)ojr;
    double vx = ocx - cx, vy = ocy - cy;
    if (rp->boundary == ABM_BOUNDARY_INFINITE) {
        if (fabs(vx) > rp->width / 2.0) ocx += cx < ocx ? -rp->width : (cx > ocx ? rp->width : 0.0);
        if (fabs(vy) > rp->height / 2.0) ocy += cy < ocy ? -rp->height : (cy > ocy ? rp->height : 0.0);
        vx = ocx - cx;
        vy = ocy - cy;
    }
    const int R = rp->n_rays;
    const double m = 0.5 * (linspace_at(b, R, rp->step) + linspace_at(b + 1, R, rp->step));
    double sp, cp;
    if (!sincos_f64_small((double)fth + m, &sp, &cp)) return -1;
    const double sv = -vy * cp - vx * sp;
    const double d2 = vx * vx + vy * vy;
    if (sv * sv <= 1e-14 * d2) return -1;
    return sv > 0.0 ? b + 1 : b;
}

static __device__ __noinline__ int tie_bin(float fx, float fy, float fth, float fr, float ojx,
                                           float ojy, float ojr, const VFRep* rp, int b) {
    return tie_bin_inl(fx, fy, fth, fr, ojx, ojy, ojr, rp, b);
}

// A pair of fast_pair's with .tie set decided in place by the float64 side
// test (an out-of-line leaf taking and returning values only, so the hot loop
// keeps the pair in registers): on success sets start / end / drawn and .ok;
// otherwise the pair is left to the float64 replica.  Returns .ok.
__device__ __forceinline__ bool tie_in_place(PairF32& p, float fx, float fy, float fth, float fr,
                                             float ojx, float ojy, float ojr, const VFRep* rp,
                                             const RepConst& C) {
    const int t32 = (p.start + p.end) / 2, h = (p.end - p.start) / 2;
    const int t = tie_bin(fx, fy, fth, fr, ojx, ojy, ojr, rp, p.dfr > 0.f ? t32 : t32 - 1);
    if (t < 0) return false;
    p.start = t - h;
    p.end = t + h;
    p.drawn = ((C.fov_px0 < p.start) & (p.start < C.fov_px1)) | ((C.fov_px0 < p.end) & (p.end < C.fov_px1));
    p.ok = true;
    return true;
}

__device__ __forceinline__ void exact_pair_inl(float fx, float fy, float fth, float fr, float ojx,
                                               float ojy, float ojr, const VFRep* rp, int* start,
                                               int* end, int* drawn) {
    const Focal f = make_focal((double)fx, (double)fy, (double)fth, (double)fr);
    const PairVF e = pair_vf_exact(f, (double)ojx, (double)ojy, (double)ojr,
                                   rp->boundary == ABM_BOUNDARY_INFINITE ? 1 : 0, rp->width,
                                   rp->height, rp->n_rays, rp->step, rp->fov_px0, rp->fov_px1);
    *start = e.start; *end = e.end; *drawn = e.skipped ? 0 : e.drawn;
}

static __device__ __noinline__ void exact_pair(float fx, float fy, float fth, float fr, float ojx, float ojy,
                                               float ojr, const VFRep* rp, int* start, int* end, int* drawn) {
    exact_pair_inl(fx, fy, fth, fr, ojx, ojy, ojr, rp, start, end, drawn);
}


// sincos for float32 headings.  Cody-Waite reduction by pi/2 with a
// three-part constant (the first two parts carry 12 significant bits, so k * part
// is exact for |k| < 4096, i.e. |x| < 6400), then polynomials on [-pi/4, pi/4]:
// sin r = r + r^3 P(r^2), cos r = 1 - r^2/2 + r^4 Q(r^2), with P and Q of
// degree 2 fitted (least squares on Chebyshev nodes) to the float64 functions.
// Absolute error <= 1.1e-7 on [0, 6000]; larger arguments take sincosf.
static __device__ __noinline__ float2 sincos_slow(float x) {
    float s, c;
    sincosf(x, &s, &c);
    return make_float2(s, c);
}

__device__ inline void sincos_f32(float x, float* s, float* c) {
    const float k = rintf(x * 0.636619772367581343f);
    float r = fmaf(-k, 1.5708008f, x);
    r = fmaf(-k, -4.4535846e-06f, r);
    r = fmaf(-k, -8.7055163e-10f, r);
    const float r2 = r * r;
    const float ps = fmaf(fmaf(-1.9513932e-04f, r2, 8.3321482e-03f), r2, -1.6666654e-01f);
    const float sn = fmaf(r * r2, ps, r);
    const float pc = fmaf(fmaf(2.4435853e-05f, r2, -1.3887342e-03f), r2, 4.1666646e-02f);
    const float cn = fmaf(r2 * r2, pc, fmaf(-0.5f, r2, 1.0f));
    const int q = (int)k;
    const bool swap = (q & 1) != 0;
    const float s0 = swap ? cn : sn, c0 = swap ? sn : cn;
    *s = (q & 2) ? -s0 : s0;
    *c = ((q + 1) & 2) ? -c0 : c0;
    const bool big = !(fabsf(x) <= 6000.f);
    if (__ballot(big)) {          // wave-uniform test; rare
        if (big) {
            const float2 sc = sincos_slow(x);
            *s = sc.x; *c = sc.y;
        }
    }
}

// Heading and speed update from the two sums (vf_agent.py:263-326): returns
// the new heading o (one-shot wrap) and speed, and dv / dpsi for the diagnostics.
__device__ inline void steer(const RepConst& C, float fth, float vel, float v0, float Sv, float Sp,
                             float* o, float* vel_new, float* dv_out, float* dpsi_out) {
    float dv = 0.f, dpsi = 0.f;
    if (C.phi_ok) {
        dv = C.gam * (v0 - vel) + Sv;
        dpsi = Sp;
    }
    float theta = dpsi;
    vel = vel + dv;
    if (C.limit) {   // uniform; kept a branch (the empty asm stops if-conversion)
        asm volatile("");
        if (fabsf(theta) > C.max_th) theta = theta < 0.f ? -C.max_th : C.max_th;
        if (fabsf(vel) > C.max_vel) vel = vel < 0.f ? -C.max_vel : C.max_vel;
    }
    *o = wrap_2pi(fth + theta);
    *vel_new = vel;
    *dv_out = dv;
    *dpsi_out = dpsi;
}

// Euler step and boundary: periodic snap (vf_agent.py:188-204) or walls with
// the quadrant turns (agent.py:347-395); o may change at a wall.
__device__ inline void euler_boundary(const RepConst& C, float fx, float fy, float fr, float vel,
                                      float so, float co, float* o_io, float* nx_out,
                                      float* ny_out) {
    float o = *o_io;
    float nx = fx + vel * co;
    float ny = fy - vel * so;
    const float xc = nx + fr, yc = ny + fr;
    const float lo_b = C.pad, hx_b = C.pad + C.fW, hy_b = C.pad + C.fH;
    if (C.torus) {
        nx = xc < lo_b ? hx_b - fr : (xc > hx_b ? lo_b + fr : nx);
        ny = yc < lo_b ? hy_b - fr : (yc > hy_b ? lo_b + fr : ny);
    } else {
        if (xc < lo_b) {
            nx = lo_b - fr;
            if (kHalfPiF <= o && o < kPiF) o -= kHalfPiF;
            else if (kPiF <= o && o <= kThreeHalfPiF) o += kHalfPiF;
            o = wrap_2pi(o);
        }
        if (xc > hx_b) {
            nx = hx_b - fr - 1.f;
            if (kThreeHalfPiF <= o && o < kTwoPiF) o -= kHalfPiF;
            else if (0.f <= o && o <= kHalfPiF) o += kHalfPiF;
            o = wrap_2pi(o);
        }
        if (yc < lo_b) {
            ny = lo_b - fr;
            if (kHalfPiF <= o && o <= kPiF) o += kHalfPiF;
            else if (0.f <= o && o < kHalfPiF) o -= kHalfPiF;
            o = wrap_2pi(o);
        }
        if (yc > hy_b) {
            ny = hy_b - fr - 1.f;
            if (kThreeHalfPiF <= o && o <= kTwoPiF) o += kHalfPiF;
            else if (kPiF <= o && o < kThreeHalfPiF) o -= kHalfPiF;
            o = wrap_2pi(o);
        }
    }
    *o_io = o;
    *nx_out = nx;
    *ny_out = ny;
}

}  // namespace f32
}  // namespace abm

namespace abm {
namespace f32 {

static __device__ __noinline__ int resolve_queued(int e, int fc, float fx, float fy, float fth, float fr,
                                                  float ojx, float ojy, float ojr, const VFRep* rp,
                                                  bool verify, int* disagree) {
    // a leaf (the float64 replica inlined): a call from the kernels then
    // clobbers only the registers this function itself uses
    const int R = rp->n_rays;
    const bool forced = (e & kEntryForced) != 0;
    int xs = 0, xe = 0, xd = 0;
    exact_pair_inl(fx, fy, fth, fr, ojx, ojy, ojr, rp, &xs, &xe, &xd);
    if (verify) {
        const int ex = xd ? fast_code(xs, xe, 0, R) : -1;
        const int got = (fc & kFastDrawn) ? (fc & ~kFastDrawn) : -1;
        *disagree = !forced && got != ex;
    } else {
        *disagree = 0;
    }
    if (!forced) return -1;
    return xd ? fast_code(xs, xe, 1, R) : -1;
}

}  // namespace f32
}  // namespace abm
