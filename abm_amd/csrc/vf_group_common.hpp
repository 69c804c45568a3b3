// vf_group_common.hpp -- wave-level helpers of the grouped fused kernels
// (G replicas per wave, L = 64 / G lanes per replica).
#pragma once

#include "vf_f32_common.hpp"

namespace abm {
namespace f32 {

// Orders the LDS accesses of different lanes of this wave: everything before
// is complete and visible to every lane of the wave before anything after.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Issue priority from progress through the n focal agents: 3 for the first
// half, 2 to three quarters, 1 to fifteen sixteenths, then 0.  At equal
// priority a SIMD's arbiter favours its oldest wave, so waves that share a
// SIMD would finish one after another and the last would run alone,
// latency-bound; lowering the priority of waves that have advanced lets the
// younger ones keep up, and the short last stage keeps the staggered finish
// short.  The focal loop runs stage by stage, so the priority changes only
// at stage boundaries.
struct ProgressStages {
    int k1, k2, k3;   // first focal agent of stages 1, 2, 3
};
__device__ inline ProgressStages progress_stages(int n) {
    // stage(k) = 0 for 16 k < 8 n, 1 for 16 k < 12 n, 2 for 16 k < 15 n, else 3
    return {(8 * n + 15) / 16, (12 * n + 15) / 16, (15 * n + 15) / 16};
}
// Sets the priority of stage s (0..3) and returns the stage's end (the first
// focal agent of the next stage, or n).
__device__ inline int stage_priority(int s, const ProgressStages& ps, int n) {
    switch (s) {
        case 0: __builtin_amdgcn_s_setprio(3); return ps.k1;
        case 1: __builtin_amdgcn_s_setprio(2); return ps.k2;
        case 2: __builtin_amdgcn_s_setprio(1); return ps.k3;
        default: __builtin_amdgcn_s_setprio(0); return n;
    }
}

// Value of lane (group base + L - 2) in every lane of the group.
template <int G>
__device__ inline float group_bcast_second_last(float v, int gl) {
    if (G == 4) {   // row_newbcast:14
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x15E, 0xf, 0xf, false));
    } else if (G == 2) {
        const float a = lane_f32(v, 30), b = lane_f32(v, 62);
        return gl == 0 ? a : b;
    } else {
        return lane_f32(v, 62);
    }
}

// ((1 << width) - 1) << offset for width, offset in [0, 31], truncated to
// 32 bits (one v_bfm_b32; the compiler does not form it from the shifts).
__device__ inline uint32_t bit_field_mask(uint32_t width, uint32_t offset) {
    uint32_t r;
    asm("v_bfm_b32 %0, %1, %2" : "=v"(r) : "v"(width), "v"(offset));
    return r;
}

// Masks of n < 32 rays starting at bit s < 32 of a word (n = 0: nothing): the
// bits in that word and the bits that spill into the next one.
__device__ inline void span_masks(int s, int n, uint32_t* lo, uint32_t* hi) {
    *lo = bit_field_mask((uint32_t)n, (uint32_t)s);
    const uint32_t e = (uint32_t)(s + n);
    *hi = bit_field_mask(e > 32u ? e - 32u : 0u, 0u);
}

// Sums over each group of L = 64 / G lanes of two values, their DPP steps
// interleaved (no hazard stalls between dependent DPP instructions).  The
// totals are valid in the last lane of every group (for G = 4 in every lane
// of the group).
template <int G>
__device__ inline void group_sum2(float& a, float& b) {
    a += dpp_f32<0xB1>(a); b += dpp_f32<0xB1>(b);
    a += dpp_f32<0x4E>(a); b += dpp_f32<0x4E>(b);
    a += dpp_f32<0x141>(a); b += dpp_f32<0x141>(b);
    a += dpp_f32<0x140>(a); b += dpp_f32<0x140>(b);
    // Masked row broadcasts as single DPP adds: rows outside row_mask are not
    // written and keep their value.  The compiler does not fuse a masked
    // broadcast move into the add (no exact float identity for the unwritten
    // lanes), so these are written out; each DPP read of a register written
    // by the VALU in the two instructions before it gets its two wait states
    // from the s_nop in front of the pair, and the trailing s_nop covers a
    // DPP read of the results right after the block.
    if (G == 1) {
        asm volatile(
            "s_nop 1\n\t"
            "v_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
            "v_add_f32_dpp %1, %1, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
            "s_nop 1\n\t"
            "v_add_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
            "v_add_f32_dpp %1, %1, %1 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
            "s_nop 1"
            : "+v"(a), "+v"(b));
    } else if (G == 2) {
        asm volatile(
            "s_nop 1\n\t"
            "v_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
            "v_add_f32_dpp %1, %1, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
            "s_nop 1"
            : "+v"(a), "+v"(b));
    }
}

// Four group sums, DPP steps interleaved.
template <int G>
__device__ inline void group_sum4(float& a, float& b, float& c, float& d) {
    group_sum2<G>(a, b);
    group_sum2<G>(c, d);
}

// Lane value moved by a DPP control (bound_ctrl: lanes without a source read 0).
template <int CTRL>
__device__ inline uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true);
}

// Bits of lane li's ray span [li * RPL, li * RPL + RPL) of the field F of R
// rays, with the neighbouring bits the edge rule needs, and the edge mask of
// the span (dPhi alignment rule of vf_supcalc.py:266-274).  When the spans of
// a group tile the R rays exactly (R = L * RPL) and a group is the wave (G = 1)
// or a DPP row (G = 4), the neighbouring bits come from the adjacent lanes'
// spans by DPP rotation instead of LDS reads.
template <int G, int RPL>
__device__ inline void span_bits(const uint32_t* F, int li, int R, uint64_t* v_out, uint64_t* sp_out) {
    constexpr int L = 64 / G;
    const int rs = li * RPL;
    const int rn = rs >= R ? 0 : (R - rs < RPL ? R - rs : RPL);
    uint32_t vlo = 0u, vhi = 0u, pbit = 0u, nbit = 0u;
    bool case1;
    if ((G == 1 || G == 4) && RPL <= 32 && R == L * RPL) {   // wave-uniform
        vlo = RPL == 32 ? F[rs >> 5] : ((F[rs >> 5] >> (rs & 31)) & low_mask(RPL));
        uint32_t prev, next, first, last;
        if (G == 1) {
            prev = dpp_u32<0x13C>(vlo);                       // wave_ror:1 (lane 0 <- lane 63)
            next = dpp_u32<0x134>(vlo);                       // wave_rol:1 (lane 63 <- lane 0)
            first = (uint32_t)__builtin_amdgcn_readlane((int)vlo, 0);
            last = (uint32_t)__builtin_amdgcn_readlane((int)vlo, 63);
        } else {
            prev = dpp_u32<0x121>(vlo);                       // row_ror:1
            next = dpp_u32<0x12F>(vlo);                       // row_ror:15
            first = dpp_u32<0x150>(vlo);                      // row_newbcast:0
            last = dpp_u32<0x15F>(vlo);                       // row_newbcast:15
        }
        pbit = (prev >> (RPL - 1)) & 1u;
        nbit = next & 1u;
        case1 = (first & 1u) == 1u && ((last >> (RPL - 1)) & 1u) == 0u;
        // every span is full (rn = RPL): 32-bit masks, constant shifts
        const uint32_t m = low_mask(RPL);
        uint32_t sp32;
        if (case1) sp32 = (vlo ^ ((vlo << 1) | pbit)) & m;   // wave-uniform
        else sp32 = vlo ^ ((vlo >> 1) | (nbit << (RPL - 1)));
        *v_out = vlo;
        *sp_out = sp32;
        return;
    } else {
        if (rn > 0) {
            if (RPL <= 32) {
                vlo = RPL == 32 ? F[rs >> 5] : ((F[rs >> 5] >> (rs & 31)) & low_mask(RPL));
            } else {
                vlo = F[rs >> 5];
                vhi = F[(rs >> 5) + 1];
            }
            pbit = field_bit(F, rs == 0 ? R - 1 : rs - 1);
            nbit = field_bit(F, rs + rn >= R ? 0 : rs + rn);
        }
        case1 = (F[0] & 1u) == 1u && field_bit(F, R - 1) == 0u;
    }
    const uint64_t m = rn >= 64 ? ~0ull : ((1ull << rn) - 1ull);
    const uint64_t v = (((uint64_t)vhi << 32) | vlo) & m;
    uint64_t sp = 0;
    if (rn > 0) {
        if (case1) sp = (v ^ ((v << 1) | pbit)) & m;
        else sp = (v ^ ((v >> 1) | ((uint64_t)nbit << (rn - 1)))) & m;
    }
    *v_out = v;
    *sp_out = sp;
}

// The four sums of one lane's span from the nibble tables (entry [q][v][li]:
// tb, ts point at lane li's column).
// Bits [off, off + 4) of x (one v_bfe_u32; left to itself the compiler turns
// the extract-then-scale of the table index into shift, and, add).
__device__ inline uint32_t nibble(uint32_t x, uint32_t off) {
    uint32_t r;
    asm("v_bfe_u32 %0, %1, %2, 4" : "=v"(r) : "v"(x), "s"(off));
    return r;
}

template <int RPL, int L>
__device__ inline void span_sums(const float2* tb, const float2* ts, uint64_t v, uint64_t sp,
                                 float* bc, float* bs, float* ec, float* es) {
    float a, b, c, d;
#pragma unroll
    for (int q = 0; q < RPL / 4; ++q) {
        const uint32_t vw = q < 8 ? (uint32_t)v : (uint32_t)(v >> 32);
        const uint32_t sw = q < 8 ? (uint32_t)sp : (uint32_t)(sp >> 32);
        const uint32_t vb = nibble(vw, 4 * (q & 7));
        const uint32_t sb = nibble(sw, 4 * (q & 7));
        const float2 tbv = (tb + q * 16 * L)[vb * L];
        const float2 tsv = (ts + q * 16 * L)[sb * L];
        if (q == 0) {   // no 0.f + x (not an identity for -0.0, so it stays an add)
            a = tbv.x; b = tbv.y; c = tsv.x; d = tsv.y;
        } else {
            a += tbv.x; b += tbv.y;
            c += tsv.x; d += tsv.y;
        }
    }
    *bc = a; *bs = b; *ec = c; *es = d;
}

// Kahan-compensated version of span_sums: each of the four sums carries its
// rounding error (Kahan's running compensation) over the lane's nibble
// entries; the results are value + compensation pairs.
template <int RPL, int L>
__device__ inline void span_sums_kahan(const float2* tb, const float2* ts, uint64_t v, uint64_t sp,
                                       float s[4], float c[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) { s[i] = 0.f; c[i] = 0.f; }
#pragma unroll
    for (int q = 0; q < RPL / 4; ++q) {
        const uint32_t vw = q < 8 ? (uint32_t)v : (uint32_t)(v >> 32);
        const uint32_t sw = q < 8 ? (uint32_t)sp : (uint32_t)(sp >> 32);
        const uint32_t vb = nibble(vw, 4 * (q & 7));
        const uint32_t sb = nibble(sw, 4 * (q & 7));
        const float2 tbv = (tb + q * 16 * L)[vb * L];
        const float2 tsv = (ts + q * 16 * L)[sb * L];
        const float x[4] = {tbv.x, tbv.y, tsv.x, tsv.y};
#pragma unroll
        for (int i = 0; i < 4; ++i) {   // Kahan step: y = x - c; t = s + y; c = (t - s) - y; s = t
            const float y = x[i] - c[i];
            const float t = s[i] + y;
            c[i] = (t - s[i]) - y;
            s[i] = t;
        }
    }
}

// Sum over each group of L = 64 / G lanes in float64 (the lane values are
// exact in double): the DPP steps of group_sum2 on both halves of each
// double; valid in the last lane of every group (every lane for G = 4).
template <int CTRL, int ROW_MASK, bool BOUND>
__device__ inline double dpp_f64_any(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, BOUND);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, BOUND);
    return __hiloint2double(hi, lo);
}
template <int G>
__device__ inline double group_sum_f64(double x) {
    x += dpp_f64_any<0xB1, 0xf, true>(x);    // quad_perm [1,0,3,2]
    x += dpp_f64_any<0x4E, 0xf, true>(x);    // quad_perm [2,3,0,1]
    x += dpp_f64_any<0x141, 0xf, true>(x);   // row_half_mirror
    x += dpp_f64_any<0x140, 0xf, true>(x);   // row_mirror: every lane holds its row's sum
    if (G <= 2) x += dpp_f64_any<0x142, 0xa, false>(x);   // row_bcast15 -> rows 1, 3
    if (G == 1) x += dpp_f64_any<0x143, 0xc, false>(x);   // row_bcast31 -> rows 2, 3
    return x;
}

}  // namespace f32
}  // namespace abm
