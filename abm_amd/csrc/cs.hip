// cs.hip -- cooperative-signalling visual projections on the GPU.
//
// The cooperative-signalling agent (abm/projects/cooperative_signaling)
// perceives others through its own projection_field
// (cs_agent/cs_supcalc.py:204-295), which differs from the VSWRM one
// (vf_supcalc.py:20-138) in five ways:
//   * object centres and visual angles use the focal agent's radius;
//   * no periodic images;
//   * visibility is the closed-angle test fov0 <= ca <= fov1, not the pixel
//     test on start / end;
//   * an object whose projection size exceeds max_proj_size is not drawn
//     (:270, validate_projection_size :298-307);
//   * a drawn row is scaled by the object's meter (:286-287), and after the
//     flip the rays with phis < fov0 or phis > fov1 are zeroed (:291-294).
// CSAgent combines the rows two ways (cs_agent.py:83-131): the crowding
// density, rows summed over objects and divided by len(agents), and the
// signalling density, the maximum over rows.  This file computes the rows,
// the sum and the maximum for many focal agents at once, each with its own
// object list (CSR), in float64 with the reference's operation order, so the
// results equal the reference's bit for bit (up to the device libm's last ulp
// in acos / atan, which moves a decision only within an ulp of a tie).
//
// One wave per focal agent.  Lanes decide a chunk of objects (exact float64
// pair decision); rows are written directly, and the two combines
// accumulate in the wave's LDS (see CSPath).
// Built with -ffp-contract=off.
#include "vf_batch.hpp"
#include "vf_exact.hpp"

namespace abm {

namespace {

constexpr int kCSMaxRaysLDS = 7936;        // one wave's LDS (segments + accumulator) within 64 KB

// One object's row: not drawn (all zero), or the pieces of rays set to 1 and
// the row then multiplied by val (so 0 * val elsewhere, as numpy's row *=).
struct CSSeg {
    int a0, b0, a1, b1;
    int drawn;
    double val;
};

// Raster decision of one object for the cooperative-signalling field
// (cs_supcalc.py:238-287).
__device__ inline CSSeg cs_segment(const Focal& f, double ox, double oy, double meter, int R,
                                   double step, double fov0, double fov1, int has_max,
                                   double max_proj) {
#pragma clang fp contract(off)
    CSSeg s{0, 0, 0, 0, 0, 0.0};
    if (ox == f.px && oy == f.py) return s;                  // coincident (:240)
    const double ocx = ox + f.radius;                         // focal radius (:244)
    const double ocy = oy + f.radius;
    const double v2x = ocx - f.cx;                            // :247
    const double v2y = ocy - f.cy;
    const double ca = closed_angle_vf(f, v2x, v2y, norm2(v2x, v2y));   // :250
    const double dist = norm2(ocx - f.cx, ocy - f.cy);       // :252
    const double vis = 2.0 * atan(f.radius / (1.0 * dist));  // :255
    const int t = linspace_argmin(ca, R, step);               // :258
    if (!(fov0 <= ca && ca <= fov1)) return s;                // :262
    const double ps = (vis / kTwoPi) * (double)R;             // :267
    if (has_max && !(ps <= max_proj)) return s;               // :270
    const int h = (int)floor(ps / 2.0);                       // :273-274
    vf_pieces(t - h, t + h, R, &s.a0, &s.b0, &s.a1, &s.b1);  // :277-284
    s.drawn = 1;
    s.val = meter;
    return s;
}

// Value of the row at raw ray r.
__device__ inline double row_value(const CSSeg& s, int r) {
#pragma clang fp contract(off)
    if (!s.drawn) return 0.0;
    const bool on = (r >= s.a0 && r < s.b0) || (r >= s.a1 && r < s.b1);
    return (on ? 1.0 : 0.0) * s.val;
}

// numpy's maximum: NaN propagates.
__device__ inline double np_maximum(double a, double v) {
    if (a != a) return a;
    if (v != v) return v;
    return v > a ? v : a;
}

// Post positions [lo, hi] of the flipped field that the FOV mask keeps
// (cs_supcalc.py:291-294 zeroes phis < fov0 and phis > fov1; linspace is
// increasing, so the kept positions are one range, found from the same
// float64 linspace values).
ABM_HD void fov_keep_range(int R, double step, double fov0, double fov1, int* lo, int* hi) {
    int a = 0, b = R - 1;
    while (a < R && linspace_at(a, R, step) < fov0) ++a;
    while (b >= 0 && linspace_at(b, R, step) > fov1) --b;
    *lo = a;
    *hi = b;
}

// How a focal agent's rows are combined (wave-uniform):
//   kRows     every row written out;
//   kCount    SUM without meters: every row is 0/1, so the sum is a count per
//             ray, exact in any order -- a difference array (+1 at each
//             piece's start, -1 at its end, LDS atomics) and a prefix scan,
//             O(objects + R) instead of O(objects x R);
//   kMaxBits  MAX when every meter is a non-negative number other than -0.0:
//             such doubles order like their bit patterns as unsigned
//             integers, so each object raises its covered rays with a 64-bit
//             LDS atomic max, O(covered rays) (faster here than lanes
//             sharing one object at a time, whose read-max-write chain
//             serialises on LDS latency);
//   kOrdered  otherwise (meters in a SUM, or a NaN / negative / -0.0 meter in
//             a MAX): per ray, the rows folded in object order as numpy's
//             reduce does.
enum CSPath { kRows, kCount, kMaxBits, kOrdered };

__device__ inline bool plain_meter(double v) { return v >= 0.0 && !signbit(v); }

// LDS of the wave's lanes before the barrier visible to every lane after it.
__device__ inline void cs_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave per focal agent, `waves` per workgroup; each wave owns seg[64]
// and an accumulator of R + 1 doubles in LDS.
__global__ void __launch_bounds__(256)
cs_projection_kernel(int R, double step, double fov0, double fov1, int has_max, double max_proj,
                     const double* __restrict__ focal, const int64_t* __restrict__ obj_offset,
                     const double* __restrict__ obj, const double* __restrict__ meters,
                     const double* __restrict__ norm, int mode, double* __restrict__ out,
                     int n_focal, int keep_lo, int keep_hi) {
#pragma clang fp contract(off)
    extern __shared__ double smem[];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int waves = blockDim.x >> 6;
    const int f = blockIdx.x * waves + wave;
    if (f >= n_focal) return;                  // no workgroup barrier below
    CSSeg* seg = reinterpret_cast<CSSeg*>(smem) + 64 * wave;
    double* acc = smem + (sizeof(CSSeg) * 64 * waves) / sizeof(double) + (size_t)(R + 1) * wave;
    int* cnt = reinterpret_cast<int*>(acc);
    unsigned long long* mbits = reinterpret_cast<unsigned long long*>(acc);
    const double* fp = focal + 4 * (int64_t)f;
    const Focal F = make_focal(fp[0], fp[1], fp[3], fp[2]);
    const int64_t o0 = obj_offset[f], o1 = obj_offset[f + 1];

    CSPath path = mode == ABM_CS_ROWS ? kRows : kOrdered;
    if (mode == ABM_CS_SUM && !meters) path = kCount;
    if (mode == ABM_CS_MAX) {
        bool plain = true;
        if (meters)
            for (int64_t i = o0 + lane; i < o1; i += 64) plain = plain && plain_meter(meters[i]);
        path = __ballot(!plain) == 0ull ? kMaxBits : kOrdered;
    }
    if (path == kCount)
        for (int r = lane; r <= R; r += 64) cnt[r] = 0;
    else if (path == kMaxBits)
        for (int r = lane; r < R; r += 64) mbits[r] = 0ull;   // +0.0
    cs_wave_sync();

    for (int64_t c = o0; c < o1; c += 64) {
        const int n = (int)(o1 - c < 64 ? o1 - c : 64);
        if (lane < n) {
            const int64_t i = c + lane;
            const CSSeg sg = cs_segment(F, obj[2 * i], obj[2 * i + 1], meters ? meters[i] : 1.0, R,
                                        step, fov0, fov1, has_max, max_proj);
            if (path == kCount) {
                if (sg.drawn) {
                    if (sg.a0 < sg.b0) { atomicAdd(&cnt[sg.a0], 1); atomicAdd(&cnt[sg.b0], -1); }
                    if (sg.a1 < sg.b1) { atomicAdd(&cnt[sg.a1], 1); atomicAdd(&cnt[sg.b1], -1); }
                }
            } else if (path == kMaxBits) {
                const unsigned long long vb = (unsigned long long)__double_as_longlong(sg.val);
                if (sg.drawn && vb != 0ull) {
                    for (int r = sg.a0; r < sg.b0; ++r) atomicMax(&mbits[r], vb);
                    for (int r = sg.a1; r < sg.b1; ++r) atomicMax(&mbits[r], vb);
                }
            } else {
                seg[lane] = sg;
            }
        }
        cs_wave_sync();
        if (path == kRows) {
            // row i, post position p <- raw ray R - 1 - p (flip), masked rays 0
            for (int i = 0; i < n; ++i) {
                double* row = out + (c + i) * R;
                for (int p = lane; p < R; p += 64)
                    row[p] = (p < keep_lo || p > keep_hi) ? 0.0 : row_value(seg[i], R - 1 - p);
            }
        } else if (path == kOrdered) {
            for (int r = lane; r < R; r += 64) {
                // object order; numpy's reduce starts from the first row
                double a = c == o0 ? row_value(seg[0], r) : acc[r];
                for (int i = c == o0 ? 1 : 0; i < n; ++i) {
                    const double v = row_value(seg[i], r);
                    a = mode == ABM_CS_MAX ? np_maximum(a, v) : a + v;
                }
                acc[r] = a;
            }
        }
        cs_wave_sync();
    }
    if (path == kRows) return;
    if (path == kCount) {
        // inclusive prefix sum of the difference array: each lane walks a
        // contiguous run of rays; the run totals are scanned across the wave
        const int per = (R + 63) / 64;
        const int lo = lane * per < R ? lane * per : R, hi = lo + per < R ? lo + per : R;
        int local = 0;
        for (int r = lo; r < hi; ++r) local += cnt[r];
        int incl = local;
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        int run = incl - local;
        for (int r = lo; r < hi; ++r) {
            run += cnt[r];
            cnt[r] = run;
        }
        cs_wave_sync();
    }
    const double dv = norm ? norm[f] : 1.0;
    const bool empty = o1 <= o0;
    double* dst = out + (int64_t)f * R;
    for (int p = lane; p < R; p += 64) {
        const int r = R - 1 - p;
        double v;
        if (path == kCount) v = (double)cnt[r];
        else if (path == kMaxBits) v = __longlong_as_double((long long)mbits[r]);
        else v = empty ? 0.0 : acc[r];
        if (p < keep_lo || p > keep_hi) v = 0.0;
        if (mode == ABM_CS_SUM) v = v / dv;
        dst[p] = v;
    }
}

}  // namespace

}  // namespace abm

extern "C" {

int abm_cs_projection(const abm_cs_params* prm, int32_t n_focal, const double* focal,
                      const int64_t* obj_offset, const double* obj, const double* meters,
                      const double* norm, int32_t mode, double* out, void* stream) {
    using namespace abm;
    if (!prm) return set_error(ABM_ERR_INVALID, "abm_cs_projection: NULL params");
    const int R = prm->n_rays;
    if (R < 2 || R > kCSMaxRaysLDS)
        return set_error(ABM_ERR_INVALID, "abm_cs_projection: n_rays=%d outside [2, %d]", R, kCSMaxRaysLDS);
    if (mode != ABM_CS_ROWS && mode != ABM_CS_SUM && mode != ABM_CS_MAX)
        return set_error(ABM_ERR_INVALID, "abm_cs_projection: mode=%d", mode);
    if (n_focal < 0) return set_error(ABM_ERR_INVALID, "abm_cs_projection: n_focal < 0");
    if (n_focal == 0) return ABM_OK;
    if (!focal || !obj_offset || !out)
        return set_error(ABM_ERR_INVALID, "abm_cs_projection: NULL focal, offset or output array");
    if (!(prm->fov0 == prm->fov0) || !(prm->fov1 == prm->fov1))
        return set_error(ABM_ERR_INVALID, "abm_cs_projection: NaN field of view");
    const double step = linspace_step(R);
    int keep_lo, keep_hi;
    fov_keep_range(R, step, prm->fov0, prm->fov1, &keep_lo, &keep_hi);
    // waves (focal agents) per workgroup: four, fewer when their LDS would
    // not fit in 64 KB
    const size_t per_wave = sizeof(CSSeg) * 64 + sizeof(double) * ((size_t)R + 1);
    int waves = 4;
    while (waves > 1 && per_wave * waves > 65536) --waves;
    const int blocks = (n_focal + waves - 1) / waves;
    hipLaunchKernelGGL(cs_projection_kernel, dim3(blocks), dim3(64 * waves), per_wave * waves,
                       (hipStream_t)stream, R, step, prm->fov0, prm->fov1,
                       prm->has_max_proj_size ? 1 : 0, prm->max_proj_size, focal, obj_offset, obj,
                       meters, norm, (int)mode, out, (int)n_focal, keep_lo, keep_hi);
    return check_hip(hipGetLastError(), "cs_projection_kernel launch");
}

}  // extern "C"
