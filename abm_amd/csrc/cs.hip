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
// One workgroup per focal agent.  Threads decide a chunk of objects (exact
// float64 pair decision into LDS segments), then threads over rays fold the
// chunk into an LDS accumulator in object order; rows are written directly.
// Built with -ffp-contract=off.
#include "vf_batch.hpp"
#include "vf_exact.hpp"

namespace abm {

namespace {

constexpr int kCSThreads = 256;            // threads per focal agent
constexpr int kCSMaxRaysLDS = 8192;        // accumulator in LDS up to this R

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

// rays of the flipped field zeroed by the FOV mask (cs_supcalc.py:291-294):
// post position p holds raw ray R - 1 - p and is zero when phis[p] is outside
// [fov0, fov1]
__device__ inline bool fov_masked(int p, int R, double step, double fov0, double fov1) {
    const double ph = linspace_at(p, R, step);
    return ph < fov0 || ph > fov1;
}

__global__ void __launch_bounds__(kCSThreads)
cs_projection_kernel(int R, double step, double fov0, double fov1, int has_max, double max_proj,
                     const double* __restrict__ focal, const int64_t* __restrict__ obj_offset,
                     const double* __restrict__ obj, const double* __restrict__ meters,
                     const double* __restrict__ norm, int mode, double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ CSSeg seg[kCSThreads];
    extern __shared__ double acc[];            // [R] (SUM / MAX)
    const int f = blockIdx.x, tid = threadIdx.x;
    const double* fp = focal + 4 * (int64_t)f;
    const Focal F = make_focal(fp[0], fp[1], fp[3], fp[2]);
    const int64_t o0 = obj_offset[f], o1 = obj_offset[f + 1];
    for (int64_t c = o0; c < o1; c += kCSThreads) {
        const int n = (int)(o1 - c < kCSThreads ? o1 - c : kCSThreads);
        __syncthreads();
        if (tid < n) {
            const int64_t i = c + tid;
            seg[tid] = cs_segment(F, obj[2 * i], obj[2 * i + 1], meters ? meters[i] : 1.0, R, step,
                                  fov0, fov1, has_max, max_proj);
        }
        __syncthreads();
        if (mode == ABM_CS_ROWS) {
            // row i, post position p <- raw ray R - 1 - p (flip), masked rays 0
            for (int64_t e = tid; e < (int64_t)n * R; e += kCSThreads) {
                const int i = (int)(e / R), p = (int)(e % R);
                double v = row_value(seg[i], R - 1 - p);
                if (fov_masked(p, R, step, fov0, fov1)) v = 0.0;
                out[(c + i) * R + p] = v;
            }
        } else {
            for (int r = tid; r < R; r += kCSThreads) {
                // object order; numpy's reduce starts from the first row
                double a = c == o0 ? row_value(seg[0], r) : acc[r];
                for (int i = c == o0 ? 1 : 0; i < n; ++i) {
                    const double v = row_value(seg[i], r);
                    a = mode == ABM_CS_MAX ? np_maximum(a, v) : a + v;
                }
                acc[r] = a;
            }
        }
    }
    if (mode == ABM_CS_ROWS) return;
    __syncthreads();
    const double dv = norm ? norm[f] : 1.0;
    const bool empty = o1 <= o0;
    for (int p = tid; p < R; p += kCSThreads) {
        double v = empty ? 0.0 : acc[R - 1 - p];
        if (fov_masked(p, R, step, fov0, fov1)) v = 0.0;
        if (mode == ABM_CS_SUM) v = v / dv;
        out[(int64_t)f * R + p] = v;
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
    const size_t lds = mode == ABM_CS_ROWS ? 0 : sizeof(double) * (size_t)R;
    hipLaunchKernelGGL(cs_projection_kernel, dim3(n_focal), dim3(kCSThreads), lds, (hipStream_t)stream, R,
                       step, prm->fov0, prm->fov1, prm->has_max_proj_size ? 1 : 0, prm->max_proj_size,
                       focal, obj_offset, obj, meters, norm, (int)mode, out);
    return check_hip(hipGetLastError(), "cs_projection_kernel launch");
}

}  // extern "C"
