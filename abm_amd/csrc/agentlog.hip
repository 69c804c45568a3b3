// agentlog.hip -- per-step agent records kept on the GPU.
//
// The reference logs every agent every step into a dict of lists
// (abm/monitoring/ifdb.py:55-96, save_agent_data_RAM): int(position), the
// orientation and velocity, the foraging state (w, u, I_priv, mode,
// collected_r, exploited_patch_id) and, instead of the visual field, its
// edges: vfield_up = np.where(np.roll(f, 1) < f)[0] and vfield_down =
// np.where(np.roll(f, 1) > f)[0] of the agent's soc_v_field f.  This kernel
// appends one step of those records for every agent of a batch into device
// arrays indexed [t][agent], the edges as bit masks in output order, so a
// batched run logs without a host round trip per step; the host turns the
// arrays into the reference's dict (and its JSON file) at the end.
#include "vf_batch.hpp"

namespace abm {

namespace {

// bit k of the agent's soc_v_field: output order, or the flipped raw raster
__device__ inline uint32_t soc_bit(const uint32_t* f, int k, int R, int raw) {
    const int r = raw ? R - 1 - k : k;
    return (f[r >> 5] >> (r & 31)) & 1u;
}

template <typename T>
__global__ void __launch_bounds__(256)
log_record_kernel(abm_log_buffers buf, int64_t t, int n, const T* __restrict__ x,
                  const T* __restrict__ y, const T* __restrict__ theta, const T* __restrict__ v,
                  const double* __restrict__ w, const double* __restrict__ u,
                  const double* __restrict__ ip, const double* __restrict__ cr,
                  const int32_t* __restrict__ mode, const int32_t* __restrict__ patch,
                  const uint32_t* __restrict__ bits, int R, int raw) {
    const int words = bits ? (R + 31) / 32 : 1;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * words) return;
    const int a = (int)(e / words), wd = (int)(e % words);
    const int64_t s = t * n + a;
    if (wd == 0) {
        // int() truncates toward zero; a non-finite position has no int
        const double px = (double)x[a], py = (double)y[a];
        buf.posx[s] = isfinite(px) ? (int32_t)trunc(px) : INT32_MIN;
        buf.posy[s] = isfinite(py) ? (int32_t)trunc(py) : INT32_MIN;
        buf.orientation[s] = (double)theta[a];
        buf.velocity[s] = (double)v[a];
        if (buf.w && w) buf.w[s] = w[a];
        if (buf.u && u) buf.u[s] = u[a];
        if (buf.I_priv && ip) buf.I_priv[s] = ip[a];
        if (buf.collected_r && cr) buf.collected_r[s] = cr[a];
        if (buf.mode && mode) buf.mode[s] = mode[a];
        if (buf.expl_patch_id && patch) buf.expl_patch_id[s] = patch[a];
    }
    if (!bits || !buf.up || !buf.down) return;
    const uint32_t* f = bits + (int64_t)a * words;
    uint32_t up = 0u, down = 0u;
    for (int b = 0; b < 32; ++b) {
        const int k = wd * 32 + b;
        if (k >= R) break;
        const uint32_t cur = soc_bit(f, k, R, raw);
        const uint32_t prev = soc_bit(f, k == 0 ? R - 1 : k - 1, R, raw);   // np.roll(f, 1)[k]
        up |= (prev < cur ? 1u : 0u) << b;
        down |= (prev > cur ? 1u : 0u) << b;
    }
    buf.up[s * words + wd] = up;
    buf.down[s * words + wd] = down;
}

}  // namespace

}  // namespace abm

extern "C" {

int abm_log_record(const abm_log_buffers* buf, int64_t t, int64_t capacity, int32_t n_agents,
                   const void* x, const void* y, const void* theta, const void* v, int32_t state_f32,
                   const double* w, const double* u, const double* I_priv, const double* collected_r,
                   const int32_t* mode, const int32_t* expl_patch_id, const uint32_t* field_bits,
                   int32_t n_rays, int32_t raw_raster, void* stream) {
    using namespace abm;
    if (!buf) return set_error(ABM_ERR_INVALID, "abm_log_record: NULL buffers");
    if (n_agents < 0) return set_error(ABM_ERR_INVALID, "abm_log_record: n_agents < 0");
    if (t < 0 || t >= capacity)
        return set_error(ABM_ERR_INVALID, "abm_log_record: step %lld outside the log's %lld steps",
                         (long long)t, (long long)capacity);
    if (n_agents == 0) return ABM_OK;
    if (!buf->posx || !buf->posy || !buf->orientation || !buf->velocity || !x || !y || !theta || !v)
        return set_error(ABM_ERR_INVALID, "abm_log_record: NULL state or record array");
    if (field_bits && (n_rays < 1 || n_rays > (1 << 16)))
        return set_error(ABM_ERR_INVALID, "abm_log_record: n_rays=%d", n_rays);
    const int words = field_bits ? (n_rays + 31) / 32 : 1;
    const int64_t work = (int64_t)n_agents * words;
    const int threads = 256;
    const int64_t blocks = (work + threads - 1) / threads;
    if (state_f32)
        hipLaunchKernelGGL(log_record_kernel<float>, dim3((unsigned)blocks), dim3(threads), 0,
                           (hipStream_t)stream, *buf, t, (int)n_agents, (const float*)x, (const float*)y,
                           (const float*)theta, (const float*)v, w, u, I_priv, collected_r, mode,
                           expl_patch_id, field_bits, (int)n_rays, raw_raster ? 1 : 0);
    else
        hipLaunchKernelGGL(log_record_kernel<double>, dim3((unsigned)blocks), dim3(threads), 0,
                           (hipStream_t)stream, *buf, t, (int)n_agents, (const double*)x, (const double*)y,
                           (const double*)theta, (const double*)v, w, u, I_priv, collected_r, mode,
                           expl_patch_id, field_bits, (int)n_rays, raw_raster ? 1 : 0);
    return check_hip(hipGetLastError(), "log_record_kernel launch");
}

}  // extern "C"
