// stats.hip -- per-step ensemble statistics of many replicas on the GPU.
//
// The reference judges runs offline from the logged trajectories
// (abm/loader/data_loader.py): polarisation (calculate_polarization,
// :1761-1836), the inter-individual distance matrix and its mean over pairs
// (calculate_interindividual_distance, :1367-1459) and the mean nearest-
// neighbour distance (calculate_mean_NN_dist, :1461-1486).  Logged positions
// are the sprite positions truncated to int (ifdb.py:83-84); orientations are
// logged as floats.  This kernel computes the same per-step, per-replica
// quantities from the state in HBM, so long batched runs are summarised
// without copying trajectories to the host:
//   [0] polarisation   |sum_i (cos th_i, sin th_i)| / N (numpy's 2-norm:
//                      sqrt(x*x + y*y), no fused multiply-add)
//   [1] mean iid       mean over pairs i < j of the distance between the
//                      int-truncated positions (supcalc.distance_coords, or
//                      supcalc.distance_torus on a torus of int(W) x int(H))
//   [2] mean NN        mean over agents of the distance to the nearest other
//                      agent (same distances)
//   [3] reference NN   what calculate_mean_NN_dist returns for this step: it
//                      takes np.nanmin over the second agent axis j of the
//                      iid matrix iid[i, j], which holds distances only for
//                      i < j (zeros below the diagonal, NaN on it), so every
//                      agent but the first contributes 0 and agent 0 its
//                      nearest neighbour: min_j>0 d(0, j) / N (NaN for N = 1)
// One wave per replica; the pair sums are reduced in a fixed order (lanes,
// then a wave tree), so results are deterministic.
// Built with -ffp-contract=off.
#include "vf_batch.hpp"

namespace abm {

namespace {

__device__ inline double wave_sum_d(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ inline double wave_min_d(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o, 64));
    return x;
}

// supcalc.distance_coords / distance_torus between two logged positions
__device__ inline double logged_distance(double x1, double y1, double x2, double y2, int torus,
                                         double dw, double dh) {
#pragma clang fp contract(off)
    double dx = x2 - x1, dy = y2 - y1;
    if (torus) {
        dx = fabs(dx);
        dy = fabs(dy);
        if (dx > 0.5 * dw) dx = dx - dw;
        if (dy > 0.5 * dh) dy = dy - dh;
    }
    return sqrt(dx * dx + dy * dy);
}

template <typename T>
__global__ void __launch_bounds__(256)
ensemble_stats_kernel(int n_rep, const int32_t* __restrict__ agent_offset, const T* __restrict__ x,
                      const T* __restrict__ y, const T* __restrict__ theta, int torus, double dw,
                      double dh, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (b >= n_rep) return;
    const int a0 = agent_offset[b], n = agent_offset[b + 1] - a0;
    double* o = out + 4 * (int64_t)b;
    if (n <= 0) {
        if (lane == 0) { o[0] = 0.0; o[1] = 0.0; o[2] = 0.0; o[3] = 0.0; }
        return;
    }
    // polarisation: numpy's nansum over agents skips NaN orientations
    double cx = 0.0, cy = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double t = (double)theta[a0 + i];
        if (t == t) { cx += cos(t); cy += sin(t); }
    }
    cx = wave_sum_d(cx);
    cy = wave_sum_d(cy);
    const double pol = sqrt(cx * cx + cy * cy) / (double)n;
    // distances between logged (int-truncated) positions
    double psum = 0.0, nnsum = 0.0, first_nn = HUGE_VAL;
    const double x0 = trunc((double)x[a0]), y0 = trunc((double)y[a0]);
    for (int i = lane; i < n; i += 64) {
        const double xi = trunc((double)x[a0 + i]), yi = trunc((double)y[a0 + i]);
        double nn = HUGE_VAL;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double d = logged_distance(xi, yi, trunc((double)x[a0 + j]), trunc((double)y[a0 + j]),
                                             torus, dw, dh);
            if (j > i) psum += d;
            nn = fmin(nn, d);
        }
        nnsum += nn;
        // iid[0, i] as the reference stores it (pair ordered 0 < i)
        if (i > 0) first_nn = fmin(first_nn, logged_distance(x0, y0, xi, yi, torus, dw, dh));
    }
    psum = wave_sum_d(psum);
    nnsum = wave_sum_d(nnsum);
    first_nn = wave_min_d(first_nn);
    if (lane == 0) {
        const double pairs = 0.5 * (double)n * (double)(n - 1);
        o[0] = pol;
        o[1] = n > 1 ? psum / pairs : NAN;
        o[2] = n > 1 ? nnsum / (double)n : NAN;
        o[3] = n > 1 ? first_nn / (double)n : NAN;
    }
}

}  // namespace

}  // namespace abm

extern "C" {

int abm_ensemble_stats(int32_t n_replicas, const int32_t* agent_offset, const void* x, const void* y,
                       const void* theta, int32_t state_f32, int32_t periodic, double width,
                       double height, double* out, void* stream) {
    using namespace abm;
    if (n_replicas < 0) return set_error(ABM_ERR_INVALID, "abm_ensemble_stats: n_replicas < 0");
    if (n_replicas == 0) return ABM_OK;
    if (!agent_offset || !x || !y || !theta || !out)
        return set_error(ABM_ERR_INVALID, "abm_ensemble_stats: NULL array");
    if (periodic && !(width > 0 && height > 0))
        return set_error(ABM_ERR_INVALID, "abm_ensemble_stats: a torus needs width and height");
    // distance_torus uses the int() of the arena size (data_loader.py:1426)
    const double dw = periodic ? (double)(int64_t)width : 0.0, dh = periodic ? (double)(int64_t)height : 0.0;
    const int waves = 4, blocks = (n_replicas + waves - 1) / waves;
    if (state_f32)
        hipLaunchKernelGGL(ensemble_stats_kernel<float>, dim3(blocks), dim3(64 * waves), 0,
                           (hipStream_t)stream, (int)n_replicas, agent_offset, (const float*)x,
                           (const float*)y, (const float*)theta, periodic ? 1 : 0, dw, dh, out);
    else
        hipLaunchKernelGGL(ensemble_stats_kernel<double>, dim3(blocks), dim3(64 * waves), 0,
                           (hipStream_t)stream, (int)n_replicas, agent_offset, (const double*)x,
                           (const double*)y, (const double*)theta, periodic ? 1 : 0, dw, dh, out);
    return check_hip(hipGetLastError(), "ensemble_stats_kernel launch");
}

}  // extern "C"
