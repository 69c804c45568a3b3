// vf_f64.hip -- plain float64 VSWRM step: the reference arithmetic on the GPU.
//
// One workgroup per replica.  The replica's poses live in LDS for the whole
// step and are updated in place, agent by agent in id order, which is the
// reference's Group.update semantics (vf_sims.py:302; agent k sees agents <k
// at their new poses).  Per agent:
//   1. lanes over objects: exact float64 pair decision (vf_exact.hpp), and
//      each drawn interval written as bytes into a per-ray LDS field;
//   2. lanes over rays: numpy's trapz terms and dPhi spike terms
//      (vf_supcalc.py:199-254, 257-277) summed with a block reduction;
//   3. one lane: VSWRM motion, Euler step and boundary (vf_agent.py:188-204,
//      263-326; agent.py:347-395, 605-610).
// Built with -ffp-contract=off.
#include "vf_batch.hpp"
#include "vf_exact.hpp"

namespace abm {

namespace {

__device__ inline double wave_sum(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Block-wide float64 sums over the raw raster field V (0/1 bytes in LDS):
// trapz(cos(PHI)*(-V), PHI), trapz(sin(PHI)*(-V), PHI), sum(cos(PHI)*dPhi_V^2),
// sum(sin(PHI)*dPhi_V^2) (vf_supcalc.py:199-238, 257-277).  Each trapz term is
// numpy's d_k*(y_{k+1}+y_k)/2 with y = cos(PHI)*G, G = -V + ALP2*0.  Leaves the
// per-wave partial sums in red[4*wave + 0..3]; the caller synchronises.
__device__ void integrals_f64(const uint8_t* V, int R, const double* cs, const double* sn,
                              const double* dp, int tid, int nthr, double* red) {
#pragma clang fp contract(off)
    const bool case1 = V[0] == 1 && V[R - 1] == 0;
    double tc = 0.0, ts = 0.0, ec = 0.0, es = 0.0;
    for (int r = tid; r < R; r += nthr) {
        const double vr = V[r] ? 1.0 : 0.0;
        const double gr = -vr + 0.0;
        if (r < R - 1) {
            const double vn = V[r + 1] ? 1.0 : 0.0;
            const double gn = -vn + 0.0;
            const double yc = cs[r] * gr, ycn = cs[r + 1] * gn;
            const double ys = sn[r] * gr, ysn = sn[r + 1] * gn;
            tc += (dp[r] * (ycn + yc)) / 2.0;
            ts += (dp[r] * (ysn + ys)) / 2.0;
        }
        double o;
        if (case1) o = vr - (V[r == 0 ? R - 1 : r - 1] ? 1.0 : 0.0);
        else o = (V[r == R - 1 ? 0 : r + 1] ? 1.0 : 0.0) - vr;
        const double sq = o * o;
        ec += cs[r] * sq;
        es += sn[r] * sq;
    }
    tc = wave_sum(tc); ts = wave_sum(ts); ec = wave_sum(ec); es = wave_sum(es);
    if ((tid & 63) == 0) {
        red[4 * (tid >> 6) + 0] = tc; red[4 * (tid >> 6) + 1] = ts;
        red[4 * (tid >> 6) + 2] = ec; red[4 * (tid >> 6) + 3] = es;
    }
}

__global__ void __launch_bounds__(256)
vf_step_f64_kernel(const VFRep* __restrict__ reps, const double* __restrict__ tcos,
                   const double* __restrict__ tsin, const double* __restrict__ tdphi,
                   abm_vf_state_f64 st, uint32_t* __restrict__ bits_out,
                   double* __restrict__ diag, uint32_t flags, unsigned long long* __restrict__ counters) {
#pragma clang fp contract(off)
    extern __shared__ double smem[];
    const VFRep P = reps[blockIdx.x];
    // near-tie margin count: pairs whose bin or half-width decision lies
    // within 1e-12 of its tie, where the device's and the host's libm (acos,
    // atan) may round the deciding value to opposite sides
    const bool count = (flags & ABM_COUNT_RECHECKS) != 0 && counters != nullptr;
    unsigned long long n_near = 0;
    const int N = P.n_agents, R = P.n_rays;
    const int tid = threadIdx.x, nthr = blockDim.x, nw = nthr >> 6;
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;

    double* sx = smem;
    double* sy = sx + N;
    double* sth = sy + N;
    double* sv = sth + N;
    double* srad = sv + N;
    double* ox = sync_order ? srad + N : sx;   // object positions
    double* oy = sync_order ? srad + 2 * N : sy;
    double* red = (sync_order ? srad + 3 * N : srad + N);
    double* shared_out = red + 4 * nw;          // 2 doubles: dv, dpsi broadcast
    uint8_t* V = reinterpret_cast<uint8_t*>(shared_out + 8);

    const int a0 = P.agent_offset;
    for (int i = tid; i < N; i += nthr) {
        sx[i] = st.x[a0 + i];
        sy[i] = st.y[a0 + i];
        sth[i] = st.theta[a0 + i];
        sv[i] = st.v[a0 + i];
        srad[i] = (double)st.radius[a0 + i];
        if (sync_order) { ox[i] = sx[i]; oy[i] = sy[i]; }
    }
    const double* cs = tcos + P.table_offset;
    const double* sn = tsin + P.table_offset;
    const double* dp = tdphi + P.table_offset;
    const bool torus = P.boundary == ABM_BOUNDARY_INFINITE;
    __syncthreads();

    for (int k = 0; k < N; ++k) {
        for (int r = tid; r < R; r += nthr) V[r] = 0;
        __syncthreads();
        const Focal f = make_focal(sx[k], sy[k], sth[k], srad[k]);
        for (int j = tid; j < N; j += nthr) {
            if (j == k) continue;
            PairVF p = pair_vf_exact(f, ox[j], oy[j], srad[j], torus, P.width, P.height,
                                     R, P.step, P.fov_px0, P.fov_px1);
            if (count && !p.skipped) {
                const double ca = p.closed_angle;
                const double dt = fabs(linspace_at(p.t, R, P.step) - ca);
                double dn = 1e300;
                if (p.t > 0) dn = fmin(dn, fabs(linspace_at(p.t - 1, R, P.step) - ca));
                if (p.t < R - 1) dn = fmin(dn, fabs(linspace_at(p.t + 1, R, P.step) - ca));
                const double hf = p.half_f, fr = hf - floor(hf);
                const bool near_bin = dn - dt < 1e-12;
                const bool near_half = hf >= 0.5 && (fr < 1e-12 * hf || 1.0 - fr < 1e-12 * hf);
                if (near_bin || near_half) ++n_near;
            }
            if (p.skipped || !p.drawn) continue;
            int a0p, b0p, a1p, b1p;
            vf_pieces(p.start, p.end, R, &a0p, &b0p, &a1p, &b1p);
            for (int r = a0p; r < b0p; ++r) V[r] = 1;
            for (int r = a1p; r < b1p; ++r) V[r] = 1;
        }
        __syncthreads();

        integrals_f64(V, R, cs, sn, dp, tid, nthr, red);
        // pack the field for output while the sums settle
        if (bits_out) {
            const int W = P.words;
            uint32_t* dst = bits_out + P.field_offset + (int64_t)k * W;
            for (int w = tid; w < W; w += nthr) {
                uint32_t m = 0;
                for (int b = 0; b < 32; ++b) {
                    int r = w * 32 + b;
                    if (r < R && V[r]) m |= 1u << b;
                }
                dst[w] = m;
            }
        }
        __syncthreads();
        if (tid == 0) {
            double Tc = 0.0, Ts = 0.0, Ec = 0.0, Es = 0.0;
            for (int w = 0; w < nw; ++w) {
                Tc += red[4 * w]; Ts += red[4 * w + 1]; Ec += red[4 * w + 2]; Es += red[4 * w + 3];
            }
            const int ag = a0 + k;
            const double alp0 = st.alp0 ? st.alp0[ag] : P.alp0;
            const double bet0 = st.bet0 ? st.bet0[ag] : P.bet0;
            const double v0 = st.v0 ? st.v0[ag] : P.v0;
            double dv = 0.0, dpsi = 0.0, ab = 0.0, ae = 0.0, bb = 0.0, be = 0.0;
            const double vel = sv[k];
            if (P.phi_ok) {
                ab = alp0 * Tc;
                ae = (alp0 * P.alp1) * Ec;
                bb = bet0 * Ts;
                be = (bet0 * P.bet1) * Es;
                dv = P.gam * (v0 - vel) + ab + ae;
                dpsi = bb + be;
            }
            Pose p{sx[k], sy[k], sth[k], vel};
            p = vf_move(p, dpsi, dv, P.limit_movement, P.max_th, P.max_vel);
            if (torus) p = teleport(p, srad[k], P.pad, P.width, P.height);
            else p = reflect_walls(p, srad[k], P.pad, P.width, P.height);
            sx[k] = p.x; sy[k] = p.y; sth[k] = p.theta; sv[k] = p.v;
            if (diag && (flags & ABM_WANT_DIAG)) {
                double* d = diag + 6 * (int64_t)ag;
                d[0] = dv; d[1] = dpsi; d[2] = ab; d[3] = ae; d[4] = bb; d[5] = be;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < N; i += nthr) {
        st.x[a0 + i] = sx[i];
        st.y[a0 + i] = sy[i];
        st.theta[a0 + i] = sth[i];
        st.v[a0 + i] = sv[i];
    }
    if (count && n_near) atomicAdd(&counters[3], n_near);
}

}  // namespace

size_t vf_f64_lds_bytes(int max_agents, int max_rays, int nthr, bool sync_order) {
    size_t d = (size_t)max_agents * (sync_order ? 7 : 5) + 4 * (nthr / 64) + 8;
    return d * sizeof(double) + (size_t)max_rays + 16;
}

int launch_vf_step_f64(abm_vf_batch* b, const abm_vf_state_f64* st, uint32_t* bits,
                       double* diag, uint32_t flags, hipStream_t s) {
    const bool sync_order = (flags & ABM_ORDER_SYNCHRONOUS) != 0;
    int nthr = 64;
    if (b->max_rays >= 1024 || b->max_agents > 128) nthr = 128;
    if (b->max_agents > 512) nthr = 256;
    size_t lds = vf_f64_lds_bytes(b->max_agents, b->max_rays, nthr, sync_order);
    if (lds > (size_t)b->dev_max_lds)
        return set_error(ABM_ERR_UNSUPPORTED,
                         "abm_vf_step_f64: %d agents need %zu B of LDS per replica (device max %d)",
                         b->max_agents, lds, b->dev_max_lds);
    if (lds > 65536) {
        if (int rc = check_hip(hipFuncSetAttribute((const void*)vf_step_f64_kernel,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)lds), "hipFuncSetAttribute"))
            return rc;
    }
    hipLaunchKernelGGL(vf_step_f64_kernel, dim3(b->n_replicas), dim3(nthr), lds, s,
                       b->reps_dev, b->tab_cos, b->tab_sin, b->tab_dphi, *st, bits, diag, flags,
                       (unsigned long long*)b->counters);
    b->last_kernels = ABM_KERNEL_F64;
    return check_hip(hipGetLastError(), "vf_step_f64_kernel launch");
}

// ---------------------------------------------------------------- probes
namespace {

__global__ void vf_projection_rows_kernel(int R, double step, int fov_px0, int fov_px1,
                                          int torus, double W, double H, double fx, double fy,
                                          double ftheta, double fradius,
                                          const double* __restrict__ ox,
                                          const double* __restrict__ oy,
                                          const double* __restrict__ osize, int n_obj,
                                          uint32_t* __restrict__ rows, int32_t* __restrict__ rec) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_obj) return;
    const Focal f = make_focal(fx, fy, ftheta, fradius);
    const PairVF p = pair_vf_exact(f, ox[i], oy[i], osize[i], torus, W, H, R, step,
                                   fov_px0, fov_px1);
    const int words = (R + 31) / 32;
    uint32_t* row = rows + (int64_t)i * words;
    int a0 = 0, b0 = 0, a1 = 0, b1 = 0;
    if (!p.skipped && p.drawn) vf_pieces(p.start, p.end, R, &a0, &b0, &a1, &b1);
    for (int w = 0; w < words; ++w) {
        uint32_t m = 0;
        for (int b = 0; b < 32; ++b) {
            const int r = w * 32 + b;
            if ((r >= a0 && r < b0) || (r >= a1 && r < b1)) m |= 1u << b;
        }
        row[w] = m;
    }
    if (rec) {
        int32_t* q = rec + 5 * (int64_t)i;
        q[0] = p.t; q[1] = p.half; q[2] = p.start; q[3] = p.end;
        q[4] = p.skipped ? -1 : p.drawn;
    }
}

__global__ void __launch_bounds__(256)
vf_state_variables_kernel(int R, const uint32_t* __restrict__ bits, const double* __restrict__ vel,
                          const double* __restrict__ alp0, const double* __restrict__ bet0,
                          const double* __restrict__ v0, double gam, double alp1, double bet1,
                          const double* __restrict__ tcos, const double* __restrict__ tsin,
                          const double* __restrict__ tdphi, int phi_ok,
                          double* __restrict__ out, int8_t* __restrict__ dphi) {
#pragma clang fp contract(off)
    extern __shared__ double smem[];
    const int row = blockIdx.x, tid = threadIdx.x, nthr = blockDim.x, nw = nthr >> 6;
    double* red = smem;
    uint8_t* V = reinterpret_cast<uint8_t*>(smem + 4 * nw);
    const int words = (R + 31) / 32;
    const uint32_t* src = bits + (int64_t)row * words;
    for (int r = tid; r < R; r += nthr) V[r] = (src[r >> 5] >> (r & 31)) & 1u;
    __syncthreads();
    integrals_f64(V, R, tcos, tsin, tdphi, tid, nthr, red);
    if (dphi) {
        const bool case1 = V[0] == 1 && V[R - 1] == 0;
        for (int r = tid; r < R; r += nthr) {
            int o;
            if (case1) o = (int)V[r] - (int)V[r == 0 ? R - 1 : r - 1];
            else o = (int)V[r == R - 1 ? 0 : r + 1] - (int)V[r];
            dphi[(int64_t)row * R + r] = (int8_t)o;
        }
    }
    __syncthreads();
    if (tid == 0) {
        double Tc = 0.0, Ts = 0.0, Ec = 0.0, Es = 0.0;
        for (int w = 0; w < nw; ++w) {
            Tc += red[4 * w]; Ts += red[4 * w + 1]; Ec += red[4 * w + 2]; Es += red[4 * w + 3];
        }
        const double a0 = alp0[row], b0 = bet0[row], vv0 = v0[row], vn = vel[row];
        double* o = out + 6 * (int64_t)row;
        if (phi_ok) {
            const double ab = a0 * Tc, ae = (a0 * alp1) * Ec;
            const double bb = b0 * Ts, be = (b0 * bet1) * Es;
            o[0] = gam * (vv0 - vn) + ab + ae;
            o[1] = bb + be;
            o[2] = ab; o[3] = ae; o[4] = bb; o[5] = be;
        } else {
            for (int q = 0; q < 6; ++q) o[q] = 0.0;
        }
    }
}

}  // namespace

int launch_vf_projection_rows(int R, double fov0, double fov1, int boundary, double W, double H,
                              double fx, double fy, double ftheta, double fradius,
                              const double* ox, const double* oy, const double* osize, int n_obj,
                              uint32_t* rows, int32_t* rec, hipStream_t s) {
    const double step = linspace_step(R);
    const int px0 = linspace_argmin(fov0, R, step), px1 = linspace_argmin(fov1, R, step);
    if (n_obj == 0) return ABM_OK;
    hipLaunchKernelGGL(vf_projection_rows_kernel, dim3((n_obj + 63) / 64), dim3(64), 0, s, R, step,
                       px0, px1, boundary == ABM_BOUNDARY_INFINITE ? 1 : 0, W, H, fx, fy, ftheta,
                       fradius, ox, oy, osize, n_obj, rows, rec);
    return check_hip(hipGetLastError(), "vf_projection_rows_kernel launch");
}

int launch_vf_state_variables(int R, int n, const uint32_t* bits, const double* vel,
                              const double* alp0, const double* bet0, const double* v0,
                              double gam, double alp1, double bet1, const double* tcos,
                              const double* tsin, const double* tdphi, int phi_ok, double* out,
                              int8_t* dphi, hipStream_t s) {
    if (n == 0) return ABM_OK;
    const int nthr = R >= 1024 ? 256 : 64;
    const size_t lds = (size_t)4 * (nthr / 64) * sizeof(double) + (size_t)R + 16;
    hipLaunchKernelGGL(vf_state_variables_kernel, dim3(n), dim3(nthr), lds, s, R, bits, vel, alp0,
                       bet0, v0, gam, alp1, bet1, tcos, tsin, tdphi, phi_ok, out, dphi);
    return check_hip(hipGetLastError(), "vf_state_variables_kernel launch");
}

}  // namespace abm
