// capi.hip -- the extern "C" entry points of libabm_hip.so: argument
// validation, batch creation (replica offsets, ray tables, upload), dispatch
// to the kernels, and the per-thread last-error string.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/abm_amd.h"
#include "forage_batch.hpp"
#include "vf_batch.hpp"
#include "vf_exact.hpp"

namespace abm {

static thread_local char g_err[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return ABM_OK;
    return set_error(ABM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// launchers defined in the kernel translation units
int launch_vf_projection_rows(int R, double fov0, double fov1, int boundary, double W, double H,
                              double fx, double fy, double ftheta, double fradius,
                              const double* ox, const double* oy, const double* osize, int n_obj,
                              uint32_t* rows, int32_t* rec, hipStream_t s);
int launch_vf_state_variables(int R, int n, const uint32_t* bits, const double* vel,
                              const double* alp0, const double* bet0, const double* v0,
                              double gam, double alp1, double bet1, const double* tcos,
                              const double* tsin, const double* tdphi, int phi_ok, double* out,
                              int8_t* dphi, hipStream_t s);

// float64 ray tables of one R: cos(PHI_k), sin(PHI_k), diff(PHI)_k (last = 0),
// PHI = np.arange(-pi, pi, 2*pi/R) (vf_agent.py:44) when its length is R.
struct HostTables {
    std::vector<double> c, s, d;
    std::vector<RayTab32> t32;
};

static void make_tables(int R, HostTables* t) {
    const double delta = arange_delta(R);
    std::vector<double> phi(R);
    for (int k = 0; k < R; ++k) phi[k] = arange_at(k, delta);
    t->c.resize(R); t->s.resize(R); t->d.assign(R, 0.0); t->t32.resize(R);
    for (int k = 0; k < R; ++k) { t->c[k] = std::cos(phi[k]); t->s[k] = std::sin(phi[k]); }
    for (int k = 0; k + 1 < R; ++k) t->d[k] = phi[k + 1] - phi[k];
    for (int k = 0; k < R; ++k) {
        // trapz weights: w_0 = d_0/2, w_k = (d_{k-1}+d_k)/2, w_{R-1} = d_{R-2}/2
        const double wl = k > 0 ? t->d[k - 1] : 0.0;
        const double wr = k + 1 < R ? t->d[k] : 0.0;
        const double w = 0.5 * (wl + wr);
        t->t32[k].wcos = (float)(w * t->c[k]);
        t->t32[k].wsin = (float)(w * t->s[k]);
        t->t32[k].cos = (float)t->c[k];
        t->t32[k].sin = (float)t->s[k];
    }
}

// Cached float64 tables per R for the probe entry points.
struct ProbeTables { double* c; double* s; double* d; };
static std::mutex g_probe_mu;
static std::map<std::pair<int, int>, ProbeTables> g_probe_tables;

static int probe_tables(int R, ProbeTables* out) {
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
    std::lock_guard<std::mutex> lk(g_probe_mu);
    auto key = std::make_pair(dev, R);
    auto it = g_probe_tables.find(key);
    if (it != g_probe_tables.end()) { *out = it->second; return ABM_OK; }
    HostTables t;
    make_tables(R, &t);
    ProbeTables p{};
    if (int rc = check_hip(hipMalloc(&p.c, 3 * sizeof(double) * R), "hipMalloc(tables)")) return rc;
    p.s = p.c + R; p.d = p.c + 2 * R;
    std::vector<double> all(3 * (size_t)R);
    std::copy(t.c.begin(), t.c.end(), all.begin());
    std::copy(t.s.begin(), t.s.end(), all.begin() + R);
    std::copy(t.d.begin(), t.d.end(), all.begin() + 2 * R);
    if (int rc = check_hip(hipMemcpy(p.c, all.data(), all.size() * sizeof(double),
                                     hipMemcpyHostToDevice), "hipMemcpy(tables)"))
        return rc;
    g_probe_tables[key] = p;
    *out = p;
    return ABM_OK;
}

// ------------------------------------------------ no-allocation probe
// abm_vf_state_variables_noalloc keeps the float64 ray tables it needs in
// this static device array instead of memory the library allocates.  Every
// GPU the library runs on holds its own copy until the process ends; host
// code reaches a device's copy only through the address hipGetSymbolAddress
// returns there (kept per device below), never through the array's name.
// Every call on a GPU rewrites that GPU's copy, so calls that overlap on one
// GPU (other streams, other host threads) race on it.
constexpr int kNoAllocMaxRays = 8192;
static __device__ __align__(256) double scratch[3 * kNoAllocMaxRays];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
static int noalloc_scratch(double** out) {
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(ABM_ERR_UNSUPPORTED, "no-allocation entry point: device %d beyond the %d "
                         "it keeps table space for", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                               "hipGetSymbolAddress(scratch)"))
            return rc;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ABM_OK;
}

// cos(PHI_k), sin(PHI_k) and diff(PHI)_k (last 0) of PHI = np.arange(-pi, pi,
// 2 pi / R), built on the device: the grid is exact (numpy's arange
// arithmetic), cos and sin are the device's (within an ulp of the host's).
__global__ void ray_tables_f64_kernel(int R, double delta, double* __restrict__ c,
                                      double* __restrict__ s, double* __restrict__ d) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= R) return;
    const double phi = arange_at(k, delta);
    c[k] = cos(phi);
    s[k] = sin(phi);
    d[k] = k + 1 < R ? arange_at(k + 1, delta) - phi : 0.0;
}

// Replicas that may share a wave of the grouped kernel: every constant the
// kernel reads from the group's first replica must agree.
static bool group_class_less(const VFRep& a, const VFRep& b) {
    const double ka[] = {(double)a.n_rays, (double)a.boundary, (double)a.fov_px0, (double)a.fov_px1,
                         (double)a.phi_ok, (double)a.limit_movement, a.width, a.height, a.pad,
                         a.gam, a.alp1, a.bet1, a.max_vel, a.max_th, a.step};
    const double kb[] = {(double)b.n_rays, (double)b.boundary, (double)b.fov_px0, (double)b.fov_px1,
                         (double)b.phi_ok, (double)b.limit_movement, b.width, b.height, b.pad,
                         b.gam, b.alp1, b.bet1, b.max_vel, b.max_th, b.step};
    for (size_t i = 0; i < sizeof(ka) / sizeof(ka[0]); ++i) {
        if (ka[i] < kb[i]) return true;
        if (ka[i] > kb[i]) return false;
    }
    return false;
}

// Groups of G replicas (same class, similar N, largest first) and the nibble
// tables of R for G replicas per wave: lane li of a group owns rays
// [li * RPL, (li + 1) * RPL); entry [q][v][li] of the first table is the sum of
// (trapz weight * cos PHI, * sin PHI) over the rays li*RPL + 4q + i whose bit
// i is set in v, the second table the same sum of (cos PHI, sin PHI).
static int build_group_plan(abm_vf_batch* bt, const std::vector<VFRep>& host, int G, int R,
                            const HostTables& t) {
    GroupPlan& gp = bt->gplan[G == 1 ? 0 : (G == 2 ? 1 : 2)];
    const int n = (int)host.size();
    std::vector<int> idx(n);
    for (int b = 0; b < n; ++b) idx[b] = b;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) {
        if (group_class_less(host[a], host[c])) return true;
        if (group_class_less(host[c], host[a])) return false;
        return host[a].n_agents > host[c].n_agents;
    });
    std::vector<std::vector<int>> groups;
    for (int i = 0; i < n;) {
        std::vector<int> g{idx[i]};
        int j = i + 1;
        while (j < n && (int)g.size() < G && !group_class_less(host[idx[i]], host[idx[j]]) &&
               !group_class_less(host[idx[j]], host[idx[i]]))
            g.push_back(idx[j++]);
        while ((int)g.size() < G) g.push_back(-1);
        groups.push_back(g);
        i = j;
    }
    std::stable_sort(groups.begin(), groups.end(), [&](const std::vector<int>& a, const std::vector<int>& c) {
        return host[a[0]].n_agents > host[c[0]].n_agents;
    });
    std::vector<int32_t> flat;
    for (auto& g : groups) flat.insert(flat.end(), g.begin(), g.end());
    const int L = 64 / G, rpl = group_rpl(G, R), nq = rpl / 4;
    const int tab = 4 * rpl * L;
    std::vector<float2> nib(2 * (size_t)tab);
    for (int li = 0; li < L; ++li)
        for (int q = 0; q < nq; ++q)
            for (int v = 0; v < 16; ++v) {
                double bx = 0, by = 0, sx = 0, sy = 0;
                for (int i = 0; i < 4; ++i) {
                    const int r = li * rpl + 4 * q + i;
                    if (!((v >> i) & 1) || r >= R) continue;
                    const double wl = r > 0 ? t.d[r - 1] : 0.0;
                    const double wr = r + 1 < R ? t.d[r] : 0.0;
                    const double w = 0.5 * (wl + wr);
                    bx += w * t.c[r]; by += w * t.s[r];
                    sx += t.c[r]; sy += t.s[r];
                }
                const size_t e = (size_t)(q * 16 + v) * L + li;
                nib[e] = make_float2((float)bx, (float)by);
                nib[tab + e] = make_float2((float)sx, (float)sy);
            }
    gp.n_groups = (int32_t)groups.size();
    gp.tab_entries = tab;
    int rc = check_hip(hipMalloc(&gp.groups, sizeof(int32_t) * flat.size()), "hipMalloc(groups)");
    if (!rc) rc = check_hip(hipMalloc(&gp.nib, sizeof(float2) * nib.size()), "hipMalloc(nibble tables)");
    if (!rc) rc = check_hip(hipMemcpy(gp.groups, flat.data(), sizeof(int32_t) * flat.size(),
                                      hipMemcpyHostToDevice), "upload groups");
    if (!rc) rc = check_hip(hipMemcpy(gp.nib, nib.data(), sizeof(float2) * nib.size(),
                                      hipMemcpyHostToDevice), "upload nibble tables");
    return rc;
}

static bool is_finite(double x) { return std::isfinite(x); }

}  // namespace abm

using namespace abm;

extern "C" {

const char* abm_last_error(void) { return g_err; }

const char* abm_build_info(void) {
    static char info[128];
    snprintf(info, sizeof(info), "abm_amd 0.1 (gfx950; HIP %d.%d.%d)", HIP_VERSION_MAJOR,
             HIP_VERSION_MINOR, HIP_VERSION_PATCH);
    return info;
}

int abm_vf_batch_create(const abm_vf_replica* reps, int32_t n, abm_vf_batch** out) {
    if (!out) return set_error(ABM_ERR_INVALID, "abm_vf_batch_create: out is NULL");
    *out = nullptr;
    if (!reps || n <= 0) return set_error(ABM_ERR_INVALID, "abm_vf_batch_create: no replicas");
    std::vector<VFRep> host(n);
    std::map<int, int> table_off;
    int ray_total = 0;
    int64_t words_total = 0;
    int64_t agents_total = 0;
    int max_n = 0, max_r = 0;
    for (int b = 0; b < n; ++b) {
        const abm_vf_replica& r = reps[b];
        if (r.n_agents < 1 || r.n_agents > (1 << 20))
            return set_error(ABM_ERR_INVALID, "replica %d: n_agents=%d out of range", b, r.n_agents);
        if (r.n_rays < 2 || r.n_rays > (1 << 16))
            return set_error(ABM_ERR_INVALID, "replica %d: n_rays=%d out of range [2, 65536]", b,
                             r.n_rays);
        if (r.boundary != ABM_BOUNDARY_WALLS && r.boundary != ABM_BOUNDARY_INFINITE)
            return set_error(ABM_ERR_INVALID, "replica %d: unknown boundary %d", b, r.boundary);
        const double vals[] = {r.fov0, r.fov1, r.width, r.height, r.window_pad, r.gam, r.v0,
                               r.alp0, r.alp1, r.bet0, r.bet1, r.max_vel, r.max_th};
        for (double v : vals)
            if (!is_finite(v)) return set_error(ABM_ERR_INVALID, "replica %d: non-finite parameter", b);
        if (!(r.width > 0) || !(r.height > 0))
            return set_error(ABM_ERR_INVALID, "replica %d: arena size must be positive", b);
        VFRep& d = host[b];
        std::memset(&d, 0, sizeof(d));
        d.agent_offset = (int32_t)agents_total;
        d.n_agents = r.n_agents;
        d.n_rays = r.n_rays;
        d.words = (r.n_rays + 31) / 32;
        if (!table_off.count(r.n_rays)) { table_off[r.n_rays] = ray_total; ray_total += r.n_rays; }
        d.table_offset = table_off[r.n_rays];
        d.boundary = r.boundary;
        d.step = linspace_step(r.n_rays);
        d.cull_c = (1.0 / std::tan(kTwoPi / r.n_rays)) * (1.0 + 2e-4);
        d.fov_px0 = linspace_argmin(r.fov0, r.n_rays, d.step);
        d.fov_px1 = linspace_argmin(r.fov1, r.n_rays, d.step);
        d.phi_ok = arange_len(r.n_rays) == r.n_rays ? 1 : 0;
        d.limit_movement = r.limit_movement ? 1 : 0;
        d.field_offset = words_total;
        d.width = r.width; d.height = r.height; d.pad = r.window_pad;
        d.gam = r.gam; d.v0 = r.v0; d.alp0 = r.alp0; d.alp1 = r.alp1;
        d.bet0 = r.bet0; d.bet1 = r.bet1; d.max_vel = r.max_vel; d.max_th = r.max_th;
        agents_total += r.n_agents;
        words_total += (int64_t)r.n_agents * d.words;
        max_n = std::max(max_n, r.n_agents);
        max_r = std::max(max_r, r.n_rays);
        if (agents_total > INT32_MAX)
            return set_error(ABM_ERR_INVALID, "abm_vf_batch_create: more than 2^31 agents");
    }
    abm_vf_batch* bt = new abm_vf_batch();
    bt->n_replicas = n;
    bt->n_agents_total = (int32_t)agents_total;
    bt->field_words = words_total;
    bt->max_agents = max_n;
    bt->min_agents = max_n;
    for (int b = 0; b < n; ++b) bt->min_agents = std::min(bt->min_agents, reps[b].n_agents);
    bt->max_rays = max_r;
    bt->ray_total = ray_total;
    bt->n_tables = (int)table_off.size();
    int rc = check_hip(hipGetDevice(&bt->device), "hipGetDevice");
    if (!rc) rc = check_hip(hipDeviceGetAttribute(&bt->dev_max_lds,
                                                  hipDeviceAttributeMaxSharedMemoryPerBlock,
                                                  bt->device), "hipDeviceGetAttribute");
    if (!rc) rc = check_hip(hipDeviceGetAttribute(&bt->n_cu, hipDeviceAttributeMultiprocessorCount,
                                                  bt->device), "hipDeviceGetAttribute");
    // tables
    std::vector<double> tc(ray_total), ts(ray_total), td(ray_total);
    std::vector<RayTab32> t32(ray_total);
    for (auto& kv : table_off) {
        HostTables t;
        make_tables(kv.first, &t);
        std::copy(t.c.begin(), t.c.end(), tc.begin() + kv.second);
        std::copy(t.s.begin(), t.s.end(), ts.begin() + kv.second);
        std::copy(t.d.begin(), t.d.end(), td.begin() + kv.second);
        std::copy(t.t32.begin(), t.t32.end(), t32.begin() + kv.second);
    }
    // replica order for the fused kernel: most pair-ray work first
    std::vector<int32_t> order(n);
    for (int b = 0; b < n; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        const double wa = (double)host[a].n_agents * host[a].n_agents * (64.0 + host[a].n_rays / 8.0);
        const double wc = (double)host[c].n_agents * host[c].n_agents * (64.0 + host[c].n_rays / 8.0);
        return wa > wc;
    });
    // the large size class first (still by descending work)
    std::stable_partition(order.begin(), order.end(),
                          [&](int b) { return host[b].n_agents > kBigClass; });
    for (int b = 0; b < n; ++b) {
        if (host[b].n_agents > kBigClass) {
            ++bt->n_big_class;
            bt->big_sum_agents += host[b].n_agents;
            bt->big_max_agents = std::max(bt->big_max_agents, host[b].n_agents);
            bt->big_max_rays = std::max(bt->big_max_rays, host[b].n_rays);
        } else {
            bt->small_max_agents = std::max(bt->small_max_agents, host[b].n_agents);
            bt->small_max_rays = std::max(bt->small_max_rays, host[b].n_rays);
        }
    }
    // the push kernel's LDS: its largest replica (all, or the large size class)
    for (int b = 0; b < n; ++b) {
        const size_t push = lds_bytes_push(host[b].n_agents, host[b].n_rays);
        bt->push_lds_all = std::max(bt->push_lds_all, push);
        if (host[b].n_agents > kBigClass) bt->push_lds_big = std::max(bt->push_lds_big, push);
    }
    if (!rc) rc = check_hip(hipMalloc(&bt->reps_dev, sizeof(VFRep) * n), "hipMalloc(replicas)");
    if (!rc) rc = check_hip(hipMalloc(&bt->tab_cos, sizeof(double) * 3 * ray_total), "hipMalloc(tables)");
    if (!rc) rc = check_hip(hipMalloc(&bt->tab32, sizeof(RayTab32) * ray_total), "hipMalloc(tab32)");
    if (!rc) rc = check_hip(hipMalloc(&bt->counters, sizeof(unsigned long long) * 4), "hipMalloc(counters)");
    if (!rc) rc = check_hip(hipMalloc(&bt->order_dev, sizeof(int32_t) * n), "hipMalloc(order)");
    if (!rc && max_n > 2048) {   // replicas too large for one CU's LDS take the wide step
        rc = check_hip(hipMalloc(&bt->wide_out, sizeof(float4) * (size_t)agents_total), "hipMalloc(wide)");
        if (!rc) rc = check_hip(hipMalloc(&bt->grid_sorted, sizeof(float4) * (size_t)max_n), "hipMalloc(grid)");
        if (!rc) rc = check_hip(hipMalloc(&bt->grid_start, sizeof(int32_t) * (kWideGridMaxCells + 1)),
                                "hipMalloc(grid cells)");
        if (!rc) rc = check_hip(hipMalloc(&bt->grid_param, 64), "hipMalloc(grid params)");
        if (!rc && max_r <= 2048) rc = check_hip(hipMalloc(&bt->chain_fields,
                                          sizeof(uint32_t) * (size_t)max_n * (size_t)((max_r + 31) / 32)),
                                "hipMalloc(chain fields)");
        if (!rc && max_r <= 2048)
            rc = check_hip(hipMalloc(&bt->chain_u, sizeof(float) * 2 * (size_t)max_n), "hipMalloc(chain u)");
    }
    if (!rc) {
        bt->tab_sin = bt->tab_cos + ray_total;
        bt->tab_dphi = bt->tab_cos + 2 * (size_t)ray_total;
        rc = check_hip(hipMemcpy(bt->reps_dev, host.data(), sizeof(VFRep) * n, hipMemcpyHostToDevice), "upload replicas");
    }
    if (!rc) rc = check_hip(hipMemcpy(bt->tab_cos, tc.data(), sizeof(double) * ray_total, hipMemcpyHostToDevice), "upload cos");
    if (!rc) rc = check_hip(hipMemcpy(bt->tab_sin, ts.data(), sizeof(double) * ray_total, hipMemcpyHostToDevice), "upload sin");
    if (!rc) rc = check_hip(hipMemcpy(bt->tab_dphi, td.data(), sizeof(double) * ray_total, hipMemcpyHostToDevice), "upload dphi");
    if (!rc) rc = check_hip(hipMemcpy(bt->tab32, t32.data(), sizeof(RayTab32) * ray_total, hipMemcpyHostToDevice), "upload tab32");
    if (!rc) rc = check_hip(hipMemcpy(bt->order_dev, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice), "upload order");
    if (!rc) rc = check_hip(hipMemset(bt->counters, 0, sizeof(unsigned long long) * 4), "hipMemset(counters)");
    // the grouped kernel: one R in the batch, replicas small enough for a wave's LDS share
    if (!rc && table_off.size() == 1 && max_r <= kGroupMaxRays && max_n <= kGroupMaxAgents) {
        HostTables t;
        make_tables(max_r, &t);
        for (int G : {1, 2, 4})
            if (!rc) rc = build_group_plan(bt, host, G, max_r, t);
        if (!rc) bt->group_ok = 1;
    }
    if (rc) { abm_vf_batch_destroy(bt); return rc; }
    bt->reps_host = new VFRep[n];
    std::copy(host.begin(), host.end(), bt->reps_host);
    bt->agent_offset_host = new int32_t[n + 1];
    bt->field_offset_host = new int64_t[n + 1];
    for (int b = 0; b < n; ++b) {
        bt->agent_offset_host[b] = host[b].agent_offset;
        bt->field_offset_host[b] = host[b].field_offset;
    }
    bt->agent_offset_host[n] = (int32_t)agents_total;
    bt->field_offset_host[n] = words_total;
    *out = bt;
    return ABM_OK;
}

int abm_vf_batch_destroy(abm_vf_batch* b) {
    if (!b) return ABM_OK;
    if (b->reps_dev) (void)hipFree(b->reps_dev);
    if (b->tab_cos) (void)hipFree(b->tab_cos);
    if (b->tab32) (void)hipFree(b->tab32);
    if (b->counters) (void)hipFree(b->counters);
    if (b->order_dev) (void)hipFree(b->order_dev);
    if (b->wide_out) (void)hipFree(b->wide_out);
    if (b->grid_sorted) (void)hipFree(b->grid_sorted);
    if (b->grid_start) (void)hipFree(b->grid_start);
    if (b->chain_fields) (void)hipFree(b->chain_fields);
    if (b->chain_u) (void)hipFree(b->chain_u);
    if (b->grid_param) (void)hipFree(b->grid_param);
    for (GroupPlan& gp : b->gplan) {
        if (gp.groups) (void)hipFree(gp.groups);
        if (gp.nib) (void)hipFree(gp.nib);
    }
    delete[] b->reps_host;
    delete[] b->agent_offset_host;
    delete[] b->field_offset_host;
    delete b;
    return ABM_OK;
}

int abm_vf_batch_get_info(const abm_vf_batch* b, abm_batch_info* out) {
    if (!b || !out) return set_error(ABM_ERR_INVALID, "abm_vf_batch_get_info: NULL argument");
    out->n_replicas = b->n_replicas;
    out->n_agents_total = b->n_agents_total;
    out->field_words = b->field_words;
    out->max_agents = b->max_agents;
    out->max_rays = b->max_rays;
    return ABM_OK;
}

int abm_vf_batch_get_offsets(const abm_vf_batch* b, int32_t* agent_offset, int64_t* field_offset) {
    if (!b) return set_error(ABM_ERR_INVALID, "abm_vf_batch_get_offsets: NULL batch");
    if (agent_offset) std::copy(b->agent_offset_host, b->agent_offset_host + b->n_replicas + 1, agent_offset);
    if (field_offset) std::copy(b->field_offset_host, b->field_offset_host + b->n_replicas + 1, field_offset);
    return ABM_OK;
}

static int check_state_ptrs(const void* x, const void* y, const void* th, const void* v,
                            const void* r, const char* who) {
    if (!x || !y || !th || !v || !r)
        return set_error(ABM_ERR_INVALID, "%s: state arrays x, y, theta, v, radius are required", who);
    return ABM_OK;
}

int abm_vf_step_f64(abm_vf_batch* b, const abm_vf_state_f64* st, uint32_t* bits, double* diag,
                    uint32_t flags, void* stream) {
    if (!b || !st) return set_error(ABM_ERR_INVALID, "abm_vf_step_f64: NULL batch or state");
    if (int rc = check_state_ptrs(st->x, st->y, st->theta, st->v, st->radius, "abm_vf_step_f64")) return rc;
    if ((flags & ABM_WANT_DIAG) && !diag)
        return set_error(ABM_ERR_INVALID, "abm_vf_step_f64: ABM_WANT_DIAG without a diag buffer");
    return launch_vf_step_f64(b, st, bits, diag, flags, (hipStream_t)stream);
}

int abm_vf_step_f32(abm_vf_batch* b, const abm_vf_state_f32* st, uint32_t* bits, float* diag,
                    uint32_t flags, void* stream) {
    if (!b || !st) return set_error(ABM_ERR_INVALID, "abm_vf_step_f32: NULL batch or state");
    if (int rc = check_state_ptrs(st->x, st->y, st->theta, st->v, st->radius, "abm_vf_step_f32")) return rc;
    if ((flags & ABM_WANT_DIAG) && !diag)
        return set_error(ABM_ERR_INVALID, "abm_vf_step_f32: ABM_WANT_DIAG without a diag buffer");
    return launch_vf_step_f32(b, st, bits, diag, flags, (hipStream_t)stream);
}

int abm_vf_f32_fallback_count(abm_vf_batch* b, int64_t* pairs_total, int64_t* pairs_rechecked,
                              int64_t* pairs_disagree, void* stream) {
    if (!b) return set_error(ABM_ERR_INVALID, "abm_vf_f32_fallback_count: NULL batch");
    unsigned long long h[4] = {0, 0, 0, 0};
    hipStream_t s = (hipStream_t)stream;
    if (int rc = check_hip(hipMemcpyAsync(h, b->counters, sizeof(h), hipMemcpyDeviceToHost, s), "copy counters")) return rc;
    if (int rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize")) return rc;
    if (pairs_total) *pairs_total = (int64_t)h[0];
    if (pairs_rechecked) *pairs_rechecked = (int64_t)h[1];
    if (pairs_disagree) *pairs_disagree = (int64_t)h[2];
    return ABM_OK;
}

int abm_vf_f64_near_ties(abm_vf_batch* b, int64_t* near_ties, void* stream) {
    if (!b || !near_ties) return set_error(ABM_ERR_INVALID, "abm_vf_f64_near_ties: NULL argument");
    unsigned long long h[4] = {0, 0, 0, 0};
    hipStream_t s = (hipStream_t)stream;
    if (int rc = check_hip(hipMemcpyAsync(h, b->counters, sizeof(h), hipMemcpyDeviceToHost, s), "copy counters")) return rc;
    if (int rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize")) return rc;
    *near_ties = (int64_t)h[3];
    return ABM_OK;
}

int abm_vf_last_kernels(const abm_vf_batch* b, uint32_t* mask) {
    if (!b || !mask) return set_error(ABM_ERR_INVALID, "abm_vf_last_kernels: NULL argument");
    *mask = b->last_kernels;
    return ABM_OK;
}

int abm_vf_set_launch(abm_vf_batch* b, int32_t waves) {
    if (!b) return set_error(ABM_ERR_INVALID, "abm_vf_set_launch: NULL batch");
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4 && waves != 16 && waves != 32 && waves != 64)
        return set_error(ABM_ERR_INVALID, "abm_vf_set_launch: waves_per_group must be 0, 1, 2, 4, 16, 32 or 64");
    if (waves == 64) {   // the chain step for every replica: its buffers, if the batch has none yet
        if (b->max_rays > 2048)
            return set_error(ABM_ERR_UNSUPPORTED, "abm_vf_set_launch: the chain step takes n_rays up to 2048");
        const size_t n = (size_t)b->max_agents, words = (size_t)((b->max_rays + 31) / 32);
        int rc = ABM_OK;
        if (!b->grid_sorted) rc = check_hip(hipMalloc(&b->grid_sorted, sizeof(float4) * n), "hipMalloc(grid)");
        if (!rc && !b->grid_start)
            rc = check_hip(hipMalloc(&b->grid_start, sizeof(int32_t) * (kWideGridMaxCells + 1)), "hipMalloc(grid cells)");
        if (!rc && !b->grid_param) rc = check_hip(hipMalloc(&b->grid_param, 64), "hipMalloc(grid params)");
        if (!rc && !b->chain_fields)
            rc = check_hip(hipMalloc(&b->chain_fields, sizeof(uint32_t) * n * words), "hipMalloc(chain fields)");
        if (!rc && !b->chain_u) rc = check_hip(hipMalloc(&b->chain_u, sizeof(float) * 2 * n), "hipMalloc(chain u)");
        if (rc) return rc;
    }
    b->waves_per_group = waves;
    b->group_g = waves > 0 ? -1 : 0;
    return ABM_OK;
}

int abm_vf_set_grouping(abm_vf_batch* b, int32_t replicas_per_wave) {
    if (!b) return set_error(ABM_ERR_INVALID, "abm_vf_set_grouping: NULL batch");
    const int g = replicas_per_wave;
    if (g != -1 && g != 0 && g != 1 && g != 2 && g != 4)
        return set_error(ABM_ERR_INVALID, "abm_vf_set_grouping: replicas_per_wave must be -1, 0, 1, 2 or 4");
    if (g > 0 && !b->group_ok)
        return set_error(ABM_ERR_UNSUPPORTED, "abm_vf_set_grouping: the grouped kernel needs one n_rays "
                         "(<= %d) in the batch and at most %d agents per replica", kGroupMaxRays,
                         kGroupMaxAgents);
    b->group_g = g;
    if (g != 0) b->waves_per_group = 0;
    return ABM_OK;
}

int abm_vf_autotune(abm_vf_batch* b, const abm_vf_state_f32* st, int32_t trial_steps,
                    int32_t* chosen, float* chosen_ms, void* stream) {
    if (!b || !st) return set_error(ABM_ERR_INVALID, "abm_vf_autotune: NULL batch or state");
    if (trial_steps < 1) trial_steps = 1;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (int rc = check_hip(hipEventCreate(&e0), "hipEventCreate")) return rc;
    if (int rc = check_hip(hipEventCreate(&e1), "hipEventCreate")) { (void)hipEventDestroy(e0); return rc; }
    // candidates: waves per replica of the per-replica kernel (> 0), or minus the
    // replicas per wave of the grouped kernel
    // (16: the 16-wave kernel for every replica; 32: the push kernel for
    // every replica; 0: the size classes, the per-replica kernel up to
    // kBigClass agents and the push or 16-wave kernel above)
    std::vector<int> cand{1, 2, 4};
    if (b->max_agents <= 2048) cand.push_back(16);
    if (push_fits(b, true)) cand.push_back(32);
    if (b->n_big_class > 0) cand.push_back(0);
    if (b->group_ok) { cand.push_back(-1); cand.push_back(-2); cand.push_back(-4); }
    int best = 1;
    float best_ms = 1e30f;
    int rc = ABM_OK;
    for (size_t c = 0; c < cand.size() && !rc; ++c) {
        if (cand[c] >= 0) { b->waves_per_group = cand[c]; b->group_g = -1; }
        else { b->waves_per_group = 0; b->group_g = -cand[c]; }
        rc = launch_vf_step_f32(b, st, nullptr, nullptr, 0, s);  // warm-up
        if (!rc) rc = check_hip(hipEventRecord(e0, s), "hipEventRecord");
        for (int i = 0; i < trial_steps && !rc; ++i) rc = launch_vf_step_f32(b, st, nullptr, nullptr, 0, s);
        if (!rc) rc = check_hip(hipEventRecord(e1, s), "hipEventRecord");
        if (!rc) rc = check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        if (!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (!rc && ms < best_ms) { best_ms = ms; best = cand[c]; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best >= 0) { b->waves_per_group = best; b->group_g = -1; }
    else { b->waves_per_group = 0; b->group_g = -best; }
    if (chosen) *chosen = best;
    if (chosen_ms) *chosen_ms = best_ms / (float)trial_steps;
    return ABM_OK;
}

int abm_vf_projection_rows(int32_t R, double fov0, double fov1, int32_t boundary, double W,
                           double H, double fx, double fy, double ftheta, double fradius,
                           const double* ox, const double* oy, const double* osize, int32_t n_obj,
                           uint32_t* rows, int32_t* rec, void* stream) {
    if (R < 2 || R > (1 << 16)) return set_error(ABM_ERR_INVALID, "abm_vf_projection_rows: n_rays=%d", R);
    if (n_obj < 0) return set_error(ABM_ERR_INVALID, "abm_vf_projection_rows: n_obj < 0");
    if (n_obj > 0 && (!ox || !oy || !osize || !rows))
        return set_error(ABM_ERR_INVALID, "abm_vf_projection_rows: NULL object or output array");
    if (boundary == ABM_BOUNDARY_INFINITE && !(W > 0 && H > 0))
        return set_error(ABM_ERR_INVALID, "abm_vf_projection_rows: torus needs arena_width/height");
    return launch_vf_projection_rows(R, fov0, fov1, boundary, W, H, fx, fy, ftheta, fradius, ox, oy,
                                     osize, n_obj, rows, rec, (hipStream_t)stream);
}

int abm_vf_state_variables(int32_t R, int32_t n, const uint32_t* bits, const double* vel,
                           const double* alp0, const double* bet0, const double* v0, double gam,
                           double alp1, double bet1, double* out, int8_t* dphi, void* stream) {
    if (R < 2 || R > (1 << 16)) return set_error(ABM_ERR_INVALID, "abm_vf_state_variables: n_rays=%d", R);
    if (n < 0) return set_error(ABM_ERR_INVALID, "abm_vf_state_variables: n < 0");
    if (n > 0 && (!bits || !vel || !alp0 || !bet0 || !v0 || !out))
        return set_error(ABM_ERR_INVALID, "abm_vf_state_variables: NULL array");
    ProbeTables t{};
    if (int rc = probe_tables(R, &t)) return rc;
    const int phi_ok = arange_len(R) == R ? 1 : 0;
    return launch_vf_state_variables(R, n, bits, vel, alp0, bet0, v0, gam, alp1, bet1, t.c, t.s, t.d,
                                     phi_ok, out, dphi, (hipStream_t)stream);
}

int abm_vf_state_variables_noalloc(int32_t R, int32_t n, const uint32_t* bits, const double* vel,
                                   const double* alp0, const double* bet0, const double* v0,
                                   double gam, double alp1, double bet1, double* out, int8_t* dphi,
                                   void* stream) {
    if (R < 2 || R > kNoAllocMaxRays)
        return set_error(ABM_ERR_INVALID, "abm_vf_state_variables_noalloc: n_rays=%d outside [2, %d] "
                         "(its static table space holds %d rays)", R, kNoAllocMaxRays, kNoAllocMaxRays);
    if (n < 0) return set_error(ABM_ERR_INVALID, "abm_vf_state_variables_noalloc: n < 0");
    if (n > 0 && (!bits || !vel || !alp0 || !bet0 || !v0 || !out))
        return set_error(ABM_ERR_INVALID, "abm_vf_state_variables_noalloc: NULL array");
    if (n == 0) return ABM_OK;
    double* t = nullptr;
    if (int rc = noalloc_scratch(&t)) return rc;
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(ray_tables_f64_kernel, dim3((R + 255) / 256), dim3(256), 0, s, R,
                       arange_delta(R), t, t + R, t + 2 * R);
    if (int rc = check_hip(hipGetLastError(), "ray_tables_f64_kernel launch")) return rc;
    const int phi_ok = arange_len(R) == R ? 1 : 0;
    return launch_vf_state_variables(R, n, bits, vel, alp0, bet0, v0, gam, alp1, bet1, t, t + R,
                                     t + 2 * R, phi_ok, out, dphi, s);
}

}  // extern "C"
