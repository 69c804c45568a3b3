// vf_exact.hpp -- float64 arithmetic of the reference's perception and motion
// step, operation for operation, for host and device.
//
// Every function here evaluates the same IEEE float64 operations, in the same
// order, as the numpy code it restates, with contraction disabled (an a*b+c
// written in Python is two roundings; a fused multiply-add appears only where
// numpy itself uses one: the 2-element BLAS dot behind np.dot and
// np.linalg.norm).  Transcendentals (acos, atan, cos, sin) come from the
// device libm and may differ from the host's numpy in the last ulp; the
// integer raster decisions they feed are unaffected except within an ulp of a
// tie.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define ABM_HD __host__ __device__ inline

namespace abm {

constexpr double kPi = 3.141592653589793;            // np.pi
constexpr double kTwoPi = 2.0 * kPi;                  // 2 * np.pi (exact doubling)
constexpr double kHalfPi = kPi / 2.0;                 // np.pi / 2
constexpr double kThreeHalfPi = (3.0 * kPi) / 2.0;    // 3 * np.pi / 2

ABM_HD double fma_exact(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fma(a, b, c);
#else
    return fma(a, b, c);
#endif
}

// np.linalg.norm of a 2-vector: sqrt of the BLAS dot x.x = fma(b, b, a*a).
ABM_HD double norm2(double a, double b) {
#pragma clang fp contract(off)
    return sqrt(fma_exact(b, b, a * a));
}

// np.dot of two 2-vectors (BLAS ddot): fma(u1, w1, u0*w0).
ABM_HD double dot2(double u0, double u1, double w0, double w1) {
#pragma clang fp contract(off)
    return fma_exact(u1, w1, u0 * w0);
}

// linspace(-pi, pi, R)[k] = fl(fl(k * step) + (-pi)), last element pi
// (numpy linspace: y = arange * step + start; y[-1] = stop).
ABM_HD double linspace_step(int R) {
#pragma clang fp contract(off)
    return (kPi - (-kPi)) / (double)(R - 1);
}
ABM_HD double linspace_at(int k, int R, double step) {
#pragma clang fp contract(off)
    if (k == R - 1) return kPi;
    double y = (double)k * step;
    return y + (-kPi);
}

// First index k minimising |linspace_k - value| (supcalc.find_nearest,
// abm/agent/supcalc.py:8-11).  linspace is strictly increasing, so
// |fl(phis_k - value)| is unimodal in k and the minimiser lies within one
// index of (value + pi) / step; a +-2 window scanned upward with a strict
// comparison returns numpy's first index on ties.  A NaN value gives index 0
// (np.argmin returns the first NaN).
ABM_HD int linspace_argmin(double value, int R, double step) {
#pragma clang fp contract(off)
    if (!(value == value)) return 0;
    double xf = (value + kPi) / step;
    int k0;
    if (xf < 0.0) k0 = 0;
    else if (xf > (double)(R - 1)) k0 = R - 1;
    else k0 = (int)floor(xf + 0.5);
    int lo = k0 - 2 < 0 ? 0 : k0 - 2;
    int hi = k0 + 2 > R - 1 ? R - 1 : k0 + 2;
    int best = lo;
    double bv = fabs(linspace_at(lo, R, step) - value);
    for (int k = lo + 1; k <= hi; ++k) {
        double d = fabs(linspace_at(k, R, step) - value);
        if (d < bv) { bv = d; best = k; }
    }
    return best;
}

// numpy float64 remainder x % (2*pi) for x in [-pi, pi] or NaN (npy_divmod):
// fmod is exact here; a negative result gets 2*pi added; a zero becomes +0.
ABM_HD double remainder_2pi(double x) {
#pragma clang fp contract(off)
    if (x < 0.0) return x + kTwoPi;
    if (x == 0.0) return 0.0;
    return x;
}

// Focal-agent quantities shared by all its pairs (vf_supcalc.py:42-49;
// agent.py:484-495 computes the same v1 with the same roundings).
struct Focal {
    double px, py;     // position (top-left)
    double cx, cy;     // centre = position + radius
    double u1x, u1y;   // unit(v1)
    double radius;
};

ABM_HD Focal make_focal(double px, double py, double theta, double radius) {
#pragma clang fp contract(off)
    Focal f;
    f.px = px; f.py = py; f.radius = radius;
    f.cx = px + radius;
    f.cy = py + radius;
    double ex = px + (1.0 + cos(theta)) * radius;
    double ey = py + (1.0 - sin(theta)) * radius;
    double v1x = ex - f.cx;
    double v1y = ey - f.cy;
    double n1 = norm2(v1x, v1y);
    f.u1x = v1x / n1;
    f.u1y = v1y / n1;
    return f;
}

// Raster decision of one object for the VSWRM field (vf_supcalc.py:55-119).
struct PairVF {
    int skipped;       // coincident positions (:57)
    int t;             // phi_target
    int half;          // floor(proj_size / 2)
    int start, end;    // before the wrap adjustment
    int drawn;         // FOV pixel test (:119)
    double closed_angle, distance, half_f;  // diagnostics
};

// closed angle of angle_between(v1, v2) (supcalc.py:19-34) mapped by the
// VF rule 0 <= a' <= pi -> -a' (vf_supcalc.py:142-158).
ABM_HD double closed_angle_vf(const Focal& f, double v2x, double v2y, double n2) {
#pragma clang fp contract(off)
    double u2x = v2x / n2;
    double u2y = v2y / n2;
    double d = dot2(f.u1x, f.u1y, u2x, u2y);
    if (d < -1.0) d = -1.0;
    else if (d > 1.0) d = 1.0;
    double a = acos(d);
    double cross = f.u1x * u2y - f.u1y * u2x;
    if (cross < 0.0) a = -a;
    a = remainder_2pi(a);
    if (0.0 <= a && a <= kPi) return -a;
    return kTwoPi - a;
}

ABM_HD PairVF pair_vf_exact(const Focal& f, double ox, double oy, double osize,
                            int torus, double W, double H, int R, double step,
                            int fov_px0, int fov_px1) {
#pragma clang fp contract(off)
    PairVF p;
    p.skipped = 0; p.t = 0; p.half = 0; p.start = 0; p.end = 0; p.drawn = 0;
    p.closed_angle = 0.0; p.distance = 0.0; p.half_f = 0.0;
    if (ox == f.px && oy == f.py) { p.skipped = 1; return p; }
    double ocx = ox + osize;
    double ocy = oy + osize;
    double v2x = ocx - f.cx;
    double v2y = ocy - f.cy;
    if (torus) {
        if (fabs(v2x) > W / 2.0) {
            if (f.cx < ocx) ocx -= W;
            else if (f.cx > ocx) ocx += W;
        }
        if (fabs(v2y) > H / 2.0) {
            if (f.cy < ocy) ocy -= H;
            else if (f.cy > ocy) ocy += H;
        }
        v2x = ocx - f.cx;
        v2y = ocy - f.cy;
    }
    double n2 = norm2(v2x, v2y);
    double ca = closed_angle_vf(f, v2x, v2y, n2);
    double vis = 2.0 * atan(osize / (1.0 * n2));
    int t = linspace_argmin(ca, R, step);
    double ps = (vis / kTwoPi) * (double)R;
    double hf = floor(ps / 2.0);
    int h = (int)hf;
    p.t = t; p.half = h; p.half_f = ps / 2.0;
    p.start = t - h; p.end = t + h;
    p.drawn = ((fov_px0 < p.start && p.start < fov_px1) ||
               (fov_px0 < p.end && p.end < fov_px1)) ? 1 : 0;
    p.closed_angle = ca; p.distance = n2;
    return p;
}

// Wrapped pieces of a drawn VF interval (vf_supcalc.py:122-129): up to two
// half-open ranges [a0,b0) and [a1,b1) of ray indices (empty when a >= b).
// For valid poses the half-width is below R/4, so at most one wrap applies
// and every piece lies in [0, R]; the final clamp only keeps garbage input
// (NaN poses, negative radii) from indexing outside the field.
ABM_HD int clamp_ray(int v, int R) { return v < 0 ? 0 : (v > R ? R : v); }
ABM_HD void vf_pieces(int start, int end, int R, int* a0, int* b0, int* a1, int* b1) {
    *a1 = 0; *b1 = 0;
    if (start < 0) { *a1 = R + start; *b1 = R; start = 0; }
    if (end >= R) { *a1 = 0; *b1 = end - (R - 1); end = R; }
    *a0 = clamp_ray(start, R); *b0 = clamp_ray(end, R);
    *a1 = clamp_ray(*a1, R); *b1 = clamp_ray(*b1, R);
}

// --------------------------------------------------------- pose update
struct Pose { double x, y, theta, v; };

// prove_orientation: one-shot wrap (agent.py:605-610).
ABM_HD double prove_orientation(double o) {
#pragma clang fp contract(off)
    if (o < 0.0) o = kTwoPi + o;
    if (o > kTwoPi) o = o - kTwoPi;
    return o;
}

ABM_HD double np_sign(double x) {
    if (x > 0.0) return 1.0;
    if (x < 0.0) return -1.0;
    if (x == 0.0) return 0.0;
    return x;  // NaN
}

// VFAgent.update_agent_position (vf_agent.py:289-326), dt = 1.
ABM_HD Pose vf_move(Pose p, double dpsi, double dv, int limit, double max_th, double max_vel) {
#pragma clang fp contract(off)
    double theta = dpsi;
    if (limit) {  // prove_turning
        double s = np_sign(theta);
        if (s == 0.0) s = 1.0;
        if (fabs(theta) > max_th) theta = max_th * s;
    }
    p.theta = prove_orientation(p.theta + theta);
    p.v = p.v + dv;
    if (limit) {  // VFAgent.prove_velocity
        double s = np_sign(p.v);
        if (s == 0.0) s = 1.0;
        if (fabs(p.v) > max_vel) p.v = max_vel * s;
    }
    double c = cos(p.theta);
    double sn = sin(p.theta);
    double nx = p.x + p.v * c;
    double ny = p.y - p.v * sn;
    p.x = nx; p.y = ny;
    return p;
}

// VFAgent.teleport_infinite_arena (vf_agent.py:188-204).
ABM_HD Pose teleport(Pose p, double r, double pad, double W, double H) {
#pragma clang fp contract(off)
    double x = p.x + r, y = p.y + r;
    double bx0 = pad, bx1 = pad + W, by0 = pad, by1 = pad + H;
    if (x < bx0) p.x = bx1 - r;
    else if (x > bx1) p.x = bx0 + r;
    if (y < by0) p.y = by1 - r;
    else if (y > by1) p.y = by0 + r;
    return p;
}

// Agent.reflect_from_walls (agent.py:347-395).
ABM_HD Pose reflect_walls(Pose p, double r, double pad, double W, double H) {
#pragma clang fp contract(off)
    double x = p.x + r, y = p.y + r;
    double bx0 = pad, bx1 = pad + W, by0 = pad, by1 = pad + H;
    if (x < bx0) {
        p.x = bx0 - r;
        if (kHalfPi <= p.theta && p.theta < kPi) p.theta -= kHalfPi;
        else if (kPi <= p.theta && p.theta <= kThreeHalfPi) p.theta += kHalfPi;
        p.theta = prove_orientation(p.theta);
    }
    if (x > bx1) {
        p.x = bx1 - r - 1.0;
        if (kThreeHalfPi <= p.theta && p.theta < kTwoPi) p.theta -= kHalfPi;
        else if (0.0 <= p.theta && p.theta <= kHalfPi) p.theta += kHalfPi;
        p.theta = prove_orientation(p.theta);
    }
    if (y < by0) {
        p.y = by0 - r;
        if (kHalfPi <= p.theta && p.theta <= kPi) p.theta += kHalfPi;
        else if (0.0 <= p.theta && p.theta < kHalfPi) p.theta -= kHalfPi;
        p.theta = prove_orientation(p.theta);
    }
    if (y > by1) {
        p.y = by1 - r - 1.0;
        if (kThreeHalfPi <= p.theta && p.theta <= kTwoPi) p.theta += kHalfPi;
        else if (kPi <= p.theta && p.theta < kThreeHalfPi) p.theta -= kHalfPi;
        p.theta = prove_orientation(p.theta);
    }
    return p;
}

// numpy arange(-pi, pi, 2*pi/R): its length and element k
// (PHI, vf_agent.py:44): buffer[1] = start + step, delta = buffer[1] - start,
// buffer[k] = start + k*delta; length = ceil((stop - start) / step).
ABM_HD long arange_len(int R) {
#pragma clang fp contract(off)
    double step = kTwoPi / (double)R;
    return (long)ceil((kPi - (-kPi)) / step);
}
ABM_HD double arange_delta(int R) {
#pragma clang fp contract(off)
    double step = kTwoPi / (double)R;
    double b1 = -kPi + step;
    return b1 - (-kPi);
}
ABM_HD double arange_at(int k, double delta) {
#pragma clang fp contract(off)
    if (k == 0) return -kPi;
    return -kPi + (double)k * delta;
}

}  // namespace abm
