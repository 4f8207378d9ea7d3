// device_common.h -- device-side sampling and evaluation primitives for the
// speckle Monte Carlo core on gfx950.
//
// Each function restates a reference MATLAB primitive; the file:line is given
// beside it.  FP64 variants keep the reference's expression order (contraction
// off) so paths traced here agree with the FP64 CPU restatement to the last few
// ulps of the transcendental functions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scmc {

constexpr double kPi = 3.14159265358979323846;
constexpr double kKillThr = 0.2;   // MCcov.m:110

struct DevPhase {
    int type;            // 1 isotropic, 2 tabulated, 3 HG
    double g;
    double iso;          // 1/sqrt(2^(dim-1)*pi)
    int n_amp;
    const double2 *amp;  // evalAmp (re, im) on device
    int n_icdf;
    const double *cos_icdf;  // cos(sampleIcdf), formed on device once per problem
};

// Problem constants passed by value to every kernel.
struct DevProblem {
    int dim;
    double sigt, albedo, lambda, k;   // k = 2*pi/lambda
    double inv_lambda;
    double bmin[3], bmax[3];
    int nl, nv;
    const double *l;                  // dim x nl (device)
    const double *v;                  // dim x nv (device)
    int is_ff_l, is_ff_v, do_cbs, smp_flg;
    DevPhase amp;
    int has_amp0;
    DevPhase amp0;                    // type == amp.type
    double unmean[3], meanl[3];
    uint64_t seed;
    uint64_t path_offset;
    double w0p_iso;                   // 1/sqrt(2^(dim-1) pi)
    double ff_sign;                   // -1 far-field views, +1 near-field
    // Far-field views: no view leaves the box farther than this through the
    // faces of the last axis (W_last / min |v_last|, with a 1e-5 margin), so
    // a point whose distance to the nearer face exceeds it on every other
    // axis leaves every view through the last axis (+inf: no such bound)
    float far_last_bound;
};

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
    uint64_t v = ((((uint64_t)a) << 32) | b) >> 11;
    return (double)v * 0x1.0p-53;
}

// Two uniforms for (seed, path, vertex, group); same slot map as the oracle.
__device__ __forceinline__ double2 uniform_pair(uint64_t seed, uint64_t path, uint32_t vertex, uint32_t group)
{
    uint4 o = philox4x32_10(make_uint4(vertex, group, (uint32_t)path, (uint32_t)(path >> 32)),
                            make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
    return make_double2(u53(o.x, o.y), u53(o.z, o.w));
}

#pragma clang fp contract(off)

template <int DIM>
__device__ __forceinline__ double dot(const double *a, const double *b)
{
    if (DIM == 3) return fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0]));
    return fma(a[1], b[1], a[0] * b[0]);
}

// cubeDist.m:4-11
template <int DIM>
__device__ __forceinline__ double cube_dist(const double *x, const double *bmin, const double *bmax, const double *d)
{
    const double thr = 1.0000000000000004e-07;   // 0.1^7 as computed by pow
    double best = INFINITY;
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
        double ad = fmax(fabs(d[j]), thr);
        double bd = (d[j] <= 0) ? (x[j] - bmin[j]) / ad : (-x[j] + bmax[j]) / ad;
        best = bd < best ? bd : best;
    }
    return best;
}

// evaluateHG.m:25-39 (cosine input)
template <int DIM>
__device__ __forceinline__ double evaluate_hg(double c, double g)
{
    if (DIM == 3) return 1.0 / 4.0 / kPi * ((1 - g * g) / pow(1 + (g * g - 2 * (g * c)), 1.5));
    return 1.0 / 2.0 / kPi * ((1 - g * g) / (1 + (g * g - 2 * (g * c))));
}

// evalampfunc_general.m:5-18 and evalampfunc_amplitude.m:5-9
template <int DIM>
__device__ __forceinline__ double2 eval_amp(double c, const DevPhase &ph)
{
    if (ph.type == 1) return make_double2(ph.iso, 0.0);
    if (ph.type == 2) {
        double cc = c > 1 ? 1 : (c < -1 ? -1 : c);
        double ang = acos(cc);
        int n = ph.n_amp;
        double scale = (DIM == 2) ? (n - 1) / (2 * kPi) : (n - 1) / kPi;
        long idx = lround(ang * scale);
        idx = idx < 0 ? 0 : (idx > n - 1 ? n - 1 : idx);
        return ph.amp[idx];
    }
    return make_double2(sqrt(evaluate_hg<DIM>(c, ph.g)), 0.0);
}

// Entrance directions at a path's first vertex x1 (MCcov.m:99-105,128-135;
// MCfield.m:121-136): in the far field the direction columns themselves; in
// the near field rv_i = (v_i - x1)/|v_i - x1| and rl_j = (x1 - l_j)/|x1 - l_j|.
// Every vertex of the path reuses them (the reference computes them once per
// path).  r must hold 3 doubles.
template <int DIM>
__device__ __forceinline__ void view_dir(const DevProblem &P, const double *x1, int i, double *r)
{
    const double *c = P.v + (size_t)i * DIM;
    r[2] = 0.0;
    if (P.is_ff_v) {
#pragma unroll
        for (int j = 0; j < DIM; ++j) r[j] = c[j];
        return;
    }
    double s = 0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
        r[j] = c[j] - x1[j];
        s += r[j] * r[j];
    }
    const double nrm = sqrt(s);
#pragma unroll
    for (int j = 0; j < DIM; ++j) r[j] = r[j] / nrm;
}

template <int DIM>
__device__ __forceinline__ void light_dir(const DevProblem &P, const double *x1, int i, double *r)
{
    const double *c = P.l + (size_t)i * DIM;
    r[2] = 0.0;
    if (P.is_ff_l) {
#pragma unroll
        for (int j = 0; j < DIM; ++j) r[j] = c[j];
        return;
    }
    double s = 0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
        r[j] = x1[j] - c[j];
        s += r[j] * r[j];
    }
    const double nrm = sqrt(s);
#pragma unroll
    for (int j = 0; j < DIM; ++j) r[j] = r[j] / nrm;
}

__device__ __forceinline__ double2 cdiv(double2 a, double2 b)
{
    // a / b with the textbook formula; b is an amplitude (never tiny).
    double den = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// sampleHG.m:3-17
template <int DIM>
__device__ __forceinline__ double sample_hg(double g, double u)
{
    if (fabs(g) < 0.001) return 2 * u - 1;
    if (DIM == 3) {
        double t = (1 - g * g) / (1 - g + 2 * g * u);
        return (1 / (2 * g)) * ((1 + g * g) - t * t);
    }
    double th = 2 * atan((1 - g) / (1 + g) * tan(kPi / 2 * (1 - 2 * u)));
    return cos(th);
}

// rotateBytheta.m:4-27
template <int DIM>
__device__ __forceinline__ void rotate_by_theta(const double *ow, double costheta, double u, double *w)
{
    double s2 = 1 - costheta * costheta;
    double sintheta = s2 > 0 ? sqrt(s2) : 0.0;
    if (DIM == 2) {
        sintheta = ((u > 0.5) * 2 - 1) * sintheta;
        w[0] = costheta * ow[0] + sintheta * ow[1];
        w[1] = -sintheta * ow[0] + costheta * ow[1];
        return;
    }
    double phi = kPi * 2 * u;
    double sinphi, cosphi;
    sincos(phi, &sinphi, &cosphi);
    if (fabs(ow[2]) > 0.999999999) {
        w[0] = sintheta * cosphi;
        w[1] = sintheta * sinphi;
        w[2] = costheta * ow[2] / fabs(ow[2]);
    } else {
        double temp = sqrt(1 - ow[2] * ow[2]);
        w[0] = sintheta * (ow[0] * ow[2] * cosphi - ow[1] * sinphi) / temp + ow[0] * costheta;
        w[1] = sintheta * (ow[1] * ow[2] * cosphi + ow[0] * sinphi) / temp + ow[1] * costheta;
        w[2] = -sintheta * cosphi * temp + ow[2] * costheta;
    }
}

__device__ __forceinline__ void box_muller(double ua, double ub, double &n0, double &n1)
{
    double r = sqrt(-2.0 * log(1.0 - ua));
    double t = 2.0 * kPi * ub, st, ct;
    sincos(t, &st, &ct);
    n0 = r * ct;
    n1 = r * st;
}

// randn(dim,1)/norm: smpampfunc_general.m:7-8, MCcov.m:148
template <int DIM>
__device__ __forceinline__ void iso_dir(double d0, double d1, double d2, double d3, double *w)
{
    double n0, n1, n2, n3;
    box_muller(d0, d1, n0, n1);
    if (DIM == 3) {
        box_muller(d2, d3, n2, n3);
        double nrm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        w[0] = n0 / nrm;
        w[1] = n1 / nrm;
        w[2] = n2 / nrm;
    } else {
        double nrm = sqrt(n0 * n0 + n1 * n1);
        w[0] = n0 / nrm;
        w[1] = n1 / nrm;
    }
}

// Path regeneration.  A walking kernel with resident lanes hands its idle
// lanes new path indices (at the wave's refills, refill_due below), instead of
// letting a wave run as long as its longest path (for the slab workloads the longest of 64 walks is
// ~6x the mean, so most lanes would idle).  A wave takes indices in blocks of
// kPathBlock from a global counter (one atomic per block) and gives them to
// its idle lanes in lane order.  A path's walk depends only on its index, so
// the results do not depend on which lane walks it.
constexpr int kPathBlock = 256;
struct WavePaths {
    int64_t next = 0, end = 0;   // the wave's current block (wave-uniform)
    bool exhausted = false;
};
// Called by every lane of the wave; `want` lanes receive the next index in
// [0, np), or -1 once the counter has passed np.  *ctr starts at 0.
__device__ __forceinline__ int64_t take_path(WavePaths &wp, bool want, int64_t np, unsigned long long *ctr)
{
    int64_t id = -1;
    const unsigned lane = __lane_id();
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    unsigned long long m = __ballot(want);
    while (m && !wp.exhausted) {
        if (wp.next >= wp.end) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(ctr, (unsigned long long)kPathBlock);
            base = __shfl(base, 0);
            if ((int64_t)base >= np) {
                wp.exhausted = true;
                break;
            }
            wp.next = (int64_t)base;
            wp.end = (int64_t)base + kPathBlock < np ? (int64_t)base + kPathBlock : np;
        }
        const int64_t avail = wp.end - wp.next;
        const int rank = __popcll(m & below);
        if (((m >> lane) & 1ull) && rank < avail) id = wp.next + rank;
        const int nwant = __popcll(m);
        wp.next += nwant < avail ? nwant : avail;
        m = __ballot(want && id < 0);
    }
    return id;
}

// A refill runs the path-start code (four Philox draws, the first vertex and
// direction) for the whole wave, masked to the lanes that start a path, so a
// wave that refilled whenever one lane was idle would pay it nearly every
// step (a slab walk of ~18 vertices ends in ~3.6 of 64 lanes per step).  A
// wave therefore refills once kRefillIdle lanes are idle, or when none is
// active, and stops asking once the counter is exhausted.
constexpr int kRefillIdle = 16;
__device__ __forceinline__ bool refill_due(const WavePaths &wp, bool active)
{
    const unsigned long long act = __ballot(active);
    const int idle = 64 - __popcll(act);
    return !wp.exhausted && idle > 0 && (idle >= kRefillIdle || act == 0ull);
}

// The walking kernels stage the tabulated direction table (the cosines of
// sampleIcdf) in LDS when it has at most kIcdfLds entries.  Every thread of
// the workgroup calls stage_icdf (a tabulated problem ends it with a
// barrier); it returns the LDS table, or null to read the table from HBM.
constexpr int kIcdfLds = 1024;
__device__ __forceinline__ const double *stage_icdf(const DevPhase &ph, double *lcos)
{
    if (ph.type != 2 || ph.n_icdf > kIcdfLds) return nullptr;
    for (int j = threadIdx.x; j < ph.n_icdf; j += blockDim.x) lcos[j] = ph.cos_icdf[j];
    __syncthreads();
    return lcos;
}

// smpampfunc_general.m:5-20, smpampfunc.m:3-7.  lcos: the LDS copy of
// ph.cos_icdf (stage_icdf), or null.
template <int DIM>
__device__ __forceinline__ void sample_dir(const double *ow, const DevPhase &ph, double d0, double d1, double d2,
                                           double d3, double *w, const double *lcos = nullptr)
{
    if (ph.type == 1) {
        iso_dir<DIM>(d0, d1, d2, d3, w);
        return;
    }
    double cosa;
    if (ph.type == 2) {
        int n = ph.n_icdf;
        double stp = 1.0 / (n - 1);
        long i = lround(d0 / stp);
        i = i > n - 1 ? n - 1 : i;
        cosa = lcos ? lcos[i] : ph.cos_icdf[i];
    } else {
        cosa = sample_hg<DIM>(ph.g, d0);
    }
    rotate_by_theta<DIM>(ow, cosa, d1, w);
}

// expSmpX.m:1-12 with (u_z, u_x, u_y)
template <int DIM>
__device__ __forceinline__ void exp_smp_x(const DevProblem &P, double uz, double ux, double uy, double *x, double &px)
{
    const int e = DIM - 1;
    double lz = fabs(P.unmean[e]);
    double sigs = P.sigt / lz;
    double dmax = P.bmax[e] - P.bmin[e];
    double rmax = 1 - exp(-dmax * sigs);
    double d = -log(-uz * rmax + 1) / sigs;
    x[0] = ux * (P.bmax[0] - P.bmin[0]) + P.bmin[0];
    if (DIM == 3) x[1] = uy * (P.bmax[1] - P.bmin[1]) + P.bmin[1];
    x[e] = (d + P.bmin[e]) * (P.unmean[e] > 0) + (-d + P.bmax[e]) * (P.unmean[e] <= 0);
    px = exp(-sigs * d) / rmax * sigs * dmax;
}

template <int DIM>
__device__ __forceinline__ bool outside(const double *x, const DevProblem &P)
{
    bool o = false;
#pragma unroll
    for (int j = 0; j < DIM; ++j) o = o || (x[j] > P.bmax[j]) || (x[j] < P.bmin[j]);
    return o;
}

// First vertex and first direction of path `path`: MCcov.m:114-153.
template <int DIM>
__device__ __forceinline__ void path_start(const DevProblem &P, uint64_t path, double *x, double &px, double *w,
                                           double2 &w0p)
{
    double2 g0 = uniform_pair(P.seed, path, 0, 0);
    double2 g1 = uniform_pair(P.seed, path, 0, 1);
    double2 g2 = uniform_pair(P.seed, path, 0, 2);
    double2 g3 = uniform_pair(P.seed, path, 0, 3);
    if (P.smp_flg == 1) {
        x[0] = g0.y * (P.bmax[0] - P.bmin[0]) + P.bmin[0];
        if (DIM == 3) x[1] = g1.x * (P.bmax[1] - P.bmin[1]) + P.bmin[1];
        x[DIM - 1] = g0.x * (P.bmax[DIM - 1] - P.bmin[DIM - 1]) + P.bmin[DIM - 1];
        px = 1;
    } else {
        exp_smp_x<DIM>(P, g0.x, g0.y, g1.x, x, px);
    }
    if (!P.has_amp0) {
        iso_dir<DIM>(g1.y, g2.x, g2.y, g3.x, w);
        w0p = make_double2(P.w0p_iso, 0.0);
    } else {
        sample_dir<DIM>(P.meanl, P.amp0, g1.y, g2.x, g2.y, g3.x, w);
        w0p = eval_amp<DIM>(dot<DIM>(P.meanl, w), P.amp0);
    }
}

// One transition after the vertex n (MCcov.m:266-288).  Returns false when
// the path ends.  x, w, weight, ow are updated in place; lcos as in sample_dir.
template <int DIM>
__device__ __forceinline__ bool path_step(const DevProblem &P, uint64_t path, uint32_t n, double *x, double *w,
                                          double *ow, double &weight, double flight_u, const double *lcos = nullptr)
{
    double d = -log(-flight_u + 1) / P.sigt;
#pragma unroll
    for (int j = 0; j < DIM; ++j) x[j] = x[j] + d * w[j];
    if (outside<DIM>(x, P)) return false;
    double2 b = uniform_pair(P.seed, path, n, 1);
    if (weight < kKillThr) {
        if (b.x > P.albedo) return false;
    } else {
        weight = weight * P.albedo;
    }
    double2 c = uniform_pair(P.seed, path, n, 2);
    // the fourth uniform of the draw is used by the isotropic phase only
    const double e3 = P.amp.type == 1 ? uniform_pair(P.seed, path, n, 3).x : 0.0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) ow[j] = w[j];
    sample_dir<DIM>(ow, P.amp, b.y, c.x, c.y, e3, w, lcos);
    return true;
}

// evalphaseatt.m:1-20 for one direction / position (FP64).  d is the raw
// column; the direction used is sign*d.
template <int DIM>
__device__ __forceinline__ double2 phase_att64(const DevProblem &P, const double *x, const double *dcol, double sign,
                                               int is_ff)
{
    double dv[3] = {0, 0, 0};
#pragma unroll
    for (int j = 0; j < DIM; ++j) dv[j] = sign * dcol[j];
    double pv, rv, bdv;
    if (is_ff) {
        double nd[3] = {-dv[0], -dv[1], -dv[2]};
        pv = dot<DIM>(x, dv);
        rv = 1.0;
        bdv = cube_dist<DIM>(x, P.bmin, P.bmax, nd);
    } else {
        double dd[3] = {0, 0, 0}, s = 0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            dd[j] = x[j] - dv[j];
            s += dd[j] * dd[j];
        }
        double ndist = sqrt(s);
        double ndir[3] = {0, 0, 0};
#pragma unroll
        for (int j = 0; j < DIM; ++j) ndir[j] = -(dd[j] / ndist);
        pv = ndist;
        rv = (DIM == 3) ? 1.0 / (ndist + 0.01 * P.lambda) : 1.0 / sqrt(ndist + 0.01 * P.lambda);
        double cd = cube_dist<DIM>(x, P.bmin, P.bmax, ndir);
        bdv = cd < ndist ? cd : ndist;
    }
    double mag = exp(-P.sigt / 2 * bdv) * rv;
    double ph = P.k * pv;
    double s, c;
    sincos(ph, &s, &c);
    return make_double2(mag * c, mag * s);
}

#pragma clang fp contract(fast)

// Mixed-precision evalphaseatt (far field): the phase k*(x.d) is formed and
// range-reduced in FP64 (fraction of a cycle), sin/cos and the attenuation
// exponential run in FP32.
template <int DIM>
__device__ __forceinline__ float2 phase_att_mixed(const DevProblem &P, const double *x, const double *dcol,
                                                  double sign, int is_ff)
{
    if (!is_ff) {
        double2 e = phase_att64<DIM>(P, x, dcol, sign, is_ff);
        return make_float2((float)e.x, (float)e.y);
    }
    double dv[3] = {0, 0, 0};
#pragma unroll
    for (int j = 0; j < DIM; ++j) dv[j] = sign * dcol[j];
    double pv = dot<DIM>(x, dv);
    double cyc = pv * P.inv_lambda;             // phase in cycles
    double fr = cyc - rint(cyc);                // [-0.5, 0.5]
    float f = (float)fr;
    float s = __builtin_amdgcn_sinf(f);         // sin(2*pi*f): input in revolutions
    float c = __builtin_amdgcn_cosf(f);
    float best = INFINITY;
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
        double nd = -dv[j];
        float ad = (float)fmax(fabs(nd), 1.0000000000000004e-07);
        float num = (float)((nd <= 0) ? (x[j] - P.bmin[j]) : (-x[j] + P.bmax[j]));
        float bd = num / ad;
        best = bd < best ? bd : best;
    }
    float mag = __expf(-(float)(P.sigt * 0.5) * best);
    return make_float2(mag * c, mag * s);
}

// Isotropic or HG amplitude (evalampfunc_general.m:5-18, evaluateHG.m:25-39)
// in FP32 for the mixed variant: iso, or hg_k (1 + g^2 - 2 g c)^mu with
// mu = -3/4 (3-D) or -1/2 (2-D).  Types other than 1 and 2 are HG, as in
// eval_amp; tabulated amplitudes (type 2) are not handled here.
struct AmpF32 {
    bool hg;
    float iso, hg_k, g2p1, g2;
};
template <int DIM>
__device__ __forceinline__ AmpF32 amp_f32_setup(const DevPhase &ph)
{
    AmpF32 a;
    a.hg = ph.type != 1;
    a.iso = (float)ph.iso;
    a.hg_k = (float)sqrt((1 - ph.g * ph.g) / ((DIM == 3) ? 4 * kPi : 2 * kPi));
    a.g2p1 = (float)(1 + ph.g * ph.g);
    a.g2 = (float)(2 * ph.g);
    return a;
}
template <int DIM>
__device__ __forceinline__ float amp_f32(float c, const AmpF32 &a)
{
    if (!a.hg) return a.iso;
    const float t = fmaxf(a.g2p1 - a.g2 * c, 1e-30f);
    return a.hg_k * __builtin_amdgcn_exp2f((DIM == 3 ? -0.75f : -0.5f) * __builtin_amdgcn_logf(t));
}

} // namespace scmc
