// hetero.h -- voxel media of code/MCCovHetro.m: cell lookup, the cell-by-cell
// optical depth of evalphaseattHetro.m, Woodcock (delta) tracking of
// smpWoodcock.m and the first-vertex samplers (voxel pick of smpicdf.m /
// invCDFvec.m, plane-patch projection of expSmpXHetro.m / cubeProj.m).
//
// FP64 with contraction off, in the reference's expression order, so the
// walks and factors agree with the CPU restatement to the last ulps of the
// transcendental functions.
//
// Random numbers (Philox keys as the homogeneous walk; a uniform one-cell
// medium walks exactly the homogeneous paths):
//   vertex 0: g0 = (u_z, u_x), g1 = (u_y, d0), g2 = (d1, d2), g3 = (d3, a);
//             a picks the cell (smpFlg 1), (g0.y, g1.x) place the plane patch
//             point and (g0.x, g3.y) are the first Woodcock trial (smpFlg 2);
//             trials k >= 1 use group 16 + k, lrad retries 1024 + 2k, 1025 + 2k.
//   vertex n: g0 = (xi, first trial distance), g1 = (u_roulette, d0),
//             g2 = (d1, d2), g3 = (d3, first trial acceptance); trials k >= 1
//             use group 16 + k.
#pragma once
#include "device_common.h"

namespace scmc {

// The medium, by value in every kernel that needs it.
struct HetMedium {
    int nb[3];              // cells per axis (1 beyond dim)
    double bin[3];          // cell size (box_bin)
    const double *sigt;     // prod(nb) values, x fastest (device)
    const double *albedo;   // prod(nb) values or null (scalar P.albedo)
    double smax;            // max(sigt(:))
    const int32_t *icdf;    // invCDFvec table, n_icdf entries (device)
    int n_icdf;
    double V;               // sum(sigt(:)) * prod(box_bin)
    int has_lrad;
    double lrad;            // patch size (default: box diagonal)
    double pushc;           // max(abs([box_max; box_min])) * dim
    double u[3][2];         // orthonormal basis of the complement of meanl
    int cap;                // bound on cells crossed by one ray
    double inv_bin[3];      // 1 / bin (the ray march locates cells by multiplication)
};

#pragma clang fp contract(off)

// findBin.m:1-21 (0-based; a point on a lower face is clamped to the first cell)
template <int DIM>
__device__ __forceinline__ int find_bin(const DevProblem &P, const HetMedium &M, const double *x, double *bmn,
                                        double *bmx)
{
    int idx = 0, stride = 1;
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
        double q = ceil((x[j] - P.bmin[j]) / M.bin[j]);
        int ii = (int)q;
        ii = ii < 1 ? 1 : (ii > M.nb[j] ? M.nb[j] : ii);
        idx += (ii - 1) * stride;
        stride *= M.nb[j];
        if (bmn) {
            bmn[j] = P.bmin[j] + M.bin[j] * (ii - 1);
            bmx[j] = bmn[j] + M.bin[j];
        }
    }
    return idx;
}

template <int DIM>
__device__ __forceinline__ double med_albedo(const DevProblem &P, const HetMedium &M, const double *x)
{
    if (!M.albedo) return P.albedo;
    return M.albedo[find_bin<DIM>(P, M, x, nullptr, nullptr)];
}

template <int DIM>
__device__ __forceinline__ bool outside_box(const double *x, const DevProblem &P)
{
#pragma unroll
    for (int j = 0; j < DIM; ++j)
        if (x[j] > P.bmax[j] || x[j] < P.bmin[j]) return true;
    return false;
}

// evalphaseattHetro.m:29-48: optical depth from x along -v, cell by cell,
// stepping 1e-4 cell sizes past each face (cubeDist.m:13-16), capped at nd.
// The same iteration with the cell held as integer indices: each axis keeps
// the distance from x to its exit face of the current cell, so a step's
// distance is (nearest exit) - (distance walked so far) and the step past
// the face ends at t = nearest exit + ep; every axis whose exit face lies
// within t moves to its next cell, as relocating the stepped point
// (findBin.m) would, and leaving the box on any axis ends the march.  The
// reference recomputes the point and its cell every step; here the face
// distances come from x alone, so no rounding accumulates (the two agree to
// ~1e-11 on random rays, and a ray nearly parallel to a face no longer
// loses its position in the rounding of the stepped point), and a step is
// a few FP64 operations instead of a cell search per axis.
template <int DIM>
__device__ __forceinline__ double march_depth(const DevProblem &P, const HetMedium &M, const double *x,
                                              const double *v, double nd)
{
    const double thr = 1.0000000000000004e-07;   // cubeDist.m:5, 0.1^7 as computed by pow
    double inv_ad[3] = {0, 0, 0}, tf[3] = {INFINITY, INFINITY, INFINITY};
    bool neg[3] = {false, false, false};
    int c[3] = {0, 0, 0};
    double ep = 0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
        const double mv = -v[j];
        inv_ad[j] = 1.0 / fmax(fabs(mv), thr);
        neg[j] = mv <= 0;
        // findBin.m (0-based, a point on a lower face clamped to the first cell)
        int ii = (int)ceil((x[j] - P.bmin[j]) * M.inv_bin[j]);
        ii = ii < 1 ? 1 : (ii > M.nb[j] ? M.nb[j] : ii);
        c[j] = ii - 1;
        const double bmn = P.bmin[j] + M.bin[j] * c[j], bmx = bmn + M.bin[j];
        tf[j] = neg[j] ? (x[j] - bmn) * inv_ad[j] : (-x[j] + bmx) * inv_ad[j];   // cubeDist.m
        ep = fmax(ep, fabs(bmx - bmn));
    }
    ep = 0.00010000000000000002 * ep;   // 0.1^4 as computed by pow
    double t = 0, bdv = 0, wbdv = 0;
    for (int it = 0; it < M.cap; ++it) {
        int b = 0, stride = 1;
        double tmin = INFINITY;
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            b += c[j] * stride;
            stride *= M.nb[j];
            tmin = tf[j] < tmin ? tf[j] : tmin;
        }
        const double dist = tmin - t;
        bdv = bdv + dist;
        const double s = M.sigt[b];
        wbdv = wbdv + dist * s;
        if (bdv >= nd) {
            const double md = bdv - nd;
            wbdv = wbdv - md * s;
            break;
        }
        t = tmin + ep;
        bool out = false;
#pragma unroll
        for (int j = 0; j < DIM; ++j)
            if (tf[j] <= t) {
                c[j] += neg[j] ? -1 : 1;
                if (c[j] < 0 || c[j] >= M.nb[j]) {
                    out = true;
                } else {
                    const double bmn = P.bmin[j] + M.bin[j] * c[j];
                    tf[j] = neg[j] ? (x[j] - bmn) * inv_ad[j] : (-x[j] + (bmn + M.bin[j])) * inv_ad[j];
                }
            }
        if (out) break;
    }
    return wbdv;
}

// evalphaseattHetro.m:1-51 for one direction / position (sign * dcol).
template <int DIM>
__device__ __forceinline__ double2 phase_att_hetro(const DevProblem &P, const HetMedium &M, const double *x,
                                                   const double *dcol, double sign, int is_ff)
{
    double v[3] = {0, 0, 0};
#pragma unroll
    for (int j = 0; j < DIM; ++j) v[j] = sign * dcol[j];
    double pv, rv, nd;
    if (is_ff) {
        pv = dot<DIM>(x, v);
        rv = 1.0;
        nd = 1e12;
    } else {
        double dd[3] = {0, 0, 0}, s = 0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            dd[j] = x[j] - v[j];
            s += dd[j] * dd[j];
        }
        nd = sqrt(s);
#pragma unroll
        for (int j = 0; j < DIM; ++j) v[j] = dd[j] / nd;
        pv = nd;
        rv = (DIM == 3) ? 1.0 / (nd + 0.01 * P.lambda) : 1.0 / sqrt(nd + 0.01 * P.lambda);
    }
    const double wbdv = march_depth<DIM>(P, M, x, v, nd);
    const double mag = exp(-0.5 * wbdv) * rv;
    const double ph = P.k * pv;
    double sn, cs;
    sincos(ph, &sn, &cs);
    return make_double2(mag * cs, mag * sn);
}

// smpWoodcock.m:1-15 from x along w; trial 0 uses (ud0, ua0), trial k >= 1
// the pair of group 16 + k at this vertex.
// E(x, d): evalphaseatt (homogeneous) or evalphaseattHetro (voxel medium).
template <int DIM, bool HET>
__device__ __forceinline__ double2 nee_e(const DevProblem &P, const HetMedium &M, const double *x, const double *dcol,
                                         double sign, int is_ff)
{
    if (HET) return phase_att_hetro<DIM>(P, M, x, dcol, sign, is_ff);
    return phase_att64<DIM>(P, x, dcol, sign, is_ff);
}

template <int DIM>
__device__ __forceinline__ double woodcock(const DevProblem &P, const HetMedium &M, const double *x, const double *w,
                                           uint64_t path, uint32_t vertex, double ud0, double ua0)
{
    double d = 0;
    for (int it = 0; it < 1000000; ++it) {
        double ud = ud0, ua = ua0;
        if (it > 0) {
            const double2 g = uniform_pair(P.seed, path, vertex, 16u + (uint32_t)it);
            ud = g.x;
            ua = g.y;
        }
        d = d - log(1 - ud) / M.smax;
        double nx[3] = {0, 0, 0};
#pragma unroll
        for (int j = 0; j < DIM; ++j) nx[j] = x[j] + d * w[j];
        if (outside_box<DIM>(nx, P)) break;
        const int b = find_bin<DIM>(P, M, nx, nullptr, nullptr);
        if (ua <= M.sigt[b] / M.smax) break;
    }
    return d;
}

// cubeProj.m:1-20: move x along v to just inside the box (false when the
// entry lies behind x, where the reference returns a scalar)
template <int DIM>
__device__ __forceinline__ bool cube_proj(const DevProblem &P, const double *x, const double *v, double *xd)
{
    const double thr = 1.0000000000000004e-07;
    double d = -INFINITY;
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
        const double ad = fmax(fabs(v[j]), thr);
        const double bd = (v[j] <= 0) ? (x[j] - P.bmax[j]) / ad : (-x[j] + P.bmin[j]) / ad;
        d = fmax(d, bd);
    }
    if (!(d >= 0)) return false;
    double ep = 0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) ep = fmax(ep, fabs(P.bmax[j] - P.bmin[j]));
    ep = 0.00010000000000000002 * ep;
#pragma unroll
    for (int j = 0; j < DIM; ++j) xd[j] = x[j] + (d + ep) * v[j];
    return true;
}

// MCCovHetro.m:101-153: the first vertex and direction.  Returns false when
// the sampled point is outside the box (the iteration is spent).
template <int DIM>
__device__ __forceinline__ bool hetero_path_start(const DevProblem &P, const HetMedium &M, uint64_t path, double *x,
                                                  double &px, double *w, double2 &w0p)
{
    const double2 g0 = uniform_pair(P.seed, path, 0, 0);
    const double2 g1 = uniform_pair(P.seed, path, 0, 1);
    const double2 g2 = uniform_pair(P.seed, path, 0, 2);
    const double2 g3 = uniform_pair(P.seed, path, 0, 3);
    if (P.smp_flg == 1) {
        // MCCovHetro.m:103-114: cell from the quantised icdf, uniform inside it
        int i = (int)ceil(g3.y * M.n_icdf);
        i = i < 1 ? 1 : i;
        int b = M.icdf[i - 1] - 1;
        b = b < 0 ? 0 : b;   // a 0 entry (one cell) is not a valid MATLAB index
        int rem = b;
        double bmn[3] = {0, 0, 0};
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            const int ii = rem % M.nb[j];
            rem /= M.nb[j];
            bmn[j] = P.bmin[j] + M.bin[j] * ii;
        }
        const double r[3] = {g0.y, DIM == 3 ? g1.x : g0.x, g0.x};
#pragma unroll
        for (int j = 0; j < DIM; ++j) x[j] = r[j] * M.bin[j] + bmn[j];
        px = 1 / M.V;
        if (M.has_lrad) {
            // MCCovHetro.m:45-48, 107-113 (one light)
            double ls[3];
#pragma unroll
            for (int j = 0; j < DIM; ++j) ls[j] = (P.bmin[j] + P.bmax[j]) / 2;
            ls[DIM - 1] = (P.l[DIM - 1] > 0) ? P.bmin[DIM - 1] : P.bmax[DIM - 1];
            for (int it = 0;; ++it) {
                const double lx = dot<DIM>(P.l, x);
                double r2 = 0;
#pragma unroll
                for (int j = 0; j < DIM; ++j) {
                    const double q = x[j] - lx * P.l[j] - ls[j];
                    r2 += q * q;
                }
                if (!(sqrt(r2) > M.lrad) || it > 100000) break;
                const double2 a = uniform_pair(P.seed, path, 0, 1024u + 2u * (uint32_t)it);
                const double2 c = uniform_pair(P.seed, path, 0, 1025u + 2u * (uint32_t)it);
                const double rr[3] = {a.x, a.y, c.x};
#pragma unroll
                for (int j = 0; j < DIM; ++j) x[j] = rr[j] * (P.bmax[j] - P.bmin[j]) + P.bmin[j];
                px = 1;
            }
        }
    } else {
        // expSmpXHetro.m:1-31
        const double r[2] = {(g0.y - 0.5) * M.lrad, (g1.x - 0.5) * M.lrad};
        double x0[3] = {0, 0, 0};
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            double s = 0;
#pragma unroll
            for (int q = 0; q < DIM - 1; ++q) s += M.u[j][q] * r[q];
            x0[j] = s - P.meanl[j] * M.pushc;
        }
        if (!cube_proj<DIM>(P, x0, P.meanl, x) || outside_box<DIM>(x, P)) return false;
        const double d = woodcock<DIM>(P, M, x, P.meanl, path, 0, g0.x, g3.y);
#pragma unroll
        for (int j = 0; j < DIM; ++j) x[j] = x[j] + d * P.meanl[j];
        if (outside_box<DIM>(x, P)) return false;
        // px = abs(evalphaseattHetro(x, lmean, 1, ...))^2 / lrad^(dim-1), the
        // modulus of the unit phase taken as exactly 1
        const double mag = exp(-0.5 * march_depth<DIM>(P, M, x, P.meanl, 1e12));
        px = mag * mag / pow(M.lrad, DIM - 1);
    }
    if (outside_box<DIM>(x, P)) return false;
    if (!P.has_amp0) {
        iso_dir<DIM>(g1.y, g2.x, g2.y, g3.x, w);
        w0p = make_double2(P.w0p_iso, 0.0);
    } else {
        sample_dir<DIM>(P.meanl, P.amp0, g1.y, g2.x, g2.y, g3.x, w);
        w0p = eval_amp<DIM>(dot<DIM>(P.meanl, w), P.amp0);
    }
    return true;
}

// MCCovHetro.m:228-243: Woodcock flight, escape, roulette with the albedo of
// the new cell, new direction.  Returns false when the path ends.
template <int DIM>
__device__ __forceinline__ bool hetero_path_step(const DevProblem &P, const HetMedium &M, uint64_t path, uint32_t n,
                                                 double *x, double *w, double *ow, double &weight, double ud0,
                                                 double ua0, const double *lcos = nullptr)
{
    const double d = woodcock<DIM>(P, M, x, w, path, n, ud0, ua0);
#pragma unroll
    for (int j = 0; j < DIM; ++j) x[j] = x[j] + d * w[j];
    if (outside_box<DIM>(x, P)) return false;
    const double2 b = uniform_pair(P.seed, path, n, 1);
    if (weight < kKillThr) {
        if (b.x > med_albedo<DIM>(P, M, x)) return false;
    } else {
        weight = weight * med_albedo<DIM>(P, M, x);
    }
    const double2 c = uniform_pair(P.seed, path, n, 2);
    // the fourth uniform of the draw is used by the isotropic phase only
    const double e3 = P.amp.type == 1 ? uniform_pair(P.seed, path, n, 3).x : 0.0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) ow[j] = w[j];
    sample_dir<DIM>(ow, P.amp, b.y, c.x, c.y, e3, w, lcos);
    return true;
}

#pragma clang fp contract(fast)

} // namespace scmc
