// quad.hip -- Fourier-accelerated field (code/MCfieldFourierQuad.m and its
// illumination-wave form code/MCfieldFourierQuadOnWave.m).
//
// Per stored vertex (pxL = sqrt(W/px) e^{2 pi i xi}, /sqrt 2 for pL > 1 with
// CBS, MCfieldFourierQuad.m:181-182) the reference adds, for every light,
//     pxL * exp(-i k (tz - z_s)) * exp(i k l.x_L - sigma/2 att(x_L))
// into the xy cell of the "scatter" position x_s in the z-slab of z_s
// (MCfieldFourierQuad.m:263-273): pass 1 x_s = x1, x_L = x_n; pass 2 (CBS)
// swaps them (:224-232).  Each slab grid is then Fourier transformed
// (fftshift/flip/crop of :276-280) and summed with the slab phase
// exp(zexp0 tz) (:281-282).
//
// GPU mapping: one lane per vertex scatters with hardware atomics into a
// per-(slab, light) grid resident in HBM (FP64 or FP32); the 2-D DFT of every
// slab is two batched complex GEMMs H = W X W^T with W[o][ix] =
// exp(+2 pi i (f_o q_ix mod n)/n) (the shift/flip/crop folded into the
// indices, twiddles reduced exactly in integers, odd or even n); a combine
// kernel applies the slab phases and sums over slabs.
#include "kernels.h"
#include "quad.h"

#include <algorithm>
#include <type_traits>

namespace scmc {

template <typename T> struct cx;
template <> struct cx<double> { typedef double2 t; };
template <> struct cx<float> { typedef float2 t; };

// (cos, sin) of 2 pi cyc: FP64 reduction to a fraction of a cycle, then the
// hardware FP32 sin/cos (whose input is in revolutions)
__device__ __forceinline__ float2 cis_cycles(double cyc)
{
    const float fr = (float)(cyc - rint(cyc));
    return make_float2(__builtin_amdgcn_cosf(fr), __builtin_amdgcn_sinf(fr));
}

// twl = sum_l Wl(l) e^{2 pi i (vsign l.x / lambda)} over the wave's lights
// (MCfieldFourierQuadOnWave.m:281-302) in FP32 after the FP64 reduction of
// each phase; wl is the LDS copy of the lights (stage_wave) or the global one.
__device__ __forceinline__ float2 wave_sum(const WaveLight *wl, int nl, const double *x)
{
    float twr = 0.f, twi = 0.f;
#pragma unroll 4
    for (int l = 0; l < nl; ++l) {
        const WaveLight w = wl[l];
        const float2 c = cis_cycles(fma(w.kz, x[2], fma(w.ky, x[1], w.kx * x[0])));
        twr = fmaf(w.wr, c.x, fmaf(-w.wi, c.y, twr));
        twi = fmaf(w.wr, c.y, fmaf(w.wi, c.x, twi));
    }
    return make_float2(twr, twi);
}

// Copies the wave's lights into LDS when there are at most kWaveLightsLds of
// them; every thread of the workgroup calls it (it ends with a barrier).
// Returns whether the LDS copy holds them.
__device__ __forceinline__ bool stage_wave(const DevProblem &P, const QuadGeom &g, WaveLight *lds)
{
    if (!g.wave || P.nl > kWaveLightsLds) return false;
    for (int l = threadIdx.x; l < P.nl; l += blockDim.x) lds[l] = g.wlk[l];
    __syncthreads();
    return true;
}

__device__ __forceinline__ void atomic_add_c(double2 *p, double re, double im)
{
    unsafeAtomicAdd(&p->x, re);
    unsafeAtomicAdd(&p->y, im);
}
__device__ __forceinline__ void atomic_add_c(float2 *p, double re, double im)
{
    unsafeAtomicAdd(&p->x, (float)re);
    unsafeAtomicAdd(&p->y, (float)im);
}

// First slab j with z < gz[j+1] (the while loop of MCfieldFourierQuad.m:249);
// -1 when z is at or above the last boundary.
__device__ __forceinline__ int slab_of(double z, const QuadGeom &g)
{
    if (!(z < g.gz[g.nz - 1])) return -1;
    int j = (int)floor((z - g.gz[0]) / g.z_bin);
    j = j < 0 ? 0 : (j > g.nz - 2 ? g.nz - 2 : j);
    while (j > 0 && z < g.gz[j]) --j;
    while (j < g.nz - 2 && !(z < g.gz[j + 1])) ++j;
    return j;
}

// One vertex's contribution in one pass (MCfieldFourierQuad.m:263-273, and
// the light-summed form of MCfieldFourierQuadOnWave.m:281-302): xs the
// scatter (view-side) position whose cell and slab receive it, xl the
// light-side position, amp = sqrt(W/px) [/sqrt 2 for later vertices with
// CBS], xi the vertex's random-phase uniform.  Pass 1 has xs = x1, xl = the
// vertex; pass 2 (CBS) swaps them (:224-232).
template <typename T>
__device__ __forceinline__ void quad_scatter_pass(const DevProblem &P, const QuadGeom &g, bool wave,
                                                  const WaveLight *wls, const double *xs, const double *xl, double amp,
                                                  double xi, typename cx<T>::t *grid)
{
    const int j = slab_of(xs[2], g);
    if (j < 0) return;
    long ix = lround((xs[0] - g.gx0) / g.box_stp), iy = lround((xs[1] - g.gy0) / g.box_stp);
    ix = ix < 0 ? 0 : (ix > g.nx - 1 ? g.nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy > g.nx - 1 ? g.nx - 1 : iy);
    const double tz = 0.5 * (g.gz[j] + g.gz[j + 1]);
    const double att = -P.sigt / 2 * ((xl[2] - P.bmin[2]) * g.l_diratt + (P.bmax[2] - xl[2]) * g.l_idiratt);
    if constexpr (std::is_same<T, float>::value) {
        // FP32 grid: every phase counted in cycles and reduced in FP64, then
        // hardware sin/cos/exp2 in FP32 (the FP64 grid keeps libm FP64)
        const float2 ex = cis_cycles(xi);   // e^{2 pi i xi}
        const float pre = (float)amp * ex.x, pim = (float)amp * ex.y;
        const float m = __builtin_amdgcn_exp2f((float)(att * 1.4426950408889634));
        if (wave) {
            const float2 tw = wave_sum(wls, P.nl, xl);
            const float twr = tw.x, twi = tw.y;
            const float2 e1 = cis_cycles(g.vsign * (tz - xs[2]) * P.inv_lambda);
            const float ar = pre * e1.x - pim * e1.y, ai = pre * e1.y + pim * e1.x;
            atomic_add_c(&grid[((size_t)j * g.nx + iy) * g.nx + ix], (ar * twr - ai * twi) * m,
                         (ar * twi + ai * twr) * m);
            return;
        }
        const float2 e1 = cis_cycles(-(tz - xs[2]) * P.inv_lambda);
        const float ar = pre * e1.x - pim * e1.y, ai = pre * e1.y + pim * e1.x;
        for (int l = 0; l < P.nl; ++l) {
            const float2 c2 = cis_cycles(dot<3>(P.l + 3 * l, xl) * P.inv_lambda);
            const float br = m * c2.x, bi = m * c2.y;
            atomic_add_c(&grid[(((size_t)j * P.nl + l) * g.nx + iy) * g.nx + ix], ar * br - ai * bi,
                         ar * bi + ai * br);
        }
        return;
    }
    double sx, cxv;
    sincos(2 * kPi * xi, &sx, &cxv);
    const double pre = amp * cxv, pim = amp * sx;
    if (wave) {
        // MCfieldFourierQuadOnWave.m:281-302: pxL exp(i vsign k (tz - z_s)) twl atl,
        // twl = sum_l Wl(l) exp(i vsign k l.x_L)
        double twr = 0, twi = 0;
        for (int l = 0; l < P.nl; ++l) {
            double s2, c2;
            sincos(g.vsign * P.k * dot<3>(P.l + 3 * l, xl), &s2, &c2);
            const double2 w = g.wl[l];
            twr += w.x * c2 - w.y * s2;
            twi += w.x * s2 + w.y * c2;
        }
        double s1, c1;
        sincos(g.vsign * P.k * (-xs[2] + tz), &s1, &c1);
        const double m = exp(att);
        double ar = pre * c1 - pim * s1, ai = pre * s1 + pim * c1;
        double tr = (ar * twr - ai * twi) * m, tim = (ar * twi + ai * twr) * m;
        atomic_add_c(&grid[((size_t)j * g.nx + iy) * g.nx + ix], tr, tim);
        return;
    }
    double s1, c1;
    sincos(-P.k * (-xs[2] + tz), &s1, &c1);
    for (int l = 0; l < P.nl; ++l) {
        const double *lc = P.l + 3 * l;
        double ph = P.k * dot<3>(lc, xl);
        double s2, c2;
        sincos(ph, &s2, &c2);
        const double m2 = exp(att);
        // pxL * e1 * e2
        double ar = pre * c1 - pim * s1, ai = pre * s1 + pim * c1;
        double br = m2 * c2, bi = m2 * s2;
        double tr = ar * br - ai * bi, tim = ar * bi + ai * br;
        atomic_add_c(&grid[(((size_t)j * P.nl + l) * g.nx + iy) * g.nx + ix], tr, tim);
    }
}

// Pass 1 of a whole path lands in one cell: its scatter position is the
// path's first vertex for every vertex.  The fused kernel therefore sums a
// path's pass-1 terms in registers -- per vertex the light-side factor
// amp e^{2 pi i xi} e^{i k l.x_n} att(x_n) (the wave: the Wl-weighted sum over
// lights) -- and adds the sum times the first vertex's slab phase with one
// atomic pair per light when the path ends (quad_flush_path).
template <typename T>
struct QuadPathSum {
    T re[kQuadFusedLights], im[kQuadFusedLights];
};

template <typename T>
__device__ __forceinline__ void quad_forward_term(const DevProblem &P, const QuadGeom &g, bool wave,
                                                  const WaveLight *wls, const double *xl, double amp, double xi,
                                                  QuadPathSum<T> &S)
{
    const double att = -P.sigt / 2 * ((xl[2] - P.bmin[2]) * g.l_diratt + (P.bmax[2] - xl[2]) * g.l_idiratt);
    if constexpr (std::is_same<T, float>::value) {
        const float2 ex = cis_cycles(xi);
        const float m = __builtin_amdgcn_exp2f((float)(att * 1.4426950408889634));
        const float pre = (float)amp * m * ex.x, pim = (float)amp * m * ex.y;
        if (wave) {
            const float2 tw = wave_sum(wls, P.nl, xl);
            const float twr = tw.x, twi = tw.y;
            S.re[0] += pre * twr - pim * twi;
            S.im[0] += pre * twi + pim * twr;
            return;
        }
#pragma unroll
        for (int l = 0; l < kQuadFusedLights; ++l)
            if (l < P.nl) {
                const float2 c2 = cis_cycles(dot<3>(P.l + 3 * l, xl) * P.inv_lambda);
                S.re[l] += pre * c2.x - pim * c2.y;
                S.im[l] += pre * c2.y + pim * c2.x;
            }
        return;
    }
    double sx, cxv;
    sincos(2 * kPi * xi, &sx, &cxv);
    const double m = exp(att);
    const double pre = amp * m * cxv, pim = amp * m * sx;
    if (wave) {
        double twr = 0, twi = 0;
        for (int l = 0; l < P.nl; ++l) {
            double s2, c2;
            sincos(g.vsign * P.k * dot<3>(P.l + 3 * l, xl), &s2, &c2);
            const double2 w = g.wl[l];
            twr += w.x * c2 - w.y * s2;
            twi += w.x * s2 + w.y * c2;
        }
        S.re[0] += pre * twr - pim * twi;
        S.im[0] += pre * twi + pim * twr;
        return;
    }
#pragma unroll
    for (int l = 0; l < kQuadFusedLights; ++l)
        if (l < P.nl) {
            double s2, c2;
            sincos(P.k * dot<3>(P.l + 3 * l, xl), &s2, &c2);
            S.re[l] += pre * c2 - pim * s2;
            S.im[l] += pre * s2 + pim * c2;
        }
}

template <typename T>
__device__ __forceinline__ void quad_flush_path(const DevProblem &P, const QuadGeom &g, bool wave, const double *x1,
                                                const QuadPathSum<T> &S, typename cx<T>::t *grid)
{
    const int j = slab_of(x1[2], g);
    if (j < 0) return;
    long ix = lround((x1[0] - g.gx0) / g.box_stp), iy = lround((x1[1] - g.gy0) / g.box_stp);
    ix = ix < 0 ? 0 : (ix > g.nx - 1 ? g.nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy > g.nx - 1 ? g.nx - 1 : iy);
    const double tz = 0.5 * (g.gz[j] + g.gz[j + 1]);
    double er, ei;
    if constexpr (std::is_same<T, float>::value) {
        const float2 e1 = cis_cycles((wave ? g.vsign : -1.0) * (tz - x1[2]) * P.inv_lambda);
        er = e1.x;
        ei = e1.y;
    } else {
        sincos((wave ? g.vsign : -1.0) * P.k * (tz - x1[2]), &ei, &er);
    }
    const int nout = wave ? 1 : P.nl;
#pragma unroll
    for (int l = 0; l < kQuadFusedLights; ++l)
        if (l < nout) {
            const T sr = S.re[l], si = S.im[l];
            atomic_add_c(&grid[(((size_t)j * nout + l) * g.nx + iy) * g.nx + ix], (double)(sr * (T)er - si * (T)ei),
                         (double)(sr * (T)ei + si * (T)er));
        }
}

// Vertex-parallel scatter over stored records (one lane per vertex): for
// many lights per vertex, where the walk's divergence would multiply the
// per-vertex work in the fused kernel below.
template <typename T>
__global__ __launch_bounds__(256) void quad_scatter_kernel(DevProblem P, QuadGeom g, int64_t nvert, const PathRec *paths,
                                                           const VertexRec *verts, const int32_t *vpath,
                                                           const int64_t *offsets, typename cx<T>::t *grid)
{
    __shared__ WaveLight wls[kWaveLightsLds];
    const bool staged = stage_wave(P, g, wls);
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nvert) return;
    const VertexRec r = verts[e];
    const int p = vpath[e];
    const PathRec pr = paths[p];
    double amp = sqrt(r.weight / pr.px);
    if (P.do_cbs && e != offsets[p]) amp *= 1 / sqrt(2.0);
    if (staged) {
        quad_scatter_pass<T>(P, g, g.wave, wls, pr.x1, r.x, amp, r.xi, grid);
        if (P.do_cbs) quad_scatter_pass<T>(P, g, g.wave, wls, r.x, pr.x1, amp, r.xi, grid);
    } else {
        quad_scatter_pass<T>(P, g, g.wave, g.wlk, pr.x1, r.x, amp, r.xi, grid);
        if (P.do_cbs) quad_scatter_pass<T>(P, g, g.wave, g.wlk, r.x, pr.x1, amp, r.xi, grid);
    }
}

// The walk and the scatter in one pass: every vertex is scattered as the walk
// reaches it, so no vertex records are stored and no counting pass is needed.
// Lanes stay resident and take a new path at the wave's next refill once
// theirs ends (refill_due, take_path): a wave no longer waits for its longest
// walk.  A lane follows
// trace_emit_kernel's walk exactly (vertex n: position, weight and xi = the
// first uniform of draw n, taken before the step); *nvert receives the vertex
// count (the normaliser), one atomic per workgroup.  *ctr (zeroed) hands out
// the path indices [0, np).
template <typename T, bool WAVE>
__global__ __launch_bounds__(256) void quad_trace_scatter_kernel(DevProblem P, QuadGeom g, uint64_t p0, int64_t np,
                                                                 typename cx<T>::t *grid,
                                                                 unsigned long long *nvert, unsigned long long *ctr)
{
    __shared__ unsigned long long ws[4];
    __shared__ double lcos[kIcdfLds];
    __shared__ WaveLight wls[WAVE ? kWaveLightsLds : 1];
    const double *tab = stage_icdf(P.amp, lcos);
    if (WAVE) stage_wave(P, g, wls);   // a fused wave has at most kWaveLightsLds lights
    unsigned long long cnt = 0;
    WavePaths wp;
    bool active = false;
    uint64_t path = 0;
    double x[3] = {0, 0, 0}, w[3] = {0, 0, 0}, ow[3] = {0, 0, 0}, x1[3] = {0, 0, 0}, px = 1, weight = 0;
    uint32_t n = 0;
    QuadPathSum<T> S;
    for (;;) {
        if (refill_due(wp, active)) {
            const int64_t id = take_path(wp, !active, np, ctr);
            if (id >= 0) {
                path = P.path_offset + p0 + (uint64_t)id;
                double2 w0p;
                path_start<3>(P, path, x, px, w, w0p);
                ow[0] = ow[1] = ow[2] = 0;   // a new path has no incoming direction
                for (int j = 0; j < 3; ++j) x1[j] = x[j];
                weight = P.albedo;
                n = 0;
#pragma unroll
                for (int l = 0; l < kQuadFusedLights; ++l) S.re[l] = S.im[l] = 0;
                active = true;
            }
        }
        if (!__ballot(active)) break;
        if (active) {
            ++n;
            const double2 a = uniform_pair(P.seed, path, n, 0);
            double amp = sqrt(weight / px);
            if (P.do_cbs && n > 1) amp *= 1 / sqrt(2.0);
            quad_forward_term<T>(P, g, WAVE, wls, x, amp, a.x, S);
            if (P.do_cbs) quad_scatter_pass<T>(P, g, WAVE, wls, x, x1, amp, a.x, grid);
            ++cnt;
            if (!path_step<3>(P, path, n, x, w, ow, weight, a.y, tab)) {
                quad_flush_path<T>(P, g, WAVE, x1, S, grid);
                active = false;
            }
        }
    }
    // workgroup sum of the vertex counts
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(nvert, ws[0] + ws[1] + ws[2] + ws[3]);
}

// Batched complex GEMM C[b] = A[b] * B[b] (column-major, lda = M, ldb = K,
// ldc = M), strides in elements (0 = shared operand).  64x64 tile per
// workgroup, 4x4 outputs per thread, 16-deep K panels in LDS.
template <typename T>
__global__ __launch_bounds__(256) void cgemm_kernel(int M, int N, int K, const typename cx<T>::t *A, int64_t sa,
                                                    const typename cx<T>::t *B, int64_t sb, typename cx<T>::t *C,
                                                    int64_t sc, bool transB)
{
    typedef typename cx<T>::t C2;
    __shared__ C2 As[16][64 + 1];
    __shared__ C2 Bs[16][64 + 1];
    const int b = blockIdx.z;
    const C2 *Ab = A + (size_t)b * sa;
    const C2 *Bb = B + (size_t)b * sb;
    C2 *Cb = C + (size_t)b * sc;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    T accr[4][4], acci[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) accr[i][j] = acci[i][j] = 0;
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int t = threadIdx.x; t < 16 * 64; t += 256) {
            const int kk = t / 64, mm = t % 64;
            const int gm = m0 + mm, gk = k0 + kk;
            C2 a;
            a.x = a.y = 0;
            if (gm < M && gk < K) a = Ab[(size_t)gk * M + gm];
            As[kk][mm] = a;
            const int gn = n0 + mm;
            C2 bb;
            bb.x = bb.y = 0;
            if (gn < N && gk < K) bb = transB ? Bb[(size_t)gk * N + gn] : Bb[(size_t)gn * K + gk];
            Bs[kk][mm] = bb;
        }
        __syncthreads();
        for (int kk = 0; kk < 16; ++kk) {
            C2 av[4], bv[4];
            for (int i = 0; i < 4; ++i) av[i] = As[kk][tx + 16 * i];
            for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][ty + 16 * j];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    accr[i][j] += av[i].x * bv[j].x - av[i].y * bv[j].y;
                    acci[i][j] += av[i].x * bv[j].y + av[i].y * bv[j].x;
                }
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const int gm = m0 + tx + 16 * i, gn = n0 + ty + 16 * j;
            if (gm < M && gn < N) {
                C2 c;
                c.x = accr[i][j];
                c.y = acci[i][j];
                Cb[(size_t)gn * M + gm] = c;
            }
        }
}

// u[o, l] = scale * boxexp0[o] * sum_s H[s, l][o] * exp(zexp0[o] * tz_s)
// (MCfieldFourierQuad.m:124-138, 281-282, 291-303), one thread per output
// pixel o = i + nv0 j and plane l; vx, vy, vz of MCfieldFourierQuad.m:103-105.
template <typename T>
__global__ void quad_combine_kernel(int nv0, int nl, int nz, const typename cx<T>::t *H, const double *gz, QuadView v,
                                    double2 *u)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nvv = (int64_t)nv0 * nv0;
    if (idx >= nvv * nl) return;
    const int64_t o = idx % nvv;
    const int l = (int)(idx / nvv);
    const int i = (int)(o % nv0), j = (int)(o / nv0);
    const double vx = -(-v.v_max + i * v.v_stp), vy = -(-v.v_max + j * v.v_stp);
    const double vz = v.vsign * sqrt(1 - vx * vx - vy * vy);
    const double dir = vz < 0 ? 1 / fabs(vz) : 0, idir = vz > 0 ? 1 / fabs(vz) : 0;
    // zexp0: MCfieldFourierQuad.m:134 (-) / MCfieldFourierQuadOnWave.m:143 (+)
    const double zr = -v.sigt / 2 * (dir - idir);
    const double zi = (v.wave ? 2 : -2) * kPi / v.lambda * vz * v.vsign;
    double ur = 0, ui = 0;
    for (int s = 0; s < nz - 1; ++s) {
        const double tz = 0.5 * (gz[s] + gz[s + 1]);
        const typename cx<T>::t h = H[((size_t)s * nl + l) * nvv + o];
        const double m = exp(zr * tz);
        double sn, cs;
        sincos(zi * tz, &sn, &cs);
        const double er = m * cs, ei = m * sn;
        ur += (double)h.x * er - (double)h.y * ei;
        ui += (double)h.x * ei + (double)h.y * er;
    }
    // boxexp0 (MCfieldFourierQuad.m:137-138)
    const double bm = exp(-v.sigt / 2 * (-v.zmin * dir + v.zmax * idir)) * v.scale;
    double bs, bc;
    sincos(2 * kPi / v.lambda * (-v.zmin * (vz < 0) * vz + v.zmax * (vz > 0) * vz), &bs, &bc);
    const double br = bm * bc, bi = bm * bs;
    u[idx].x = ur * br - ui * bi;
    u[idx].y = ur * bi + ui * br;
}

template <typename T>
__global__ void quad_dft_matrix_kernel(int nv0, int nx, int n, int nva_pad, int nx_pad, const double2 *tw,
                                       typename cx<T>::t *W)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nv0 * nx) return;
    const int o = (int)(idx % nv0), ix = (int)(idx / nv0);
    const int h = n / 2, am = n - 1 - h;
    const int64_t f = o + nva_pad - am, q = ix + nx_pad - h;
    const int64_t r = ((f * q) % n + n) % n;
    const double2 t = tw[r];
    W[idx].x = (T)t.x;
    W[idx].y = (T)t.y;
}

void launch_quad_dft_matrix(const QuadGeom &g, const double2 *tw, void *W, bool fp32, hipStream_t s)
{
    const int64_t n = (int64_t)g.nv0 * g.nx;
    dim3 gr((unsigned)((n + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL(quad_dft_matrix_kernel<float>, gr, dim3(256), 0, s, g.nv0, g.nx, g.nfft, g.nva_pad, g.nx_pad,
                           tw, (float2 *)W);
    else
        hipLaunchKernelGGL(quad_dft_matrix_kernel<double>, gr, dim3(256), 0, s, g.nv0, g.nx, g.nfft, g.nva_pad,
                           g.nx_pad, tw, (double2 *)W);
}

// One pass of the slab transform as a factored DFT (four-step Cooley-Tukey):
// for n = n1 n2, x[q1 + n1 q2] -> Y[f2 + n2 f1] =
//   sum_q1 w^(q1 f2) w1^(q1 f1) sum_q2 w2^(q2 f2) x[q1 + n1 q2],
// w = e^{2 pi i/n}, w1 = w^n2, w2 = w^n1: n (n1 + n2) products per vector
// instead of the n^2 of the DFT matrix.  Vector j of batch b holds nin input
// elements at in[b in_bs + j in_vs + e in_es], placed at q = (e + qoff) mod n
// (the other positions are zero); output o < nout is Y[(o + foff) mod n],
// stored at out[b out_bs + j out_vs + o out_es].  With qoff = nx_pad - h and
// foff = nva_pad - a this is the row (or column) transform of H = W X W^T with
// W[o][ix] = w^((o + nva_pad - a)(ix + nx_pad - h) mod n).
// A workgroup takes nv consecutive vectors of one batch in LDS (two buffers of
// nv x n) with the n twiddles and a table of the index split k = c n2 + f
// (both stages walk the outputs k in that form; stage 1's twiddle exponent is
// c f < n).  Every per-element index is stepped, not divided: a thread's
// elements are t, t + 256, ..., and (t / m, t mod m) advances by a constant.
struct DftPass {
    int n, n1, n2, nin, nout, nvec, nv, qoff, foff;
    int64_t in_bs, in_vs, in_es, out_bs, out_vs, out_es;
};

// (t / m, t mod m) for t = t0, t0 + stride, ...
struct DivStep {
    int q, r, dq, dr, m;
    __device__ DivStep(int t0, int m_, int stride) : m(m_)
    {
        q = t0 / m;
        r = t0 - q * m;
        dq = stride / m;
        dr = stride - dq * m;
    }
    __device__ void next()
    {
        q += dq;
        r += dr;
        if (r >= m) {
            r -= m;
            ++q;
        }
    }
};

// (e + off) mod n for 0 <= e < n and -n < off < n
__device__ __forceinline__ int wrap_index(int e, int off, int n)
{
    int q = e + off;
    q += q < 0 ? n : 0;
    q -= q >= n ? n : 0;
    return q;
}

template <typename T>
__global__ __launch_bounds__(256) void dft_pass_kernel(DftPass d, const typename cx<T>::t *__restrict__ in,
                                                       typename cx<T>::t *__restrict__ out,
                                                       const double2 *__restrict__ tw)
{
    typedef typename cx<T>::t C2;
    extern __shared__ __attribute__((aligned(16))) unsigned char dft_lds[];
    C2 *tws = (C2 *)dft_lds;          // n twiddles
    C2 *xa = tws + d.n;               // nv x n: input, then stage 2 output
    C2 *xb = xa + (size_t)d.nv * d.n; // nv x n: stage 1 output
    int *kd = (int *)(xb + (size_t)d.nv * d.n);   // n entries: c | f << 16
    const int b = blockIdx.y, j0 = blockIdx.x * d.nv;
    const int nv = min(d.nv, d.nvec - j0);
    const int n = d.n, n1 = d.n1, n2 = d.n2;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        tws[r].x = (T)tw[r].x;
        tws[r].y = (T)tw[r].y;
        const int c = r / n2;
        kd[r] = c | ((r - c * n2) << 16);
    }
    for (int e = threadIdx.x; e < d.nv * n; e += blockDim.x) {
        xa[e].x = 0;
        xa[e].y = 0;
    }
    __syncthreads();
    // load: consecutive threads read consecutive addresses
    const C2 *ib = in + (size_t)b * d.in_bs;
    if (d.in_es == 1) {
        DivStep it(threadIdx.x, d.nin, blockDim.x);   // (v, e)
        for (int t = threadIdx.x; t < nv * d.nin; t += blockDim.x, it.next())
            xa[it.q * n + wrap_index(it.r, d.qoff, n)] = ib[(size_t)(j0 + it.q) * d.in_vs + it.r];
    } else {
        DivStep it(threadIdx.x, nv, blockDim.x);   // (e, v)
        for (int t = threadIdx.x; t < nv * d.nin; t += blockDim.x, it.next())
            xa[it.r * n + wrap_index(it.q, d.qoff, n)] = ib[(size_t)(j0 + it.r) * d.in_vs + (size_t)it.q * d.in_es];
    }
    __syncthreads();
    // stage 1: Z[q1 n2 + f2] = w^(q1 f2) sum_q2 w2^(q2 f2) x[q1 + n1 q2]
    {
        DivStep it(threadIdx.x, n, blockDim.x);   // (v, k)
        for (int t = threadIdx.x; t < nv * n; t += blockDim.x, it.next()) {
            const int dk = kd[it.r], q1 = dk & 0xffff, f2 = dk >> 16;
            const C2 *x = xa + it.q * n + q1;
            const int inc = f2 * n1;
            T sr = 0, si = 0;
            int rn = 0;   // ((q2 f2) mod n2) n1
            for (int q2 = 0; q2 < n2; ++q2, x += n1) {
                const C2 w = tws[rn];
                const C2 a = *x;
                sr += a.x * w.x - a.y * w.y;
                si += a.x * w.y + a.y * w.x;
                rn += inc;
                rn -= rn >= n ? n : 0;
            }
            const C2 w = tws[q1 * f2];
            xb[t].x = sr * w.x - si * w.y;
            xb[t].y = sr * w.y + si * w.x;
        }
    }
    __syncthreads();
    // stage 2: Y[f2 + n2 f1] = sum_q1 w1^(q1 f1) Z[q1 n2 + f2]
    {
        DivStep it(threadIdx.x, n, blockDim.x);   // (v, k)
        for (int t = threadIdx.x; t < nv * n; t += blockDim.x, it.next()) {
            const int dk = kd[it.r], f1 = dk & 0xffff, f2 = dk >> 16;
            const C2 *z = xb + it.q * n + f2;
            const int inc = f1 * n2;
            T sr = 0, si = 0;
            int rn = 0;   // ((q1 f1) mod n1) n2
            for (int q1 = 0; q1 < n1; ++q1, z += n2) {
                const C2 w = tws[rn];
                const C2 a = *z;
                sr += a.x * w.x - a.y * w.y;
                si += a.x * w.y + a.y * w.x;
                rn += inc;
                rn -= rn >= n ? n : 0;
            }
            xa[it.q * n + f2 + n2 * f1].x = sr;
            xa[it.q * n + f2 + n2 * f1].y = si;
        }
    }
    __syncthreads();
    C2 *ob = out + (size_t)b * d.out_bs;
    if (d.out_es == 1) {
        DivStep it(threadIdx.x, d.nout, blockDim.x);   // (v, o)
        for (int t = threadIdx.x; t < nv * d.nout; t += blockDim.x, it.next())
            ob[(size_t)(j0 + it.q) * d.out_vs + it.r] = xa[it.q * n + wrap_index(it.r, d.foff, n)];
    } else {
        DivStep it(threadIdx.x, nv, blockDim.x);   // (o, v)
        for (int t = threadIdx.x; t < nv * d.nout; t += blockDim.x, it.next())
            ob[(size_t)(j0 + it.r) * d.out_vs + (size_t)it.q * d.out_es] = xa[it.r * n + wrap_index(it.q, d.foff, n)];
    }
}

// n = n1 n2 with n1 + n2 smallest (n1 >= n2 >= 2); false if n is prime or the
// factored form (n (n1 + n2) products per vector against n^2, with more index
// work per product) saves less than a factor 3.
bool dft_factor(int n, int *n1, int *n2)
{
    int best = 0;
    for (int f = 2; f * f <= n; ++f)
        if (n % f == 0) best = f;
    if (!best) return false;
    *n2 = best;
    *n1 = n / best;
    return 3 * (*n1 + *n2) <= n;
}

// H[b] = W X[b] W^T for every slab (batch b) by two factored-DFT passes:
// along ix into T (nv0 x nx per batch), then along iy into H (nv0 x nv0).
// Returns false (nothing launched) when n does not factor usefully or one
// vector of it does not fit in LDS.
bool launch_quad_dft2(const QuadGeom &g, int batch, const void *X, void *T1, void *H, const double2 *tw, bool fp32,
                      hipStream_t s)
{
    int n1, n2;
    const int n = g.nfft;
    if (!dft_factor(n, &n1, &n2)) return false;
    const int h = n / 2, am = n - 1 - h;
    const size_t cs = fp32 ? sizeof(float2) : sizeof(double2);
    // two nv x n buffers, the twiddles and the index table within 64 KB of
    // LDS; a transform too long for one vector goes to the DFT-matrix GEMMs
    const int64_t fixed = (int64_t)(cs + sizeof(int)) * n, per = 2 * (int64_t)cs * n;
    if (fixed + per > 65536) return false;
    const int nv = (int)std::min<int64_t>(16, (65536 - fixed) / per);
    const size_t lds = (size_t)(fixed + per * nv);
    DftPass a;
    a.n = n;
    a.n1 = n1;
    a.n2 = n2;
    a.qoff = g.nx_pad - h;
    a.foff = g.nva_pad - am;
    a.nv = nv;
    // pass 1: vectors = columns iy of X (ix contiguous), out T[o + nv0 iy]
    a.nin = g.nx;
    a.nout = g.nv0;
    a.nvec = g.nx;
    a.in_bs = (int64_t)g.nx * g.nx;
    a.in_vs = g.nx;
    a.in_es = 1;
    a.out_bs = (int64_t)g.nv0 * g.nx;
    a.out_vs = g.nv0;
    a.out_es = 1;
    DftPass c = a;
    // pass 2: vectors = rows o of T (iy strided by nv0), out H[o + nv0 o']
    c.nvec = g.nv0;
    c.in_bs = (int64_t)g.nv0 * g.nx;
    c.in_vs = 1;
    c.in_es = g.nv0;
    c.out_bs = (int64_t)g.nv0 * g.nv0;
    c.out_vs = 1;
    c.out_es = g.nv0;
    dim3 ga((unsigned)((a.nvec + nv - 1) / nv), (unsigned)batch), gc((unsigned)((c.nvec + nv - 1) / nv), (unsigned)batch);
    if (fp32) {
        hipLaunchKernelGGL(dft_pass_kernel<float>, ga, dim3(256), lds, s, a, (const float2 *)X, (float2 *)T1, tw);
        hipLaunchKernelGGL(dft_pass_kernel<float>, gc, dim3(256), lds, s, c, (const float2 *)T1, (float2 *)H, tw);
    } else {
        hipLaunchKernelGGL(dft_pass_kernel<double>, ga, dim3(256), lds, s, a, (const double2 *)X, (double2 *)T1, tw);
        hipLaunchKernelGGL(dft_pass_kernel<double>, gc, dim3(256), lds, s, c, (const double2 *)T1, (double2 *)H, tw);
    }
    return true;
}

void launch_quad_scatter(const DevProblem &P, const QuadGeom &g, int64_t nvert, const PathRec *paths,
                         const VertexRec *verts, const int32_t *vpath, const int64_t *offsets, void *grid, bool fp32,
                         hipStream_t s)
{
    if (nvert <= 0) return;
    dim3 gr((unsigned)((nvert + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL(quad_scatter_kernel<float>, gr, dim3(256), 0, s, P, g, nvert, paths, verts, vpath, offsets,
                           (float2 *)grid);
    else
        hipLaunchKernelGGL(quad_scatter_kernel<double>, gr, dim3(256), 0, s, P, g, nvert, paths, verts, vpath,
                           offsets, (double2 *)grid);
}

bool launch_quad_trace_scatter(const DevProblem &P, const QuadGeom &g, uint64_t p0, int64_t np, void *grid, bool fp32,
                               unsigned long long *nvert, unsigned long long *ctr, hipStream_t s)
{
    if (g.wave ? P.nl > kWaveLightsLds : P.nl > kQuadFusedLights) return false;
    if (np <= 0) return true;
    auto go = [&](auto kernel, auto *gr) {
        hipLaunchKernelGGL(kernel, dim3(resident_blocks((const void *)kernel, np)), dim3(256), 0, s, P, g, p0, np, gr,
                           nvert, ctr);
    };
    if (fp32) {
        if (g.wave)
            go(quad_trace_scatter_kernel<float, true>, (float2 *)grid);
        else
            go(quad_trace_scatter_kernel<float, false>, (float2 *)grid);
    } else {
        if (g.wave)
            go(quad_trace_scatter_kernel<double, true>, (double2 *)grid);
        else
            go(quad_trace_scatter_kernel<double, false>, (double2 *)grid);
    }
    return true;
}
void launch_cgemm(int M, int N, int K, const void *A, int64_t sa, const void *B, int64_t sb, void *C, int64_t sc,
                  int batch, bool transB, bool fp32, hipStream_t s)
{
    dim3 gr((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), (unsigned)batch);
    if (fp32)
        hipLaunchKernelGGL(cgemm_kernel<float>, gr, dim3(256), 0, s, M, N, K, (const float2 *)A, sa, (const float2 *)B,
                           sb, (float2 *)C, sc, transB);
    else
        hipLaunchKernelGGL(cgemm_kernel<double>, gr, dim3(256), 0, s, M, N, K, (const double2 *)A, sa,
                           (const double2 *)B, sb, (double2 *)C, sc, transB);
}

void launch_quad_combine(int nv0, int nl, int nz, const void *H, const double *gz, const QuadView &v, double2 *u,
                         bool fp32, hipStream_t s)
{
    const int64_t n = (int64_t)nv0 * nv0 * nl;
    dim3 gr((unsigned)((n + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL(quad_combine_kernel<float>, gr, dim3(256), 0, s, nv0, nl, nz, (const float2 *)H, gz, v, u);
    else
        hipLaunchKernelGGL(quad_combine_kernel<double>, gr, dim3(256), 0, s, nv0, nl, nz, (const double2 *)H, gz, v,
                           u);
}

} // namespace scmc
