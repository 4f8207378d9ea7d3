// nee_cov.hip -- next-event estimation factors for the covariance update.
//
// For every path and vertex this evaluates the complex volumetric
// throughputs of code/MCcov.m:157-222 (evalphaseatt x evalampfunc_general)
// and writes them as the Kronecker factors of the rank-1 term
//   u(v,l) = sqrt(W/px) * [ e_v_ms(v) e_l0_ms(l) + e_v0_ms(v) e_l_ms(l) ]
// whose outer product u u^H expands to the four M00/Mnn/Mn0/M0n terms of
// MCcov.m:244-254 (the second term only with CBS; MCcov.m:256-262 otherwise).
// Single scattering (MCcov.m:234-241) is u_s(v,l) = sqrt(albedo/px) e_v0(v)
// e_l0(l) af_ang_vl(v,l); in the far field the amplitude table is a per-row
// factor applied when the covariance is finalised, in the near field it
// depends on the first vertex and the rows are written out whole.
//
// Evaluation is FP64 in both precision modes (it is <1% of the work at the
// headline size); the mixed mode stores FP32 factors and records the largest
// |u| bound of the chunk for the FP16 power-of-two scaling.
#include "hetero.h"
#include "kernels.h"

#include <type_traits>

#include <algorithm>

namespace scmc {

template <typename T> __device__ __forceinline__ T to_c(double2 v);
template <> __device__ __forceinline__ float2 to_c<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
template <> __device__ __forceinline__ double2 to_c<double2>(double2 v) { return v; }

__device__ __forceinline__ float cmag(double2 v) { return (float)fmax(fabs(v.x), fabs(v.y)); }
__device__ __forceinline__ bool cfinite(double2 v) { return isfinite(v.x) && isfinite(v.y); }

__device__ __forceinline__ float block_max(float v, float *sh)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    float r = 0.f;
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) r = fmaxf(r, sh[i]);
    return r;
}

// Positive floats order like their bit patterns.
__device__ __forceinline__ void atomic_max_pos(unsigned *p, float v)
{
    if (v > 0.f) atomicMax(p, __float_as_uint(v));
}

// MCcov.m:157-179 : path-level factors at the first vertex.  The entrance
// directions rl/rv are the source columns in the far field and depend on x1
// in the near field (MCcov.m:128-135).  Far field only, the single-scatter
// bound is the product of the factor maxima (the amplitude table is applied
// when finalising); near field the single-scatter rows carry the amplitude
// and nee_single_near_kernel records their bound.

// The single-scattering weight is the first vertex's W (albedo, or the
// albedo of its cell in a voxel medium, MCCovHetro.m:164).  With a medium,
// EV0 / UL0 receive W/px e_v0 and the unattenuated u_l0 for the mean field.
template <int DIM, typename T, bool HET>
__global__ __launch_bounds__(128) void nee_cov_path_kernel(DevProblem P, HetMedium M, int64_t np, const PathRec *paths,
                                                           const VertexRec *verts, const int64_t *offsets, T *Gp,
                                                           T *Fp2, T *SA, T *SB, float *maxG, float *maxF2,
                                                           unsigned *bound, double2 *EV0, double2 *UL0,
                                                           unsigned long long *nonfinite)
{
    __shared__ float sh[4];
    const int64_t p = blockIdx.x;
    if (p >= np) return;
    bool bad = false;   // a NaN/Inf factor (path weight, position, direction)
    const PathRec pr = paths[p];
    const double W1 = verts[offsets[p]].weight;
    const double ssw = sqrt(W1 / pr.px);
    float mG = 0.f, mF = 0.f, mSA = 0.f, mSB = 0.f;
    for (int i = threadIdx.x; i < P.nl; i += blockDim.x) {
        const double *lc = P.l + (size_t)i * DIM;
        double rl[3];
        light_dir<DIM>(P, pr.x1, i, rl);
        double2 e = nee_e<DIM, HET>(P, M, pr.x1, lc, 1.0, P.is_ff_l);
        double2 af = cdiv(eval_amp<DIM>(dot<DIM>(pr.w1, rl), P.amp), pr.w0p);
        double2 g = cmul(e, af);
        Gp[p * P.nl + i] = to_c<T>(g);
        SB[p * P.nl + i] = to_c<T>(e);
        bad |= !cfinite(g) || !cfinite(e);
        if (HET && UL0) {
            DevProblem P0 = P;
            P0.sigt = 0.0;   // MCCovHetro.m:152: evalphaseatt(x, l, is_ff_l, 0, ...)
            UL0[p * P.nl + i] = phase_att64<DIM>(P0, pr.x1, lc, 1.0, P.is_ff_l);
        }
        mG = fmaxf(mG, cmag(g));
        mSB = fmaxf(mSB, cmag(e));
    }
    for (int i = threadIdx.x; i < P.nv; i += blockDim.x) {
        const double *vc = P.v + (size_t)i * DIM;
        double2 e = nee_e<DIM, HET>(P, M, pr.x1, vc, P.ff_sign, P.is_ff_v);
        double2 se = make_double2(ssw * e.x, ssw * e.y);
        SA[p * P.nv + i] = to_c<T>(se);
        bad |= !cfinite(se);
        if (HET && EV0) {
            const double wp = W1 / pr.px;
            EV0[p * P.nv + i] = make_double2(wp * e.x, wp * e.y);
        }
        mSA = fmaxf(mSA, cmag(se));
        if (P.do_cbs) {
            double rv[3];
            view_dir<DIM>(P, pr.x1, i, rv);
            double mw[3] = {-pr.w1[0], -pr.w1[1], -pr.w1[2]};
            double2 af = cdiv(eval_amp<DIM>(dot<DIM>(mw, rv), P.amp), pr.w0p);
            double2 f = cmul(e, af);
            Fp2[p * P.nv + i] = to_c<T>(f);
            mF = fmaxf(mF, cmag(f));
            bad |= !cfinite(f);
        }
    }
    mG = block_max(mG, sh);
    mF = block_max(mF, sh);
    mSA = block_max(mSA, sh);
    mSB = block_max(mSB, sh);
    if (threadIdx.x == 0) {
        maxG[p] = mG;
        maxF2[p] = mF;
        if (P.is_ff_l && P.is_ff_v) atomic_max_pos(&bound[1], 2.f * mSA * mSB);
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0 && nonfinite) atomicAdd(nonfinite, 1ULL);
}

// Near-field single scattering (MCcov.m:138-142, 234-241): the amplitude
// af(v,l) = a(rl_l . rv_v) depends on x1, so the rank-1 term is not a
// Kronecker product; the rows u_s[p][v + nv*l] = SA[p][v] SB[p][l] af(v,l)
// are written out and fed to the update as an (n x 1) "Kronecker" pair with
// B = 1 (ones[p]).
template <int DIM, typename T>
__global__ __launch_bounds__(256) void nee_single_near_kernel(DevProblem P, int64_t np, const PathRec *paths,
                                                              const T *SA, const T *SB, T *U, T *ones,
                                                              unsigned *bound)
{
    __shared__ float sh[4];
    const int64_t p = blockIdx.x;
    if (p >= np) return;
    const PathRec pr = paths[p];
    const int n = P.nv * P.nl;
    float mU = 0.f;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const int v = r % P.nv, l = r / P.nv;
        double rl[3], rv[3];
        light_dir<DIM>(P, pr.x1, l, rl);
        view_dir<DIM>(P, pr.x1, v, rv);
        double2 af = eval_amp<DIM>(dot<DIM>(rl, rv), P.amp);
        T a = SA[p * P.nv + v], b = SB[p * P.nl + l];
        double2 u = cmul(cmul(make_double2(a.x, a.y), make_double2(b.x, b.y)), af);
        U[p * n + r] = to_c<T>(u);
        mU = fmaxf(mU, cmag(u));
    }
    mU = block_max(mU, sh);
    if (threadIdx.x == 0) {
        ones[p] = to_c<T>(make_double2(1.0, 0.0));
        atomic_max_pos(&bound[1], mU);
    }
}

// MCcov.m:205-222 : vertex-level factors for pL > 1, written at the
// multiple-scattering index m = vertex - path - 1 (path-major order).  A
// workgroup takes vpb consecutive vertices and spreads their nv view items
// and nl light items over its threads; the per-vertex maxima meet in LDS and
// the workgroup adds one value to the chunk bound.
constexpr int kNeeVertexThreads = 256, kNeeMaxVpb = 64;
template <int DIM, typename T, bool HET, bool FAST = false>
__global__ __launch_bounds__(kNeeVertexThreads) void nee_cov_vertex_kernel(
    DevProblem P, HetMedium M, int64_t nvert, int vpb, const PathRec *paths, const VertexRec *verts,
    const int32_t *vpath, const int64_t *offsets, const T *Gp, const T *Fp2, const float *maxG, const float *maxF2,
    T *A1, T *B1, T *A2, T *B2, unsigned *bound, unsigned long long *nonfinite)
{
    __shared__ unsigned mAv[kNeeMaxVpb], mBv[kNeeMaxVpb], badv[kNeeMaxVpb];
    __shared__ float sh[kNeeVertexThreads / 64];
    const int64_t v0 = (int64_t)blockIdx.x * vpb;
    for (int j = threadIdx.x; j < vpb; j += blockDim.x) mAv[j] = mBv[j] = badv[j] = 0u;
    __syncthreads();
    const int w = P.nv + P.nl;
    // FAST (FP32 factors, far-field views and lights, homogeneous medium,
    // isotropic or HG phase; checked by the launcher): E by the FP64 cycle
    // reduction and FP32 sin/cos/exp, the amplitude in FP32 (the factors
    // are rounded to FP32 anyway).  Its own instance, so the FP64 path's
    // registers do not limit its occupancy.
    static_assert(!FAST || (std::is_same<T, float2>::value && !HET), "FAST needs FP32 factors, no voxels");
    const AmpF32 af32 = amp_f32_setup<DIM>(P.amp);
    for (int j = threadIdx.x; j < vpb * w; j += blockDim.x) {
        const int vl = j / w, i = j - vl * w;
        const int64_t vi = v0 + vl;
        if (vi >= nvert) continue;
        const int p = vpath[vi];
        if (vi == offsets[p]) continue;   // first vertex: single scattering
        const int64_t m = vi - p - 1;
        const VertexRec r = verts[vi];
        const PathRec pr = paths[p];
        const double sw = sqrt(r.weight / pr.px);
        if constexpr (FAST) {
            const float swf = (float)sw;
            float2 f;
            if (i < P.nv) {
                const double *vc = P.v + (size_t)i * DIM;   // far field: rv = v_i
                f = phase_att_mixed<DIM>(P, r.x, vc, P.ff_sign, 1);
                const float a = amp_f32<DIM>((float)dot<DIM>(r.ow, vc), af32) * swf;
                f.x *= a;
                f.y *= a;
                A1[m * P.nv + i] = to_c<T>(make_double2(f.x, f.y));
                const float mg = sqrtf(f.x * f.x + f.y * f.y);
                if (mg > 0.f) atomicMax(&mAv[vl], __float_as_uint(mg));
                if (!(isfinite(f.x) && isfinite(f.y))) atomicOr(&badv[vl], 1u);
                if (P.do_cbs) A2[m * P.nv + i] = Fp2[(int64_t)p * P.nv + i];
            } else {
                const int li = i - P.nv;
                B1[m * P.nl + li] = Gp[(int64_t)p * P.nl + li];
                if (P.do_cbs) {
                    const double *lc = P.l + (size_t)li * DIM;   // far field: rl = l_j
                    f = phase_att_mixed<DIM>(P, r.x, lc, 1.0, 1);
                    const double mow[3] = {-r.ow[0], -r.ow[1], -r.ow[2]};
                    const float a = amp_f32<DIM>((float)dot<DIM>(mow, lc), af32) * swf;
                    f.x *= a;
                    f.y *= a;
                    B2[m * P.nl + li] = to_c<T>(make_double2(f.x, f.y));
                    const float mg = sqrtf(f.x * f.x + f.y * f.y);
                    if (mg > 0.f) atomicMax(&mBv[vl], __float_as_uint(mg));
                    if (!(isfinite(f.x) && isfinite(f.y))) atomicOr(&badv[vl], 1u);
                }
            }
            continue;
        }
        if (i < P.nv) {
            const double *vc = P.v + (size_t)i * DIM;
            double rv[3];
            view_dir<DIM>(P, pr.x1, i, rv);
            double2 e = nee_e<DIM, HET>(P, M, r.x, vc, P.ff_sign, P.is_ff_v);
            double2 af = eval_amp<DIM>(dot<DIM>(r.ow, rv), P.amp);
            double2 f = cmul(e, af);
            f.x *= sw;
            f.y *= sw;
            A1[m * P.nv + i] = to_c<T>(f);
            const float mg = cmag(f);
            if (mg > 0.f) atomicMax(&mAv[vl], __float_as_uint(mg));
            if (!cfinite(f)) atomicOr(&badv[vl], 1u);
            if (P.do_cbs) A2[m * P.nv + i] = Fp2[(int64_t)p * P.nv + i];
        } else {
            const int li = i - P.nv;
            B1[m * P.nl + li] = Gp[(int64_t)p * P.nl + li];
            if (P.do_cbs) {
                const double *lc = P.l + (size_t)li * DIM;
                double rl[3];
                light_dir<DIM>(P, pr.x1, li, rl);
                double2 e = nee_e<DIM, HET>(P, M, r.x, lc, 1.0, P.is_ff_l);
                double mow[3] = {-r.ow[0], -r.ow[1], -r.ow[2]};
                double2 af = eval_amp<DIM>(dot<DIM>(mow, rl), P.amp);
                double2 g = cmul(e, af);
                g.x *= sw;
                g.y *= sw;
                B2[m * P.nl + li] = to_c<T>(g);
                const float mg = cmag(g);
                if (mg > 0.f) atomicMax(&mBv[vl], __float_as_uint(mg));
                if (!cfinite(g)) atomicOr(&badv[vl], 1u);
            }
        }
    }
    __syncthreads();
    float bv = 0.f;
    for (int j = threadIdx.x; j < vpb; j += blockDim.x) {
        const int64_t vi = v0 + j;
        if (vi >= nvert) continue;
        const int p = vpath[vi];
        if (vi == offsets[p]) continue;
        bv = fmaxf(bv, 2.f * (__uint_as_float(mAv[j]) * maxG[p] + maxF2[p] * __uint_as_float(mBv[j])));
        if (badv[j] && nonfinite) atomicAdd(nonfinite, 1ULL);
    }
    bv = block_max(bv, sh);
    if (threadIdx.x == 0) atomic_max_pos(&bound[0], bv);
}

template <int DIM, typename T, bool HET>
static void nee_cov_dispatch(const DevProblem &P, const HetMedium &M, int64_t np, int64_t nvert, const PathRec *paths,
                             const VertexRec *verts, const int32_t *vpath, const int64_t *offsets,
                             const CovFactors &f, hipStream_t s)
{
    if (np > 0)
        hipLaunchKernelGGL((nee_cov_path_kernel<DIM, T, HET>), dim3((unsigned)np), dim3(128), 0, s, P, M, np, paths,
                           verts, offsets, (T *)f.Gp, (T *)f.Fp2, (T *)f.SA, (T *)f.SB, f.maxG, f.maxF2, f.bound,
                           f.EV0, f.UL0, f.nonfinite);
    if (np > 0 && !(P.is_ff_l && P.is_ff_v))
        hipLaunchKernelGGL((nee_single_near_kernel<DIM, T>), dim3((unsigned)np), dim3(256), 0, s, P, np, paths,
                           (const T *)f.SA, (const T *)f.SB, (T *)f.U, (T *)f.ones, f.bound);
    if (nvert > 0) {
        // about four items per thread
        const int vpb = std::max(1, std::min(kNeeMaxVpb, 4 * kNeeVertexThreads / (P.nv + P.nl)));
        const dim3 grid((unsigned)((nvert + vpb - 1) / vpb));
        constexpr bool kCanFast = std::is_same<T, float2>::value && !HET;
        if (kCanFast && P.is_ff_v && P.is_ff_l && P.amp.type != 2)
            hipLaunchKernelGGL((nee_cov_vertex_kernel<DIM, T, HET, kCanFast>), grid, dim3(kNeeVertexThreads), 0, s, P,
                               M, nvert, vpb, paths, verts, vpath, offsets, (const T *)f.Gp, (const T *)f.Fp2,
                               f.maxG, f.maxF2, (T *)f.A1, (T *)f.B1, (T *)f.A2, (T *)f.B2, f.bound, f.nonfinite);
        else
            hipLaunchKernelGGL((nee_cov_vertex_kernel<DIM, T, HET, false>), grid, dim3(kNeeVertexThreads), 0, s, P, M,
                               nvert, vpb, paths, verts, vpath, offsets, (const T *)f.Gp, (const T *)f.Fp2, f.maxG,
                               f.maxF2, (T *)f.A1, (T *)f.B1, (T *)f.A2, (T *)f.B2, f.bound, f.nonfinite);
    }
}

void launch_nee_cov(const DevProblem &P, const HetMedium *M, int64_t np, int64_t nvert, const PathRec *paths,
                    const VertexRec *verts, const int32_t *vpath, const int64_t *offsets, const CovFactors &f,
                    bool mixed, hipStream_t s)
{
    HetMedium none{};
#define SCMC_NEE(D, T)                                                                                  \
    do {                                                                                                \
        if (M)                                                                                          \
            nee_cov_dispatch<D, T, true>(P, *M, np, nvert, paths, verts, vpath, offsets, f, s);         \
        else                                                                                            \
            nee_cov_dispatch<D, T, false>(P, none, np, nvert, paths, verts, vpath, offsets, f, s);      \
    } while (0)
    if (P.dim == 3) {
        if (mixed)
            SCMC_NEE(3, float2);
        else
            SCMC_NEE(3, double2);
    } else {
        if (mixed)
            SCMC_NEE(2, float2);
        else
            SCMC_NEE(2, double2);
    }
#undef SCMC_NEE
}

// af_ang_vl(v,l) = evalampfunc_general(l(:,j)'*v) (MCcov.m:91-96), far field.
template <int DIM>
__global__ void af_table_kernel(DevProblem P, double2 *af)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)P.nv * P.nl) return;
    int v = (int)(idx % P.nv), l = (int)(idx / P.nv);
    af[idx] = eval_amp<DIM>(dot<DIM>(P.l + (size_t)l * DIM, P.v + (size_t)v * DIM), P.amp);
}

void launch_af_table(const DevProblem &P, double2 *af, hipStream_t s)
{
    int64_t n = (int64_t)P.nv * P.nl;
    dim3 g((unsigned)((n + 255) / 256));
    if (P.dim == 3)
        hipLaunchKernelGGL(af_table_kernel<3>, g, dim3(256), 0, s, P, af);
    else
        hipLaunchKernelGGL(af_table_kernel<2>, g, dim3(256), 0, s, P, af);
}

} // namespace scmc
