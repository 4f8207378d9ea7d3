// field.hip -- MCfield field accumulation (code/MCfield.m:190-222).
//
// The field is a sum over "view-side entries" e of
//     u(v,l) += E(x_e, v) * a_e(v,l) * coef_e(l),   E = evalphaseatt(x, -v)
// with three entry kinds:
//   first vertex (pL == 1, MCfield.m:209): x = x1, a = af_ang_vl(v,l),
//       coef = c_1 e_l0(l)
//   vertex pL > 1 (MCfield.m:211):         x = x_n, a = a(ow . v),
//       coef = c_n e_l0_ms(l) [/sqrt 2 with CBS]
//   reversed path, CBS (MCfield.m:213):    x = x1, a = a(-w1 . v)/w0p,
//       coef = sum_n c_n e_l_ms(x_n)(l) / sqrt 2
// where c_n = sqrt(W_n/px) exp(2 pi i xi_n) (MCfield.m:218-220).  The CBS
// term of every vertex of a path shares x1, so it is summed per path first.
//
// Illumination wave (code/MCfieldOnWave.m): the lights are combined with the
// weights Wl before the outer product, so every light-side coefficient becomes
// a sum over lights and u has one column; the first-vertex entry keeps its
// per-light coefficients because af_ang_vl(v,l) couples view and light
// (MCfieldOnWave.m:214).
//
// Layout: one lane per view, a workgroup per 256 views x a slice of the
// entries; entries are staged through LDS in batches and broadcast to all
// lanes; FP64 accumulators in registers; per-slice partial sums are reduced
// in a fixed order (deterministic).  The mixed variant forms the phase k x.v
// and its range reduction in FP64 and evaluates sin/cos (hardware, input in
// revolutions), the attenuation exponential and the amplitude in FP32.
#include "hetero.h"
#include "kernels.h"

#include <type_traits>

namespace scmc {

constexpr int kFieldViews = 256;   // views per workgroup
constexpr int kFieldBatch = 64;    // entries per LDS batch
constexpr int kFieldNL = 4;        // lights per launch pass

// Per-path / per-vertex coefficients.  Block per path, lanes over lights.
// HET: voxel medium (MCsampleHetro.m:141-176, evalphaseattHetro).
// nonfinite += paths with a NaN/Inf coefficient (path weights, positions).
// FAST (mixed precision, far-field lights, no voxels, isotropic or HG phase;
// chosen by the launcher): the per-vertex e^{2 pi i xi} and the CBS light
// terms in FP32 after the FP64 cycle reduction (the accumulation that uses
// the coefficients stays FP64).
template <int DIM>
__device__ __forceinline__ double2 cn_f32(const VertexRec &r, const PathRec &pr)
{
    const float amp = sqrtf((float)(r.weight / pr.px));
    const float fr = (float)(r.xi - rint(r.xi));   // xi in revolutions
    return make_double2(amp * __builtin_amdgcn_cosf(fr), amp * __builtin_amdgcn_sinf(fr));
}

template <int DIM, bool HET, bool FAST = false>
__global__ __launch_bounds__(64) void field_coef_kernel(DevProblem P, HetMedium M, int64_t np, const PathRec *paths,
                                                        const VertexRec *verts, const int64_t *offsets,
                                                        double2 *coef, double2 *coefR, unsigned long long *nonfinite)
{
    static_assert(!FAST || !HET, "FAST: homogeneous media only");
    const AmpF32 af32 = amp_f32_setup<DIM>(P.amp);
    const int64_t p = blockIdx.x;
    if (p >= np) return;
    bool bad = false;
    const PathRec pr = paths[p];
    const int64_t o = offsets[p], e = offsets[p + 1];
    const double rs2 = 1 / sqrt(2.0);
    for (int l = threadIdx.x; l < P.nl; l += blockDim.x) {
        const double *lc = P.l + (size_t)l * DIM;
        double rl[3];
        light_dir<DIM>(P, pr.x1, l, rl);   // MCfield.m:134-137 (near field: from x1)
        double2 el0 = nee_e<DIM, HET>(P, M, pr.x1, lc, 1.0, P.is_ff_l);
        double2 af = cdiv(eval_amp<DIM>(dot<DIM>(pr.w1, rl), P.amp), pr.w0p);
        double2 el0ms = cmul(el0, af);
        double2 R = make_double2(0.0, 0.0);
        for (int64_t vi = o; vi < e; ++vi) {
            const VertexRec r = verts[vi];
            double2 cn;
            if constexpr (FAST) {
                cn = cn_f32<DIM>(r, pr);
            } else {
                double amp = sqrt(r.weight / pr.px);
                double s, c;
                sincos(2 * kPi * r.xi, &s, &c);
                cn = make_double2(amp * c, amp * s);
            }
            double2 t;
            if (vi == o) {
                t = cmul(el0, cn);
            } else {
                t = cmul(el0ms, cn);
                if (P.do_cbs) {
                    t.x *= rs2;
                    t.y *= rs2;
                    double mow[3] = {-r.ow[0], -r.ow[1], -r.ow[2]};
                    double2 el;
                    if constexpr (FAST) {
                        const float2 e32 = phase_att_mixed<DIM>(P, r.x, lc, 1.0, 1);
                        const float a = amp_f32<DIM>((float)dot<DIM>(mow, rl), af32);
                        el = make_double2(e32.x * a, e32.y * a);
                    } else {
                        el = cmul(nee_e<DIM, HET>(P, M, r.x, lc, 1.0, P.is_ff_l),
                                  eval_amp<DIM>(dot<DIM>(mow, rl), P.amp));
                    }
                    double2 q = cmul(el, cn);
                    R.x += q.x * rs2;
                    R.y += q.y * rs2;
                }
            }
            coef[vi * P.nl + l] = t;
            bad |= !(isfinite(t.x) && isfinite(t.y));
        }
        if (P.do_cbs) coefR[p * P.nl + l] = R;
        bad |= !(isfinite(R.x) && isfinite(R.y));
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0 && nonfinite) atomicAdd(nonfinite, 1ULL);
}

__device__ __forceinline__ double2 wave_sum(double2 v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o);
        v.y += __shfl_xor(v.y, o);
    }
    return v;
}

// MCfieldOnWave.m:165-221 per path (one wave, lanes over lights):
//   S0 = sum_l e_l0(l) af_l(l) Wl(l)                       (:177)
//   coef[n] = c_n S0 [/sqrt 2]      for vertices n >= 2     (:211, :213)
//   coefR   = sum_n c_n SL_n / sqrt 2, SL_n = sum_l e_l(x_n, l) a(-ow_n . rl_l) Wl(l)   (:199, :213)
//   cl0[l]  = c_1 e_l0(l) Wl(l)     (first vertex, contracted with af_ang_vl in the
//                                    view kernel, :214)
template <int DIM, bool FAST = false>
__global__ __launch_bounds__(64) void field_wave_coef_kernel(DevProblem P, int64_t np, const PathRec *paths,
                                                             const VertexRec *verts, const int64_t *offsets,
                                                             const double2 *wl, double2 *coef, double2 *coefR,
                                                             double2 *cl0, unsigned long long *nonfinite)
{
    const AmpF32 af32 = amp_f32_setup<DIM>(P.amp);
    const int64_t p = blockIdx.x;
    if (p >= np) return;
    bool bad = false;
    const PathRec pr = paths[p];
    const int64_t o = offsets[p], e = offsets[p + 1];
    const double rs2 = 1 / sqrt(2.0);
    const int lane = threadIdx.x;
    auto cn_of = [&](int64_t vi) {
        const VertexRec r = verts[vi];
        if constexpr (FAST) return cn_f32<DIM>(r, pr);
        double amp = sqrt(r.weight / pr.px);
        double sn, cs;
        sincos(2 * kPi * r.xi, &sn, &cs);
        return make_double2(amp * cs, amp * sn);
    };
    const double2 c1 = cn_of(o);
    double2 s0 = make_double2(0.0, 0.0);
    for (int l = lane; l < P.nl; l += 64) {
        const double *lc = P.l + (size_t)l * DIM;
        double rl[3];
        light_dir<DIM>(P, pr.x1, l, rl);
        double2 el0 = phase_att64<DIM>(P, pr.x1, lc, 1.0, P.is_ff_l);
        double2 af = cdiv(eval_amp<DIM>(dot<DIM>(pr.w1, rl), P.amp), pr.w0p);
        double2 t = cmul(cmul(el0, af), wl[l]);
        s0.x += t.x;
        s0.y += t.y;
        cl0[p * P.nl + l] = cmul(cmul(el0, wl[l]), c1);
        bad |= !(isfinite(cl0[p * P.nl + l].x) && isfinite(cl0[p * P.nl + l].y));
    }
    s0 = wave_sum(s0);
    for (int64_t vi = o + lane; vi < e; vi += 64) {
        double2 t = make_double2(0.0, 0.0);
        if (vi != o) {
            t = cmul(cn_of(vi), s0);
            if (P.do_cbs) {
                t.x *= rs2;
                t.y *= rs2;
            }
        }
        coef[vi] = t;
        bad |= !(isfinite(t.x) && isfinite(t.y));
    }
    if (P.do_cbs) {
        double2 R = make_double2(0.0, 0.0);
        for (int64_t vi = o + 1; vi < e; ++vi) {
            const VertexRec r = verts[vi];
            const double mow[3] = {-r.ow[0], -r.ow[1], -r.ow[2]};
            double2 sl = make_double2(0.0, 0.0);
            for (int l = lane; l < P.nl; l += 64) {
                const double *lc = P.l + (size_t)l * DIM;
                double rl[3];
                light_dir<DIM>(P, pr.x1, l, rl);
                double2 el;
                if constexpr (FAST) {
                    const float2 e32 = phase_att_mixed<DIM>(P, r.x, lc, 1.0, 1);
                    const float a = amp_f32<DIM>((float)dot<DIM>(mow, rl), af32);
                    el = make_double2(e32.x * a, e32.y * a);
                } else {
                    el = cmul(phase_att64<DIM>(P, r.x, lc, 1.0, P.is_ff_l), eval_amp<DIM>(dot<DIM>(mow, rl), P.amp));
                }
                double2 t = cmul(el, wl[l]);
                sl.x += t.x;
                sl.y += t.y;
            }
            sl = wave_sum(sl);
            double2 q = cmul(cn_of(vi), sl);
            R.x += q.x * rs2;
            R.y += q.y * rs2;
        }
        if (lane == 0) coefR[p] = R;
        bad |= !(isfinite(R.x) && isfinite(R.y));
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0 && nonfinite) atomicAdd(nonfinite, 1ULL);
}

// The entry side of the exit distance: the signed distance to the nearer
// face (the exit distance of the mixed far-field E), from the FP64 position.  A point on the max face gives
// -0.0, whose sign bit still selects fbn.
__device__ __forceinline__ float near_face(double x, double lo, double hi)
{
    const double dlo = x - lo, dhi = hi - x;
    return dlo <= dhi ? (float)dlo : -(float)dhi;
}

// Bits 0-2: the sign bits of q.  Bit 3 (kLastAxisOnly): the point is farther
// than P.far_last_bound from the nearer face on every axis but the last, so
// every far-field view leaves it through the last axis's faces (the exit
// distance through axis j is at least |q_j| since |v_j| <= 1, through the last
// axis at most that bound); the minimum over the faces is then that face's
// distance, exactly as the full minimum would find it (the bound carries a
// margin over FP32 rounding).  The kernels test the bit for a whole batch of
// entries (one barrier) and run the batch with only the last axis when every
// entry has it -- in a slab whose lateral extent is far larger than its
// thickness, nearly every batch.
constexpr int kLastAxisOnly = 8;
template <int DIM>
__device__ __forceinline__ int qneg_bits(const DevProblem &P, const float *q)
{
    int b = (int)((unsigned)__float_as_int(q[0]) >> 31 | ((unsigned)__float_as_int(q[1]) >> 31) << 1 |
                  ((unsigned)__float_as_int(q[2]) >> 31) << 2);
    bool last = true;
    for (int j = 0; j + 1 < DIM; ++j) last = last && fabsf(q[j]) > P.far_last_bound;
    return last ? b | kLastAxisOnly : b;
}

struct FEntry {
    double x[3];
    double ad[3];     // amplitude direction: ow (vertex) or -w1 (reversed)
    double2 div;      // 1/w0p for the reversed entry, (1,0) otherwise
    float q[3];       // signed distance to the nearer face in FP32 (mixed
                      // attenuation): x - box_min, or -(box_max - x) when
                      // the max face is nearer (near_face)
    int qneg;         // bit j: the sign bit of q[j]
    float adf[3];     // ad rounded to FP32 (mixed HG amplitude)
    int kind;         // 0 first vertex, 1 later vertex, 2 reversed path
    int p;            // path of the entry
};

template <int DIM>
__device__ __forceinline__ void load_entry(const DevProblem &P, int64_t e, int64_t nvert, const PathRec *paths,
                                           const VertexRec *verts, const int32_t *vpath, const int64_t *offsets,
                                           FEntry &f)
{
    if (e < nvert) {
        const VertexRec r = verts[e];
        const int p = vpath[e];
        for (int j = 0; j < 3; ++j) {
            f.x[j] = r.x[j];
            f.ad[j] = r.ow[j];
        }
        f.kind = (e == offsets[p]) ? 0 : 1;
        f.p = p;
        f.div = make_double2(1.0, 0.0);
    } else {
        const PathRec pr = paths[e - nvert];
        for (int j = 0; j < 3; ++j) {
            f.x[j] = pr.x1[j];
            f.ad[j] = -pr.w1[j];
        }
        f.kind = 2;
        f.p = (int)(e - nvert);
        double den = pr.w0p.x * pr.w0p.x + pr.w0p.y * pr.w0p.y;
        f.div = make_double2(pr.w0p.x / den, -pr.w0p.y / den);
    }
    for (int j = 0; j < 3; ++j) {
        f.q[j] = j < DIM ? near_face(f.x[j], P.bmin[j], P.bmax[j]) : 0.f;
        f.adf[j] = (float)f.ad[j];
    }
    f.qneg = qneg_bits<DIM>(P, f.q);
}

// Tabulated amplitude with the nearest-bin index formed in FP64, so the bin
// matches evalampfunc_amplitude.m:8 exactly (an FP32 acos flips bins near
// their edges).
template <int DIM>
__device__ __forceinline__ float2 amp_tab(double c, const DevPhase &ph)
{
    double2 a = eval_amp<DIM>(c, ph);
    return make_float2((float)a.x, (float)a.y);
}

// Per-view constants of the mixed far-field E: the direction times
// -1/lambda (the phase in cycles is then x . vn: the far-field phase is
// k x.(-v)), and the distance over which the view leaves the box
// (cubeDist.m:4-11 for a point inside the box), D_j = dlo_j / a_j when
// v_j <= 0 and dhi_j / a_j when v_j > 0, with dlo = x - box_min,
// dhi = box_max - x, a_j = max(|v_j|, 1e-7) and W_j = dlo_j + dhi_j the box
// width.  An entry carries q_j = dlo_j when that face is nearer and -dhi_j
// otherwise (near_face), and D_j = fa_j q_j + (q_j < 0 ? fbn_j : fbp_j):
//   q >= 0, v_j <= 0:  q/a             q >= 0, v_j > 0:  (W - q)/a
//   q <  0, v_j <= 0:  (W + q)/a       q <  0, v_j > 0:  -q/a
// The only difference of two FP32 terms, W/a -|q|/a, has |q| <= W/2, so it
// never cancels; writing every case from dlo alone (W/a - dlo/a) loses all
// bits of dhi next to the max face of a wide box (a slab of lateral width
// 1e5: an absolute error of ~8 distance units at |v_j| ~ 1e-3).
struct FarView {
    double vn[3];
    float fa[3], fbp[3], fbn[3];
};
template <int DIM>
__device__ __forceinline__ FarView far_view(const DevProblem &P, const double *v)
{
    FarView f;
    for (int j = 0; j < 3; ++j) {
        const double nd = j < DIM ? v[j] : 0.0;
        f.vn[j] = nd * -P.inv_lambda;
        const float inv = (float)(1.0 / fmax(fabs(nd), 1.0000000000000004e-07));
        const float wi = j < DIM ? (float)(P.bmax[j] - P.bmin[j]) * inv : 0.f;
        f.fa[j] = nd > 0 ? -inv : inv;
        f.fbp[j] = nd > 0 ? wi : 0.f;
        f.fbn[j] = nd > 0 ? 0.f : wi;
    }
    return f;
}

// qneg: bit j set when q_j carries a sign bit.  Every lane of a wave reads
// the same entry, so the caller passes qneg as a wave-uniform (scalar)
// value and each face costs one select with a scalar mask and one FMA.
__device__ __forceinline__ float exit_dist(const FarView &fv, int j, float q, int qneg)
{
    return fmaf(fv.fa[j], q, ((qneg >> j) & 1) ? fv.fbn[j] : fv.fbp[j]);
}

// Far-field E(x, v) of the mixed variant: FP64 x.vn and range reduction to a
// fraction of a cycle, hardware sin/cos (input in revolutions), FP32
// attenuation over the distance to the face the view leaves by (a select and
// an FMA per face from the entry's q).  nhs = -sigma/2 log2(e).  lamp: log2
// of a real amplitude factor folded into the same exp2 (0 for none), so an
// isotropic or HG amplitude costs no extra exponential and no complex
// multiply.
template <int DIM, bool LAST = false>
__device__ __forceinline__ float2 e_far_mixed(const double *x, const FarView &fv, const float *q, int qneg, float nhs,
                                              float lamp = 0.f)
{
    const double cyc = dot<DIM>(x, fv.vn);
    const float fr = (float)(cyc - rint(cyc));
    const float sn = __builtin_amdgcn_sinf(fr), cs = __builtin_amdgcn_cosf(fr);
    float best;
    if (LAST) {
        // the point is farther from the other axes' faces than any view's exit
        // through the last axis (kLastAxisOnly, for every entry of the batch)
        best = exit_dist(fv, DIM - 1, q[DIM - 1], qneg);
    } else {
        // the first face starts the minimum (a start at +inf costs a min per call)
        best = exit_dist(fv, 0, q[0], qneg);
        for (int j = 1; j < DIM; ++j) best = fminf(best, exit_dist(fv, j, q[j], qneg));
    }
    const float mag = __builtin_amdgcn_exp2f(fmaf(nhs, best, lamp));
    return make_float2(mag * cs, mag * sn);
}

// The same for a voxel medium (far field): the optical depth by the FP64
// march of evalphaseattHetro, the phase reduced in FP64, sin/cos/exp in FP32.
template <int DIM>
__device__ __forceinline__ float2 e_far_mixed_het(const DevProblem &P, const HetMedium &M, const double *x,
                                                  const double *mv, float lamp = 0.f)
{
    const double cyc = dot<DIM>(x, mv) * P.inv_lambda;
    const float fr = (float)(cyc - rint(cyc));
    const float sn = __builtin_amdgcn_sinf(fr), cs = __builtin_amdgcn_cosf(fr);
    const float depth = (float)march_depth<DIM>(P, M, x, mv, 1e12);
    const float mag = __builtin_amdgcn_exp2f(fmaf(-0.5f * 1.4426950408889634f, depth, lamp));
    return make_float2(mag * cs, mag * sn);
}

// WAVE: one output column; coef/coefR hold the light-summed coefficients and
// the first-vertex entries are left to field_wave_first_kernel.  NLQ:
// light columns accumulated per pass (1 for one column, else kFieldNL).
// VPT views per lane (v, v + 256, ...): the entry reads from LDS and the
// per-entry scalar work are shared between them.  AMPT: 0 reads the
// amplitude type at run time; 3 is an instance for HG amplitudes only (the
// uniform type tests then fold away).
// Waves per SIMD the compiler must leave room for: five for the mixed HG
// one-column instance with two views per lane (MCfield), whose batch loop
// exists twice (all faces / last axis only) and would otherwise settle at four.
constexpr int field_min_waves(bool mixed, bool het, int vpt, int ampt) { return mixed && !het && vpt == 2 && ampt == 3 ? 5 : 1; }

template <int DIM, bool MIXED, bool WAVE, int NLQ, bool HET = false, int VPT = 1, int AMPT = 0>
__global__ __launch_bounds__(kFieldViews) __attribute__((amdgpu_waves_per_eu(field_min_waves(MIXED, HET, VPT, AMPT))))
void field_kernel(DevProblem P, HetMedium M, int64_t np, int64_t nvert, int64_t nent,
                                                            int nsplit, int l0, int nlp, const PathRec *paths,
                                                            const VertexRec *verts, const int32_t *vpath,
                                                            const int64_t *offsets, const double2 *coef,
                                                            const double2 *coefR, const double2 *aftab,
                                                            const double2 *cl0, double2 *part)
{
    const int ncol = WAVE ? 1 : P.nl;   // columns of coef / coefR / part
    __shared__ FEntry se[kFieldBatch];
    __shared__ double2 sc[kFieldBatch][kFieldNL];
    // F32ACC (mixed): FP32 copies of the coefficients; each batch's products
    // are summed in FP32 (64 entries at most) and then added into the FP64
    // accumulators.  Not for the two-views-per-lane one-column instance
    // (MCfield), which would then exceed 128 registers (four waves per SIMD)
    // and measured slower.
    constexpr bool F32ACC = MIXED && (WAVE || HET || VPT == 1);
    __shared__ float2 scf[F32ACC ? kFieldBatch : 1][F32ACC ? kFieldNL : 1];
    const int split = blockIdx.y;
    const int64_t e0 = nent * split / nsplit, e1 = nent * (split + 1) / nsplit;

    // per-view constants
    int vv[VPT];
    bool vok[VPT];
    double vd[VPT][3], mv[VPT][3];   // mv = -v: the far-field march direction (HET)
    float vf[VPT][3];
    FarView fv[VPT];
    double2 afv[VPT][NLQ], acc[VPT][NLQ];
#pragma unroll
    for (int t = 0; t < VPT; ++t) {
        const int v = (blockIdx.x * VPT + t) * kFieldViews + threadIdx.x;
        vv[t] = v;
        vok[t] = v < P.nv;
        for (int j = 0; j < 3; ++j) vd[t][j] = (vok[t] && j < DIM) ? P.v[(size_t)v * DIM + j] : 0.0;
        for (int j = 0; j < 3; ++j) {
            mv[t][j] = -vd[t][j];
            vf[t][j] = (float)vd[t][j];
        }
        fv[t] = far_view<DIM>(P, vd[t]);
        for (int q = 0; q < NLQ; ++q) {
            afv[t][q] = (!WAVE && vok[t] && q < nlp) ? aftab[(size_t)v + (size_t)P.nv * (l0 + q)]
                                                      : make_double2(0.0, 0.0);
            acc[t][q] = make_double2(0.0, 0.0);
        }
    }
    const float g = (float)P.amp.g;
    const float g2p1 = 1.f + g * g, g2 = 2.f * g;
    const float nhs = (float)(P.sigt * 0.5) * -1.4426950408889634f;
    // FP32 amplitude in the log2 domain (evalampfunc_general.m:5-18,
    // evaluateHG.m:25-39): isotropic a = iso; HG a = hg_k t^mu with
    // t = 1 + g^2 - 2 g c, mu = -3/4 (3-D) or -1/2 (2-D).  Tabulated
    // amplitudes are complex (amp_tab).  Types other than 1 and 2 are HG,
    // as in eval_amp.
    const bool tab = AMPT == 0 ? P.amp.type == 2 : false;
    const bool hg = AMPT == 0 ? P.amp.type != 1 && !tab : AMPT == 3;
    // uniform: kept in a scalar register
    const float lk = __builtin_bit_cast(
        float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(
                   int, hg ? (float)log2(sqrt((1 - P.amp.g * P.amp.g) / ((DIM == 3) ? 4 * kPi : 2 * kPi)))
                           : (float)log2(P.amp.iso))));
    const float mu = (DIM == 3) ? -0.75f : -0.5f;

    for (int64_t eb = e0; eb < e1; eb += kFieldBatch) {
        const int cnt = (int)((e1 - eb) < kFieldBatch ? (e1 - eb) : kFieldBatch);
        __syncthreads();
        if (threadIdx.x < cnt)
            load_entry<DIM>(P, eb + threadIdx.x, nvert, paths, verts, vpath, offsets, se[threadIdx.x]);
        for (int t = threadIdx.x; t < cnt * NLQ; t += blockDim.x) {
            const int ei = t / NLQ, q = t % NLQ;
            const int64_t e = eb + ei;
            double2 c = make_double2(0.0, 0.0);
            if (q < nlp)
                c = (e < nvert) ? coef[e * ncol + l0 + q] : coefR[(e - nvert) * ncol + l0 + q];
            sc[ei][q] = c;
            if (F32ACC) scf[ei][q] = make_float2((float)c.x, (float)c.y);
        }
        // mixed far field: the batch leaves every view through the last axis
        int mine = 0;
        if (MIXED && !HET) mine = threadIdx.x >= cnt || (se[threadIdx.x].qneg & kLastAxisOnly) != 0;
        const bool batch_last = __syncthreads_and(mine) != 0;
        if (!vok[0]) continue;
        float2 accf[VPT][NLQ];
#pragma unroll
        for (int t = 0; t < VPT; ++t)
#pragma unroll
            for (int q = 0; q < NLQ; ++q) accf[t][q] = make_float2(0.f, 0.f);
        auto entries = [&](auto last_tag) {
        constexpr bool LAST = decltype(last_tag)::value;
        for (int ei = 0; ei < cnt; ++ei) {
            const FEntry &f = se[ei];
            // every lane reads the same entry: branch on its kind as a scalar
            const int kind = __builtin_amdgcn_readfirstlane(f.kind);
            // wave first vertices: field_wave_first_kernel
            if (WAVE && kind == 0) continue;
            const double x[3] = {f.x[0], f.x[1], f.x[2]};
            float q[3], adf[3];
            const int qneg = __builtin_amdgcn_readfirstlane(f.qneg);
            for (int j = 0; j < 3; ++j) {
                q[j] = f.q[j];
                adf[j] = f.adf[j];
            }
            double2 sce[NLQ];
            for (int q = 0; q < NLQ; ++q) sce[q] = sc[ei][q];
#pragma unroll
            for (int t = 0; t < VPT; ++t) {
                double2 ev;
                if (MIXED) {
                    // isotropic / HG: a real amplitude, log2 a = lk + mu log2(1 + g^2 - 2 g c)
                    float lamp = 0.f;
                    if (kind != 0 && !tab) {
                        lamp = lk;
                        if (hg) {
                            float c = adf[0] * vf[t][0] + adf[1] * vf[t][1] + (DIM == 3 ? adf[2] * vf[t][2] : 0.f);
                            lamp = fmaf(mu, __builtin_amdgcn_logf(fmaxf(g2p1 - g2 * c, 1e-30f)), lk);
                        }
                    }
                    float2 evf;
                    if (HET) {
                        evf = e_far_mixed_het<DIM>(P, M, x, mv[t], lamp);
                    } else {
                        evf = e_far_mixed<DIM, LAST>(x, fv[t], q, qneg, nhs, lamp);
                    }
                    if (kind != 0 && tab) {
                        const float2 am = amp_tab<DIM>(dot<DIM>(f.ad, vd[t]), P.amp);
                        evf = make_float2(evf.x * am.x - evf.y * am.y, evf.x * am.y + evf.y * am.x);
                    }
                    if (F32ACC && kind != 0) {
                        if (kind == 2) {
                            const float dr = (float)f.div.x, di = (float)f.div.y;
                            evf = make_float2(evf.x * dr - evf.y * di, evf.x * di + evf.y * dr);
                        }
#pragma unroll
                        for (int q = 0; q < NLQ; ++q) {
                            // the complex product added as four FMAs into the sum
                            const float2 c = scf[ei][q];
                            accf[t][q].x = fmaf(-evf.y, c.y, fmaf(evf.x, c.x, accf[t][q].x));
                            accf[t][q].y = fmaf(evf.y, c.x, fmaf(evf.x, c.y, accf[t][q].y));
                        }
                        continue;
                    }
                    ev = make_double2(evf.x, evf.y);
                } else {
                    ev = nee_e<DIM, HET>(P, M, x, P.v + (size_t)vv[t] * DIM, P.ff_sign, P.is_ff_v);
                    if (kind != 0) {
                        double2 a = eval_amp<DIM>(dot<DIM>(f.ad, P.v + (size_t)vv[t] * DIM), P.amp);
                        ev = cmul(ev, a);
                    }
                }
                if (!F32ACC && kind == 2) ev = cmul(ev, f.div);
                // one column: the first vertex's amplitude factor goes into ev,
                // so every kind adds with the same four FMAs (one code path
                // keeps the accumulators in place across the loop)
                if (NLQ == 1 && kind == 0) ev = cmul(ev, afv[t][0]);
                if (NLQ > 1 && kind == 0) {
                    for (int q = 0; q < NLQ; ++q) {
                        const double2 tt = cmul(ev, afv[t][q]), c = sce[q];
                        acc[t][q].x += tt.x * c.x - tt.y * c.y;
                        acc[t][q].y += tt.x * c.y + tt.y * c.x;
                    }
                } else {
                    for (int q = 0; q < NLQ; ++q) {
                        // the complex product added as four FMAs into the sum
                        const double2 c = sce[q];
                        acc[t][q].x = fma(-ev.y, c.y, fma(ev.x, c.x, acc[t][q].x));
                        acc[t][q].y = fma(ev.y, c.x, fma(ev.x, c.y, acc[t][q].y));
                    }
                }
            }
        }
        };
        if constexpr (MIXED && !HET) {
            if (batch_last)
                entries(std::true_type{});
            else
                entries(std::false_type{});
        } else {
            entries(std::false_type{});
        }
        if (F32ACC) {
#pragma unroll
            for (int t = 0; t < VPT; ++t)
#pragma unroll
                for (int q = 0; q < NLQ; ++q) {
                    acc[t][q].x += (double)accf[t][q].x;
                    acc[t][q].y += (double)accf[t][q].y;
                }
        }
    }
#pragma unroll
    for (int t = 0; t < VPT; ++t)
        if (vok[t])
#pragma unroll
            for (int q = 0; q < NLQ; ++q)   // static indices keep acc in registers
                if (q < nlp) part[((size_t)split * ncol + l0 + q) * P.nv + vv[t]] = acc[t][q];
}

// Near-field sources (either side): the entrance directions rl/rv depend on
// the path's first vertex x1 (MCfield.m:129-143), so the amplitude of every
// entry is evaluated per (entry, view) from x1, and the first-vertex
// amplitude table af_ang_vl per (entry, view, light).  Evaluated in FP64 in
// both precision modes.
struct NEntry {
    double x[3];
    double x1[3];
    double ad[3];
    double2 div;
    int kind;
    int p;
};

template <int DIM, bool WAVE, bool HET = false>
__global__ __launch_bounds__(kFieldViews) void field_near_kernel(DevProblem P, HetMedium M, int64_t np, int64_t nvert,
                                                                 int64_t nent,
                                                                 int nsplit, int l0, int nlp, const PathRec *paths,
                                                                 const VertexRec *verts, const int32_t *vpath,
                                                                 const int64_t *offsets, const double2 *coef,
                                                                 const double2 *coefR, const double2 *cl0,
                                                                 double2 *part)
{
    const int ncol = WAVE ? 1 : P.nl;
    __shared__ NEntry se[kFieldBatch];
    __shared__ double2 sc[kFieldBatch][kFieldNL];
    __shared__ double srl[kFieldBatch][kFieldNL][3];
    const int v = blockIdx.x * kFieldViews + threadIdx.x;
    const bool vok = v < P.nv;
    const int split = blockIdx.y;
    const int64_t e0 = nent * split / nsplit, e1 = nent * (split + 1) / nsplit;
    double2 acc[kFieldNL];
    for (int q = 0; q < kFieldNL; ++q) acc[q] = make_double2(0.0, 0.0);

    for (int64_t eb = e0; eb < e1; eb += kFieldBatch) {
        const int cnt = (int)((e1 - eb) < kFieldBatch ? (e1 - eb) : kFieldBatch);
        __syncthreads();
        if (threadIdx.x < cnt) {
            const int64_t e = eb + threadIdx.x;
            NEntry &f = se[threadIdx.x];
            if (e < nvert) {
                const VertexRec r = verts[e];
                const int p = vpath[e];
                for (int j = 0; j < 3; ++j) {
                    f.x[j] = r.x[j];
                    f.x1[j] = paths[p].x1[j];
                    f.ad[j] = r.ow[j];
                }
                f.kind = (e == offsets[p]) ? 0 : 1;
                f.p = p;
                f.div = make_double2(1.0, 0.0);
            } else {
                const PathRec pr = paths[e - nvert];
                for (int j = 0; j < 3; ++j) {
                    f.x[j] = pr.x1[j];
                    f.x1[j] = pr.x1[j];
                    f.ad[j] = -pr.w1[j];
                }
                f.kind = 2;
                f.p = (int)(e - nvert);
                double den = pr.w0p.x * pr.w0p.x + pr.w0p.y * pr.w0p.y;
                f.div = make_double2(pr.w0p.x / den, -pr.w0p.y / den);
            }
        }
        for (int t = threadIdx.x; t < cnt * kFieldNL; t += blockDim.x) {
            const int ei = t / kFieldNL, q = t % kFieldNL;
            const int64_t e = eb + ei;
            double2 c = make_double2(0.0, 0.0);
            if (q < nlp)
                c = (e < nvert) ? coef[e * ncol + l0 + q] : coefR[(e - nvert) * ncol + l0 + q];
            sc[ei][q] = c;
        }
        __syncthreads();
        // light entrance directions at each entry's x1 (used by first vertices)
        for (int t = threadIdx.x; t < cnt * kFieldNL; t += blockDim.x) {
            const int ei = t / kFieldNL, q = t % kFieldNL;
            double rl[3] = {0.0, 0.0, 0.0};
            if (!WAVE && q < nlp && se[ei].kind == 0) light_dir<DIM>(P, se[ei].x1, l0 + q, rl);
            for (int j = 0; j < 3; ++j) srl[ei][q][j] = rl[j];
        }
        __syncthreads();
        if (!vok) continue;
        for (int ei = 0; ei < cnt; ++ei) {
            const NEntry &f = se[ei];
            double x1[3] = {f.x1[0], f.x1[1], f.x1[2]};
            double rv[3];
            view_dir<DIM>(P, x1, v, rv);
            double xx[3] = {f.x[0], f.x[1], f.x[2]};
            double2 ev = nee_e<DIM, HET>(P, M, xx, P.v + (size_t)v * DIM, P.ff_sign, P.is_ff_v);
            if (WAVE && f.kind == 0) {
                // sum_l a(rl_l . rv) cl0[p, l] (MCfieldOnWave.m:143-148, 214)
                double2 a = make_double2(0.0, 0.0);
                for (int l = 0; l < P.nl; ++l) {
                    double rl[3];
                    light_dir<DIM>(P, x1, l, rl);
                    double2 t = cmul(eval_amp<DIM>(dot<DIM>(rl, rv), P.amp), cl0[(size_t)f.p * P.nl + l]);
                    a.x += t.x;
                    a.y += t.y;
                }
                double2 t = cmul(ev, a);
                acc[0].x += t.x;
                acc[0].y += t.y;
            } else if (f.kind == 0) {
                for (int q = 0; q < kFieldNL; ++q) {
                    double rl[3] = {srl[ei][q][0], srl[ei][q][1], srl[ei][q][2]};
                    double2 t = cmul(ev, eval_amp<DIM>(dot<DIM>(rl, rv), P.amp));
                    double2 c = sc[ei][q];
                    acc[q].x += t.x * c.x - t.y * c.y;
                    acc[q].y += t.x * c.y + t.y * c.x;
                }
            } else {
                double ad[3] = {f.ad[0], f.ad[1], f.ad[2]};
                double2 t = cmul(ev, eval_amp<DIM>(dot<DIM>(ad, rv), P.amp));
                if (f.kind == 2) t = cmul(t, f.div);
                for (int q = 0; q < kFieldNL; ++q) {
                    double2 c = sc[ei][q];
                    acc[q].x += t.x * c.x - t.y * c.y;
                    acc[q].y += t.x * c.y + t.y * c.x;
                }
            }
        }
    }
    if (vok)
#pragma unroll
        for (int q = 0; q < kFieldNL; ++q)   // static indices keep acc in registers
            if (q < nlp) part[((size_t)split * ncol + l0 + q) * P.nv + v] = acc[q];
}

// MCfieldOnWave first vertices, far field (MCfieldOnWave.m:143-148, 214):
//   sum_p E(x1_p, v) sum_l af_ang_vl(v, l) cl0[p, l] = sum_l af_ang_vl(v, l) H[v, l],
//   H[v, l] = sum_p E(x1_p, v) cl0[p, l],
// so the table is applied once per (view, light) instead of once per (path,
// view, light).  A workgroup holds 256 views x kWaveLG lights of H over one
// range of paths, in FP64; the ranges' partials are reduced in a fixed order.
constexpr int kWaveLG = 16;        // lights per workgroup
constexpr int kWavePB = 32;        // paths per LDS batch
constexpr int kWaveMaxSplit = 16;  // path ranges
template <int DIM, bool MIXED>
__global__ __launch_bounds__(kFieldViews) void field_wave_first_kernel(DevProblem P, int64_t np, int nsplit,
                                                                       const PathRec *paths, const double2 *cl0,
                                                                       double2 *hpart)
{
    __shared__ double sx[kWavePB][3];
    __shared__ float sq[kWavePB][3];
    __shared__ int sneg[kWavePB];
    __shared__ double2 scl[kWavePB][kWaveLG];
    const int v = blockIdx.x * kFieldViews + threadIdx.x;
    const bool vok = v < P.nv;
    double vd[3] = {0, 0, 0};
    for (int j = 0; j < DIM && vok; ++j) vd[j] = P.v[(size_t)v * DIM + j];
    const FarView fv = far_view<DIM>(P, vd);
    const float nhs = (float)(P.sigt * 0.5) * -1.4426950408889634f;
    const int l0 = blockIdx.y * kWaveLG;
    const int nlg = (P.nl - l0) < kWaveLG ? (P.nl - l0) : kWaveLG;
    const int split = blockIdx.z;
    const int64_t p0 = np * split / nsplit, p1 = np * (split + 1) / nsplit;
    double2 acc[kWaveLG];
#pragma unroll
    for (int q = 0; q < kWaveLG; ++q) acc[q] = make_double2(0.0, 0.0);
    for (int64_t pb = p0; pb < p1; pb += kWavePB) {
        const int cnt = (int)((p1 - pb) < kWavePB ? (p1 - pb) : kWavePB);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt * 3; t += blockDim.x) {
            const int i = t / 3, j = t % 3;
            const double xj = paths[pb + i].x1[j];
            sx[i][j] = xj;
            sq[i][j] = j < DIM ? near_face(xj, P.bmin[j], P.bmax[j]) : 0.f;
        }
        int mine = 1;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            float q[3];
            for (int j = 0; j < 3; ++j) q[j] = j < DIM ? near_face(paths[pb + i].x1[j], P.bmin[j], P.bmax[j]) : 0.f;
            sneg[i] = qneg_bits<DIM>(P, q);
            mine &= (sneg[i] & kLastAxisOnly) != 0;
        }
        for (int t = threadIdx.x; t < cnt * kWaveLG; t += blockDim.x) {
            const int i = t / kWaveLG, q = t % kWaveLG;
            scl[i][q] = q < nlg ? cl0[(size_t)(pb + i) * P.nl + l0 + q] : make_double2(0.0, 0.0);
        }
        // the batch's first vertices leave every view through the last axis
        const bool batch_last = __syncthreads_and(mine) != 0;
        if (!vok) continue;
        auto paths_of_batch = [&](auto last_tag) {
            constexpr bool LAST = decltype(last_tag)::value;
            for (int i = 0; i < cnt; ++i) {
                const double x[3] = {sx[i][0], sx[i][1], sx[i][2]};
                double2 ev;
                if (MIXED) {
                    const float2 e =
                        e_far_mixed<DIM, LAST>(x, fv, sq[i], __builtin_amdgcn_readfirstlane(sneg[i]), nhs);
                    ev = make_double2(e.x, e.y);
                } else {
                    ev = phase_att64<DIM>(P, x, P.v + (size_t)v * DIM, P.ff_sign, P.is_ff_v);
                }
#pragma unroll
                for (int q = 0; q < kWaveLG; ++q) {
                    const double2 c = scl[i][q];
                    acc[q].x += ev.x * c.x - ev.y * c.y;
                    acc[q].y += ev.x * c.y + ev.y * c.x;
                }
            }
        };
        if (MIXED && batch_last)
            paths_of_batch(std::true_type{});
        else
            paths_of_batch(std::false_type{});
    }
    if (vok)
#pragma unroll
        for (int q = 0; q < kWaveLG; ++q)   // static indices keep acc in registers
            if (q < nlg) hpart[((size_t)split * P.nl + l0 + q) * P.nv + v] = acc[q];
}

// The same contraction on the matrix cores (mixed precision): per path,
// H[32 views, 32 lights] += E[views] (x) cl0[lights] as two real rank-2
// products with v_mfma_f32_32x32x2f32, K = (re, im):
//   Re H = [Er, -Ei] . [Cr; Ci],   Im H = [Ei, Er] . [Cr; Ci].
// Lane l of a wave computes E(x1_p, v) for view l % 32 (A row) and supplies
// K index l / 32; B is the staged cl0 row of the path.  A workgroup holds 128
// views (32 per wave) x 32 NB lights; the FP32 accumulators of each 32-path
// batch are added into FP64 ones, and the ranges' partials are reduced as
// before.
constexpr int kWmViews = 128, kWmPB = 32;
constexpr int kWmBlocks = 1;   // 32-light blocks per workgroup (1: 148 registers, three waves per SIMD; 2 measured 9 % slower)
typedef float wm_f32x16 __attribute__((ext_vector_type(16)));
template <int DIM, int NB>
__global__ __launch_bounds__(256) void field_wave_first_mfma_kernel(DevProblem P, int64_t np, int nsplit,
                                                                    const PathRec *paths, const double2 *cl0,
                                                                    double2 *hpart)
{
    __shared__ double sx[kWmPB][3];
    __shared__ float sq[kWmPB][3];
    __shared__ int sneg[kWmPB];
    constexpr int kWmLights = 32 * NB;
    __shared__ float2 scl[kWmPB][kWmLights];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 31, kk = lane >> 5;
    const int vbase = blockIdx.x * kWmViews + wave * 32;
    const int v = vbase + m;
    const bool vok = v < P.nv;
    double vd[3] = {0, 0, 0};
    for (int j = 0; j < DIM && vok; ++j) vd[j] = P.v[(size_t)v * DIM + j];
    const FarView fv = far_view<DIM>(P, vd);
    const float nhs = (float)(P.sigt * 0.5) * -1.4426950408889634f;
    const int l0 = blockIdx.y * kWmLights;
    const int split = blockIdx.z;
    const int64_t p0 = np * split / nsplit, p1 = np * (split + 1) / nsplit;
    wm_f32x16 cr[NB], ci[NB];
    double ar[NB][16], ai[NB][16];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            cr[b][e] = ci[b][e] = 0.f;
            ar[b][e] = ai[b][e] = 0.0;
        }
    for (int64_t pb = p0; pb < p1; pb += kWmPB) {
        const int cnt = (int)((p1 - pb) < kWmPB ? (p1 - pb) : kWmPB);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt * 3; t += blockDim.x) {
            const int i = t / 3, j = t % 3;
            const double xj = paths[pb + i].x1[j];
            sx[i][j] = xj;
            sq[i][j] = j < DIM ? near_face(xj, P.bmin[j], P.bmax[j]) : 0.f;
        }
        int mine = 1;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            float q[3];
            for (int j = 0; j < 3; ++j) q[j] = j < DIM ? near_face(paths[pb + i].x1[j], P.bmin[j], P.bmax[j]) : 0.f;
            sneg[i] = qneg_bits<DIM>(P, q);
            mine &= (sneg[i] & kLastAxisOnly) != 0;
        }
        for (int t = threadIdx.x; t < kWmPB * kWmLights; t += blockDim.x) {
            const int i = t / kWmLights, q = t % kWmLights;
            float2 c = make_float2(0.f, 0.f);
            if (i < cnt && l0 + q < P.nl) {
                const double2 d = cl0[(size_t)(pb + i) * P.nl + l0 + q];
                c = make_float2((float)d.x, (float)d.y);
            }
            scl[i][q] = c;
        }
        // the batch's first vertices leave every view through the last axis
        const bool batch_last = __syncthreads_and(mine) != 0;
        auto paths_of_batch = [&](auto last_tag) {
            constexpr bool LAST = decltype(last_tag)::value;
            for (int i = 0; i < cnt; ++i) {
                const double x[3] = {sx[i][0], sx[i][1], sx[i][2]};
                float2 ev = make_float2(0.f, 0.f);
                if (vok) ev = e_far_mixed<DIM, LAST>(x, fv, sq[i], __builtin_amdgcn_readfirstlane(sneg[i]), nhs);
                const float a1 = kk == 0 ? ev.x : -ev.y;
                const float a2 = kk == 0 ? ev.y : ev.x;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const float2 c = scl[i][b * 32 + m];
                    const float bv = kk == 0 ? c.x : c.y;
                    cr[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bv, cr[b], 0, 0, 0);
                    ci[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, bv, ci[b], 0, 0, 0);
                }
            }
        };
        if (batch_last)
            paths_of_batch(std::true_type{});
        else
            paths_of_batch(std::false_type{});
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                ar[b][e] += (double)cr[b][e];
                ai[b][e] += (double)ci[b][e];
                cr[b][e] = ci[b][e] = 0.f;
            }
    }
    // accumulator element e of lane l: view row 8 (e / 4) + 4 (l / 32) + e % 4,
    // light column l % 32
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int vv = vbase + 8 * (e >> 2) + 4 * kk + (e & 3);
            const int l = l0 + b * 32 + m;
            if (vv < P.nv && l < P.nl) hpart[((size_t)split * P.nl + l) * P.nv + vv] = make_double2(ar[b][e], ai[b][e]);
        }
}

__global__ void field_wave_first_reduce_kernel(int nv, int nl, int nsplit, const double2 *aftab, const double2 *hpart,
                                               double2 *acc)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double2 s = make_double2(0.0, 0.0);
    for (int l = 0; l < nl; ++l) {
        double2 h = make_double2(0.0, 0.0);
        for (int k = 0; k < nsplit; ++k) {
            const double2 t = hpart[((size_t)k * nl + l) * nv + v];
            h.x += t.x;
            h.y += t.y;
        }
        const double2 t = cmul(aftab[(size_t)v + (size_t)nv * l], h);
        s.x += t.x;
        s.y += t.y;
    }
    acc[v].x += s.x;
    acc[v].y += s.y;
}

__global__ void field_reduce_kernel(const double2 *part, int nsplit, int64_t n, double2 *acc)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 s = acc[i];
    for (int k = 0; k < nsplit; ++k) {
        double2 t = part[(size_t)k * n + i];
        s.x += t.x;
        s.y += t.y;
    }
    acc[i] = s;
}

template <int DIM>
static void field_dispatch(const DevProblem &P, const HetMedium *Mp, int64_t np, int64_t nvert, const PathRec *paths,
                           const VertexRec *verts, const int32_t *vpath, const int64_t *offsets, double2 *work,
                           double2 *acc, bool mixed, const double2 *wl, unsigned long long *nonfinite, hipStream_t s)
{
    // a voxel medium (MCsampleHetro) has no wave variant and marches in FP64
    const bool het = Mp != nullptr;
    HetMedium M{};
    if (het) M = *Mp;
    const bool wave = wl != nullptr && !het;
    const int ncol = wave ? 1 : P.nl;
    // work layout: coef[nvert*ncol] | coefR[np*ncol] | aftab[nv*nl] | cl0[np*nl] (wave) | part[64*nv*ncol]
    //              | hpart[kWaveMaxSplit*nl*nv] (wave)
    double2 *coef = work;
    double2 *coefR = coef + (size_t)nvert * ncol;
    double2 *aftab = coefR + (size_t)np * ncol;
    double2 *cl0 = aftab + (size_t)P.nv * P.nl;
    double2 *part = cl0 + (wave ? (size_t)np * P.nl : 0);
    // mixed precision, far-field lights, isotropic/HG phase: FP32 per-vertex terms
    const bool fastc = mixed && P.is_ff_l && P.amp.type != 2;
    if (wave && fastc)
        hipLaunchKernelGGL((field_wave_coef_kernel<DIM, true>), dim3((unsigned)np), dim3(64), 0, s, P, np, paths, verts,
                           offsets, wl, coef, coefR, cl0, nonfinite);
    else if (wave)
        hipLaunchKernelGGL((field_wave_coef_kernel<DIM, false>), dim3((unsigned)np), dim3(64), 0, s, P, np, paths,
                           verts, offsets, wl, coef, coefR, cl0, nonfinite);
    else if (het)
        hipLaunchKernelGGL((field_coef_kernel<DIM, true>), dim3((unsigned)np), dim3(64), 0, s, P, M, np, paths, verts,
                           offsets, coef, coefR, nonfinite);
    else if (fastc)
        hipLaunchKernelGGL((field_coef_kernel<DIM, false, true>), dim3((unsigned)np), dim3(64), 0, s, P, M, np, paths,
                           verts, offsets, coef, coefR, nonfinite);
    else
        hipLaunchKernelGGL((field_coef_kernel<DIM, false>), dim3((unsigned)np), dim3(64), 0, s, P, M, np, paths, verts,
                           offsets, coef, coefR, nonfinite);
    const bool near = !(P.is_ff_l && P.is_ff_v);
    if (!near) launch_af_table(P, aftab, s);
    const int64_t nent = nvert + (P.do_cbs ? np : 0);
    const int vt = (P.nv + kFieldViews - 1) / kFieldViews;
    // the mixed one-column kernel takes two views per lane: its workgroups
    // cover 512 views, so size the entry splits on that
    const int vpt = (mixed && ncol == 1 && !het && !near) ? 2 : 1;
    const int vtw = (P.nv + vpt * kFieldViews - 1) / (vpt * kFieldViews);
    int nsplit = (2048 + vtw - 1) / vtw;
    const int64_t maxsplit = (nent + kFieldBatch - 1) / kFieldBatch;
    if (nsplit > maxsplit) nsplit = (int)maxsplit;
    if (nsplit > 64) nsplit = 64;
    if (nsplit < 1) nsplit = 1;
    for (int l0 = 0; l0 < ncol; l0 += kFieldNL) {
        const int nlp = (ncol - l0) < kFieldNL ? (ncol - l0) : kFieldNL;
        dim3 g((unsigned)vt, (unsigned)nsplit);
#define SCMC_FIELD(KERNEL)                                                                                         \
    hipLaunchKernelGGL((KERNEL), g, dim3(kFieldViews), 0, s, P, M, np, nvert, nent, nsplit, l0, nlp, paths, verts, \
                       vpath, offsets, coef, coefR, aftab, cl0, part)
#define SCMC_FIELD_NEAR(KERNEL)                                                                                    \
    hipLaunchKernelGGL((KERNEL), g, dim3(kFieldViews), 0, s, P, M, np, nvert, nent, nsplit, l0, nlp, paths, verts, \
                       vpath, offsets, coef, coefR, cl0, part)
        if (het) {
            if (near)
                SCMC_FIELD_NEAR((field_near_kernel<DIM, false, true>));
            else if (mixed && ncol == 1)
                SCMC_FIELD((field_kernel<DIM, true, false, 1, true>));
            else if (mixed)
                SCMC_FIELD((field_kernel<DIM, true, false, kFieldNL, true>));
            else if (ncol == 1)
                SCMC_FIELD((field_kernel<DIM, false, false, 1, true>));
            else
                SCMC_FIELD((field_kernel<DIM, false, false, kFieldNL, true>));
        } else if (near) {
            if (wave)
                SCMC_FIELD_NEAR((field_near_kernel<DIM, true>));
            else
                SCMC_FIELD_NEAR((field_near_kernel<DIM, false>));
        } else if (mixed) {
            dim3 g2((unsigned)((P.nv + 2 * kFieldViews - 1) / (2 * kFieldViews)), (unsigned)nsplit);
            // types other than isotropic (1) and tabulated (2) are HG, as in eval_amp
            const bool hg_only = P.amp.type != 1 && P.amp.type != 2;
            if (wave && hg_only)
                hipLaunchKernelGGL((field_kernel<DIM, true, true, 1, false, 2, 3>), g2, dim3(kFieldViews), 0, s, P, M,
                                   np, nvert, nent, nsplit, l0, nlp, paths, verts, vpath, offsets, coef, coefR, aftab,
                                   cl0, part);
            else if (ncol == 1 && !wave && hg_only)
                hipLaunchKernelGGL((field_kernel<DIM, true, false, 1, false, 2, 3>), g2, dim3(kFieldViews), 0, s, P, M,
                                   np, nvert, nent, nsplit, l0, nlp, paths, verts, vpath, offsets, coef, coefR, aftab,
                                   cl0, part);
            else if (wave)
                hipLaunchKernelGGL((field_kernel<DIM, true, true, 1, false, 2>), g2, dim3(kFieldViews), 0, s, P, M, np,
                                   nvert, nent, nsplit, l0, nlp, paths, verts, vpath, offsets, coef, coefR, aftab, cl0,
                                   part);
            else if (ncol == 1)
                hipLaunchKernelGGL((field_kernel<DIM, true, false, 1, false, 2>), g2, dim3(kFieldViews), 0, s, P, M,
                                   np, nvert, nent, nsplit, l0, nlp, paths, verts, vpath, offsets, coef, coefR, aftab,
                                   cl0, part);
            else
                SCMC_FIELD((field_kernel<DIM, true, false, kFieldNL>));
        } else {
            if (wave)
                SCMC_FIELD((field_kernel<DIM, false, true, 1>));
            else if (ncol == 1)
                SCMC_FIELD((field_kernel<DIM, false, false, 1>));
            else
                SCMC_FIELD((field_kernel<DIM, false, false, kFieldNL>));
        }
#undef SCMC_FIELD
#undef SCMC_FIELD_NEAR
    }
    const int64_t n = (int64_t)P.nv * ncol;
    hipLaunchKernelGGL(field_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, nsplit, n, acc);
    if (wave && !near && np > 0) {
        double2 *hpart = part + (size_t)64 * P.nv * ncol;
        const int ng = (P.nl + kWaveLG - 1) / kWaveLG;
        int nsf = (2048 + vt * ng - 1) / (vt * ng);
        const int64_t maxsf = (np + kWavePB - 1) / kWavePB;
        if (nsf > maxsf) nsf = (int)maxsf;
        if (nsf > kWaveMaxSplit) nsf = kWaveMaxSplit;
        if (nsf < 1) nsf = 1;
        dim3 gf((unsigned)vt, (unsigned)ng, (unsigned)nsf);
        if (mixed) {
            // matrix cores: 128 views x 64 lights per workgroup
            const int vtm = (P.nv + kWmViews - 1) / kWmViews, ngm = (P.nl + 32 * kWmBlocks - 1) / (32 * kWmBlocks);
            int nsm = (2048 + vtm * ngm - 1) / (vtm * ngm);
            if (nsm > maxsf) nsm = (int)maxsf;
            if (nsm > kWaveMaxSplit) nsm = kWaveMaxSplit;
            if (nsm < 1) nsm = 1;
            nsf = nsm;
            hipLaunchKernelGGL((field_wave_first_mfma_kernel<DIM, kWmBlocks>), dim3((unsigned)vtm, (unsigned)ngm, (unsigned)nsm),
                               dim3(256), 0, s, P, np, nsm, paths, cl0, hpart);
        } else
            hipLaunchKernelGGL((field_wave_first_kernel<DIM, false>), gf, dim3(kFieldViews), 0, s, P, np, nsf, paths,
                               cl0, hpart);
        hipLaunchKernelGGL(field_wave_first_reduce_kernel, dim3((unsigned)((P.nv + 255) / 256)), dim3(256), 0, s, P.nv,
                           P.nl, nsf, aftab, hpart, acc);
    }
}

size_t field_work_elems(int64_t np, int64_t nvert, int nv, int nl, bool wave)
{
    const int ncol = wave ? 1 : nl;
    return (size_t)nvert * ncol + (size_t)np * ncol + (size_t)nv * nl + (wave ? (size_t)np * nl : 0) +
           (size_t)64 * nv * ncol + (wave ? (size_t)kWaveMaxSplit * nv * nl : 0);
}

void launch_field(const DevProblem &P, const HetMedium *M, int64_t np, int64_t nvert, const PathRec *paths,
                  const VertexRec *verts, const int32_t *vpath, const int64_t *offsets, double2 *work, double2 *acc,
                  bool mixed, const double2 *wl, unsigned long long *nonfinite, hipStream_t s)
{
    if (P.dim == 3)
        field_dispatch<3>(P, M, np, nvert, paths, verts, vpath, offsets, work, acc, mixed, wl, nonfinite, s);
    else
        field_dispatch<2>(P, M, np, nvert, paths, verts, vpath, offsets, work, acc, mixed, wl, nonfinite, s);
}

} // namespace scmc
