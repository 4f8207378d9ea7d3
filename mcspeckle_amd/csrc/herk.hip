// herk.hip -- complex Hermitian rank-K covariance update on the gfx950 matrix
// cores, with the Kronecker factors expanded on load.
//
// Computes C += sum_k u_k u_k^H for the covariance of code/MCcov.m:232-264
// (the four-term CBS update of :244-254 equals (f+r)(f+r)^H with
// u = A1 (x) B1 + A2 (x) B2).  Row index i of C is (v,l) with i = v + nv*l,
// the "matrix view" of Ms(v1,v2,l1,l2).
//
// Real formulation.  Interleave (re, im) along the reduction axis: with
// Z[i][2k] = Re u_k(i), Z[i][2k+1] = Im u_k(i) and the rotated rows
// Z'[i] = (Im, -Re) (i.e. -i*u):  Re C = Z Z^T and Im C = Z' Z^T.  One tile
// of C therefore needs two real GEMMs sharing the B operand.
//
// Mixed kernel: u is formed in FP32 from the factors, scaled by a power of
// two so that max|u| < 2^15, split as u = hi + lo with FP16 hi and lo, and
// accumulated as hi*hi + hi*lo + lo*hi with v_mfma_f32_32x32x16_f16 into FP32
// registers (~22-bit products, FP32 sums).  Each (tile, k-split) work item then
// adds its FP32 tile, unscaled, into its own FP64 accumulator copy, so the
// result is deterministic and the per-item FP32 sum length stays short.
//
// FP64 kernel: the same structure with v_mfma_f64_16x16x4f64 and no split.
//
// Only lower-triangle tiles (ti >= tj) are computed; the finaliser restores
// the upper triangle by Hermitian symmetry.
#include "kernels.h"

#include <initializer_list>

#include <algorithm>
#include <vector>

namespace scmc {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Work item -> (k-split s, tile rank t).  Workgroups are dispatched to the 8
// XCDs round robin (XCD = blockIdx % 8).  The splits run one after another
// across the whole chip, so the expanded rows being read at any time are one
// split's (about 225 MB at N = 4096, within the 256 MB MALL); inside a split
// each XCD walks its own contiguous run of ntile/8 tiles in the blocked order
// of `order` (an 8x8 block of tiles), so its rows are shared in its L2.
__device__ __forceinline__ void item_decode(int b, int ksplit, int ntile, int &s, int &t)
{
    (void)ksplit;
    if ((ntile & 7) == 0) {
        const int x = b & 7, g = b >> 3, per = ntile >> 3;
        s = g / per;
        t = x * per + g % per;
    } else {
        s = b / ntile;
        t = b % ntile;
    }
}

__device__ __forceinline__ float2 fcmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// (re, im) pairs -> (im, -re): the rows of -i*u.
__device__ __forceinline__ half8 rot_neg_i(half8 a)
{
    half8 r;
    r[0] = a[1]; r[1] = -a[0];
    r[2] = a[3]; r[3] = -a[2];
    r[4] = a[5]; r[5] = -a[4];
    r[6] = a[7]; r[7] = -a[6];
    return r;
}

// ------------------------------------------------------------ mixed
namespace mixed {
constexpr int TM = kTileMixed;   // 128x128 tile per workgroup, 4 waves of 64x64
constexpr int NBUF = 5;          // LDS stages (16 KB each, 80 KB: two workgroups per CU)
// LDS stage: [tile 2 (rows, columns)][plane 2 (hi, lo)][TM rows][32 B]
constexpr int STAGEB = 2 * 2 * TM * 32;
__device__ __forceinline__ int loff(int tile, int plane, int row, int q)
{
    return ((tile * 2 + plane) * TM + row) * 32 + q * 16;
}
// The expanded buffer is padded with zeros to whole tiles and to a multiple
// of four K-steps so the update kernel's loads are unconditional.
__host__ __device__ __forceinline__ int pad_rows(int n) { return (n + TM - 1) / TM * TM; }
__host__ __device__ __forceinline__ int64_t pad_ksteps(int64_t k) { return ((k + 7) / 8 + 3) / 4 * 4; }
} // namespace mixed

// Expansion: for K-step ks (vertices 8ks..8ks+7) and row i = v + nv*l, plane
// 0 holds hi and plane 1 lo, 32 bytes each (8 vertices x (re, im) FP16), at
// Z[ks][plane][i], scaled by the chunk's power of two.  A workgroup covers
// 128 rows x kExpandSteps K-steps; two threads per row, each forming four
// vertices, so every store instruction of a wave writes 1 KB contiguously.
constexpr int kExpandSteps = 16;
template <bool CBS>
__global__ __launch_bounds__(256) void expand_mixed_kernel(int nv, int nl, int n, int64_t k, const float2 *__restrict__ A1,
                                                           const float2 *__restrict__ B1, const float2 *__restrict__ A2,
                                                           const float2 *__restrict__ B2, const unsigned *__restrict__ bound,
                                                           uint4 *__restrict__ Z)
{
    const int npad = mixed::pad_rows(n);
    const int64_t nks = mixed::pad_ksteps(k);
    const int i = blockIdx.x * 128 + (threadIdx.x >> 1), hf = threadIdx.x & 1;
    if (i >= npad) return;
    const bool ok = i < n;
    const int v = ok ? i % nv : 0, l = ok ? i / nv : 0;
    float b = bound ? __uint_as_float(*bound) : 0.f;
    int ex = 0;
    if (b > 0.f) frexpf(b, &ex);
    const float scale = ldexpf(1.f, 15 - ex);
    const int64_t ks0 = (int64_t)blockIdx.y * kExpandSteps;
    const int64_t ks1 = ks0 + kExpandSteps < nks ? ks0 + kExpandSteps : nks;
    for (int64_t ks = ks0; ks < ks1; ++ks) {
        half8 h, lo;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t kk = ks * 8 + hf * 4 + q;
            float2 u = make_float2(0.f, 0.f);
            if (ok && kk < k) {
                u = fcmul(A1[kk * nv + v], B1[kk * nl + l]);
                if (CBS) {
                    float2 u2 = fcmul(A2[kk * nv + v], B2[kk * nl + l]);
                    u.x += u2.x;
                    u.y += u2.y;
                }
                u.x *= scale;
                u.y *= scale;
            }
            const _Float16 hr = (_Float16)u.x, hi_ = (_Float16)u.y;
            h[2 * q] = hr;
            h[2 * q + 1] = hi_;
            lo[2 * q] = (_Float16)(u.x - (float)hr);
            lo[2 * q + 1] = (_Float16)(u.y - (float)hi_);
        }
        uint4 *d = Z + ((size_t)ks * 2 * npad + i) * 2 + hf;
        d[0] = __builtin_bit_cast(uint4, h);
        d[(size_t)npad * 2] = __builtin_bit_cast(uint4, lo);
    }
}

// The same expansion, one K-step per workgroup over all rows: the eight
// vertices' factor vectors are staged in LDS once, then the K-step's block of
// both planes is written front to back.  For factor vectors that fit in
// kExpandStepLds bytes (launch_expand_mixed picks the kernel).
constexpr int kExpandStepLds = 64 * 1024;
template <bool CBS>
__global__ __launch_bounds__(256) void expand_step_kernel(int nv, int nl, int n, int64_t k, const float2 *__restrict__ A1,
                                                          const float2 *__restrict__ B1, const float2 *__restrict__ A2,
                                                          const float2 *__restrict__ B2, const unsigned *__restrict__ bound,
                                                          uint4 *__restrict__ Z)
{
    extern __shared__ float2 fs[];
    float2 *a1 = fs, *b1 = a1 + 8 * nv, *a2 = b1 + 8 * nl, *b2 = a2 + 8 * nv;
    const int npad = mixed::pad_rows(n);
    const int64_t ks = blockIdx.x;
    for (int j = threadIdx.x; j < 8 * nv; j += blockDim.x) {
        const int q = j / nv, v = j - q * nv;
        const int64_t kk = ks * 8 + q;
        a1[j] = kk < k ? A1[kk * nv + v] : make_float2(0.f, 0.f);
        if (CBS) a2[j] = kk < k ? A2[kk * nv + v] : make_float2(0.f, 0.f);
    }
    for (int j = threadIdx.x; j < 8 * nl; j += blockDim.x) {
        const int q = j / nl, l = j - q * nl;
        const int64_t kk = ks * 8 + q;
        b1[j] = kk < k ? B1[kk * nl + l] : make_float2(0.f, 0.f);
        if (CBS) b2[j] = kk < k ? B2[kk * nl + l] : make_float2(0.f, 0.f);
    }
    float b = bound ? __uint_as_float(*bound) : 0.f;
    int ex = 0;
    if (b > 0.f) frexpf(b, &ex);
    const float scale = ldexpf(1.f, 15 - ex);
    __syncthreads();
    const int hf = threadIdx.x & 1;
    uint4 *zk = Z + (size_t)ks * 4 * npad + hf;
    for (int i = threadIdx.x >> 1; i < npad; i += blockDim.x >> 1) {
        half8 h, lo;
        const bool ok = i < n;
        const int l = ok ? i / nv : 0, v = ok ? i - l * nv : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int qq = hf * 4 + q;
            float2 u = make_float2(0.f, 0.f);
            if (ok) {
                u = fcmul(a1[qq * nv + v], b1[qq * nl + l]);
                if (CBS) {
                    float2 u2 = fcmul(a2[qq * nv + v], b2[qq * nl + l]);
                    u.x += u2.x;
                    u.y += u2.y;
                }
                u.x *= scale;
                u.y *= scale;
            }
            const _Float16 hr = (_Float16)u.x, hi_ = (_Float16)u.y;
            h[2 * q] = hr;
            h[2 * q + 1] = hi_;
            lo[2 * q] = (_Float16)(u.x - (float)hr);
            lo[2 * q + 1] = (_Float16)(u.y - (float)hi_);
        }
        zk[(size_t)i * 2] = __builtin_bit_cast(uint4, h);
        zk[((size_t)npad + i) * 2] = __builtin_bit_cast(uint4, lo);
    }
}

// The Hermitian update over pre-expanded rows: a streaming GEMM.  The rows
// of each K-step go from HBM straight into LDS (global_load_lds_dwordx4, no
// register staging), NBUF - 1 K-steps ahead of the matrix cores; each wave
// reads its 64-row A and 64-column B fragments one K-step ahead and issues
// 24 MFMAs per K-step (Re and Im, hi*hi + hi*lo + lo*hi).  Each K-step costs
// one workgroup barrier.
//
// One launch carries several rank-K problems: the two of a chunk (multiple
// and single scattering, whose short items come last and fill the workgroup
// slots the long ones leave free at the end of the grid), or, in serving,
// the chunks of several requests of different N.  Work items are numbered
// problem after problem.
struct MixedUpdate {
    int64_t nks;               // K-steps of the expanded rows (pad_ksteps)
    const uint4 *Z;            // expanded rows
    double2 *acc;              // [ksplit][n][n] accumulator copies
    const unsigned *bound;     // float bits of max |u| (scale), may be null
    const int2 *order;         // lower-triangle tiles in blocked order
    int n, ntile, ksplit;
    int nitems;                // ntile * ksplit
};
template <int MAXG>
struct MixedGroup {
    MixedUpdate u[MAXG];
    int count;
};

// MAXG = 2 for a chunk's pair of problems, kMaxUpdateGroup for shared
// launches (a smaller argument block and select chain for the common case)
// One work item (k-split s, tile t) of a mixed update, on a workgroup of 256
// threads with NBUF * STAGEB bytes of LDS at lds.  The update kernel runs one
// item per workgroup; the resident loop kernel runs items one after another.
__device__ __forceinline__ void mixed_item(const uint4 *__restrict__ Z, double2 *__restrict__ acc,
                                           const unsigned *__restrict__ bound, const int2 *__restrict__ order, int n,
                                           int ntile, int ksplit, int64_t nks_total, int item, unsigned char *lds)
{
    using namespace mixed;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int s, t;
    item_decode(item, ksplit, ntile, s, t);
    const int ti = order[t].x, tj = order[t].y;
    const int npad = pad_rows(n);
    const int64_t npair = nks_total / 2;   // K-steps come in pairs (padded to 4)
    const int64_t sb = 2 * (npair * s / ksplit), se = 2 * (npair * (s + 1) / ksplit);
    const int nstage = (int)(se - sb);

    float b = bound ? __uint_as_float(*bound) : 0.f;
    int ex = 0;
    if (b > 0.f) frexpf(b, &ex);
    const double unscale = ldexp(1.0, -2 * (15 - ex));

    f32x16 accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                accR[a][c][e] = 0.f;
                accI[a][c][e] = 0.f;
            }

    // Loading role: wave w moves tile T = w/2 (rows of ti, then columns of
    // tj; a diagonal tile loads the same rows twice), plane P = w%2, four
    // chunks of 32 rows x 32 B per K-step.  Lane l covers row l/2, 16 B
    // half l%2: each chunk is 1 KB contiguous in the plane and lands at
    // M0 + 16 l, which is exactly the [row][32 B] layout of the stage.
    const int T = wave >> 1, P = wave & 1;
    const uint4 *src = Z + ((size_t)P * npad + (size_t)((T ? tj : ti) * TM) + (lane >> 1)) * 2 + (lane & 1);
    const size_t kstride = (size_t)npad * 4;   // uint4 per K-step (two planes)
    const int ldst = loff(T, P, 0, 0);
    typedef __attribute__((address_space(3))) void *lds_ptr;
#define SCMC_ISSUE(kstep, buf)                                                                           \
    do {                                                                                                 \
        const uint4 *g = src + (size_t)(kstep) * kstride;                                                \
        unsigned char *d = lds + (buf) * STAGEB + ldst;                                                  \
        __builtin_amdgcn_global_load_lds((const void *)g, (lds_ptr)d, 16, 0, 0);                         \
        __builtin_amdgcn_global_load_lds((const void *)(g + 64), (lds_ptr)(d + 1024), 16, 0, 0);        \
        __builtin_amdgcn_global_load_lds((const void *)(g + 128), (lds_ptr)(d + 2048), 16, 0, 0);       \
        __builtin_amdgcn_global_load_lds((const void *)(g + 192), (lds_ptr)(d + 3072), 16, 0, 0);       \
    } while (0)

    const int wr = wave >> 1, wc = wave & 1;
    const int half = lane >> 5, lrow = lane & 31;
    const int oa0 = loff(0, 0, wr * 64 + lrow, half), oa1 = oa0 + 32 * 32;   // A rows, blocks 0/1
    const int ob0 = loff(1, 0, wc * 64 + lrow, half), ob1 = ob0 + 32 * 32;   // B columns
    constexpr int PL = TM * 32;                                               // hi -> lo plane
    // Fragments of one K-step: A hi/lo of the two 32-row blocks, B hi/lo of
    // the two 32-column blocks.  Two sets (F, G) alternate so the next
    // K-step's fragments are read while the current one's MFMAs run.
    half8 Fah0, Fah1, Fal0, Fal1, Fbh0, Fbh1, Fbl0, Fbl1;
    half8 Gah0, Gah1, Gal0, Gal1, Gbh0, Gbh1, Gbl0, Gbl1;
#define SCMC_READ(X, bufi)                                                  \
    do {                                                                    \
        const unsigned char *bp = lds + (bufi) * STAGEB;                    \
        X##ah0 = *(const half8 *)(bp + oa0);                                \
        X##ah1 = *(const half8 *)(bp + oa1);                                \
        X##bh0 = *(const half8 *)(bp + ob0);                                \
        X##bh1 = *(const half8 *)(bp + ob1);                                \
        X##al0 = *(const half8 *)(bp + oa0 + PL);                           \
        X##al1 = *(const half8 *)(bp + oa1 + PL);                           \
        X##bl0 = *(const half8 *)(bp + ob0 + PL);                           \
        X##bl1 = *(const half8 *)(bp + ob1 + PL);                           \
    } while (0)
#define SCMC_MFMA_ROW(X, rb)                                                                               \
    do {                                                                                                   \
        const half8 rh = rot_neg_i(X##ah##rb), rl = rot_neg_i(X##al##rb);                                  \
        accR[rb][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(X##ah##rb, X##bh0, accR[rb][0], 0, 0, 0);      \
        accI[rb][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(rh, X##bh0, accI[rb][0], 0, 0, 0);            \
        accR[rb][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(X##ah##rb, X##bh1, accR[rb][1], 0, 0, 0);      \
        accI[rb][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(rh, X##bh1, accI[rb][1], 0, 0, 0);            \
        accR[rb][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(X##ah##rb, X##bl0, accR[rb][0], 0, 0, 0);      \
        accI[rb][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(rh, X##bl0, accI[rb][0], 0, 0, 0);            \
        accR[rb][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(X##ah##rb, X##bl1, accR[rb][1], 0, 0, 0);      \
        accI[rb][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(rh, X##bl1, accI[rb][1], 0, 0, 0);            \
        accR[rb][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(X##al##rb, X##bh0, accR[rb][0], 0, 0, 0);      \
        accI[rb][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(rl, X##bh0, accI[rb][0], 0, 0, 0);            \
        accR[rb][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(X##al##rb, X##bh1, accR[rb][1], 0, 0, 0);      \
        accI[rb][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(rl, X##bh1, accI[rb][1], 0, 0, 0);            \
    } while (0)
    // s_barrier orders execution only: the compiler may move LDS reads across
    // it.  The empty asm with a memory clobber (no instruction) keeps every
    // LDS read of a stage after the barrier that publishes that stage (its
    // global_load_lds writes come from all four waves).  No s_waitcnt is
    // added, so the loads of later stages stay in flight.
#define SCMC_COMPILER_FENCE() asm volatile("" ::: "memory")
    // nstage is even.  Every load stays inside the item: stages past the
    // last reload the last one into a buffer that is never read again.
    const int last = nstage - 1;
    if (nstage > 0) {
        SCMC_ISSUE(sb, 0);
        SCMC_ISSUE(sb + (1 < last ? 1 : last), 1);
        SCMC_ISSUE(sb + (2 < last ? 2 : last), 2);
        SCMC_ISSUE(sb + (3 < last ? 3 : last), 3);
        __builtin_amdgcn_s_waitcnt(0xF7C);   // vmcnt(12): stage 0
        __builtin_amdgcn_s_barrier();
        SCMC_COMPILER_FENCE();
        SCMC_READ(F, 0);
    }
    // K-step st with fragments X: publish stage st + 1 (its loads were issued
    // three steps ago), issue stage st + 4 into the buffer of stage st - 1
    // (its fragments were read during step st - 2 and consumed by step
    // st - 1, which every wave finished before this barrier), run the MFMAs
    // of stage st and read stage st + 1's fragments into Y between them.
#define SCMC_STEP(X, Y, st)                                                                              \
    do {                                                                                                 \
        __builtin_amdgcn_s_waitcnt(0xF78); /* vmcnt(8): stage st+1 landed */                            \
        __builtin_amdgcn_s_barrier();                                                                    \
        SCMC_COMPILER_FENCE();                                                                           \
        SCMC_ISSUE(sb + ((st) + 4 < last ? (st) + 4 : last), ((st) + 4) % NBUF);                         \
        SCMC_MFMA_ROW(X, 0);                                                                             \
        SCMC_READ(Y, ((st) + 1) % NBUF);                                                                 \
        SCMC_MFMA_ROW(X, 1);                                                                             \
    } while (0)
    for (int st = 0; st < nstage; st += 2) {
        SCMC_STEP(F, G, st);
        SCMC_STEP(G, F, st + 1);
    }
#undef SCMC_STEP
#undef SCMC_COMPILER_FENCE
#undef SCMC_MFMA_ROW
#undef SCMC_READ
#undef SCMC_ISSUE
    // the clamped loads past the last stage may still be writing LDS: drain
    // them before the workgroup can end and its LDS be handed to another
    __builtin_amdgcn_s_waitcnt(0xF70);   // vmcnt(0)

    // epilogue: add the unscaled FP32 tile into accumulator copy s
    double2 *C = acc + (size_t)s * n * n;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int i = ti * TM + wr * 64 + rb * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
                const int j = tj * TM + wc * 64 + cb * 32 + lrow;
                if (i < n && j < n) {
                    double2 *p = &C[(size_t)i * n + j];
                    double2 o = *p;
                    o.x += (double)accR[rb][cb][e] * unscale;
                    o.y += (double)accI[rb][cb][e] * unscale;
                    *p = o;
                }
            }
}

template <int MAXG>
__global__ __launch_bounds__(256, 2) void herk_mixed_kernel(MixedGroup<MAXG> G)
{
    using namespace mixed;
    __shared__ __attribute__((aligned(16))) unsigned char lds[NBUF * STAGEB];
    // the problem of this work item (uniform); fields are picked with
    // selects so the argument block is never indexed dynamically
    int item = (int)blockIdx.x, g = 0;
#pragma unroll
    for (int q = 0; q + 1 < MAXG; ++q)
        if (g == q && q + 1 < G.count && item >= G.u[q].nitems) {
            item -= G.u[q].nitems;
            g = q + 1;
        }
    MixedUpdate U = G.u[0];
#pragma unroll
    for (int q = 1; q < MAXG; ++q)
        if (g == q) U = G.u[q];
    // a chunk's pair shares N and the tile order (launch_herk_mixed_update)
    const int n = MAXG == 2 ? G.u[0].n : U.n, ntile = MAXG == 2 ? G.u[0].ntile : U.ntile;
    const int2 *order = MAXG == 2 ? G.u[0].order : U.order;
    mixed_item(U.Z, U.acc, U.bound, order, n, ntile, U.ksplit, U.nks, item, lds);
}

// ------------------------------------------------------- resident loop
// Serving: a persistent update kernel that takes the work items of the jobs
// the host publishes (one mixed update each: one problem of one chunk of one
// request), so the updates of every request in flight share one resident
// launch instead of one launch each.
//
// Items are dealt to the eight XCDs as the update kernel's grid deals them:
// item b of a job goes to XCD label b % 8 (a job's item count is padded to a
// multiple of 8 with empty items), so item_decode gives each XCD the same
// contiguous run of tiles per k-split as in a launch of its own.  The
// workgroups of label x = blockIdx.x % 8 take items 8c + x from a per-job,
// per-label counter, and a per-label cursor moves past jobs whose share is
// taken.  A job's descriptor is read from host memory (system-scope loads,
// uncached).
//
// Memory: the job's rows, accumulators and bound were written by other
// kernels, some on other XCDs, so every wave starts an item with an
// agent-scope acquire (L1/L2 invalidate) and ends it with an agent-scope
// release (L2 write-back) before the item is counted; the last item of a job
// sets done[j] with a system-scope release.
//
// Every wait is bounded.  A workgroup that finds no published work for
// idle_ticks of the wall clock leaves, after announcing it: store
// closing[inst & 1]; fence; load published.  The host publishes: store
// published; fence; load closing[inst & 1].  So either this workgroup sees the
// new job, or the host sees the announcement and queues another instance
// behind this one on the loop's stream (the next instance continues from the
// same counters).  `stop` ends idle workgroups at once.
//
// Helper instances (inst < 0) take items only of the jobs help_lo..help_hi
// (a chunk's two updates), leave when those are all taken and never announce
// it: launched on a request's own stream when it publishes a chunk, they add
// workgroups wherever the hardware finds room, as a launch of the update
// would, while the resident instance keeps one workgroup per CU.
__device__ __forceinline__ unsigned sys_peek(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ LoopJob read_job(const LoopJob *j)
{
    const uint64_t *w = (const uint64_t *)j;
    uint64_t v[sizeof(LoopJob) / 8];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(LoopJob) / 8); ++i)
        v[i] = __hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return __builtin_bit_cast(LoopJob, v);
}

__global__ __launch_bounds__(256, 2) void herk_mixed_loop_kernel(LoopCtl *H, LoopDev *D, int inst, uint64_t idle_ticks,
                                                                 unsigned help_lo, unsigned help_hi)
{
    using namespace mixed;
    __shared__ __attribute__((aligned(16))) unsigned char lds[NBUF * STAGEB];
    // the claimed (job, item) and the job's descriptor go to the other
    // threads through the first bytes of the last stage, which an item writes
    // only after its first barrier (two workgroups per CU leave no LDS for
    // anything beside the stages); a barrier after reading them keeps them
    // from being overwritten early
    volatile int *claim = (volatile int *)(lds + (NBUF - 1) * STAGEB);
    volatile uint64_t *jw = (volatile uint64_t *)(lds + (NBUF - 1) * STAGEB + 16);
    constexpr int kJobWords = (int)(sizeof(LoopJob) / 8);
    const unsigned x = blockIdx.x & 7;
    // thread 0: this workgroup's cursor (a helper walks only its own jobs)
    unsigned p = help_lo;
    if (threadIdx.x == 0 && inst >= 0) p = __hip_atomic_load(&D->active[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        // every thread is done with the previous item and its claim
        __syncthreads();
        if (threadIdx.x == 0) {
            int job = -1, item = 0;
            uint64_t t0 = wall_clock64();
            for (;;) {
                if (inst < 0 && p > help_hi) break;
                if (p < sys_peek(&H->published)) {
                    const int n8 = (int)__hip_atomic_load(&H->jobs[p].nitems8, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_SYSTEM);
                    const unsigned c = atomicAdd(&D->next[(size_t)p * 8 + x], 1u);
                    if ((int64_t)c * 8 + x < n8) {
                        job = (int)p;
                        item = (int)(c * 8 + x);
                        break;
                    }
                    // this label's share of job p is taken
                    ++p;
                    if (inst >= 0) atomicMax(&D->active[x], p);
                    t0 = wall_clock64();
                    continue;
                }
                // a helper's jobs are published before it is launched
                if (inst < 0 || sys_peek(&H->stop)) break;
                if (wall_clock64() - t0 > idle_ticks) {
                    __hip_atomic_store(&H->closing[inst & 1], 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_SYSTEM);
                    __atomic_thread_fence(__ATOMIC_SEQ_CST);
                    if (p < __hip_atomic_load(&H->published, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_SYSTEM)) {
                        t0 = wall_clock64();
                        continue;
                    }
                    break;
                }
                // about 20 us between looks at host memory
                for (int q = 0; q < 6; ++q) __builtin_amdgcn_s_sleep(127);
            }
            claim[0] = job;
            claim[1] = item;
            if (job >= 0) {
                const LoopJob J = read_job(&H->jobs[job]);
                const uint64_t *w = (const uint64_t *)&J;
#pragma unroll
                for (int i = 0; i < kJobWords; ++i) jw[i] = w[i];
            }
        }
        __syncthreads();
        const int job = claim[0], item = claim[1];
        if (job < 0) break;
        uint64_t w[kJobWords];
#pragma unroll
        for (int i = 0; i < kJobWords; ++i) w[i] = jw[i];
        const LoopJob J = __builtin_bit_cast(LoopJob, w);
        __syncthreads();
        if (item < J.nitems) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            mixed_item((const uint4 *)J.Z, J.acc, J.bound, (const int2 *)J.order, J.n, J.ntile, J.ksplit, J.nks, item,
                       lds);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned f = atomicAdd(&D->finished[job], 1u);
            if ((int)f == J.nitems8 - 1)
                __hip_atomic_store(&H->done[job], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// ------------------------------------------------------------- fp64
// The FP64 update in two steps, as the mixed one: the rows u_i = a_v b_l
// (+ a2_v b2_l with CBS) of a chunk are expanded once into HBM, then the
// Hermitian update streams them through LDS into v_mfma_f64_16x16x4f64.
// Expanded layout: K-step t (vertices 2t, 2t+1), row i: 32 bytes
// (Re u_i(2t), Im u_i(2t), Re u_i(2t+1), Im u_i(2t+1)); rows padded with
// zeros to whole 64-row tiles, K-steps to whole stages of 8 vertices.
namespace f64 {
constexpr int TM = kTileF64;    // 64x64 tile per workgroup, 4 waves of 32x32
constexpr int SKS = 4;          // K-steps per stage
constexpr int VPS = 2 * SKS;    // 8 vertices per stage
// LDS stage: [tile 2 (rows, columns)][K-step SKS][TM rows][32 B]
constexpr int STAGEB = 2 * SKS * TM * 32;
// Two LDS stages of 16 KB: a stage's loads are issued as the MFMAs of the
// stage before start and have those MFMAs' time to land.  At 32 KB of LDS
// five workgroups fit on a CU, and with at least four waves per SIMD
// required the compiler keeps the accumulators in VGPRs (91 in all), so
// five waves per SIMD take turns on the matrix core.  Measured at N = 4096
// (20 000 paths): 122 TFLOP/s algorithmic, 83 % MFMA-busy; three stages
// (48 KB, three waves per SIMD) 101, four 86.
constexpr int NBUF = 2;
__host__ __device__ __forceinline__ int pad_rows(int n) { return (n + TM - 1) / TM * TM; }
__host__ __device__ __forceinline__ int64_t nstages(int64_t k) { return (k + VPS - 1) / VPS; }
} // namespace f64

// One workgroup per stage and 256 rows: each thread forms its row's eight
// products (the factors of a stage are a few KB and stay in L2) and writes
// them as four 32-byte K-step entries; consecutive rows are consecutive in
// every K-step, so the stores of a wave are contiguous.
template <bool CBS>
__global__ __launch_bounds__(256) void expand_fp64_kernel(int nv, int nl, int n, int64_t k,
                                                          const double2 *__restrict__ A1,
                                                          const double2 *__restrict__ B1,
                                                          const double2 *__restrict__ A2,
                                                          const double2 *__restrict__ B2, double2 *__restrict__ Z)
{
    using namespace f64;
    const int npad = pad_rows(n);
    const int i = blockIdx.y * 256 + threadIdx.x;
    if (i >= npad) return;
    const int64_t st = blockIdx.x;
    const bool row = i < n;
    const int v = row ? i % nv : 0, l = row ? i / nv : 0;
#pragma unroll
    for (int t = 0; t < SKS; ++t) {
        double2 u[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t kk = st * VPS + 2 * t + h;
            u[h] = make_double2(0.0, 0.0);
            if (row && kk < k) {
                u[h] = cmul(A1[kk * nv + v], B1[kk * nl + l]);
                if (CBS) {
                    const double2 w = cmul(A2[kk * nv + v], B2[kk * nl + l]);
                    u[h].x += w.x;
                    u[h].y += w.y;
                }
            }
        }
        double2 *z = Z + ((st * SKS + t) * (int64_t)npad + i) * 2;
        z[0] = u[0];
        z[1] = u[1];
    }
}

// The same with one workgroup per stage for all rows: the stage's factors
// (8 vertices x (Nv + Nl), both orders with CBS) are staged in LDS once, so
// the rows' products read LDS instead of L2.  For factor sets of up to 64 KB.
constexpr int kExpandFp64Lds = 64 * 1024;
template <bool CBS>
__global__ __launch_bounds__(256) void expand_fp64_stage_kernel(int nv, int nl, int n, int64_t k,
                                                                const double2 *__restrict__ A1,
                                                                const double2 *__restrict__ B1,
                                                                const double2 *__restrict__ A2,
                                                                const double2 *__restrict__ B2,
                                                                double2 *__restrict__ Z)
{
    using namespace f64;
    extern __shared__ double2 fd[];
    double2 *a1 = fd, *b1 = a1 + VPS * nv, *a2 = b1 + VPS * nl, *b2 = a2 + VPS * nv;
    const int npad = pad_rows(n);
    const int64_t st = blockIdx.x;
    for (int j = threadIdx.x; j < VPS * nv; j += blockDim.x) {
        const int q = j / nv, v = j - q * nv;
        const int64_t kk = st * VPS + q;
        a1[j] = kk < k ? A1[kk * nv + v] : make_double2(0.0, 0.0);
        if (CBS) a2[j] = kk < k ? A2[kk * nv + v] : make_double2(0.0, 0.0);
    }
    for (int j = threadIdx.x; j < VPS * nl; j += blockDim.x) {
        const int q = j / nl, l = j - q * nl;
        const int64_t kk = st * VPS + q;
        b1[j] = kk < k ? B1[kk * nl + l] : make_double2(0.0, 0.0);
        if (CBS) b2[j] = kk < k ? B2[kk * nl + l] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npad; i += blockDim.x) {
        const bool row = i < n;
        const int l = row ? i / nv : 0, v = row ? i - l * nv : 0;
#pragma unroll
        for (int t = 0; t < SKS; ++t) {
            double2 u[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = 2 * t + h;
                u[h] = make_double2(0.0, 0.0);
                if (row) {
                    u[h] = cmul(a1[q * nv + v], b1[q * nl + l]);
                    if (CBS) {
                        const double2 w = cmul(a2[q * nv + v], b2[q * nl + l]);
                        u[h].x += w.x;
                        u[h].y += w.y;
                    }
                }
            }
            double2 *z = Z + ((st * SKS + t) * (int64_t)npad + i) * 2;
            z[0] = u[0];
            z[1] = u[1];
        }
    }
}

// The Hermitian update over the expanded rows.  Work item = (k-split s,
// lower-triangle tile); s takes a contiguous run of stages.  Items are laid
// out over the XCDs as the mixed update's (item_decode over the blocked tile
// order), so the workgroups on one XCD share row panels in its L2.  Each stage's
// rows go from HBM straight into LDS (global_load_lds_dwordx4: wave w moves
// K-steps 2(w%2), 2(w%2)+1 of tile w/2 (rows, then columns), two 1 KB
// chunks of 32 rows per K-step, lane order = LDS layout), NBUF - 1 stages
// ahead of the matrix cores, with one workgroup barrier per stage.  Each
// wave computes a 32x32 block: Re C += Z Z^T and Im C += Z' Z^T with
// Z' = (Im, -Re) formed from the partner K element.
// One launch carries a chunk's two problems (multiple scattering, then the
// short single-scattering items, which fill the workgroup slots the long
// ones leave free at the end of the grid), each with its own rows,
// accumulators and k-splits.
struct F64Update {
    const double *Z;       // expanded rows (fp64_zbuf_bytes)
    double2 *acc;          // [ksplit][n][n] accumulator copies
    const int2 *order;     // lower-triangle tiles in blocked order
    int64_t nst;           // stages of the expanded rows
    int n, ntile, ksplit;
    int nitems;            // ntile * ksplit
};
struct F64Group {
    F64Update u[2];
    int count;
};

template <int NBUF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void herk_fp64_kernel(F64Group G)
{
    using namespace f64;
    __shared__ __attribute__((aligned(16))) unsigned char lds[NBUF * STAGEB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int item = blockIdx.x;
    const bool second = G.count > 1 && item >= G.u[0].nitems;
    if (second) item -= G.u[0].nitems;
    const F64Update U = second ? G.u[1] : G.u[0];
    const double *__restrict__ Z = U.Z;
    double2 *__restrict__ acc = U.acc;
    const int n = U.n, ksplit = U.ksplit;
    const int64_t nst = U.nst;
    int s, t;
    item_decode(item, ksplit, U.ntile, s, t);
    const int ti = U.order[t].x, tj = U.order[t].y;
    const int64_t sb = nst * s / ksplit, se = nst * (s + 1) / ksplit;
    const int nstage = (int)(se - sb);
    const int npad = pad_rows(n);

    f64x4 accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                accR[a][c][e] = 0.0;
                accI[a][c][e] = 0.0;
            }

    // loading role
    const int T = wave >> 1, KP = (wave & 1) * 2;
    const double *src = Z + ((size_t)KP * npad + (size_t)((T ? tj : ti) * TM) + (lane >> 1)) * 4 + (lane & 1) * 2;
    const size_t kstride = (size_t)npad * 4;   // doubles per K-step
    unsigned char *ldst = lds + T * (SKS * TM * 32) + KP * (TM * 32);
    typedef __attribute__((address_space(3))) void *lds_ptr;
#define SCMC_ISSUE64(stage, buf)                                                                              \
    do {                                                                                                      \
        const double *g = src + (size_t)(stage) * SKS * kstride;                                              \
        unsigned char *d = ldst + (buf) * STAGEB;                                                             \
        __builtin_amdgcn_global_load_lds((const void *)g, (lds_ptr)d, 16, 0, 0);                              \
        __builtin_amdgcn_global_load_lds((const void *)(g + 32 * 4), (lds_ptr)(d + 1024), 16, 0, 0);          \
        __builtin_amdgcn_global_load_lds((const void *)(g + kstride), (lds_ptr)(d + TM * 32), 16, 0, 0);      \
        __builtin_amdgcn_global_load_lds((const void *)(g + kstride + 32 * 4), (lds_ptr)(d + TM * 32 + 1024), \
                                         16, 0, 0);                                                           \
    } while (0)
    // s_barrier orders execution only: the empty asm statements with a
    // memory clobber on both sides keep every LDS read of a stage between the
    // barriers that publish and retire it, without waiting for the loads of
    // later stages (no s_waitcnt is added).
#define SCMC_BARRIER64()                  \
    do {                                  \
        asm volatile("" ::: "memory");    \
        __builtin_amdgcn_s_barrier();     \
        asm volatile("" ::: "memory");    \
    } while (0)
    // vmcnt(4 (NBUF - 2)): every stage but the newest NBUF - 2 has landed
    constexpr int kWait = 0xF70 | (4 * (NBUF - 2));
    static_assert(4 * (NBUF - 2) <= 15, "vmcnt field");

    const int wr = wave >> 1, wc = wave & 1;
    const int kk = lane >> 4, l16 = lane & 15;
    const int last = nstage - 1;
    if (nstage > 0) {
        // every load stays inside the item: stages past the last reload the
        // last one into a buffer that is never read again
#pragma unroll
        for (int p = 0; p < NBUF - 1; ++p) SCMC_ISSUE64(sb + (p < last ? p : last), p);
        __builtin_amdgcn_s_waitcnt(kWait);
        SCMC_BARRIER64();
    }
    for (int st = 0; st < nstage; ++st) {
        const int nx = st + NBUF - 1;
        // into the buffer of stage st - 1, which every wave finished before
        // the last barrier
        SCMC_ISSUE64(sb + (nx < last ? nx : last), nx % NBUF);
        const double *bp = (const double *)(lds + (st % NBUF) * STAGEB);
#pragma unroll
        for (int ks = 0; ks < SKS; ++ks) {
            double a[2], ar[2], bb[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                const double *ra = bp + ((size_t)ks * TM + wr * 32 + rb * 16 + l16) * 4;
                a[rb] = ra[kk];
                const double partner = ra[kk ^ 1];
                ar[rb] = (kk & 1) ? -partner : partner;
            }
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) bb[cb] = bp[((size_t)(SKS + ks) * TM + wc * 32 + cb * 16 + l16) * 4 + kk];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) {
                    accR[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], bb[cb], accR[rb][cb], 0, 0, 0);
                    accI[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[rb], bb[cb], accI[rb][cb], 0, 0, 0);
                }
        }
        __builtin_amdgcn_s_waitcnt(kWait);   // stage st + 1 landed
        SCMC_BARRIER64();
    }
#undef SCMC_BARRIER64
#undef SCMC_ISSUE64
    // the clamped loads past the last stage may still be writing LDS: drain
    // them before the workgroup can end and its LDS be handed to another
    __builtin_amdgcn_s_waitcnt(0xF70);   // vmcnt(0)

    double2 *C = acc + (size_t)s * n * n;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                // v_mfma_f64_16x16x4f64 result: register e of lane l holds row 4e + l/16, column l%16
                const int i = ti * TM + wr * 32 + rb * 16 + 4 * e + kk;
                const int j = tj * TM + wc * 32 + cb * 16 + l16;
                if (i < n && j < n) {
                    double2 *p = &C[(size_t)i * n + j];
                    double2 o = *p;
                    o.x += accR[rb][cb][e];
                    o.y += accI[rb][cb][e];
                    *p = o;
                }
            }
}

static int ntiles(int n, int tm) { return (n + tm - 1) / tm; }

void launch_expand_mixed(const HerkProblem &hp, hipStream_t s)
{
    if (hp.k <= 0) return;
    const int64_t nks = mixed::pad_ksteps(hp.k);
    const size_t lds = sizeof(float2) * 8 * (size_t)(hp.nv + hp.nl) * (hp.A2 ? 2 : 1);
    if (lds <= (size_t)kExpandStepLds) {
        if (hp.A2)
            hipLaunchKernelGGL(expand_step_kernel<true>, dim3((unsigned)nks), dim3(256), lds, s, hp.nv, hp.nl, hp.n,
                               hp.k, (const float2 *)hp.A1, (const float2 *)hp.B1, (const float2 *)hp.A2,
                               (const float2 *)hp.B2, hp.bound, (uint4 *)hp.zbuf);
        else
            hipLaunchKernelGGL(expand_step_kernel<false>, dim3((unsigned)nks), dim3(256), lds, s, hp.nv, hp.nl, hp.n,
                               hp.k, (const float2 *)hp.A1, (const float2 *)hp.B1, nullptr, nullptr, hp.bound,
                               (uint4 *)hp.zbuf);
        return;
    }
    dim3 ge((unsigned)(mixed::pad_rows(hp.n) / 128), (unsigned)((nks + kExpandSteps - 1) / kExpandSteps));
    if (hp.A2)
        hipLaunchKernelGGL(expand_mixed_kernel<true>, ge, dim3(256), 0, s, hp.nv, hp.nl, hp.n, hp.k,
                           (const float2 *)hp.A1, (const float2 *)hp.B1, (const float2 *)hp.A2,
                           (const float2 *)hp.B2, hp.bound, (uint4 *)hp.zbuf);
    else
        hipLaunchKernelGGL(expand_mixed_kernel<false>, ge, dim3(256), 0, s, hp.nv, hp.nl, hp.n, hp.k,
                           (const float2 *)hp.A1, (const float2 *)hp.B1, nullptr, nullptr, hp.bound,
                           (uint4 *)hp.zbuf);
}

static MixedUpdate mixed_desc(const HerkProblem &p)
{
    const int nt = ntiles(p.n, kTileMixed);
    MixedUpdate u{};
    u.nks = mixed::pad_ksteps(p.k);
    u.Z = (const uint4 *)p.zbuf;
    u.acc = p.acc;
    u.bound = p.bound;
    u.order = (const int2 *)p.order;
    u.n = p.n;
    u.ntile = nt * (nt + 1) / 2;
    u.ksplit = p.ksplit;
    u.nitems = u.ntile * p.ksplit;
    return u;
}

template <int MAXG>
static void launch_group(const HerkProblem *const *hp, int count, hipStream_t s)
{
    MixedGroup<MAXG> G{};
    int64_t grid = 0;
    for (int i = 0; i < count; ++i) {
        const HerkProblem *p = hp[i];
        if (!p || p->k <= 0) continue;
        if (G.count == MAXG) {   // full: launch and start another group
            hipLaunchKernelGGL(herk_mixed_kernel<MAXG>, dim3((unsigned)grid), dim3(256), 0, s, G);
            G = MixedGroup<MAXG>{};
            grid = 0;
        }
        G.u[G.count] = mixed_desc(*p);
        grid += G.u[G.count++].nitems;
    }
    if (grid > 0) hipLaunchKernelGGL(herk_mixed_kernel<MAXG>, dim3((unsigned)grid), dim3(256), 0, s, G);
}

void loop_job_from(const HerkProblem &hp, LoopJob *out)
{
    const MixedUpdate u = mixed_desc(hp);
    LoopJob j{};
    j.nks = u.nks;
    j.Z = u.Z;
    j.acc = u.acc;
    j.bound = u.bound;
    j.order = u.order;
    j.n = u.n;
    j.ntile = u.ntile;
    j.ksplit = u.ksplit;
    j.nitems = hp.k > 0 ? u.nitems : 0;
    j.nitems8 = (j.nitems + 7) / 8 * 8;
    *out = j;
}

void launch_herk_mixed_loop(LoopCtl *ctl_dev, LoopDev *dev, int inst, uint64_t idle_ticks, int grid, hipStream_t s,
                            unsigned help_lo, unsigned help_hi)
{
    hipLaunchKernelGGL(herk_mixed_loop_kernel, dim3((unsigned)grid), dim3(256), 0, s, ctl_dev, dev, inst, idle_ticks,
                       help_lo, help_hi);
}

void launch_herk_mixed_group(const HerkProblem *const *hp, int count, hipStream_t s)
{
    launch_group<kMaxUpdateGroup>(hp, count, s);
}

void launch_herk_mixed_update(const HerkProblem &a, const HerkProblem *b, hipStream_t s)
{
    // the two-entry kernel takes N and the tile order from its first entry
    if (a.k > 0 && b && b->k > 0 && (b->n != a.n || b->order != a.order)) {
        launch_group<kMaxUpdateGroup>(std::initializer_list<const HerkProblem *>{&a, b}.begin(), 2, s);
        return;
    }
    const HerkProblem *list[2] = {&a, b};
    launch_group<2>(list, 2, s);
}

// Lower-triangle tiles (ti >= tj) enumerated by 8x8 super-blocks.
std::vector<int> blocked_tile_order(int nt)
{
    const int SB = 8;
    std::vector<int> o;
    const int ns = (nt + SB - 1) / SB;
    for (int I = 0; I < ns; ++I)
        for (int J = 0; J <= I; ++J)
            for (int ti = I * SB; ti < std::min(nt, I * SB + SB); ++ti)
                for (int tj = J * SB; tj < std::min(nt, J * SB + SB); ++tj)
                    if (ti >= tj) {
                        o.push_back(ti);
                        o.push_back(tj);
                    }
    return o;
}

size_t mixed_zbuf_bytes(int64_t k, int n) { return (size_t)mixed::pad_ksteps(k) * (size_t)mixed::pad_rows(n) * 64; }

size_t fp64_zbuf_bytes(int64_t k, int n) { return (size_t)f64::nstages(k) * (size_t)f64::pad_rows(n) * 128; }

void launch_expand_fp64(const HerkProblem &hp, hipStream_t s)
{
    if (hp.k <= 0) return;
    const size_t lds = sizeof(double2) * f64::VPS * (size_t)(hp.nv + hp.nl) * (hp.A2 ? 2 : 1);
    if (lds <= (size_t)kExpandFp64Lds) {
        const dim3 g1((unsigned)f64::nstages(hp.k));
        if (hp.A2)
            hipLaunchKernelGGL(expand_fp64_stage_kernel<true>, g1, dim3(256), lds, s, hp.nv, hp.nl, hp.n, hp.k,
                               (const double2 *)hp.A1, (const double2 *)hp.B1, (const double2 *)hp.A2,
                               (const double2 *)hp.B2, (double2 *)hp.zbuf);
        else
            hipLaunchKernelGGL(expand_fp64_stage_kernel<false>, g1, dim3(256), lds, s, hp.nv, hp.nl, hp.n, hp.k,
                               (const double2 *)hp.A1, (const double2 *)hp.B1, nullptr, nullptr, (double2 *)hp.zbuf);
        return;
    }
    const dim3 g((unsigned)f64::nstages(hp.k), (unsigned)((f64::pad_rows(hp.n) + 255) / 256));
    if (hp.A2)
        hipLaunchKernelGGL(expand_fp64_kernel<true>, g, dim3(256), 0, s, hp.nv, hp.nl, hp.n, hp.k,
                           (const double2 *)hp.A1, (const double2 *)hp.B1, (const double2 *)hp.A2,
                           (const double2 *)hp.B2, (double2 *)hp.zbuf);
    else
        hipLaunchKernelGGL(expand_fp64_kernel<false>, g, dim3(256), 0, s, hp.nv, hp.nl, hp.n, hp.k,
                           (const double2 *)hp.A1, (const double2 *)hp.B1, nullptr, nullptr, (double2 *)hp.zbuf);
}

void launch_herk_fp64(const HerkProblem &a, const HerkProblem *b, hipStream_t s)
{
    F64Group G{};
    int64_t grid = 0;
    for (const HerkProblem *p : {&a, b}) {
        if (!p || p->k <= 0) continue;
        F64Update &u = G.u[G.count++];
        const int nt = ntiles(p->n, kTileF64);
        u.Z = (const double *)p->zbuf;
        u.acc = p->acc;
        u.order = (const int2 *)p->order;
        u.nst = f64::nstages(p->k);
        u.n = p->n;
        u.ntile = nt * (nt + 1) / 2;
        u.ksplit = p->ksplit;
        u.nitems = u.ntile * u.ksplit;
        grid += u.nitems;
    }
    if (G.count == 0) return;
    hipLaunchKernelGGL(herk_fp64_kernel<f64::NBUF>, dim3((unsigned)grid), dim3(256), 0, s, G);
}

// ------------------------------------------------------------ finaliser
__global__ void cov_finalize_kernel(const double2 *acc, int ksplit, int nv, int nl, double scale,
                                    const double2 *rowf, double *out_re, double *out_im)
{
    const int n = nv * nl;
    const int64_t total = (int64_t)n * n;
    int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= total) return;
    // o = v1 + nv*(v2 + nv*(l1 + nl*l2))  (MATLAB column-major Nv x Nv x Nl x Nl)
    int64_t t = o;
    const int v1 = (int)(t % nv); t /= nv;
    const int v2 = (int)(t % nv); t /= nv;
    const int l1 = (int)(t % nl); t /= nl;
    const int l2 = (int)t;
    const int i = v1 + nv * l1, j = v2 + nv * l2;
    // Every entry is read from the lower triangle of the matrix (a >= b lies
    // in a computed tile, ti >= tj), the upper one as its conjugate, and the
    // diagonal is real: the result is exactly Hermitian even though a
    // diagonal tile's two triangles were summed independently.
    const bool lower = i >= j;
    const int a = lower ? i : j, b = lower ? j : i;
    double re = 0.0, im = 0.0;
    for (int s = 0; s < ksplit; ++s) {
        double2 c = acc[(size_t)s * n * n + (size_t)a * n + b];
        re += c.x;
        im += c.y;
    }
    if (!lower) im = -im;
    if (i == j) im = 0.0;
    if (rowf) {
        // * rowf[i] * conj(rowf[j])
        double2 fi = rowf[i], fj = rowf[j];
        double pr = fi.x * fj.x + fi.y * fj.y, pi = fi.y * fj.x - fi.x * fj.y;
        double nr = re * pr - im * pi, ni = re * pi + im * pr;
        re = nr;
        im = ni;
    }
    re *= scale;
    im *= scale;
    if (out_im) {
        out_re[o] = re;
        out_im[o] = im;
    } else {
        out_re[2 * o] = re;
        out_re[2 * o + 1] = im;
    }
}

void launch_cov_finalize(const double2 *acc, int ksplit, int nv, int nl, double scale, const double2 *rowf,
                         double *out_re, double *out_im, hipStream_t s)
{
    const int64_t total = (int64_t)nv * nl * nv * nl;
    dim3 g((unsigned)((total + 255) / 256));
    hipLaunchKernelGGL(cov_finalize_kernel, g, dim3(256), 0, s, acc, ksplit, nv, nl, scale, rowf, out_re, out_im);
}

} // namespace scmc
