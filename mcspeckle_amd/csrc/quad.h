// quad.h -- geometry of the Fourier-accelerated field (MCfieldFourierQuad.m:80-138).
#pragma once
#include "kernels.h"

namespace scmc {

constexpr int kQuadMaxSlabs = 4096;
constexpr int kQuadFusedLights = 4;   // lights per vertex the fused walk-and-scatter kernel handles

// One light of an illumination wave as the FP32-grid kernels read it: the
// phase gradient vsign l / lambda (cycles per unit length) and the weight Wl.
struct WaveLight {
    double kx, ky, kz;
    float wr, wi;
};
constexpr int kWaveLightsLds = 512;   // a workgroup stages up to this many lights in LDS

// Grid geometry shared by the host set-up and the scatter kernel.
struct QuadGeom {
    int nv0;          // views per axis
    int nx;           // lateral cells per axis (odd; OnWave: the parity of nv0)
    int nfft;         // padded transform size, max(nx, nv0)
    int nx_pad;       // zero padding of the cell grid on each side
    int nva_pad;      // crop of the transform on each side
    int nz;           // number of slab boundaries (slabs = nz - 1)
    double box_stp, z_bin, gx0, gy0;
    double l_diratt, l_idiratt;   // 1/|l_z| by sign of l(end) (MCfieldFourierQuad.m:127-128)
    const double *gz;             // nz slab boundaries (device)
    // illumination wave (MCfieldFourierQuadOnWave.m): lights combined per
    // vertex with the weights wl, one output plane, vsign conventions
    int wave = 0;
    double vsign = -1.0;
    const double2 *wl = nullptr;  // nl weights (device)
    const WaveLight *wlk = nullptr;   // nl lights with their weights (device)
};

// Host restatement of MCfieldFourierQuad.m:80-121 (with parity_fix, of
// MCfieldFourierQuadOnWave.m:84-131); fills the enlarged box.
int quad_grid_setup(const double *box_min, const double *box_max, double v_max, double v_stp, double lambda,
                    bool parity_fix, QuadGeom &g, double *bmin_out, double *bmax_out, std::vector<double> *gz_out);

void launch_quad_scatter(const DevProblem &P, const QuadGeom &g, int64_t nvert, const PathRec *paths,
                         const VertexRec *verts, const int32_t *vpath, const int64_t *offsets, void *grid, bool fp32,
                         hipStream_t s);
// Walk paths [p0, p0 + np) and scatter every vertex into the slab grids in
// the same pass; *nvert (device) is incremented by the vertex count; *ctr
// (device, zeroed before the launch) hands out the path indices.  False (and
// nothing launched) for more lights than the pass holds: kQuadFusedLights
// separate lights, or a wave of more than kWaveLightsLds.
bool launch_quad_trace_scatter(const DevProblem &P, const QuadGeom &g, uint64_t p0, int64_t np, void *grid, bool fp32,
                               unsigned long long *nvert, unsigned long long *ctr, hipStream_t s);
void launch_cgemm(int M, int N, int K, const void *A, int64_t sa, const void *B, int64_t sb, void *C, int64_t sc,
                  int batch, bool transB, bool fp32, hipStream_t s);
// The view grid of MCfieldFourierQuad.m:80-138 as the combine kernel needs
// it: slab phases zexp0, the box factor boxexp0 and the final scale.
struct QuadView {
    double v_max, v_stp, vsign, sigt, lambda, zmin, zmax;
    double scale;   // sqrt(V sigt / stored vertices), MCfieldFourierQuad.m:302-303
    int wave;       // MCfieldFourierQuadOnWave.m:143 slab-phase sign
};
// W[ix * nv0 + o] = tw[((o + nva_pad - a)(ix + nx_pad - h)) mod n], the DFT
// matrix with the shifts, flips and crop folded in (tw: n twiddles e^{2 pi i r/n}).
void launch_quad_dft_matrix(const QuadGeom &g, const double2 *tw, void *W, bool fp32, hipStream_t s);
// H[b] = W X[b] W^T for batch slabs by two factored-DFT passes in LDS (n =
// n1 n2); false when n is prime, factors too unevenly or one vector of it
// does not fit in LDS (then nothing is launched and the DFT-matrix GEMMs are
// used).  tw: the n twiddles.
bool launch_quad_dft2(const QuadGeom &g, int batch, const void *X, void *T1, void *H, const double2 *tw, bool fp32,
                      hipStream_t s);
// u[o, l] = scale * boxexp0[o] * sum_s H[s, l][o] exp(zexp0[o] tz_s)
void launch_quad_combine(int nv0, int nl, int nz, const void *H, const double *gz, const QuadView &v, double2 *u,
                         bool fp32, hipStream_t s);

} // namespace scmc
