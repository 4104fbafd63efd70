// kernels.hpp -- launch wrappers of the alignment kernels (device code in *.hip).
#pragma once
#include "common.hpp"

namespace ettk {

// (a, b, c) of the flat index h into an [n0][n1][n2] array (h < 2^32): 32-bit divisions only
__device__ __forceinline__ void split3(unsigned h, unsigned n1, unsigned n2, int& a, int& b, int& c) {
    const unsigned q = h / n2;
    c = (int)(h - q * n2);
    const unsigned a_ = q / n1;
    b = (int)(q - a_ * n1);
    a = (int)a_;
}

// Tile of workgroup `b` out of `nb` per evaluation, grouped by XCD: the dispatcher deals consecutive workgroups
// of a grid round-robin to the 8 XCDs (each with its own L2), so workgroup b runs on XCD b % 8.  Giving XCD k
// the contiguous tiles [k nb/8, (k+1) nb/8) keeps the tiles that share cache lines (neighbours along the
// fastest axis) in one L2.  nb must be a multiple of 8 (otherwise the identity).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// (evaluation, tile) of workgroup (bx, by) of an nb x ne grid of per-evaluation tiles, one evaluation per XCD:
// when both extents are multiples of 8, the workgroups dealt to XCD k (linear id = k mod 8) cover whole
// evaluations (e = k mod 8 within each group of 8), so a particle volume is fetched into one L2 only.
__device__ __forceinline__ void xcd_eval_tile(int bx, int by, int nb, int ne, int& e, int& tile) {
    if ((nb & 7) == 0 && (ne & 7) == 0) {
        e = (by & ~7) | (bx & 7);
        tile = (by & 7) * (nb >> 3) + (bx >> 3);
    } else {
        e = by;
        tile = xcd_tile(bx, nb);
    }
}

// per-evaluation geometry computed on the host in FP64
struct EvalGeom {
    float m[12];      // reslice map: input index = m[0..2]*i + m[3..5]*j + m[6..8]*k + m[9..11]  (rows x,y,z)
    float v1[3];      // rotated wedge plane normals of the particle (nbfWedge3D::getImageHalf)
    float v2[3];
    int full;         // particle wedge spans >= 180 degrees (W2 == 1)
    int particle;     // index of the particle volume
    float cut;        // cut offset of the sub-volume along z (nbfWedgedSubImage3D::setCutOffset)
    int pad;
    double md[12];    // the reslice map in FP64: decides validity for samples within FP32 rounding of the border
};

// The CCC steps between the particle's r2c plane passes and the fused plane kernels, on pencils along i
// (n <= 64, fft3_fused_supported): S2 holds the plane-pass output of the resliced particles and receives the
// filtered particle spectra X2; S3 receives X2 and S1 the filtered reference spectrum X1, both inverse-transformed
// along i (the inputs of fft3_op_r2c_after_ipass).  Partials: fft3_ccc_ipass_blocks(n) per evaluation.
struct CccIpassArgs {
    float2* S2;
    float2* S3;
    float2* S1;
    const float2* F1mcf;   // reference spectrum, MCF applied (one volume)
    const float* filter;   // half grid
    const float* W1;       // reference wedge on the half grid, or null (no missing wedge)
    const EvalGeom* geom;  // per evaluation: the particle wedge at its transform
    int mcf, apply;
    float* p_n1;
    float* p_wf;
    float* p_n2;
};
int fft3_ccc_ipass_blocks(int n);
void fft3_ccc_ipass(const FftTables& T, const CccIpassArgs& a, int n, int batch, hipStream_t s);
// the k/j plane passes of an r2c only (n <= 64): the i pass is left to fft3_ccc_ipass
void fft3_r2c_planes(const FftTables& T, const R2CArgs& a, hipStream_t s);
// fft3_c2r_op_r2c whose inverse i pass has already run (io holds the i-inverse-transformed spectra); without
// `forward_i` its forward i pass is left to launch_combine_fftA
void fft3_op_r2c_after_ipass(const FftTables& T, float2* io, int n, int batch, int op, const float* aux, float* partials,
                             bool forward_i, hipStream_t s);

// smoothed analytic wedge parameters of the reference
struct WedgeRef {
    float v1[3], v2[3];
    int full;
};

struct CccSizes {
    int N, H, E;          // box, half length, crop edge of the upsampled search window
    int I0;               // first crop index on the (2N) grid
    float restr2;         // restriction^2 for the spherical search mask
    int ub2;              // (E-1)/2
};
// launch_combine_dftA's result with the first DFT axis evaluated by FFTs (box size 64, any window); with
// `forward_i`, F1m and Q are passed before their forward i transform, which the kernel runs itself
bool ccc_combine_fft_supported(const CccSizes& z);
void launch_combine_fftA(const FftTables& T, const float2* F1m, const float2* F2, const float2* Q, const float2* Mfft,
                         const float* scalars, const CccSizes& z, int n_eval, float2* G1, bool forward_i, hipStream_t s);
// launch_dftBC_argmax's result with the second DFT axis evaluated by FFTs (box size 64, any window)
void launch_dftBC_fft(const FftTables& T, const float2* G1, const float2* tabK, const CccSizes& z, int n_eval,
                      float* best_val, int* best_idx, hipStream_t s);

// reslice (tricubic Keys a=-1/2, clamp border, -FLT_MAX outside): out [e][i][j][k]; `nb` must equal
// reslice_blocks(N) (one workgroup per 8^3 output tile), partials hold 2 * nb floats per evaluation
int reslice_blocks(int N);
void launch_reslice(const float* parts, const EvalGeom* geom, int n_eval, int N, float* out, float* partials,
                    int nb, hipStream_t s,
                    hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
// the same reslice for the S^2 projection of particles in their own frame (its own name in kernel traces)
void launch_reslice_projection(const float* parts, const EvalGeom* geom, int n, int N, float* out, float* partials,
                               int nb, hipStream_t s);
// mean of the valid voxels per evaluation, from the reslice partials (nb per eval): stats[e] = mean
void launch_reslice_mean(const float* partials, int nb, int n_eval, float* stats, hipStream_t s);
// spectral weighting of both CCC sides: X1 = F1 filter W1 W2(e) (F1 shared), X2 = mcf(X2) filter W1 W2(e) in
// place (mcf: F / |F|^.5), DC = 0; per-block partials of the Hermitian weighted |X1|^2, sum(W filter), |X2|^2
void launch_spec_filter2(const float2* F1, float2* X1, float2* X2, const float* filter, const float* W1,
                         const EvalGeom* geom, int n_eval, int N, int mcf, int apply, float* p_n1, float* p_wf,
                         float* p_n2, int nb, hipStream_t s);
// r1 *= m1, partial sums of r1^2 m1
void launch_mask_ref(float* R, const float* m1, int n_eval, int N, float* partials, int nb, hipStream_t s);
// per-evaluation scalars: norm1, norm2, k1_m, sum(W filter)
void launch_ccc_scalars(const float* p_n1, const float* p_n2, const float* p_k1, const float* p_wf, int nb_spec,
                        int nb_real, int n_eval, int N, float* scalars, hipStream_t s);
// C = (-M conj(Q)/n2^2 + 2 F1m/n1 conj(F2/n2 N^3)) (-1)^(i+j+k), then G1[e][x][j][k] = sum_i A[x][i] C(i,j,k)
void launch_combine_dftA(const float2* F1m, const float2* F2, const float2* Q, const float2* Mfft,
                         const float* scalars, const float2* tabA, const CccSizes& z, int n_eval, float2* G1,
                         hipStream_t s);
// G2 = sum_j A[y][j] G1, T = sum_k Re(G2 K[z][k]) on the crop, spherical mask, first argmax per (e, x)
void launch_dftBC_argmax(const float2* G1, const float2* tabA, const float2* tabK, const CccSizes& z, int n_eval,
                         float* best_val, int* best_idx, hipStream_t s);
// continuous CCC and adjoint in the particle frame (refine.hip)
struct RefineGeom {
    float G[12];          // q = G[r*4 + 0..2] (x, y, z) - N/2 + G[r*4 + 3]: reference index of particle voxel (x,y,z)
    float n1[3], n2[3];   // particle wedge normals, lab frame
    float u1[3], u2[3];   // reference wedge normals in the particle frame at the seed pose
    float A[9];           // particle-frame frequency -> reference-frame frequency at the seed (rows), rw == 2
    int pw, rw;           // wedge tests on (particle, reference): rw 1 = analytic normals u1/u2, 2 = lookup of
                          // the binarised accumulated wedge (full grid, k fastest) at A k
    int spec;             // slot of the particle's spectrum
    int pad[3];
};

// H at half-spectrum element (a, b, c) = (z, y, x) of the particle frame: filter x binary wedges
__device__ __forceinline__ float refine_hweight(const float* __restrict__ filter, const float* __restrict__ wbin,
                                         const RefineGeom& g, long h, int a, int b, int c, int N, int H) {
    float w = filter[h];  // radial table: identical in any axis order
    const float kx = (float)(c - (H - 1)), ky = (float)(b - N / 2), kz = (float)(a - N / 2);
    if (g.pw) {
        float s1 = g.n1[0] * kx + g.n1[1] * ky + g.n1[2] * kz, s2 = g.n2[0] * kx + g.n2[1] * ky + g.n2[2] * kz;
        if (s1 * s2 > 0.f) w = 0.f;
    }
    if (g.rw == 1) {
        float s1 = g.u1[0] * kx + g.u1[1] * ky + g.u1[2] * kz, s2 = g.u2[0] * kx + g.u2[1] * ky + g.u2[2] * kz;
        if (s1 * s2 > 0.f) w = 0.f;
    } else if (g.rw == 2) {
        // nearest voxel of the reference-frame frequency, clamped to the grid
        int r[3];
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const float p = g.A[t * 3] * kx + g.A[t * 3 + 1] * ky + g.A[t * 3 + 2] * kz;
            r[t] = min(max((int)rintf(p) + N / 2, 0), N - 1);
        }
        if (wbin[(r[0] * N + r[1]) * N + r[2]] == 0.f) w = 0.f;
    }
    return w;
}
int refine_moment_blocks(int N);
void launch_refine_scalars(const double* part, int nb, int E, double* sc, hipStream_t s);
// the refinement's transforms with the score and adjoint steps in their i passes (box sizes 16, 32, 48, 64):
//   fft3_refine_dot: forward i pass of X1 in place (after fft3_r2c_planes) with the FP64 score partials of
//     launch_refine_dot, fft3_refine_ipass_blocks(n) per evaluation;
//   fft3_refine_adj_c2r: the adjoint spectrum of launch_refine_adj formed on load, its inverse i pass into
//     `tmp`, then the c2r plane passes into `out` (X1 unchanged)
int fft3_refine_ipass_blocks(int n);
void fft3_refine_dot(const FftTables& T, float2* X1, const float2* FP, const float* filter, const float* wbin,
                     const RefineGeom* geom, int n, int batch, double* part, hipStream_t s);
void fft3_refine_adj_c2r(const FftTables& T, const float2* X1, const float2* FP, const float* filter,
                         const float* wbin, const RefineGeom* geom, const double* sc, int n, int batch, float2* tmp,
                         float* out, hipStream_t s);

int refine_dot_blocks(int N);
void launch_vol_mean(const float* vols, int count, int N, float* mean, hipStream_t s);
void launch_reslice_ref(const float* ref, const RefineGeom* geom, int E, int N, float* out, hipStream_t s);
void launch_refine_dot(const float2* X1, const float2* FP, const float* filter, const float* wbin, const RefineGeom* geom,
                       int E, int N, double* part, double* sc, hipStream_t s);
void launch_refine_adj(float2* X1, const float2* FP, const float* filter, const float* wbin, const RefineGeom* geom,
                       int E, int N, const double* sc, hipStream_t s);
void launch_refine_moments(const float* ref, const RefineGeom* geom, const float* G, const float* maskfft, int E, int N,
                           float* mpart, double* mom, hipStream_t s);
void launch_wedge_ref(const WedgeRef& w, int N, float* W, hipStream_t s);
// out[e] = sum over the half grid of W1 W2(e) filter (FP64); `part` holds 64 doubles per evaluation
void launch_wedge_overlap(const float* W1, const float* filter, const EvalGeom* geom, int E, int N, double* part,
                          double* out, hipStream_t s);
void launch_xfast_to_internal(const float* in, float* out, int N, hipStream_t s);
void launch_internal_to_xfast(const float* in, float* out, int N, hipStream_t s);
void launch_mcf_inplace(float2* F, long n, hipStream_t s);
// zero n elements with a kernel (a fill node, not a memset node, when the stream is captured into a graph)
void launch_fill_zero(int* p, size_t n, hipStream_t s);
void launch_fill_zero(double* p, size_t n, hipStream_t s);
// final per evaluation: first maximum over x, distance and shift
void launch_peak_final(const float* best_val, const int* best_idx, const float* scalars, float sum_filter,
                       int overlap, const CccSizes& z, int n_eval, float* out /* [e][4] = dist, sx, sy, sz */,
                       hipStream_t s);

}  // namespace ettk
