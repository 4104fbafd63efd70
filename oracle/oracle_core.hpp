// oracle_core.hpp -- FP64 CPU restatement of the ETTK sub-volume alignment path.
//
// TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it.  The product (ettk_amd/libettk_align.so) never links,
// loads or calls anything under oracle/.
//
// Internal arrays use the reference's blitz index order A(i,j,k) with i = VTK x, k fastest in memory
// (inc/io/nbfVTKInterface.cpp:77,134-135 maps VTK x to blitz's first index), so the half spectrum is
// taken along k exactly as fftw_plan_dft_r2c_3d does on blitzFFTreal (inc/em/nbfFourierFilter.h:345-351).
#pragma once
#include <complex>
#include <cstdint>
#include <vector>
#include "../include/ettk_align.h"

namespace orc {

using cd = std::complex<double>;

// ---------------------------------------------------------------- 4x4 transforms (vtkTransform)
struct M4 {
    double a[16];
    double& operator()(int r, int c) { return a[r * 4 + c]; }
    double operator()(int r, int c) const { return a[r * 4 + c]; }
};
M4 m4_ident();
M4 m4_from(const double* p);
M4 m4_mul(const M4& A, const M4& B);
M4 m4_inv(const M4& A);
double m4_det(const M4& A);
// vtkTransformConcatenation::Rotate (VTK Common/vtkAbstractTransform.cxx:548-598): quaternion built matrix,
// identity for a zero angle.
M4 m4_rot(double deg, double x, double y, double z);
inline M4 m4_rx(double d) { return m4_rot(d, 1, 0, 0); }
inline M4 m4_ry(double d) { return m4_rot(d, 0, 1, 0); }
inline M4 m4_rz(double d) { return m4_rot(d, 0, 0, 1); }
M4 m4_translate(double x, double y, double z);
M4 m4_scale(double x, double y, double z);
inline void pre(M4& M, const M4& R) { M = m4_mul(M, R); }  // vtkTransform PreMultiply
M4 m4_strip_translation(const M4& t);                      // setSeedTransform / wedge "rotation"
// transform of candidate angles: RotateZ(-g) RotateY(-b) RotateZ(-a) (ProjRot.h:710-713)
M4 m4_zyz(double a, double b, double g);
bool transform_valid(const M4& t, double trans_restrict, double rot_restrict);  // nbfImageMetric.h:416-459

// ---------------------------------------------------------------- FFT (FP64)
void fft1d(cd* x, int n, int sign);                        // in place, unnormalised, sign -1 = forward
// FFTW semantics on an n0 x n1 x n2 row-major real array (k fastest), half spectrum along the last axis
void r2c3(const double* in, cd* out, int n0, int n1, int n2);
// c2r: complex transforms along the two leading axes, then a real-output transform along the last axis
// in which only the real part of the k=0 and k=n2/2 terms contributes (the half-complex convention).
void c2r3(const cd* in, double* out, int n0, int n1, int n2);

// ---------------------------------------------------------------- image helpers
// separable vtkImageGaussianSmooth on an A(i,j,k) float array (axes 2,1,0 in that order, float rounding
// after each pass, kernel clipped and renormalised at the border): Imaging/vtkImageGaussianSmooth.cxx
void gauss_smooth3(std::vector<float>& A, int n0, int n1, int n2, double stdv, double radius_factor);
std::vector<float> sphere_mask(int N, float radius, float sigma);           // IFilt.h:384-472
std::vector<float> square_fft_mask(int N, int square);                      // FIM.h:642-664
// resliced particle image (WSub.h:603-828): x-fastest float volume in, A(i,j,k) double out
void reslice_image(const float* vol, int N, const ettk_particle& d, const M4* t, std::vector<double>& out);
// the same sub-volume cut from an nx x ny x nz volume around position pos (WSub.h:603-802)
void extract_image(const float* vol, int nx, int ny, int nz, const double* pos, int N, const ettk_particle& d,
                   const M4* t, std::vector<double>& out);
// transpose an x-fastest float volume into A(i,j,k) double
void to_internal(const float* vol, int N, std::vector<double>& out);

// ---------------------------------------------------------------- alignment state
struct Plan {
    ettk_params p;
    int N, H;                     // H = N/2+1
    std::vector<float> shift;     // (-1)^(i+j+k), N^3
    std::vector<float> maskfft;   // FFT window
    std::vector<float> refmask;   // m1 (reference mask == image mask)
    std::vector<float> imgmask;   // nbfImageFilter::mask used by execute()
    std::vector<float> filter;    // half spectrum band-pass x B-factor, DC = 0
    bool dummy_filter;
    std::vector<cd> Mfft;         // r2c(m1 * shift)
    // reference side
    std::vector<double> ref;      // I1
    ettk_particle refd;
    std::vector<float> W1;        // half-spectrum wedge of the reference
    // reference that is an average: 0 = analytic wedge of refd, 1 = accumulated wedge (effective),
    // 2 = accumulated wedge not effective (ones); binarised full grid and spherical window of kind 1
    int ref_wedge_kind = 0;
    std::vector<float> Wbin_acc, W1s_acc, W1acc_half;
    std::vector<cd> F1m;
    double k1_m;
    std::vector<float> Pp1;       // S^2 projection of the reference
    int ccc_count = 0;
    float last_peak = 0.0f;       // correlationPeak of the last ccc_eval
};
void plan_init(Plan& P, const ettk_params& p);
void plan_set_reference(Plan& P, const float* ref, const ettk_particle& d);
std::vector<float> wedge_half(int N, float lw, float uw, const M4& rot);   // Wedge.h:313-360
std::vector<float> particle_wedge_half(int N, const ettk_particle& d, const M4* t);  // WSub.h:941-979
std::vector<float> wedge_full(int N, float lw, float uw, const M4& rot);               // Wedge.h:195-290
// accumulated wedge of an average (WAvg.h:49-70, 496-538, 614-667)
struct AvgWedge {
    std::vector<float> accum, bin, wedge, sph;  // N^3 sum, binarised, getWedgeImage; (2B)^2 binarised
    bool effective = false;
};
AvgWedge avg_wedge(int N, int B, int n, const ettk_particle* d, const float* w, const double* mats);
void plan_set_reference_average(Plan& P, const float* ref, const AvgWedge& aw);
// nbfCorrelationImageMetric::getWedgeOverlap(t) (CorrMetric.h:533-566) of the reference against particle d
float wedge_overlap(const Plan& P, const ettk_particle& d, const M4& t);
// what ccc_eval leaves behind for the checks of its arithmetic: the search window T of the upsampled
// correlation (ext^3 values, before the spherical mask) and k1_m
struct CccTrace {
    std::vector<double> T;
    int ext = 0;
    double k1_m = 0;
};
// executeFourierNewHalf: returns correlationPeak, updates t (t <- t * Translate(-shift))
float ccc_eval(Plan& P, const float* part, const ettk_particle& d, M4& t, CccTrace* trace = nullptr);

// ---------------------------------------------------------------- S^2 / SO(3)
void project_image(Plan& P, std::vector<double> img, std::vector<float>& Pp);  // cartesianToSpherical
std::vector<float> spherical_wedge(int B, const ettk_particle& d);             // WSub.h:1007-1041
void condition_spherical_window(std::vector<float>& W, int B);                 // ProjRot.h:1322-1345
// S^2 forward transform with SOFT's coefficient scaling: coefficient (l,m), |m| <= l < B
void s2_forward(int B, const double* f, std::vector<cd>& coef);                // FST_semi_memo
void s2_inverse(int B, const std::vector<cd>& coef, double* f);                // InvFST_semi_memo (real part)
inline int s2_index(int l, int m) { return l * l + l + m; }
// correlation on SO(3): C[a][b][g] (alpha slowest) as nbfProjectionRotationMetric3D::soft returns it
void so3_correlate(int B, const double* sig, const double* pat, std::vector<double>& C);
void s2_rotate(int B, const double* in, double* out, double alpha, double beta, double gamma);  // rotateFct
void so3_correlate_scaled(int B, const double* F1, const double* F2, std::vector<double>& C);  // softScaled
// d^l_{m1 m2}(beta), FP64, for all l < B at one beta: table[l][m1+l][m2+l] flattened
void wigner_d_all(int B, double beta, std::vector<double>& d);

// ---------------------------------------------------------------- driver
struct Cand { float a, b, g, v; };
// the candidate chain of one correlation volume C ((2B)^3, [alpha][beta][gamma]): restriction of beta to the
// rotation restriction, 26-neighbour local maxima with a -FLT_MAX border, decoding, validity and the
// elimination of equivalent rotations; the surviving candidates in priority-queue order (ProjRot.h:590-640,
// :1593-1702)
std::vector<Cand> candidates_of(const Plan& P, std::vector<double> C, float tol);
void align_pair(Plan& P, const float* part, const ettk_particle& d, ettk_result& out);
void powell_refine(Plan& P, const float* part, const ettk_particle& d, M4& t, bool run);

}  // namespace orc
