// so3.hip -- S^2 projection of |FFT| and the SO(3) rotational search (gfx950).
//
// Device side of nbfProjectionRotationMetric3D (inc/em/nbfProjectionRotationMetric3D.h):
//   cartesianToSpherical (:265-413)  -> reslice, mask statistics, constant-padded (2N)^3 r2c, |F| (or
//                                       |F|^(1/2) with MCF), tricubic wrap-around sampling on the
//                                       Driscoll-Healy grid, radial sum (FP32, like the float Pp of the reference)
//   soft / softScaled (:1137-1319)   -> S^2 forward transforms, coefficient combination and inverse SO(3)
//                                       synthesis (FP64, as REAL in the reference), extrema, rotateFct of the
//                                       pattern (Wigner-D rotation of its coefficients) and the affine rescale
//   execute (:526-619)               -> branch combination, rotation-range restriction, 3x3x3 local maxima
// The S^2/SO(3) mathematics is re-derived from the published method (Driscoll-Healy quadrature, Kostelec &
// Rockmore 2008): direct quadrature with FP64 Legendre tables and FP64 Wigner-d three-term recurrences.
#include <cfloat>
#include <cmath>

#include "so3.hpp"

namespace ettk {

static constexpr int kT = 256;

// the fused synthesis + LDS FFT path holds one (2B)^2 FP64 complex slice in LDS: 2B in {16, 32, 64}
static inline bool fused_synthesis(int B) { return B == 8 || B == 16 || B == 32; }
// the synthesis as per-order-pair FP64 matrix products on the matrix cores (B = 16, 32), then the 2-D DFT per slice
static inline bool mfma_synthesis(int B) { return B == 16 || B == 32; }

// ============================================================== tables (host, FP64)
static double lfact(int n) { return std::lgamma((double)n + 1.0); }

static void legendre_host(int B, double theta, std::vector<double>& P) {
    P.assign((size_t)B * B, 0.0);
    double x = std::cos(theta), s = std::sin(theta);
    double pmm = std::sqrt(1.0 / (4.0 * M_PI));
    for (int m = 0; m < B; m++) {
        if (m > 0) pmm *= -std::sqrt((2.0 * m + 1) / (2.0 * m)) * s;
        P[(size_t)m * B + m] = pmm;
        if (m + 1 < B) P[(size_t)(m + 1) * B + m] = x * std::sqrt(2.0 * m + 3) * pmm;
        for (int l = m + 2; l < B; l++) {
            double a = std::sqrt((4.0 * l * l - 1) / ((double)l * l - (double)m * m));
            double b = std::sqrt(((double)(l - 1) * (l - 1) - (double)m * m) / (4.0 * (l - 1) * (l - 1) - 1));
            P[(size_t)l * B + m] = a * (x * P[(size_t)(l - 1) * B + m] - b * P[(size_t)(l - 2) * B + m]);
        }
    }
}

__host__ __device__ static inline double wig_seed(int m1, int m2, double beta) {
    int j = abs(m1) > abs(m2) ? abs(m1) : abs(m2);
    double c = cos(beta / 2), s = sin(beta / 2);
    int m, sgn = 1;
    bool top;  // d^j_{j,m} (top) or d^j_{-j,m}
    if (m1 == j) { top = true; m = m2; }
    else if (m1 == -j) { top = false; m = m2; }
    else if (m2 == j) { top = true; m = m1; sgn = ((m1 - j) & 1) ? -1 : 1; }
    else { top = false; m = m1; sgn = ((m1 + j) & 1) ? -1 : 1; }
    double lg = 0.5 * (lgamma(2.0 * j + 1.0) - lgamma((double)(j + m) + 1.0) - lgamma((double)(j - m) + 1.0));
    double v = top ? exp(lg) * pow(c, (double)(j + m)) * pow(-s, (double)(j - m))
                   : exp(lg) * pow(c, (double)(j - m)) * pow(s, (double)(j + m));
    return sgn * v;
}
// one step of the three-term recurrence in l: d^{l+1} from d^l, d^{l-1}
__host__ __device__ static inline double wig_next(int l, int m1, int m2, double cb, double dcur, double dprev) {
    if (l == 0) return cb;
    double a = sqrt(((double)(l + 1) * (l + 1) - (double)m1 * m1) * ((double)(l + 1) * (l + 1) - (double)m2 * m2));
    double bt = sqrt(((double)l * l - (double)m1 * m1) * ((double)l * l - (double)m2 * m2));
    return ((2.0 * l + 1) * ((double)l * (l + 1) * cb - (double)m1 * m2) * dcur - (l + 1.0) * bt * dprev) / (l * a);
}

void so3_tables(ettk_plan* P) {
    So3State& S = *P->so3;
    const int B = S.B, n = 2 * B, W = 2 * B - 1;
    hipStream_t s = P->stream;
    // Driscoll-Healy weights
    std::vector<double> w(n);
    for (int j = 0; j < n; j++) {
        double th = M_PI * (2 * j + 1) / (4.0 * B), acc = 0;
        for (int k = 0; k < B; k++) acc += std::sin((2 * k + 1) * th) / (2 * k + 1);
        w[j] = (2.0 / B) * std::sin(th) * acc;
    }
    // Legendre tables [m][l][j]
    std::vector<double> LF((size_t)B * B * n), LI((size_t)B * B * n), Pl;
    for (int j = 0; j < n; j++) {
        legendre_host(B, M_PI * (2 * j + 1) / (4.0 * B), Pl);
        for (int m = 0; m < B; m++)
            for (int l = m; l < B; l++) {
                LI[((size_t)m * B + l) * n + j] = Pl[(size_t)l * B + m];
                LF[((size_t)m * B + l) * n + j] = Pl[(size_t)l * B + m] * w[j] * 2.0 * M_PI / n;
            }
    }
    S.legF.upload(LF.data(), LF.size(), s);
    S.legI.upload(LI.data(), LI.size(), s);
    // Wigner-d table [beta][pair][l - l0] with sqrt((2l+1)/(8 pi^2))
    std::vector<int> off((size_t)W * W);
    int tot = 0;
    for (int m1 = -(B - 1); m1 <= B - 1; m1++)
        for (int m2 = -(B - 1); m2 <= B - 1; m2++) {
            off[(size_t)(m1 + B - 1) * W + (m2 + B - 1)] = tot;
            tot += B - std::max(std::abs(m1), std::abs(m2));
        }
    S.d_slice = tot;
    std::vector<double> D((size_t)tot * n);
    for (int kb = 0; kb < n; kb++) {
        double beta = M_PI * (2 * kb + 1) / (4.0 * B), cb = std::cos(beta);
        for (int m1 = -(B - 1); m1 <= B - 1; m1++)
            for (int m2 = -(B - 1); m2 <= B - 1; m2++) {
                int l0 = std::max(std::abs(m1), std::abs(m2));
                double dprev = 0, dcur = wig_seed(m1, m2, beta);
                size_t o = (size_t)kb * tot + off[(size_t)(m1 + B - 1) * W + (m2 + B - 1)];
                for (int l = l0; l < B; l++) {
                    D[o + (l - l0)] = std::sqrt((2.0 * l + 1) / (8.0 * M_PI * M_PI)) * dcur;
                    double dn = wig_next(l, m1, m2, cb, dcur, dprev);
                    dprev = dcur;
                    dcur = dn;
                }
            }
    }
    S.dtab.upload(D.data(), D.size(), s);
    S.doff.upload(off.data(), off.size(), s);
    if (fused_synthesis(B)) {
        // [beta][c1][l][c2] with wrapped order indices c (m = c < B ? c : c - n), zero where l < l0 or c == B:
        // consecutive c2 are consecutive addresses for the synthesis loop over l
        std::vector<double> D2((size_t)n * n * B * n, 0.0);
        for (int kb = 0; kb < n; kb++)
            for (int c1 = 0; c1 < n; c1++)
                for (int c2 = 0; c2 < n; c2++) {
                    if (c1 == B || c2 == B) continue;
                    int m1 = c1 < B ? c1 : c1 - n, m2 = c2 < B ? c2 : c2 - n;
                    int l0 = std::max(std::abs(m1), std::abs(m2));
                    size_t o = (size_t)kb * tot + off[(size_t)(m1 + B - 1) * W + (m2 + B - 1)];
                    for (int l = l0; l < B; l++) D2[(((size_t)kb * n + c1) * B + l) * n + c2] = D[o + (l - l0)];
                }
        S.dtab2.upload(D2.data(), D2.size(), s);
        if (mfma_synthesis(B)) {
            // [c1 n + c2][l][beta]: for one order pair, the columns l of the (beta x l) matrix the synthesis
            // multiplies, beta contiguous
            std::vector<double> DP((size_t)n * n * B * n, 0.0);
            for (int kb = 0; kb < n; kb++)
                for (int c1 = 0; c1 < n; c1++)
                    for (int l = 0; l < B; l++)
                        for (int c2 = 0; c2 < n; c2++)
                            DP[(((size_t)c1 * n + c2) * B + l) * n + kb] = D2[(((size_t)kb * n + c1) * B + l) * n + c2];
            S.dtabP.upload(DP.data(), DP.size(), s);
        }
    }
    std::vector<double2> tw(n);
    for (int t = 0; t < n; t++) tw[t] = make_double2(std::cos(2 * M_PI * t / n), -std::sin(2 * M_PI * t / n));
    S.tw.upload(tw.data(), tw.size(), s);
    ETTK_HIP(hipStreamSynchronize(s));
}
static void so3_kernel_attributes(int B);

// ============================================================== projection kernels
// statistics of the masked image: st[0] = valid mean (from reslice), st[1] = mean of the image, st[2] = pad
// constant (value of voxel (0,0,0) after masking), st[3] = 1 when the image mask is on.  Two kernels: kProjParts
// workgroups per image sum (x - valid mean) over contiguous parts in FP64, one wave per image adds the parts in
// a fixed order.
static constexpr int kProjParts = 64;
__device__ __forceinline__ double wave_dsum64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// mean of the valid voxels from the reslice partials (sum, count per tile), the same for every thread
__device__ __forceinline__ float proj_valid_mean(const float* __restrict__ res_partials, int b, int nb_res, double* sh) {
    double s = 0, c = 0;
    for (int q = threadIdx.x; q < nb_res; q += kT) {
        s += res_partials[((size_t)b * nb_res + q) * 2];
        c += res_partials[((size_t)b * nb_res + q) * 2 + 1];
    }
    s = wave_dsum64(s);
    c = wave_dsum64(c);
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = s;
        sh[kT / 64 + (threadIdx.x >> 6)] = c;
    }
    __syncthreads();
    double ts = 0, tc = 0;
    for (int w = 0; w < kT / 64; w++) {
        ts += sh[w];
        tc += sh[kT / 64 + w];
    }
    __syncthreads();
    return (float)(ts / (double)(float)tc);
}
__global__ void __launch_bounds__(kT) k_proj_partial(const float* __restrict__ R, const float* __restrict__ res_partials,
                                                     int nb_res, int N, int is_ref, double* __restrict__ part,
                                                     float* __restrict__ st) {
    __shared__ double sh[2 * (kT / 64)];
    const int q = blockIdx.x, b = blockIdx.y;
    const long n3 = (long)N * N * N;
    const float* r = R + (size_t)b * n3;
    const float mv = is_ref ? 0.f : proj_valid_mean(res_partials, b, nb_res, sh);
    const long chunk = (n3 + kProjParts - 1) / kProjParts, o0 = q * chunk, o1 = min(n3, o0 + chunk);
    double acc = 0;
    for (long o = o0 + threadIdx.x; o < o1; o += kT) {
        const float x = r[o];
        acc += (x == -FLT_MAX) ? 0.0 : (double)(x - mv);
    }
    acc = wave_dsum64(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kT / 64; w++) t += sh[w];
        part[(size_t)b * kProjParts + q] = t;
        if (q == 0) st[b * 4 + 0] = mv;
    }
}
__global__ void __launch_bounds__(64) k_proj_final(const float* __restrict__ R, const double* __restrict__ part,
                                                   const float* __restrict__ mask, int N, int mask_on,
                                                   float* __restrict__ st) {
    const int b = blockIdx.x;
    const long n3 = (long)N * N * N;
    double acc = 0;
    for (int q = threadIdx.x; q < kProjParts; q += 64) acc += part[(size_t)b * kProjParts + q];
    acc = wave_dsum64(acc);
    if (threadIdx.x == 0) {
        const float mv = st[b * 4 + 0];
        const double mean = acc / (double)n3;
        const float x0 = R[(size_t)b * n3];
        const double img0 = (x0 == -FLT_MAX) ? 0.0 : (double)(x0 - mv);
        const double c = mask_on ? (img0 - mean) * mask[0] + mean : img0;
        st[b * 4 + 1] = (float)mean;
        st[b * 4 + 2] = (float)c;
        st[b * 4 + 3] = mask_on ? 1.f : 0.f;
    }
}

// |F| (|F|^(1/2) with MCF) of the (2N)^3 half spectrum
__global__ void k_magnitude(const float2* __restrict__ F, float* __restrict__ mag, long n, int mcf) {
    for (long h = blockIdx.x * (long)blockDim.x + threadIdx.x; h < n; h += (long)gridDim.x * blockDim.x) {
        float2 z = F[h];
        float m2 = z.x * z.x + z.y * z.y;
        mag[h] = mcf ? sqrtf(sqrtf(m2)) : sqrtf(m2);
    }
}

// full-size magnitude volume as ProjRot.h:327-338 leaves it: A(i,j,k) for k <= M/2 is |F| except the i = 0
// plane (0); for k > M/2: i = 0 -> 0, j = 0 -> the pad constant (the loop never overwrites those voxels of
// the padded image), otherwise the mirrored value |F|(M-i, M-j, M-k)
__device__ __forceinline__ float mag_at(const float* __restrict__ mag, int M, int i, int j, int k, float padc) {
    const int Hm = M / 2 + 1;
    if (i == 0) return 0.f;
    if (k <= M / 2) return mag[((size_t)i * M + j) * Hm + k];
    if (j == 0) return padc;
    return mag[((size_t)(M - i) * M + (M - j)) * Hm + (M - k)];
}

// One S^2 sample per group of kLanes consecutive threads; the radial shells r = bandL..bandU are split over the
// lanes and summed with wavefront shuffles.  Tricubic Keys weights (a = -1/2) with the 2^-17 floor tolerance
// and wrap-around taps, FP64 like the double image of the reference.
static constexpr int kLanes = 8;
__device__ __forceinline__ int wrap_idx(int x, int M) { return x < 0 ? x + M : (x >= M ? x - M : x); }

__global__ void __launch_bounds__(kT) k_sph_sample(const float* __restrict__ mag, const float* __restrict__ st, int M,
                                                    int B, int bandL, int bandU, float* __restrict__ Pp) {
    // one particle per XCD at a time: its magnitude volume is gathered through that XCD's L2 only
    int b, tile;
    xcd_eval_tile(blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, b, tile);
    const int n = 2 * B;
    const int t = tile * (kT / kLanes) + threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    const bool live = t < n * n;
    const int jp = live ? t / n : 0, kt = live ? t % n : 0;
    const float* mg = mag + (size_t)b * M * M * (M / 2 + 1);
    const float padc = st[b * 4 + 2];
    const float Bf = (float)B;
    const double phi = M_PI / 4.0 / Bf + jp * (2.0 * M_PI / 4.0 / Bf);
    const double th = kt * (M_PI / Bf);
    const double sp = sin(phi), cp = cos(phi), ct = cos(th), stt = sin(th);
    const double origin = -(M - 1) / 2.0 - 0.5;
    double sum = 0;
    for (int r = bandL + lane; live && r <= bandU; r += kLanes) {
        double p[3] = {r * sp * ct - origin, r * sp * stt - origin, r * cp - origin};
        int idx[3];
        double w[3][4];
        for (int a = 0; a < 3; a++) {
            double xx = p[a] + 7.62939453125e-06;
            idx[a] = (int)floor(xx);
            double f = xx - idx[a];
            double fm1 = f - 1, fd2 = f * 0.5, ft3 = f * 3;
            w[a][0] = -fd2 * fm1 * fm1;
            w[a][1] = ((ft3 - 2) * fd2 - 1) * fm1;
            w[a][2] = -((ft3 - 4) * f - 1) * fd2;
            w[a][3] = f * fd2 * fm1;
        }
        // the 4x4x4 stencil lies inside the stored half (k <= M/2) or entirely in the mirrored half, away from the
        // i = 0 plane, the j = 0 column of the mirror and the wrap: then its taps sit at fixed strides (the values
        // mag_at would return); otherwise every tap goes through mag_at
        const int i0 = idx[0] - 1, j0 = idx[1] - 1, k0 = idx[2] - 1, Hm = M / 2 + 1;
        const bool ij_in = i0 >= 1 && i0 + 3 <= M - 1 && j0 >= 0 && j0 + 3 <= M - 1;
        const bool direct = ij_in && k0 >= 0 && k0 + 3 <= M / 2;
        const bool mirror = ij_in && j0 >= 1 && k0 >= M / 2 + 1 && k0 + 3 <= M - 1;
        double v = 0;
        if (direct || mirror) {
            const long si = direct ? (long)M * Hm : -(long)M * Hm, sj = direct ? Hm : -Hm;
            const float* pb = direct ? mg + ((long)i0 * M + j0) * Hm + k0
                                     : mg + ((long)(M - i0) * M + (M - j0)) * Hm + (M - k0);
            // the 4 taps along k of row (cx, cy) are consecutive floats (ascending in the stored half, descending
            // in the mirror): one 16-byte load at their 4-byte-aligned start (global loads of four dwords need
            // only dword alignment on this architecture)
            float tap[4][4][4];  // [cz][cy][cx]
#pragma unroll
            for (int cy = 0; cy < 4; cy++)
#pragma unroll
                for (int cx = 0; cx < 4; cx++) {
                    const float* lo4 = pb + cy * sj + cx * si - (direct ? 0 : 3);
                    float r[4];
                    __builtin_memcpy(r, lo4, sizeof(r));
#pragma unroll
                    for (int cz = 0; cz < 4; cz++) tap[cz][cy][cx] = direct ? r[cz] : r[3 - cz];
                }
#pragma unroll
            for (int cz = 0; cz < 4; cz++) {
                double vy = 0;
#pragma unroll
                for (int cy = 0; cy < 4; cy++) {
                    double vx = 0;
#pragma unroll
                    for (int cx = 0; cx < 4; cx++) vx += w[0][cx] * (double)tap[cz][cy][cx];
                    vy += w[1][cy] * vx;
                }
                v += w[2][cz] * vy;
            }
        } else {
            for (int cz = 0; cz < 4; cz++) {
                const int kz = wrap_idx(idx[2] - 1 + cz, M);
                double vy = 0;
                for (int cy = 0; cy < 4; cy++) {
                    const int jy = wrap_idx(idx[1] - 1 + cy, M);
                    double vx = 0;
                    for (int cx = 0; cx < 4; cx++)
                        vx += w[0][cx] * (double)mag_at(mg, M, wrap_idx(idx[0] - 1 + cx, M), jy, kz, padc);
                    vy += w[1][cy] * vx;
                }
                v += w[2][cz] * vy;
            }
        }
        sum += v;
    }
    for (int o = kLanes / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, kLanes);
    if (live && lane == 0) Pp[(size_t)b * n * n + t] = (float)sum;
}

// ============================================================== soft instances: assemble S^2 samples
// recipe codes (the F1/F2 arrays of ProjRot.h:526-584, evaluated in float like the Pixel arrays there)
__device__ __forceinline__ float recipe_val(int code, float p1, float p2, float w1, float w2) {
    switch (code) {
        case RC_P1SQ_W1: return p1 * p1 * w1;
        case RC_W1: return w1;
        case RC_P1_W1: return p1 * w1;
        case RC_P2SQ_W2: return p2 * p2 * w2;
        case RC_W2: return w2;
        case RC_P2_W2: return w2 * p2;
        case RC_P1: return p1;
        case RC_P2: return p2;
        case RC_P1SQ: return p1 * p1;
        case RC_P2SQ: return p2 * p2;
    }
    return 0.f;
}

__global__ void k_assemble(const SoftInst* __restrict__ inst, const float* __restrict__ Pp1, const float* __restrict__ Pp2,
                           const float* __restrict__ W1, const float* __restrict__ W2, const float* __restrict__ Mx,
                           int B, double* __restrict__ sig, double* __restrict__ pat) {
    const int q = blockIdx.y;
    const int n2 = 4 * B * B;
    const SoftInst I = inst[q];
    const float M = Mx[I.particle];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += gridDim.x * blockDim.x) {
        float p1 = Pp1[t] / M, p2 = Pp2[(size_t)I.particle * n2 + t] / M;
        float w1 = W1[t], w2 = W2[(size_t)I.particle * n2 + t];
        sig[(size_t)q * n2 + t] = (double)recipe_val(I.sig, p1, p2, w1, w2);
        pat[(size_t)q * n2 + t] = (double)recipe_val(I.pat, p1, p2, w1, w2);
    }
}

// M = max(max Pp1, max Pp2) per particle (ProjRot.h:467)
__global__ void k_ppmax(const float* __restrict__ Pp1, const float* __restrict__ Pp2, int n2, float* __restrict__ Mx) {
    __shared__ float sh[kT];
    const int b = blockIdx.x;
    float m = -FLT_MAX;
    for (int t = threadIdx.x; t < n2; t += blockDim.x) m = fmaxf(m, fmaxf(Pp1[t], Pp2[(size_t)b * n2 + t]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) Mx[b] = sh[0];
}

// ============================================================== S^2 transforms (FP64)
// ring DFT: fm[q][j][m] = sum_k f[q][j][k] e^{-i m phi_k}, m = 0..2B-1 (negative orders at 2B+m)
__global__ void k_ring_dft(const double* __restrict__ f, const double2* __restrict__ tw, int B, double2* __restrict__ fm) {
    const int n = 2 * B, q = blockIdx.y;
    __shared__ double2 stw[256];
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    __syncthreads();
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n * n; t += gridDim.x * blockDim.x) {
        int j = t / n, m = t % n;
        const double* row = f + ((size_t)q * n + j) * n;
        double re = 0, im = 0;
        for (int k = 0; k < n; k++) {
            double2 w = stw[(m * k) % n];
            re += row[k] * w.x;
            im += row[k] * w.y;
        }
        fm[((size_t)q * n + j) * n + m] = make_double2(re, im);
    }
}
// Legendre quadrature: c[q][l^2+l+m] = sum_j LF[|m|][l][j] fm[q][j][m] ((-1)^m for m < 0)
__global__ void k_legendre_fwd(const double2* __restrict__ fm, const double* __restrict__ LF, int B, double2* __restrict__ c) {
    const int n = 2 * B, q = blockIdx.y;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < B * B; t += gridDim.x * blockDim.x) {
        int l = (int)sqrt((double)t);
        while (l * l > t) l--;
        while ((l + 1) * (l + 1) <= t) l++;
        int m = t - l * l - l;
        int am = m < 0 ? -m : m;
        const double* tab = LF + ((size_t)am * B + l) * n;
        const double2* src = fm + (size_t)q * n * n;
        int col = m < 0 ? n + m : m;
        double re = 0, im = 0;
        for (int j = 0; j < n; j++) {
            double2 z = src[(size_t)j * n + col];
            re += tab[j] * z.x;
            im += tab[j] * z.y;
        }
        double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;
        c[(size_t)q * B * B + t] = make_double2(sg * re, sg * im);
    }
}
// inverse: g[q][j][m] = sum_l c(l,m) Pbar_l^|m|(theta_j) ((-1)^m for m < 0)
__global__ void k_legendre_inv(const double2* __restrict__ c, const double* __restrict__ LI, int B, double2* __restrict__ g) {
    const int n = 2 * B, q = blockIdx.y;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n * n; t += gridDim.x * blockDim.x) {
        int j = t / n, col = t % n;
        int m = col < B ? col : col - n;  // col == B holds order -B: zero
        double re = 0, im = 0;
        if (col != B) {
            int am = m < 0 ? -m : m;
            double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;
            for (int l = am; l < B; l++) {
                double pv = LI[((size_t)am * B + l) * n + j] * sg;
                double2 z = c[(size_t)q * B * B + l * l + l + m];
                re += pv * z.x;
                im += pv * z.y;
            }
        }
        g[((size_t)q * n + j) * n + col] = make_double2(re, im);
    }
}
// f[q][j][k] = Re sum_m g[q][j][m] e^{+i m phi_k}
__global__ void k_ring_idft(const double2* __restrict__ g, const double2* __restrict__ tw, int B, double* __restrict__ f) {
    const int n = 2 * B, q = blockIdx.y;
    __shared__ double2 stw[256];
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    __syncthreads();
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n * n; t += gridDim.x * blockDim.x) {
        int j = t / n, k = t % n;
        const double2* row = g + ((size_t)q * n + j) * n;
        double re = 0;
        for (int m = 0; m < n; m++) {
            double2 w = stw[(m * k) % n];  // e^{-i}, conjugated below
            re += row[m].x * w.x + row[m].y * w.y;
        }
        f[((size_t)q * n + j) * n + k] = re;
    }
}

// ============================================================== SO(3) synthesis
// T1[q][beta][m1][m2] = sum_l sig(l,-m1) (-1)^(m1+m2) conj(pat(l,-m2)) sqrt((2l+1)/8pi^2) d^l_{m1m2}(beta)
__global__ void __launch_bounds__(kT) k_so3_synth(const double2* __restrict__ cs, const double2* __restrict__ cp,
                                                   const double* __restrict__ dtab, const int* __restrict__ doff,
                                                   int d_slice, int B, double2* __restrict__ T1) {
    // instances fastest: the workgroups that run together read the same beta slice of the table (2.8 MB at B = 64)
    // through the L2 instead of one slice each from HBM
    const int q = blockIdx.x, kb = blockIdx.y;
    const int n = 2 * B, W = 2 * B - 1;
    const double2* S = cs + (size_t)q * B * B;
    const double2* Pt = cp + (size_t)q * B * B;
    const double* D = dtab + (size_t)kb * d_slice;
    double2* out = T1 + ((size_t)q * n + kb) * n * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        int c1 = t / n, c2 = t % n;
        int m1 = c1 < B ? c1 : c1 - n, m2 = c2 < B ? c2 : c2 - n;
        double re = 0, im = 0;
        if (c1 != B && c2 != B) {
            int am1 = m1 < 0 ? -m1 : m1, am2 = m2 < 0 ? -m2 : m2;
            int l0 = am1 > am2 ? am1 : am2;
            const double* d = D + doff[(m1 + B - 1) * W + (m2 + B - 1)];
            for (int l = l0; l < B; l++) {
                double2 a = S[l * l + l - m1], b = Pt[l * l + l - m2];
                // a * conj(b)
                double pr = a.x * b.x + a.y * b.y, pi = a.y * b.x - a.x * b.y;
                double dv = d[l - l0];
                re += pr * dv;
                im += pi * dv;
            }
            if ((m1 + m2) & 1) { re = -re; im = -im; }
        }
        out[t] = make_double2(re, im);
    }
}
// T2[q][beta][alpha][m2] = sum_m1 T1[q][beta][m1][m2] e^{-i m1 alpha}
__global__ void k_dft_m1(const double2* __restrict__ T1, const double2* __restrict__ tw, int B, double2* __restrict__ T2) {
    const int n = 2 * B, kb = blockIdx.x, q = blockIdx.y;
    __shared__ double2 stw[256];
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    __syncthreads();
    const double2* in = T1 + ((size_t)q * n + kb) * n * n;
    double2* out = T2 + ((size_t)q * n + kb) * n * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        int a = t / n, m2 = t % n;
        double re = 0, im = 0;
        for (int m1 = 0; m1 < n; m1++) {
            double2 z = in[(size_t)m1 * n + m2], w = stw[(m1 * a) % n];
            re += z.x * w.x - z.y * w.y;
            im += z.x * w.y + z.y * w.x;
        }
        out[t] = make_double2(re, im);
    }
}
// C[q][alpha][beta][gamma] = Re sum_m2 T2[q][beta][alpha][m2] e^{-i m2 gamma}
__global__ void k_dft_m2(const double2* __restrict__ T2, const double2* __restrict__ tw, int B, double* __restrict__ C) {
    const int n = 2 * B, kb = blockIdx.x, q = blockIdx.y;
    __shared__ double2 stw[256];
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    __syncthreads();
    const double2* in = T2 + ((size_t)q * n + kb) * n * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        int a = t / n, g = t % n;
        double re = 0;
        for (int m2 = 0; m2 < n; m2++) {
            double2 z = in[(size_t)a * n + m2], w = stw[(m2 * g) % n];
            re += z.x * w.x - z.y * w.y;
        }
        C[(((size_t)q * n + a) * n + kb) * n + g] = re;
    }
}

// In-place radix-2 FFT (e^{-2 pi i jk/n}) of V vectors of length n = 2^LOGN in LDS, element e of vector v at
// X[v (n + 1) + e]; tw[j] = e^{-2 pi i j / n}
template <int LOGN, int V>
__device__ __forceinline__ void fft_lds_f64_vecs(double2* X, const double2* tw) {
    constexpr int n = 1 << LOGN, LD = n + 1;
    for (int t = threadIdx.x; t < n * V; t += blockDim.x) {
        const int v = t / n, i = t % n;
        const int r = (int)(__brev((unsigned)i) >> (32 - LOGN));
        if (r > i) {
            const double2 a = X[v * LD + i], b = X[v * LD + r];
            X[v * LD + i] = b;
            X[v * LD + r] = a;
        }
    }
    __syncthreads();
    for (int st = 1; st <= LOGN; st++) {
        const int half = 1 << (st - 1);
        for (int t = threadIdx.x; t < V * n / 2; t += blockDim.x) {
            const int v = t / (n / 2), b = t % (n / 2);
            const int j = b % half, k = (b / half) * 2 * half;
            const double2 w = tw[j * (n >> st)];
            const int i0 = v * LD + k + j, i1 = i0 + half;
            const double2 u = X[i0], x = X[i1];
            const double2 y = make_double2(w.x * x.x - w.y * x.y, w.x * x.y + w.y * x.x);
            X[i0] = make_double2(u.x + y.x, u.y + y.y);
            X[i1] = make_double2(u.x - y.x, u.y - y.y);
        }
        __syncthreads();
    }
}

// The two DFTs of the SO(3) synthesis as FFTs through global memory, for slices too large for one workgroup's LDS
// (B = 64: a 128 x 128 FP64 slice is 256 KB).  k_fft_m1: T2[q][beta][alpha][m2] = sum_m1 T1[q][beta][m1][m2]
// e^{-i m1 alpha}, one workgroup per 16 columns m2 of a slice; k_fft_m2: C[q][alpha][beta][gamma] =
// Re sum_m2 T2[q][beta][alpha][m2] e^{-i m2 gamma}, one workgroup per 16 rows alpha.  The same sums as k_dft_m1 and
// k_dft_m2 in n log n operations (the negative orders m = c - n carry the same phases as c on the n-point grid).
static constexpr int kFftV = 16;
template <int LOGN>
__global__ void __launch_bounds__(kT) k_fft_m1(const double2* __restrict__ T1, const double2* __restrict__ tw,
                                               double2* __restrict__ T2) {
    constexpr int n = 1 << LOGN, LD = n + 1, V = kFftV;
    __shared__ double2 X[V * LD];
    __shared__ double2 stw[n];
    const int g = blockIdx.x, kb = blockIdx.y, q = blockIdx.z;
    const size_t base = ((size_t)q * n + kb) * n * n;
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    for (int t = threadIdx.x; t < n * V; t += blockDim.x) {
        const int m1 = t / V, v = t % V;
        X[v * LD + m1] = T1[base + (size_t)m1 * n + g * V + v];
    }
    __syncthreads();
    fft_lds_f64_vecs<LOGN, V>(X, stw);
    for (int t = threadIdx.x; t < n * V; t += blockDim.x) {
        const int a = t / V, v = t % V;
        T2[base + (size_t)a * n + g * V + v] = X[v * LD + a];
    }
}
template <int LOGN>
__global__ void __launch_bounds__(kT) k_fft_m2(const double2* __restrict__ T2, const double2* __restrict__ tw,
                                               double* __restrict__ C) {
    constexpr int n = 1 << LOGN, LD = n + 1, V = kFftV;
    __shared__ double2 X[V * LD];
    __shared__ double2 stw[n];
    const int g = blockIdx.x, kb = blockIdx.y, q = blockIdx.z;
    const size_t base = ((size_t)q * n + kb) * n * n;
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    for (int t = threadIdx.x; t < n * V; t += blockDim.x) {
        const int v = t / n, m2 = t % n;
        X[v * LD + m2] = T2[base + (size_t)(g * V + v) * n + m2];
    }
    __syncthreads();
    fft_lds_f64_vecs<LOGN, V>(X, stw);
    for (int t = threadIdx.x; t < n * V; t += blockDim.x) {
        const int v = t / n, gm = t % n;
        C[(((size_t)q * n + g * V + v) * n + kb) * n + gm] = X[v * LD + gm].x;
    }
}

// In-place radix-2 FFT (e^{-2 pi i jk/n}) of n vectors of length n = 2^LOGN held in LDS: element e of vector v
// at X[v * vs + e * es]; consecutive threads take consecutive vectors.
template <int LOGN>
__device__ __forceinline__ void fft_lds_f64(double2* X, int es, int vs, const double2* tw) {
    constexpr int n = 1 << LOGN;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int v = t % n, i = t / n;
        const int r = (int)(__brev((unsigned)i) >> (32 - LOGN));
        if (r > i) {
            double2 a = X[v * vs + i * es], b = X[v * vs + r * es];
            X[v * vs + i * es] = b;
            X[v * vs + r * es] = a;
        }
    }
    __syncthreads();
    for (int st = 1; st <= LOGN; st++) {
        const int half = 1 << (st - 1);
        for (int t = threadIdx.x; t < n * n / 2; t += blockDim.x) {
            const int v = t % n, b = t / n;
            const int j = b % half, k = (b / half) * 2 * half;
            const double2 w = tw[j * (n >> st)];
            const int i0 = v * vs + (k + j) * es, i1 = i0 + half * es;
            const double2 u = X[i0], x = X[i1];
            const double2 y = make_double2(w.x * x.x - w.y * x.y, w.x * x.y + w.y * x.x);
            X[i0] = make_double2(u.x + y.x, u.y + y.y);
            X[i1] = make_double2(u.x - y.x, u.y - y.y);
        }
        __syncthreads();
    }
}

// Synthesis of one beta slice of one instance and its two DFTs, all in LDS (n = 2B <= 64):
//   T1[m1][m2] = sum_l sig(l,-m1) (-1)^(m1+m2) conj(pat(l,-m2)) sqrt((2l+1)/8pi^2) d^l_{m1m2}(beta)
//   C[q][alpha][beta][gamma] = Re sum_{m1,m2} T1[m1][m2] e^{-i m1 alpha} e^{-i m2 gamma}
// 512 threads per workgroup: the n x (n+1) FP64 LDS tile limits a CU to two workgroups, so the wavefronts that hide
// the synthesis's load latency come from the workgroup size
static constexpr int kSynT = 512;
template <int LOGN>
__global__ void __launch_bounds__(kSynT) k_so3_synth_fft(const double2* __restrict__ cs, const double2* __restrict__ cp,
                                                       const double* __restrict__ dtab2, int B,
                                                       const double2* __restrict__ tw, double* __restrict__ C) {
    constexpr int n = 1 << LOGN, LD = n + 1;
    extern __shared__ double2 X[];  // n * LD
    __shared__ double2 stw[n];
    const int q = blockIdx.x, kb = blockIdx.y;
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    const double2* S = cs + (size_t)q * B * B;
    const double2* Pt = cp + (size_t)q * B * B;
    const double* D = dtab2 + (size_t)kb * n * B * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int c1 = t / n, c2 = t % n;
        double re = 0, im = 0;
        if (c1 != B && c2 != B) {
            const int m1 = c1 < B ? c1 : c1 - n, m2 = c2 < B ? c2 : c2 - n;
            const int am1 = m1 < 0 ? -m1 : m1;
            const double* d = D + (size_t)c1 * B * n + c2;
            // degrees from |m1| rounded down to a multiple of 8, eight at a time with their 24 loads in flight:
            // the table is zero for l < max(|m1|, |m2|), so the extra leading terms add exact zeros (their
            // coefficient indices clamped into the arrays) and the sum keeps its order of l.  The start depends
            // on m1 only, which is uniform over a wavefront when n = 64, so its lanes run the same trip count.
            int l = am1 & ~7;
            if constexpr (LOGN == 6) {
                // n = 64, 512 threads: c1 = t / 64 is uniform over a wavefront, so the signal coefficients of a
                // degree are one scalar load per wavefront; sixteen degrees per round trip from |m1| rounded
                // down to a multiple of 16 (B = 32: at most two round trips per element)
                const int m1u = __builtin_amdgcn_readfirstlane(m1);
                const int am1u = m1u < 0 ? -m1u : m1u;
                for (l = am1u & ~15; l + 15 < B; l += 16) {
                    double2 a[16], b[16];
                    double dv[16];
#pragma unroll
                    for (int u = 0; u < 16; u++) {
                        const int lu = l + u;
                        a[u] = S[max(lu * lu + lu - m1u, 0)];
                        b[u] = Pt[max(lu * lu + lu - m2, 0)];
                        dv[u] = d[(size_t)lu * n];
                    }
#pragma unroll
                    for (int u = 0; u < 16; u++) {
                        const double pr = a[u].x * b[u].x + a[u].y * b[u].y, pi = a[u].y * b[u].x - a[u].x * b[u].y;
                        re += pr * dv[u];
                        im += pi * dv[u];
                    }
                }
            }
            for (; l + 7 < B; l += 8) {
                double2 a[8], b[8];
                double dv[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int lu = l + u;
                    a[u] = S[max(lu * lu + lu - m1, 0)];
                    b[u] = Pt[max(lu * lu + lu - m2, 0)];
                    dv[u] = d[(size_t)lu * n];
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const double pr = a[u].x * b[u].x + a[u].y * b[u].y, pi = a[u].y * b[u].x - a[u].x * b[u].y;
                    re += pr * dv[u];
                    im += pi * dv[u];
                }
            }
            for (; l < B; l++) {
                const double2 a = S[max(l * l + l - m1, 0)], b = Pt[max(l * l + l - m2, 0)];
                const double pr = a.x * b.x + a.y * b.y, pi = a.y * b.x - a.x * b.y;
                const double dv = d[(size_t)l * n];
                re += pr * dv;
                im += pi * dv;
            }
            if ((m1 + m2) & 1) {
                re = -re;
                im = -im;
            }
        }
        X[c1 * LD + c2] = make_double2(re, im);
    }
    __syncthreads();
    fft_lds_f64<LOGN>(X, LD, 1, stw);  // over m1 for every m2 -> alpha
    fft_lds_f64<LOGN>(X, 1, LD, stw);  // over m2 for every alpha -> gamma
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int a = t / n, g = t % n;
        C[(((size_t)q * n + a) * n + kb) * n + g] = X[a * LD + g].x;
    }
}

// The synthesis as matrix products on the FP64 matrix cores.  For one order pair (m1, m2) and a group of up to 64
// instances q the synthesis is
//   T1[m1, m2][beta][q] = (-1)^(m1+m2) sum_l D(beta, l) X(l, q),  X(l, q) = sig_q(l, -m1) conj(pat_q(l, -m2)),
// a (2B x B) by (B x 64) product per pair (real D, complex X: two products).  One workgroup per (pair, instance
// group); wave w owns the 16 beta rows [16 w, 16 w + 16) and up to four 16-column instance tiles, and issues
// v_mfma_f64_16x16x4f64 over l in steps of four from max(|m1|, |m2|) rounded down to a multiple of four (the table is
// zero below, so the extra terms are exact zeros).  Operand layout of the instruction on gfx950: lane L holds
// A[L % 16][L / 16] and B[L / 16][L % 16]; result element i of lane L is D[4 i + L / 16][L % 16].  The coefficients
// arrive instance-fastest (csT, cpT: [coefficient][qp], qp = nq rounded up to 16, zero beyond nq) and T1 is written
// instance-fastest, so the sixteen lanes of a row load and store 256 contiguous bytes.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
template <int B>
__global__ void __launch_bounds__(2 * B * 4) k_so3_synth_mfma(const double2* __restrict__ csT, const double2* __restrict__ cpT,
                                                         const double* __restrict__ Dp, int nqp, double2* __restrict__ T1) {
    constexpr int n = 2 * B;
    const int pair = blockIdx.x, c1 = pair / n, c2 = pair % n;
    const int w = threadIdx.x >> 6, L = threadIdx.x & 63, lr = L & 15, lk = L >> 4;
    const int q0 = blockIdx.y * 64;
    const int ntile = min(4, (nqp - q0) >> 4);  // instance tiles of this group (uniform)
    mfma_d4 re[4], im[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        re[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
        im[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    }
    double sg = 1.0;
    if (c1 != B && c2 != B) {
        const int m1 = c1 < B ? c1 : c1 - n, m2 = c2 < B ? c2 : c2 - n;
        const int am1 = m1 < 0 ? -m1 : m1, am2 = m2 < 0 ? -m2 : m2;
        const int l0 = am1 > am2 ? am1 : am2;
        const double* D = Dp + (size_t)pair * B * n + 16 * w + lr;
        for (int l = l0 & ~3; l < B; l += 4) {
            const int ll = l + lk;
            const double a = D[(size_t)ll * n];
            const double2* sp = csT + (size_t)max(ll * ll + ll - m1, 0) * nqp + q0 + lr;
            const double2* pp = cpT + (size_t)max(ll * ll + ll - m2, 0) * nqp + q0 + lr;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (t < ntile) {
                    const double2 sv = sp[16 * t], pv = pp[16 * t];
                    const double xr = sv.x * pv.x + sv.y * pv.y, xi = sv.y * pv.x - sv.x * pv.y;
                    re[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xr, re[t], 0, 0, 0);
                    im[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xi, im[t], 0, 0, 0);
                }
            }
        }
        if ((m1 + m2) & 1) sg = -1.0;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (t >= ntile) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int kb = 16 * w + 4 * i + lk;
            T1[((size_t)pair * n + kb) * nqp + q0 + 16 * t + lr] = make_double2(sg * re[t][i], sg * im[t][i]);
        }
    }
}

// coefficients [q][c] -> [c][qp] for the matrix-core synthesis, zero for q >= nq (c < BB = B^2)
__global__ void k_coef_transpose(const double2* __restrict__ cs, const double2* __restrict__ cp, int nq, int nqp, int BB,
                                 double2* __restrict__ csT, double2* __restrict__ cpT) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)BB * nqp) return;
    const int c = (int)(t / nqp), q = (int)(t - (long)c * nqp);
    const double2 z = make_double2(0.0, 0.0);
    csT[t] = q < nq ? cs[(size_t)q * BB + c] : z;
    cpT[t] = q < nq ? cp[(size_t)q * BB + c] : z;
}

// the two DFTs of one (instance, beta) slice of T1 in LDS: C[q][alpha][beta][gamma] = Re sum T1 e^{-i m1 alpha}
// e^{-i m2 gamma}.  The slice is gathered from T1[pair][beta][qp] (one 16-byte element per pair); the eight
// instances whose elements share a cache line run back to back on one XCD (workgroup b runs on XCD b mod 8), so
// each line of T1 is fetched into that L2 once: workgroup b takes XCD x = b mod 8, step s = b / 8, line group
// c = 8 (s / 8) + x = beta * Q8 + g (Q8 = ceil(nq / 8) groups of eight instances) and instance q = 8 g + s mod 8.
template <int LOGN>
__global__ void __launch_bounds__(kSynT) k_so3_fft2(const double2* __restrict__ T1, const double2* __restrict__ tw,
                                                   int nq, int nqp, int Q8, double* __restrict__ C) {
    constexpr int n = 1 << LOGN, LD = n + 1;
    extern __shared__ double2 X[];  // n * LD
    __shared__ double2 stw[n];
    const int b = blockIdx.x, s = b >> 3;
    const int c = ((s >> 3) << 3) | (b & 7);
    const int kb = c / Q8, q = (c - kb * Q8) * 8 + (s & 7);
    if (q >= nq) return;  // the whole workgroup: before any barrier
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    const double2* src = T1 + (size_t)kb * nqp + q;
    block_copy<kSynT, n * n>([&](int t) { return src[(size_t)t * n * nqp]; },
                             [&](int t, double2 v) { X[(t / n) * LD + (t % n)] = v; });
    __syncthreads();
    fft_lds_f64<LOGN>(X, LD, 1, stw);  // over m1 for every m2 -> alpha
    fft_lds_f64<LOGN>(X, 1, LD, stw);  // over m2 for every alpha -> gamma
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int a = t / n, g = t % n;
        C[(((size_t)q * n + a) * n + kb) * n + g] = X[a * LD + g].x;
    }
}

// ============================================================== softScaled pieces
// first maximum and first minimum in row-major order (ProjRot.h:1562-1576)
// min / max of each correlation volume and the first position of each (blitz minIndex / maxIndex order):
// kExtParts workgroups per volume scan contiguous parts, k_extrema_final combines them with the same rule
static constexpr int kExtParts = 16;
__global__ void k_extrema(const double* __restrict__ C, long n3, double* __restrict__ pvals, int* __restrict__ ppos) {
    __shared__ double smax[kT], smin[kT];
    __shared__ int imax[kT], imin[kT];
    const int q = blockIdx.x, part = blockIdx.y;
    const double* c = C + (size_t)q * n3;
    const long o0 = n3 * part / kExtParts, o1 = n3 * (part + 1) / kExtParts;
    double bx = -DBL_MAX, bn = DBL_MAX;
    int ix = 0x7fffffff, in = 0x7fffffff;
    for (long o = o0 + threadIdx.x; o < o1; o += blockDim.x) {
        double v = c[o];
        if (v > bx) { bx = v; ix = (int)o; }
        if (v < bn) { bn = v; in = (int)o; }
    }
    smax[threadIdx.x] = bx; imax[threadIdx.x] = ix;
    smin[threadIdx.x] = bn; imin[threadIdx.x] = in;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            int o = threadIdx.x + s;
            if (smax[o] > smax[threadIdx.x] || (smax[o] == smax[threadIdx.x] && imax[o] < imax[threadIdx.x])) {
                smax[threadIdx.x] = smax[o]; imax[threadIdx.x] = imax[o];
            }
            if (smin[o] < smin[threadIdx.x] || (smin[o] == smin[threadIdx.x] && imin[o] < imin[threadIdx.x])) {
                smin[threadIdx.x] = smin[o]; imin[threadIdx.x] = imin[o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t k = (size_t)q * kExtParts + part;
        pvals[k * 2 + 0] = smin[0];
        pvals[k * 2 + 1] = smax[0];
        ppos[k * 2 + 0] = imin[0];
        ppos[k * 2 + 1] = imax[0];
    }
}
__global__ void k_extrema_final(const double* __restrict__ pvals, const int* __restrict__ ppos, int nq,
                                double* __restrict__ vals, int* __restrict__ pos) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    double bn = DBL_MAX, bx = -DBL_MAX;
    int in = 0x7fffffff, ix = 0x7fffffff;
    for (int part = 0; part < kExtParts; part++) {
        const size_t k = (size_t)q * kExtParts + part;
        const double vn = pvals[k * 2 + 0], vx = pvals[k * 2 + 1];
        const int jn = ppos[k * 2 + 0], jx = ppos[k * 2 + 1];
        if (vn < bn || (vn == bn && jn < in)) { bn = vn; in = jn; }
        if (vx > bx || (vx == bx && jx < ix)) { bx = vx; ix = jx; }
    }
    vals[q * 2 + 0] = bn;
    vals[q * 2 + 1] = bx;
    pos[q * 2 + 0] = in;
    pos[q * 2 + 1] = ix;
}

// rotateFct (rotate_so3.c:695-751) on the coefficients of one S^2 function:
// r(l,m) = sum_m' e^{-i m alpha} d^l_{m m'}(beta) e^{-i m' gamma} c(l,m').  `acc` is B (2B-1) doubles2 of LDS.
__device__ void rotate_coef_block(const double2* __restrict__ cc, double alpha, double beta, double gamma, int B,
                                  double2* __restrict__ out, double2* acc) {
    const int W = 2 * B - 1;
    for (int t = threadIdx.x; t < B * W; t += blockDim.x) acc[t] = make_double2(0, 0);
    __syncthreads();
    const double cb = cos(beta);
    for (int m = -(B - 1) + (int)threadIdx.x; m <= B - 1; m += blockDim.x) {
        for (int mp = -(B - 1); mp <= B - 1; mp++) {
            int l0 = max(abs(m), abs(mp));
            double dprev = 0, dcur = wig_seed(m, mp, beta);
            double ph = -(m * alpha + mp * gamma);
            double cr = cos(ph), ci = sin(ph);
            for (int l = l0; l < B; l++) {
                double2 z = cc[l * l + l + mp];
                double Dr = cr * dcur, Di = ci * dcur;
                double2 o = acc[l * W + (m + B - 1)];
                o.x += Dr * z.x - Di * z.y;
                o.y += Dr * z.y + Di * z.x;
                acc[l * W + (m + B - 1)] = o;
                double dn = wig_next(l, m, mp, cb, dcur, dprev);
                dprev = dcur;
                dcur = dn;
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < B * B; t += blockDim.x) {
        int l = (int)sqrt((double)t);
        while (l * l > t) l--;
        while ((l + 1) * (l + 1) <= t) l++;
        int m = t - l * l - l;
        out[t] = acc[l * W + (m + B - 1)];
    }
}

// d^l_{m m'}(beta_b) for l < B, |m|, |m'| <= l at the angles rotateFct receives for the extrema's beta bins,
// beta_b = (double)(float)((2b + 1) pi / 4B) (Pixel angles in the reference); stored [b][l][m + l][m' + l]
// after sum_{l' < l} (2l' + 1)^2 entries per l.  The same recurrence as the direct evaluation.
__device__ __forceinline__ size_t rot_tab_l(int l) { return (size_t)l * (2 * l - 1) * (2 * l + 1) / 3; }
__global__ void __launch_bounds__(kT) k_wigner_rot_table(int B, double* __restrict__ tab) {
    const int b = blockIdx.x, W = 2 * B - 1;
    const double beta = (double)(float)((2 * b + 1.0) / (4. * B) * M_PI);
    const double cb = cos(beta);
    double* T = tab + (size_t)b * rot_tab_l(B);
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        const int m = t / W - (B - 1), mp = t % W - (B - 1);
        const int l0 = max(abs(m), abs(mp));
        double dprev = 0, dcur = wig_seed(m, mp, beta);
        for (int l = l0; l < B; l++) {
            T[rot_tab_l(l) + (size_t)(m + l) * (2 * l + 1) + (mp + l)] = dcur;
            double dn = wig_next(l, m, mp, cb, dcur, dprev);
            dprev = dcur;
            dcur = dn;
        }
    }
}

// k_rotate_coef with the d values from the table: one thread per output (l, m), the sum over m' in increasing
// order as the direct evaluation accumulates it; the phases (cos, sin) of -(m alpha + m' gamma) in LDS
__global__ void __launch_bounds__(kT) k_rotate_coef_tab(const double2* __restrict__ c, const int* __restrict__ pos,
                                                         int B, const double* __restrict__ tab, double2* __restrict__ r) {
    extern __shared__ double2 cs[];  // [m + B - 1][m' + B - 1]
    const int q = blockIdx.x, which = blockIdx.y;
    const int n = 2 * B, W = 2 * B - 1;
    const int p = pos[q * 2 + which];
    const int a = p / (n * n), b = (p / n) % n, g = p % n;
    const double alpha = (double)(float)(a / (double)B * M_PI);
    const double gamma = (double)(float)(g / (double)B * M_PI);
    for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
        const int m = t / W - (B - 1), mp = t % W - (B - 1);
        const double ph = -(m * alpha + mp * gamma);
        cs[t] = make_double2(cos(ph), sin(ph));
    }
    __syncthreads();
    const double* T = tab + (size_t)b * rot_tab_l(B);
    const double2* cc = c + (size_t)q * B * B;
    double2* out = r + ((size_t)q * 2 + which) * B * B;
    for (int t = threadIdx.x; t < B * B; t += blockDim.x) {
        int l = (int)sqrt((double)t);
        while (l * l > t) l--;
        while ((l + 1) * (l + 1) <= t) l++;
        const int m = t - l * l - l;
        const double* d = T + rot_tab_l(l) + (size_t)(m + l) * (2 * l + 1) + l;
        const double2* e = cs + (m + B - 1) * W + (B - 1);
        double2 o = make_double2(0, 0);
        for (int mp = -l; mp <= l; mp++) {
            const double dcur = d[mp];
            const double2 ph = e[mp];
            const double2 z = cc[l * l + l + mp];
            double Dr = ph.x * dcur, Di = ph.y * dcur;
            o.x += Dr * z.x - Di * z.y;
            o.y += Dr * z.y + Di * z.x;
        }
        out[t] = o;
    }
}

// the pattern's coefficients rotated by the Euler angles of the extrema of C (rotatedSphericalImageMinMax,
// ProjRot.h:1558-1589).  Block (q, which): which = 0 min, 1 max.
__global__ void k_rotate_coef(const double2* __restrict__ c, const int* __restrict__ pos, int B, double2* __restrict__ r) {
    extern __shared__ double2 acc[];  // [l][m + B - 1]
    const int q = blockIdx.x, which = blockIdx.y;
    const int n = 2 * B;
    const int p = pos[q * 2 + which];
    const int a = p / (n * n), b = (p / n) % n, g = p % n;
    // angles are Pixel (float) in the reference
    const double alpha = (double)(float)(a / (double)B * M_PI);
    const double beta = (double)(float)((2 * b + 1.0) / (4. * B) * M_PI);
    const double gamma = (double)(float)(g / (double)B * M_PI);
    rotate_coef_block(c + (size_t)q * B * B, alpha, beta, gamma, B, r + ((size_t)q * 2 + which) * B * B, acc);
}

// coefficients of image q rotated by the angles ang[q*3 .. q*3+2] (harmonics(A, alpha, beta, gamma, R),
// ProjRot.h:1457-1554)
__global__ void k_rotate_coef_angles(const double2* __restrict__ c, const double* __restrict__ ang, int B,
                                     double2* __restrict__ r) {
    extern __shared__ double2 acc[];
    const int q = blockIdx.x;
    rotate_coef_block(c + (size_t)q * B * B, ang[q * 3], ang[q * 3 + 1], ang[q * 3 + 2], B, r + (size_t)q * B * B, acc);
}

// wMin / wMax = sum F1 * R * sin(theta_i) and the affine rescale of C (ProjRot.h:1303-1318)
__global__ void k_wsum(const double* __restrict__ F1, const double* __restrict__ Rr, int B, double* __restrict__ wv) {
    __shared__ double sh[kT];
    const int q = blockIdx.x, which = blockIdx.y;
    const int n = 2 * B;
    const double* f = F1 + (size_t)q * n * n;
    const double* rr = Rr + ((size_t)q * 2 + which) * n * n;
    double acc = 0;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        int i = t / n;
        acc += f[t] * rr[t] * sin(M_PI * (2 * i + 1) / 4 / (double)(float)B);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) wv[q * 2 + which] = sh[0];
}
__global__ void k_rescale(double* __restrict__ C, long n3, const double* __restrict__ vals, const double* __restrict__ wv,
                          const SoftInst* __restrict__ inst) {
    const int q = blockIdx.y;
    if (!inst[q].scaled) return;
    const double mn = vals[q * 2], mx = vals[q * 2 + 1], wmin = wv[q * 2], wmax = wv[q * 2 + 1];
    double* c = C + (size_t)q * n3;
    for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n3; o += (long)gridDim.x * blockDim.x)
        c[o] = (c[o] - mn) / (mx - mn) * (wmax - wmin) + wmin;
}

// ============================================================== branch combination, restriction, local maxima
// out[p] = op(C[i0], C[i1], ...) per branch (ProjRot.h:526-584); restriction of beta rows (:602)
__global__ void k_combine(const double* __restrict__ C, const ComboOp* __restrict__ ops, long n3, int n, int b0,
                          double* __restrict__ out) {
    const int p = blockIdx.y;
    const ComboOp op = ops[p];
    double* o = out + (size_t)p * n3;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n3; t += (long)gridDim.x * blockDim.x) {
        int b = (int)((t / n) % n);
        double v;
        const double* c0 = C + (size_t)op.i[0] * n3;
        switch (op.kind) {
            case CO_COPY: v = c0[t]; break;
            case CO_DIV_SQRT: v = c0[t] / sqrt(C[(size_t)op.i[1] * n3 + t]); break;
            case CO_BOTH: {
                double a = sqrt(c0[t]) * sqrt(C[(size_t)op.i[1] * n3 + t]);
                a = 2.0 * (1.0 - C[(size_t)op.i[2] * n3 + t] / a);
                v = -a / C[(size_t)op.i[3] * n3 + t];
                break;
            }
            default: v = 0;
        }
        if (b >= b0) v = -(double)FLT_MAX;
        o[t] = v;
    }
}

// 26-neighbour >= test with a -FLT_MAX border (localMaxima3D3x3 + BordStrategyConst, ProjRot.h:67-79, 605-610);
// nonzero results are appended to a per-volume list (order restored on the host by sorting the index)
// local maxima of each correlation volume (26 neighbours, none outside the volume; ProjRot.h:67-79): one 8^3 tile
// per workgroup, staged with its one-voxel halo in LDS (a neighbour outside the volume reads -FLT_MAX)
static constexpr int kLmT = 8, kLmH = kLmT + 2;
__global__ void __launch_bounds__(kT) k_localmax_tiled(const double* __restrict__ C, int n, int cap,
                                                       int* __restrict__ count, int* __restrict__ idx,
                                                       double* __restrict__ val) {
    __shared__ double tile[kLmH * kLmH * kLmH];
    __shared__ int s_idx[kLmT * kLmT * kLmT];
    __shared__ double s_val[kLmT * kLmT * kLmT];
    __shared__ int s_n, s_base;
    const int p = blockIdx.y, nt = n / kLmT, tb = blockIdx.x;
    if (threadIdx.x == 0) s_n = 0;
    const int a0 = (tb / (nt * nt)) * kLmT, b0 = ((tb / nt) % nt) * kLmT, g0 = (tb % nt) * kLmT;
    const double* c = C + (size_t)p * n * n * n;
    block_copy<kT, kLmH * kLmH * kLmH>(
        [&](int t) {
            const int aa = a0 + t / (kLmH * kLmH) - 1, bb = b0 + (t / kLmH) % kLmH - 1, gg = g0 + t % kLmH - 1;
            const bool in = aa >= 0 && bb >= 0 && gg >= 0 && aa < n && bb < n && gg < n;
            const double x = c[in ? ((size_t)aa * n + bb) * n + gg : ((size_t)a0 * n + b0) * n + g0];
            return in ? x : -(double)FLT_MAX;
        },
        [&](int t, double x) { tile[t] = x; });
    __syncthreads();
    for (int t = threadIdx.x; t < kLmT * kLmT * kLmT; t += kT) {
        const int la = t / (kLmT * kLmT), lb = (t / kLmT) % kLmT, lg = t % kLmT;
        const int ci = ((la + 1) * kLmH + lb + 1) * kLmH + lg + 1;
        const double v = tile[ci];
        bool ok = true;
#pragma unroll
        for (int da = -1; da <= 1; da++)
#pragma unroll
            for (int db = -1; db <= 1; db++)
#pragma unroll
                for (int dg = -1; dg <= 1; dg++) {
                    if (!da && !db && !dg) continue;
                    ok = ok && (v >= tile[ci + (da * kLmH + db) * kLmH + dg]);
                }
        if (ok && v != 0.0) {
            const int k = atomicAdd(&s_n, 1);
            s_idx[k] = ((a0 + la) * n + b0 + lb) * n + g0 + lg;
            s_val[k] = v;
        }
    }
    // one global append per workgroup
    __syncthreads();
    if (threadIdx.x == 0 && s_n > 0) s_base = atomicAdd(&count[(size_t)p * kLmCountStride], s_n);
    __syncthreads();
    for (int k = threadIdx.x; k < s_n; k += kT) {
        const int slot = s_base + k;
        if (slot < cap) {
            idx[(size_t)p * cap + slot] = s_idx[k];
            val[(size_t)p * cap + slot] = s_val[k];
        }
    }
}

// the same predicate for volume edges that are not a multiple of the tile
__global__ void k_localmax(const double* __restrict__ C, int n, int cap, int* __restrict__ count,
                           int* __restrict__ idx, double* __restrict__ val) {
    const int p = blockIdx.y;
    const long n3 = (long)n * n * n;
    const double* c = C + (size_t)p * n3;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n3; t += (long)gridDim.x * blockDim.x) {
        int a = (int)(t / ((long)n * n)), b = (int)((t / n) % n), g = (int)(t % n);
        const double v = c[t];
        // all 26 neighbours loaded before the first comparison (one memory latency); a neighbour outside the
        // volume counts as -FLT_MAX (its load reads the centre and is discarded)
        double nv[26];
        unsigned inb = 0;
        int q = 0;
#pragma unroll
        for (int da = -1; da <= 1; da++)
#pragma unroll
            for (int db = -1; db <= 1; db++)
#pragma unroll
                for (int dg = -1; dg <= 1; dg++) {
                    if (!da && !db && !dg) continue;
                    const int aa = a + da, bb = b + db, gg = g + dg;
                    const bool in = aa >= 0 && bb >= 0 && gg >= 0 && aa < n && bb < n && gg < n;
                    nv[q] = c[in ? ((size_t)aa * n + bb) * n + gg : (size_t)t];
                    inb |= (in ? 1u : 0u) << q;
                    q++;
                }
        bool ok = true;
#pragma unroll
        for (int r = 0; r < 26; r++) {
            const double w = ((inb >> r) & 1u) ? nv[r] : -(double)FLT_MAX;
            ok = ok && (v >= w);
        }
        if (ok && v != 0.0) {
            int slot = atomicAdd(&count[(size_t)p * kLmCountStride], 1);
            if (slot < cap) {
                idx[(size_t)p * cap + slot] = (int)t;
                val[(size_t)p * cap + slot] = v;
            }
        }
    }
}

// ============================================================== host drivers
void so3_state_destroy(ettk_plan* P) {
    delete P->so3;
    P->so3 = nullptr;
}

So3State& so3_state(ettk_plan* P);
static So3State& so3(ettk_plan* P) { return so3_state(P); }
So3State& so3_state(ettk_plan* P) {
    if (!P->so3) {
        P->so3 = new So3State();
        P->so3->B = P->p.B;
        so3_tables(P);
        so3_kernel_attributes(P->p.B);
    }
    return *P->so3;
}

static int band_lo(const ettk_plan* P) {
    const ettk_params& p = P->p;
    bool bp = (p.bp_lo > 0.0f) || (p.bp_hi < 1.0f);
    float lo = bp ? p.bp_lo : 0.0f;
    return std::max(4, (int)std::ceil(lo * (2 * P->N - 1)));
}
static int band_hi(const ettk_plan* P) {
    const ettk_params& p = P->p;
    bool bp = (p.bp_lo > 0.0f) || (p.bp_hi < 1.0f);
    float hi = bp ? p.bp_hi : 1.0f;
    return (int)std::ceil(std::min(0.5f, hi) * (2 * P->N - 1));
}

// projections of `n` images: is_ref -> R already holds the internal image (no reslice)
void project_images(ettk_plan* P, const float* d_parts, const ettk_particle* descs, const int* pidx, int n, bool is_ref,
                    float* d_Pp, const EvalGeom* h_geom) {
    So3State& S = so3(P);
    const int N = P->N, M = 2 * N, B = S.B, Hm = N + 1;
    const size_t n3 = (size_t)N * N * N, nbig = (size_t)M * M * Hm;
    hipStream_t s = P->stream;
    const int CH = 32;
    S.projR.ensure(n3 * CH);
    S.projF.ensure(nbig * CH);
    S.projMag.ensure(nbig * CH);
    S.projStats.ensure(4 * CH);
    S.projSum.ensure((size_t)kProjParts * CH);
    int nb_res = reslice_blocks(N);
    S.projPartials.ensure((size_t)2 * nb_res * CH);
    S.projGeom.ensure(CH);
    std::vector<EvalGeom> g(CH);
    for (int c0 = 0; c0 < n; c0 += CH) {
        int cn = std::min(CH, n - c0);
        if (is_ref) {
            ETTK_HIP(hipMemcpyAsync(S.projR.p, P->ref.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
        } else {
            const EvalGeom* src = h_geom ? h_geom + c0 : g.data();
            if (!h_geom) {
                for (int q = 0; q < cn; q++) {
                    int pi = pidx ? pidx[c0 + q] : c0 + q;
                    g[q] = make_geom_public(P, descs[pi], Mat4::ident(), pi);
                }
            }
            ETTK_HIP(hipMemcpyAsync(S.projGeom.p, src, cn * sizeof(EvalGeom), hipMemcpyHostToDevice, s));
            if (!h_geom) ETTK_HIP(hipStreamSynchronize(s));  // g is reused by the next sub-chunk
            launch_reslice_projection(d_parts, S.projGeom.p, cn, N, S.projR.p, S.projPartials.p, nb_res, s);
        }
        k_proj_partial<<<dim3(kProjParts, cn), kT, 0, s>>>(S.projR.p, S.projPartials.p, nb_res, N, is_ref ? 1 : 0,
                                                           S.projSum.p, S.projStats.p);
        k_proj_final<<<cn, 64, 0, s>>>(S.projR.p, S.projSum.p, P->m1.p, N, P->p.mask_radius != 0.0f ? 1 : 0,
                                       S.projStats.p);
        ETTK_HIP(hipGetLastError());
        R2CArgs a{S.projR.p, S.projF.p, M, cn, PRO_PAD_MASKED, P->m1.p, S.projStats.p, N};
        fft3_r2c(P->ctx->fft, a, nullptr, s);
        k_magnitude<<<2048, kT, 0, s>>>(S.projF.p, S.projMag.p, (long)nbig * cn, P->p.mcf);
        ETTK_HIP(hipGetLastError());
        dim3 grid((4 * B * B + kT / kLanes - 1) / (kT / kLanes), cn);
        ProfScope ps(P->ctx->fft.prof, "k_sph_sample", (double)cn * ((double)nbig * 4 + 16.0 * B * B), s);
        k_sph_sample<<<grid, kT, 0, s>>>(S.projMag.p, S.projStats.p, M, B, band_lo(P), band_hi(P),
                                         d_Pp + (size_t)c0 * 4 * B * B);
        ETTK_HIP(hipGetLastError());
    }
}

void so3_reference_set(ettk_plan* P) {
    So3State& S = so3(P);
    const int B = S.B;
    S.Pp1.ensure((size_t)4 * B * B);
    project_images(P, nullptr, nullptr, nullptr, 1, true, S.Pp1.p);
    // spherical wedge of the reference (host, FP64), conditioned: analytic, or for an average its binarised
    // accumulated spherical wedge (all ones when the accumulated wedge is not effective, WAvg.h:670-690)
    std::vector<float> W1;
    if (P->ref_wedge_kind == 0) {
        W1 = spherical_wedge_host(B, P->refd, P->p.missing_wedge != 0);
    } else {
        W1 = P->ref_wedge_kind == 1 ? P->sph_acc_bin : std::vector<float>((size_t)4 * B * B, 1.0f);
        if (P->p.missing_wedge) smooth3(W1, 2 * B, 2 * B, 1, 2.0, 1.0);
    }
    S.W1.upload(W1.data(), W1.size(), P->stream);
    double sw = 0;
    for (float v : W1) sw += v;
    S.use_window1 = P->p.missing_wedge && (sw <= 0.95f * (float)W1.size());
    ETTK_HIP(hipStreamSynchronize(P->stream));
}

// soft / softScaled for a list of instances whose samples are already in S.sig / S.pat; C is [q][a][b][g]
void soft_core(ettk_plan* P, const std::vector<SoftInst>& inst, double* d_C) {
    So3State& S = so3(P);
    const int B = S.B, n = 2 * B, nq = (int)inst.size();
    const long n3 = (long)n * n * n;
    hipStream_t s = P->stream;
    S.fm.ensure((size_t)nq * n * n);
    S.cs.ensure((size_t)nq * B * B);
    S.cp.ensure((size_t)nq * B * B);
    S.vals.ensure((size_t)2 * nq);
    S.pos.ensure((size_t)2 * nq);
    S.rot.ensure((size_t)2 * nq * B * B);
    S.g.ensure((size_t)2 * nq * n * n);
    S.Rr.ensure((size_t)2 * nq * n * n);
    S.wv.ensure((size_t)2 * nq);
    k_ring_dft<<<dim3(16, nq), kT, 0, s>>>(S.sig.p, S.tw.p, B, S.fm.p);
    k_legendre_fwd<<<dim3(8, nq), kT, 0, s>>>(S.fm.p, S.legF.p, B, S.cs.p);
    k_ring_dft<<<dim3(16, nq), kT, 0, s>>>(S.pat.p, S.tw.p, B, S.fm.p);
    k_legendre_fwd<<<dim3(8, nq), kT, 0, s>>>(S.fm.p, S.legF.p, B, S.cp.p);
    if (mfma_synthesis(B)) {
        const int nqp = (nq + 15) & ~15, Q8 = (nq + 7) / 8, BB = B * B;
        S.csT.ensure((size_t)BB * nqp);
        S.cpT.ensure((size_t)BB * nqp);
        S.T1.ensure((size_t)nqp * n3);
        k_coef_transpose<<<(unsigned)(((long)BB * nqp + kT - 1) / kT), kT, 0, s>>>(S.cs.p, S.cp.p, nq, nqp, BB, S.csT.p,
                                                                                  S.cpT.p);
        {
            // algorithmic bytes: the coefficients and the T1 slices written
            ProfScope ps(P->ctx->fft.prof, "k_so3_synth_mfma", (double)nq * ((double)n3 * 16 + 32.0 * B * B), s);
            const dim3 grid(n * n, (nqp + 63) / 64);
            if (B == 16) k_so3_synth_mfma<16><<<grid, 128, 0, s>>>(S.csT.p, S.cpT.p, S.dtabP.p, nqp, S.T1.p);
            else k_so3_synth_mfma<32><<<grid, 256, 0, s>>>(S.csT.p, S.cpT.p, S.dtabP.p, nqp, S.T1.p);
        }
        {
            ProfScope ps(P->ctx->fft.prof, "k_so3_fft2", (double)nq * n3 * (16 + 8), s);
            const size_t shm = (size_t)n * (n + 1) * sizeof(double2);
            const unsigned nb = (unsigned)(Q8 * 8 * n);  // Q8 n is a multiple of 8 (n = 32, 64)
            if (B == 16) k_so3_fft2<5><<<nb, kSynT, shm, s>>>(S.T1.p, S.tw.p, nq, nqp, Q8, d_C);
            else k_so3_fft2<6><<<nb, kSynT, shm, s>>>(S.T1.p, S.tw.p, nq, nqp, Q8, d_C);
        }
    } else if (fused_synthesis(B)) {
        ProfScope ps(P->ctx->fft.prof, "k_so3_synth_fft", (double)nq * ((double)n3 * 8 + 32.0 * B * B), s);
        const size_t shm = (size_t)n * (n + 1) * sizeof(double2);
        switch (n) {
            case 16: k_so3_synth_fft<4><<<dim3(nq, n), kSynT, shm, s>>>(S.cs.p, S.cp.p, S.dtab2.p, B, S.tw.p, d_C); break;
            case 32: k_so3_synth_fft<5><<<dim3(nq, n), kSynT, shm, s>>>(S.cs.p, S.cp.p, S.dtab2.p, B, S.tw.p, d_C); break;
            default: k_so3_synth_fft<6><<<dim3(nq, n), kSynT, shm, s>>>(S.cs.p, S.cp.p, S.dtab2.p, B, S.tw.p, d_C); break;
        }
    } else {
        S.T1.ensure((size_t)nq * n3);
        S.T2.ensure((size_t)nq * n3);
        k_so3_synth<<<dim3(nq, n), kT, 0, s>>>(S.cs.p, S.cp.p, S.dtab.p, S.doff.p, S.d_slice, B, S.T1.p);
        if (n == 128) {
            const dim3 grid(n / kFftV, n, nq);
            k_fft_m1<7><<<grid, kT, 0, s>>>(S.T1.p, S.tw.p, S.T2.p);
            k_fft_m2<7><<<grid, kT, 0, s>>>(S.T2.p, S.tw.p, d_C);
        } else {
            k_dft_m1<<<dim3(n, nq), kT, 0, s>>>(S.T1.p, S.tw.p, B, S.T2.p);
            k_dft_m2<<<dim3(n, nq), kT, 0, s>>>(S.T2.p, S.tw.p, B, d_C);
        }
    }
    ETTK_HIP(hipGetLastError());
    bool any_scaled = false;
    for (auto& q : inst) any_scaled |= q.scaled != 0;
    if (any_scaled) {
        {
            ProfScope ps(P->ctx->fft.prof, "k_extrema", (double)nq * n3 * 8, s);
            S.pvals.ensure((size_t)2 * kExtParts * nq);
            S.ppos.ensure((size_t)2 * kExtParts * nq);
            k_extrema<<<dim3(nq, kExtParts), kT, 0, s>>>(d_C, n3, S.pvals.p, S.ppos.p);
            k_extrema_final<<<(nq + 63) / 64, 64, 0, s>>>(S.pvals.p, S.ppos.p, nq, S.vals.p, S.pos.p);
        }
        {
            ProfScope ps(P->ctx->fft.prof, "k_rotate_coef", (double)2 * nq * (B * B * 32.0), s);
            if (B <= 32) {
                // table of d at the 2B beta bins (22 MB at B = 32), phases in LDS ((2B-1)^2 complex)
                if (S.rtab.n == 0) {
                    const size_t per_b = (size_t)B * (2 * B - 1) * (2 * B + 1) / 3;
                    S.rtab.ensure(per_b * 2 * B);
                    k_wigner_rot_table<<<2 * B, kT, 0, s>>>(B, S.rtab.p);
                }
                const size_t shm = (size_t)(2 * B - 1) * (2 * B - 1) * sizeof(double2);
                k_rotate_coef_tab<<<dim3(nq, 2), kT, shm, s>>>(S.cp.p, S.pos.p, B, S.rtab.p, S.rot.p);
            } else {
                const size_t shm = (size_t)B * (2 * B - 1) * sizeof(double2);
                k_rotate_coef<<<dim3(nq, 2), kT, shm, s>>>(S.cp.p, S.pos.p, B, S.rot.p);
            }
        }
        k_legendre_inv<<<dim3(16, 2 * nq), kT, 0, s>>>(S.rot.p, S.legI.p, B, S.g.p);
        k_ring_idft<<<dim3(16, 2 * nq), kT, 0, s>>>(S.g.p, S.tw.p, B, S.Rr.p);
        k_wsum<<<dim3(nq, 2), kT, 0, s>>>(S.sig.p, S.Rr.p, B, S.wv.p);
        k_rescale<<<dim3(256, nq), kT, 0, s>>>(d_C, n3, S.vals.p, S.wv.p, S.inst.p);
        ETTK_HIP(hipGetLastError());
    }
}

void so3_search_reserve(ettk_plan* P, int cn) {
    So3State& S = so3(P);
    const int B = S.B, n = 2 * B;
    const size_t n3 = (size_t)n * n * n, nq = (size_t)5 * cn, nn = (size_t)n * n;
    S.inst.ensure(nq);
    S.sig.ensure(nq * nn);
    S.pat.ensure(nq * nn);
    S.fm.ensure(nq * nn);
    S.cs.ensure(nq * B * B);
    S.cp.ensure(nq * B * B);
    S.vals.ensure(2 * nq);
    S.pos.ensure(2 * nq);
    S.rot.ensure(2 * nq * B * B);
    S.g.ensure(2 * nq * nn);
    S.Rr.ensure(2 * nq * nn);
    S.wv.ensure(2 * nq);
    S.pvals.ensure((size_t)2 * kExtParts * nq);
    S.ppos.ensure((size_t)2 * kExtParts * nq);
    if (!fused_synthesis(B)) {
        S.T1.ensure(nq * n3);
        S.T2.ensure(nq * n3);
    }
    S.Pp2.ensure((size_t)cn * nn);
    S.Mx.ensure(cn);
    S.W2.ensure((size_t)cn * nn);
    S.C.ensure(nq * n3);
    S.Cc.ensure((size_t)2 * cn * n3);
    S.ops.ensure((size_t)2 * cn);
}

void run_soft(ettk_plan* P, const std::vector<SoftInst>& inst, const float* d_Pp2, const float* d_W2, const float* d_Mx,
              double* d_C, const SoftInst* h_inst) {
    So3State& S = so3(P);
    const int B = S.B, n = 2 * B, nq = (int)inst.size();
    hipStream_t s = P->stream;
    S.inst.ensure(nq);
    ETTK_HIP(hipMemcpyAsync(S.inst.p, h_inst ? h_inst : inst.data(), nq * sizeof(SoftInst), hipMemcpyHostToDevice, s));
    S.sig.ensure((size_t)nq * n * n);
    S.pat.ensure((size_t)nq * n * n);
    k_assemble<<<dim3(16, nq), kT, 0, s>>>(S.inst.p, S.Pp1.p, d_Pp2, S.W1.p, d_W2, d_Mx, B, S.sig.p, S.pat.p);
    ETTK_HIP(hipGetLastError());
    soft_core(P, inst, d_C);
}

// Euler angles of harmonics(A, t, R) (ProjRot.h:1359-1448): psi, phi from atan of element ratios of the rotation
// part of t, theta = acos(M22), the quadrant and the sign of theta chosen from the signs of M21, M12 and
// cos(psi) M20, in Pixel (float) arithmetic; returns the rotateFct angles (-psi, -theta, -phi)
static void harmonics_angles(const double* t, double ang[3]) {
    const double m20 = t[8], m21 = t[9], m12 = t[6], m02 = t[2], m22 = t[10];
    const double pmin = (double)FLT_MIN;
    float psi, phi, theta;
    if (std::fabs(m20) < pmin) {
        if (std::fabs(m21) < pmin) psi = 0;
        else if (m21 < -pmin) psi = (float)(M_PI / 2.0);
        else psi = (float)(-M_PI / 2.0);
    } else {
        psi = (float)std::atan(-m21 / m20);
    }
    if (std::fabs(m02) < pmin) {
        if (std::fabs(m12) < pmin) phi = 0;
        else if (m12 > pmin) phi = (float)(M_PI / 2.0);
        else phi = (float)(-M_PI / 2.0);
    } else {
        phi = (float)std::atan(m12 / m02);
    }
    theta = (float)std::acos(m22);
    const float psi_1 = psi, psi_2 = psi > 0 ? (float)(psi - M_PI) : (float)(psi + M_PI);
    const float phi_1 = phi, phi_2 = phi > 0 ? (float)(phi - M_PI) : (float)(phi + M_PI);
    psi = m21 * psi_1 > 0 ? psi_1 : psi_2;
    phi = m12 * phi_1 > 0 ? phi_1 : phi_2;
    if (std::cos(psi) * m20 > 0) {
        theta = -theta;
        psi = m21 * psi_1 < 0 ? psi_1 : psi_2;
        phi = m12 * phi_1 < 0 ? phi_1 : phi_2;
    }
    ang[0] = -(double)psi;
    ang[1] = -(double)theta;
    ang[2] = -(double)phi;
}

// harmonics(A, t, R) for n (2B x 2B) float images: S^2 forward transform, rotateFct(B, B, B-1) of the
// coefficients, inverse transform; the real part is returned as float.  (The reference's t == NULL case, R = A,
// is left to the caller: every image here goes through the transform, an identity t by zero angles.)
void sph_rotate(ettk_plan* P, int n, const float* h_img, const double* h_t, float* h_out) {
    So3State& S = so3(P);
    const int B = S.B, ns = 2 * B;
    const size_t nn = (size_t)ns * ns;
    hipStream_t s = P->stream;
    std::vector<double> img((size_t)n * nn), ang((size_t)n * 3);
    for (size_t q = 0; q < img.size(); q++) img[q] = (double)h_img[q];
    for (int i = 0; i < n; i++) harmonics_angles(h_t + (size_t)i * 16, &ang[(size_t)i * 3]);
    DevBuf<double> f, a, out;
    DevBuf<double2> fm, c, r, g;
    f.upload(img.data(), img.size(), s);
    a.upload(ang.data(), ang.size(), s);
    fm.ensure((size_t)n * nn);
    c.ensure((size_t)n * B * B);
    r.ensure((size_t)n * B * B);
    g.ensure((size_t)n * nn);
    out.ensure((size_t)n * nn);
    k_ring_dft<<<dim3(16, n), kT, 0, s>>>(f.p, S.tw.p, B, fm.p);
    k_legendre_fwd<<<dim3(8, n), kT, 0, s>>>(fm.p, S.legF.p, B, c.p);
    k_rotate_coef_angles<<<n, kT, (size_t)B * (2 * B - 1) * sizeof(double2), s>>>(c.p, a.p, B, r.p);
    k_legendre_inv<<<dim3(16, n), kT, 0, s>>>(r.p, S.legI.p, B, g.p);
    k_ring_idft<<<dim3(16, n), kT, 0, s>>>(g.p, S.tw.p, B, out.p);
    ETTK_HIP(hipGetLastError());
    std::vector<double> h((size_t)n * nn);
    ETTK_HIP(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    ETTK_HIP(hipStreamSynchronize(s));
    for (size_t q = 0; q < h.size(); q++) h_out[q] = (float)h[q];
}

void soft_given(ettk_plan* P, int nq, const double* h_sig, const double* h_pat, double* d_C, int scaled) {
    So3State& S = so3(P);
    const int B = S.B, n = 2 * B;
    hipStream_t s = P->stream;
    std::vector<SoftInst> inst(nq, SoftInst{0, 0, 0, scaled ? 1 : 0});
    S.inst.ensure(nq);
    ETTK_HIP(hipMemcpyAsync(S.inst.p, inst.data(), nq * sizeof(SoftInst), hipMemcpyHostToDevice, s));
    S.sig.upload(h_sig, (size_t)nq * n * n, s);
    S.pat.upload(h_pat, (size_t)nq * n * n, s);
    soft_core(P, inst, d_C);
}

// The dynamic-LDS limit is a per-function, process-wide attribute: plans of different bandwidths (run from
// several host threads by the size buckets) only ever raise it, so a smaller plan never lowers it under a
// larger one's launches.
static void so3_kernel_attributes(int B) {
    const size_t shm = (size_t)B * (2 * B - 1) * sizeof(double2);
    ensure_dynamic_lds((const void*)k_rotate_coef, shm);
    ensure_dynamic_lds((const void*)k_rotate_coef_angles, shm);
    if (B <= 32) ensure_dynamic_lds((const void*)k_rotate_coef_tab, (size_t)(2 * B - 1) * (2 * B - 1) * sizeof(double2));
    if (fused_synthesis(B)) {
        const int n = 2 * B;
        const void* f = n == 16 ? (const void*)k_so3_synth_fft<4> : n == 32 ? (const void*)k_so3_synth_fft<5>
                                                                            : (const void*)k_so3_synth_fft<6>;
        ensure_dynamic_lds(f, (size_t)n * (n + 1) * sizeof(double2));
    }
    if (mfma_synthesis(B)) {
        const int n = 2 * B;
        const void* f = n == 32 ? (const void*)k_so3_fft2<5> : (const void*)k_so3_fft2<6>;
        ensure_dynamic_lds(f, (size_t)n * (n + 1) * sizeof(double2));
    }
}

void launch_combine(const double* C, const ComboOp* d_ops, int np, int B, int b0, double* out, hipStream_t s) {
    const int n = 2 * B;
    k_combine<<<dim3(256, np), kT, 0, s>>>(C, d_ops, (long)n * n * n, n, b0, out);
    ETTK_HIP(hipGetLastError());
}
void launch_localmax(const double* C, int np, int B, int cap, int* count, int* idx, double* val, hipStream_t s) {
    const int n = 2 * B;
    if (n % kLmT == 0) {
        const int nt = n / kLmT;
        k_localmax_tiled<<<dim3(nt * nt * nt, np), kT, 0, s>>>(C, n, cap, count, idx, val);
    } else {
        k_localmax<<<dim3(256, np), kT, 0, s>>>(C, n, cap, count, idx, val);
    }
    ETTK_HIP(hipGetLastError());
}
// block v: offset of list v = sum of the clamped counts before it; copies its entries in slot order
// packs the per-volume lists back to back for one download: head = [offsets (np + 1), raw counts (np), modes (np),
// partner offsets (np + 1)]; a volume of mode 0 contributes its kept[v] sorted valid maxima with their partner
// counts and its nadj[v] partners, one of mode 1 its raw list of maxima
__global__ void k_lm_compact(const int* __restrict__ count, const int* __restrict__ idx, const double* __restrict__ val,
                             const int* __restrict__ kept, const int* __restrict__ mode, const int* __restrict__ deg,
                             const int* __restrict__ nadj, const int* __restrict__ adj, int np, int cap,
                             int* __restrict__ head, int* __restrict__ pidx, double* __restrict__ pval,
                             int* __restrict__ pdeg, int* __restrict__ padj) {
    const int v = blockIdx.x;
    auto packed = [&](int u) { return mode[u] ? min(count[(size_t)u * kLmCountStride], cap) : kept[u]; };
    auto partners = [&](int u) { return mode[u] ? 0 : nadj[u]; };
    int off = 0, aoff = 0;
    for (int u = 0; u < v; u++) {
        off += packed(u);
        aoff += partners(u);
    }
    const int c = packed(v), na = partners(v), m = mode[v];
    if (threadIdx.x == 0) {
        head[v] = off;
        head[np + 1 + v] = count[(size_t)v * kLmCountStride];
        head[2 * np + 1 + v] = m;
        head[3 * np + 1 + v] = aoff;
        if (v == np - 1) {
            head[np] = off + c;
            head[4 * np + 1] = aoff + na;
        }
    }
    for (int k = threadIdx.x; k < c; k += blockDim.x) {
        pidx[off + k] = idx[(size_t)v * cap + k];
        pval[off + k] = val[(size_t)v * cap + k];
        pdeg[off + k] = m ? 0 : deg[(size_t)v * kCandMax + k];
    }
    if (m || na == 0) return;
    // the partner rows back to back: row offsets by a running scan over blocks of rows
    __shared__ int sc[kT];
    int base = aoff;
    for (int r0 = 0; r0 < c; r0 += blockDim.x) {
        const int r = r0 + threadIdx.x;
        const int d = r < c ? deg[(size_t)v * kCandMax + r] : 0;
        sc[threadIdx.x] = d;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {  // inclusive scan in LDS
            const int y = threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
            __syncthreads();
            sc[threadIdx.x] += y;
            __syncthreads();
        }
        const int start = base + sc[threadIdx.x] - d;
        const int* src = adj + ((size_t)v * kCandMax + (r < c ? r : 0)) * kCandSlots;
        for (int k = 0; k < d; k++) padj[start + k] = src[k];
        const int blk = sc[blockDim.x - 1];
        __syncthreads();
        base += blk;
    }
}
void launch_lm_compact(const int* count, const int* idx, const double* val, const int* kept, const int* mode,
                       const int* deg, const int* nadj, const int* adj, int np, int cap, int* head, int* pidx,
                       double* pval, int* pdeg, int* padj, hipStream_t s) {
    k_lm_compact<<<np, kT, 0, s>>>(count, idx, val, kept, mode, deg, nadj, adj, np, cap, head, pidx, pval, pdeg, padj);
    ETTK_HIP(hipGetLastError());
}
void launch_ppmax(const float* Pp1, const float* Pp2, int np, int B, float* Mx, hipStream_t s) {
    k_ppmax<<<np, kT, 0, s>>>(Pp1, Pp2, 4 * B * B, Mx);
    ETTK_HIP(hipGetLastError());
}

}  // namespace ettk
