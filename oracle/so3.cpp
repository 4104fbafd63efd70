// so3.cpp -- S^2 projection, S^2 harmonic transforms and SO(3) correlation of the oracle.
// TEST INFRASTRUCTURE ONLY (the checker); see oracle_core.hpp.
//
// Restates nbfProjectionRotationMetric3D::cartesianToSpherical (ProjRot.h:265-413), soft (:1137-1295),
// softScaled (:1298-1319), rotatedSphericalImageMinMax (:1558-1589), conditionSphericalWindow (:1322-1345),
// nbfWedge3D::getSphericalImage (Wedge.h:364-429) and the soft-1.0 mathematics those call
// (FST_semi_memo.c:134-360, so3_correlate_fftw.c:103-165, soft_fftw.c:1045-1923, rotate_so3.c:695-751),
// re-derived from the published method (Driscoll-Healy sampling; Kostelec & Rockmore, J. Fourier Anal.
// Appl. 14 (2008)) rather than from SOFT's code: direct quadrature with FP64 Legendre functions and
// FP64 Wigner-d recurrences.
#include "oracle_core.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace orc {

// ============================================================== spherical projection
static inline int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

void project_image(Plan& P, std::vector<double> img, std::vector<float>& Pp) {
    const int N = P.N, B = P.p.B;
    const size_t n3 = (size_t)N * N * N;
    // imageFilter->execute (IFilt.h:146-293): (image - mean) * mask + mean when the mask is on
    if (P.p.mask_radius != 0.0f) {
        double s = 0;
        for (size_t o = 0; o < n3; o++) s += img[o];
        double mean = s / (double)n3;
        for (size_t o = 0; o < n3; o++) img[o] = (img[o] - mean) * P.imgmask[o] + mean;
    }
    const int M = 2 * N, HM = N + 1;
    const double c = img[0];  // pad constant = value at voxel (0,0,0) (ProjRot.h:285)
    std::vector<double> A((size_t)M * M * M);
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++)
            for (int k = 0; k < M; k++) {
                double v = (i < N && j < N && k < N) ? img[((size_t)i * N + j) * N + k] : c;
                A[((size_t)i * M + j) * M + k] = v;
            }
    std::vector<double> As(A.size());
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++)
            for (int k = 0; k < M; k++)
                As[((size_t)i * M + j) * M + k] = A[((size_t)i * M + j) * M + k] * (((i + j + k) & 1) ? -1.0 : 1.0);
    std::vector<cd> F((size_t)M * M * HM);
    r2c3(As.data(), F.data(), M, M, M);
    // compose the full-size magnitude volume (ProjRot.h:321-338), keeping the reference's fill pattern
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++)
            for (int k = 0; k < HM; k++) {
                double m2 = std::real(F[((size_t)i * M + j) * HM + k] * std::conj(F[((size_t)i * M + j) * HM + k]));
                A[((size_t)i * M + j) * M + k] = P.p.mcf ? std::sqrt(std::sqrt(m2)) : std::sqrt(m2);
            }
    for (int j = 0; j < M; j++)
        for (int k = 0; k < M; k++) A[(size_t)j * M + k] = 0.0;
    int r2 = M / 2, c2 = M / 2, d2 = M / 2;
    for (int i = -r2 + 1; i < r2; i++)
        for (int j = -c2 + 1; j < c2; j++)
            for (int k = 1; k < d2; k++)
                A[((size_t)(r2 + i) * M + (c2 + j)) * M + (d2 + k)] = A[((size_t)(r2 - i) * M + (c2 - j)) * M + (d2 - k)];
    // spherical resampling with cubic interpolation and wrap-around border (ProjRot.h:340-406)
    float lo = (P.p.bp_lo > 0.0f || P.p.bp_hi < 1.0f) ? P.p.bp_lo : 0.0f;
    float hi = (P.p.bp_lo > 0.0f || P.p.bp_hi < 1.0f) ? P.p.bp_hi : 1.0f;
    int bandL = std::max(4, (int)std::ceil(lo * (M - 1)));
    int bandU = (int)std::ceil(std::min(0.5f, hi) * (M - 1));
    const float Bf = (float)B;
    const double phi0 = M_PI / 4.0 / Bf, dphi = 2.0 * M_PI / 4.0 / Bf, dth = M_PI / Bf;
    const double origin = -(M - 1) / 2.0 - 0.5;
    Pp.assign((size_t)4 * B * B, 0.0f);
    for (int jp = 0; jp < 2 * B; jp++)
        for (int kt = 0; kt < 2 * B; kt++) {
            double phi = phi0 + jp * dphi, th = kt * dth;
            double sum = 0;
            for (int r = bandL; r <= bandU; r++) {
                double x = r * std::sin(phi) * std::cos(th), y = r * std::sin(phi) * std::sin(th), z = r * std::cos(phi);
                double p[3] = {x - origin, y - origin, z - origin};
                int idx[3];
                double w[3][4];
                for (int a = 0; a < 3; a++) {
                    double f;
                    double xx = p[a] + 7.62939453125e-06;
                    idx[a] = (int)std::floor(xx);
                    f = xx - idx[a];
                    double fm1 = f - 1, fd2 = f * 0.5, ft3 = f * 3;
                    w[a][0] = -fd2 * fm1 * fm1;
                    w[a][1] = ((ft3 - 2) * fd2 - 1) * fm1;
                    w[a][2] = -((ft3 - 4) * f - 1) * fd2;
                    w[a][3] = f * fd2 * fm1;
                }
                double v = 0;
                for (int cz = 0; cz < 4; cz++) {
                    int kz = wrapi(idx[2] - 1 + cz, M);
                    double vy = 0;
                    for (int cy = 0; cy < 4; cy++) {
                        int jy = wrapi(idx[1] - 1 + cy, M);
                        double vx = 0;
                        for (int cx = 0; cx < 4; cx++) {
                            int ix = wrapi(idx[0] - 1 + cx, M);
                            vx += w[0][cx] * A[((size_t)ix * M + jy) * M + kz];
                        }
                        vy += w[1][cy] * vx;
                    }
                    v += w[2][cz] * vy;
                }
                sum += v;
            }
            Pp[(size_t)jp * 2 * B + kt] = (float)sum;
        }
}

// ============================================================== spherical wedge
std::vector<float> spherical_wedge(int B, const ettk_particle& d) {
    const int n = 2 * B;
    std::vector<float> W((size_t)n * n, 1.0f);
    double l = d.lwedge < -90 ? -90.0 : d.lwedge;
    double u = d.uwedge > 90 ? 90.0 : d.uwedge;
    if (u - l >= 180) return W;
    M4 euler = m4_ident();
    pre(euler, m4_rz(-d.normal[2]));
    pre(euler, m4_rx(-d.normal[0]));
    pre(euler, m4_rz(-d.normal[1]));
    M4 fin = m4_inv(m4_mul(euler, m4_from(d.matrix)));
    M4 R = m4_strip_translation(fin);
    double n1[3] = {std::cos(-(90.0 - u) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - u) * M_PI / 180.0)};
    double n2[3] = {std::cos(-(90.0 - l) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - l) * M_PI / 180.0)};
    double v1[3], v2[3];
    for (int r = 0; r < 3; r++) {
        v1[r] = R(r, 0) * n1[0] + R(r, 1) * n1[1] + R(r, 2) * n1[2];
        v2[r] = R(r, 0) * n2[0] + R(r, 1) * n2[1] + R(r, 2) * n2[2];
    }
    float phiK = (float)(M_PI / (n - 1));
    float thetaK = (float)(2.0 * M_PI / (n - 1));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            float si = std::sin((float)(i * phiK)), ci = std::cos((float)(i * phiK));
            float sj = std::sin((float)(j * thetaK)), cj = std::cos((float)(j * thetaK));
            double a = v1[0] * (double)(cj * si) + v1[1] * (double)(sj * si) + v1[2] * (double)ci;
            double b = v2[0] * (double)(cj * si) + v2[1] * (double)(sj * si) + v2[2] * (double)ci;
            W[(size_t)i * n + j] = (a * b <= 0.0) ? 1.0f : 0.0f;
        }
    return W;
}

void condition_spherical_window(std::vector<float>& W, int B) {
    gauss_smooth3(W, 2 * B, 2 * B, 1, 2.0, 1.0);
}

// ============================================================== S^2 transforms
// Driscoll-Healy quadrature weights on theta_j = pi (2j+1) / (4B): exact for int_0^pi g(theta) sin(theta)
// for trigonometric polynomials g of degree < 2B.
static std::vector<double> dh_weights(int B) {
    std::vector<double> w(2 * B);
    for (int j = 0; j < 2 * B; j++) {
        double th = M_PI * (2 * j + 1) / (4.0 * B), s = 0;
        for (int k = 0; k < B; k++) s += std::sin((2 * k + 1) * th) / (2 * k + 1);
        w[j] = (2.0 / B) * std::sin(th) * s;
    }
    return w;
}
// orthonormal associated Legendre functions (Condon-Shortley phase) Pbar_l^m(cos theta), m >= 0
static void legendre_all(int B, double theta, std::vector<double>& P) {
    P.assign((size_t)B * B, 0.0);
    double x = std::cos(theta), s = std::sin(theta);
    double pmm = std::sqrt(1.0 / (4.0 * M_PI));
    for (int m = 0; m < B; m++) {
        if (m > 0) pmm *= -std::sqrt((2.0 * m + 1) / (2.0 * m)) * s;
        P[(size_t)m * B + m] = pmm;
        if (m + 1 < B) P[(size_t)(m + 1) * B + m] = x * std::sqrt(2.0 * m + 3) * pmm;
        for (int l = m + 2; l < B; l++) {
            double a = std::sqrt((4.0 * l * l - 1) / ((double)l * l - (double)m * m));
            double b = std::sqrt((((double)(l - 1) * (l - 1)) - (double)m * m) / (4.0 * (l - 1) * (l - 1) - 1));
            P[(size_t)l * B + m] = a * (x * P[(size_t)(l - 1) * B + m] - b * P[(size_t)(l - 2) * B + m]);
        }
    }
}

void s2_forward(int B, const double* f, std::vector<cd>& coef) {
    const int n = 2 * B;
    coef.assign((size_t)B * B, cd(0, 0));
    std::vector<double> w = dh_weights(B), P;
    std::vector<cd> row(n);
    for (int j = 0; j < n; j++) {
        for (int k = 0; k < n; k++) row[k] = cd(f[(size_t)j * n + k], 0.0);
        fft1d(row.data(), n, -1);  // sum_k f e^{-i m phi_k}
        double th = M_PI * (2 * j + 1) / (4.0 * B);
        legendre_all(B, th, P);
        double q = w[j] * 2.0 * M_PI / n;
        for (int l = 0; l < B; l++)
            for (int m = 0; m <= l; m++) {
                cd fm = row[m] * q * P[(size_t)l * B + m];
                coef[s2_index(l, m)] += fm;
                if (m > 0) {
                    cd fmn = row[n - m] * q * P[(size_t)l * B + m] * ((m & 1) ? -1.0 : 1.0);
                    coef[s2_index(l, -m)] += fmn;
                }
            }
    }
}

void s2_inverse(int B, const std::vector<cd>& coef, double* f) {
    const int n = 2 * B;
    std::vector<double> P;
    std::vector<cd> row(n);
    for (int j = 0; j < n; j++) {
        double th = M_PI * (2 * j + 1) / (4.0 * B);
        legendre_all(B, th, P);
        std::fill(row.begin(), row.end(), cd(0, 0));
        for (int l = 0; l < B; l++)
            for (int m = 0; m <= l; m++) {
                row[m] += coef[s2_index(l, m)] * P[(size_t)l * B + m];
                if (m > 0) row[n - m] += coef[s2_index(l, -m)] * P[(size_t)l * B + m] * ((m & 1) ? -1.0 : 1.0);
            }
        fft1d(row.data(), n, +1);  // sum_m c_m e^{+i m phi_k}
        for (int k = 0; k < n; k++) f[(size_t)j * n + k] = row[k].real();
    }
}

// ============================================================== Wigner d
static double lfact(int n) { return std::lgamma((double)n + 1.0); }
// d^l_{m1 m2}(beta) for l0 = max(|m1|,|m2|)
static double wigner_seed(int m1, int m2, double beta) {
    int j = std::max(std::abs(m1), std::abs(m2));
    double c = std::cos(beta / 2), s = std::sin(beta / 2);
    auto dj_jm = [&](int m) {  // d^j_{j,m}
        double lg = 0.5 * (lfact(2 * j) - lfact(j + m) - lfact(j - m));
        return std::exp(lg) * std::pow(c, j + m) * std::pow(-s, j - m);
    };
    auto dj_mjm = [&](int m) {  // d^j_{-j,m}
        double lg = 0.5 * (lfact(2 * j) - lfact(j + m) - lfact(j - m));
        return std::exp(lg) * std::pow(c, j - m) * std::pow(s, j + m);
    };
    if (m1 == j) return dj_jm(m2);
    if (m1 == -j) return dj_mjm(m2);
    if (m2 == j) return (((m1 - j) & 1) ? -1.0 : 1.0) * dj_jm(m1);
    return (((m1 + j) & 1) ? -1.0 : 1.0) * dj_mjm(m1);  // m2 == -j
}

void wigner_d_all(int B, double beta, std::vector<double>& d) {
    // layout: for each (m1, m2) block, values for l = l0..B-1; addressed through a dense cube
    // d[((l * (2B-1)) + (m1 + B - 1)) * (2B-1) + (m2 + B - 1)]
    const int W = 2 * B - 1;
    d.assign((size_t)B * W * W, 0.0);
    double cb = std::cos(beta);
    for (int m1 = -(B - 1); m1 <= B - 1; m1++)
        for (int m2 = -(B - 1); m2 <= B - 1; m2++) {
            int l0 = std::max(std::abs(m1), std::abs(m2));
            double dprev = 0.0, dcur = wigner_seed(m1, m2, beta);
            for (int l = l0; l < B; l++) {
                d[((size_t)l * W + (m1 + B - 1)) * W + (m2 + B - 1)] = dcur;
                if (l + 1 >= B) break;
                double dnext;
                if (l == 0) {
                    dnext = cb;  // d^1_00
                } else {
                    double a = std::sqrt(((double)(l + 1) * (l + 1) - (double)m1 * m1) * ((double)(l + 1) * (l + 1) - (double)m2 * m2));
                    double bterm = std::sqrt(((double)l * l - (double)m1 * m1) * ((double)l * l - (double)m2 * m2));
                    dnext = ((2.0 * l + 1) * ((double)l * (l + 1) * cb - (double)m1 * m2) * dcur - (l + 1.0) * bterm * dprev) / (l * a);
                }
                dprev = dcur;
                dcur = dnext;
            }
        }
}

// ============================================================== SO(3) correlation
// C(alpha, beta, gamma) = Re sum_l sum_{m1,m2} sig(l,-m1) (-1)^{m1+m2} conj(pat(l,-m2))
//                             sqrt((2l+1)/(8 pi^2)) e^{-i m1 alpha} d^l_{m1 m2}(beta) e^{-i m2 gamma}
// on alpha_j = 2 pi j/(2B), beta_k = pi (2k+1)/(4B), gamma_j = 2 pi j/(2B)  (so3CombineCoef + Inverse_SO3).
void so3_correlate(int B, const double* sig, const double* pat, std::vector<double>& C) {
    const int n = 2 * B, W = 2 * B - 1;
    std::vector<cd> cs, cp;
    s2_forward(B, sig, cs);
    s2_forward(B, pat, cp);
    C.assign((size_t)n * n * n, 0.0);
    std::vector<double> d;
    std::vector<cd> S((size_t)n * n);
    for (int kb = 0; kb < n; kb++) {
        double beta = M_PI * (2 * kb + 1) / (4.0 * B);
        wigner_d_all(B, beta, d);
        std::fill(S.begin(), S.end(), cd(0, 0));
        for (int m1 = -(B - 1); m1 <= B - 1; m1++)
            for (int m2 = -(B - 1); m2 <= B - 1; m2++) {
                int l0 = std::max(std::abs(m1), std::abs(m2));
                cd acc = 0;
                for (int l = l0; l < B; l++) {
                    cd coef = cs[s2_index(l, -m1)] * std::conj(cp[s2_index(l, -m2)]) * ((((m1 + m2) & 1) != 0) ? -1.0 : 1.0);
                    acc += coef * std::sqrt((2.0 * l + 1) / (8.0 * M_PI * M_PI)) * d[((size_t)l * W + (m1 + B - 1)) * W + (m2 + B - 1)];
                }
                S[(size_t)wrapi(m1, n) * n + wrapi(m2, n)] = acc;
            }
        // 2-D DFT with e^{-i m alpha}
        std::vector<cd> buf(n);
        for (int a = 0; a < n; a++) {
            for (int m = 0; m < n; m++) buf[m] = S[(size_t)m * n + a];
            fft1d(buf.data(), n, -1);
            for (int m = 0; m < n; m++) S[(size_t)m * n + a] = buf[m];
        }
        for (int a = 0; a < n; a++) {
            for (int m = 0; m < n; m++) buf[m] = S[(size_t)a * n + m];
            fft1d(buf.data(), n, -1);
            for (int g = 0; g < n; g++) C[((size_t)a * n + kb) * n + g] = buf[g].real();
        }
    }
}

// rotateFct(B, B, B-1, ...): g = Lambda(alpha,beta,gamma) f, coefficients rotated by Wigner D matrices
void s2_rotate(int B, const double* in, double* out, double alpha, double beta, double gamma) {
    const int W = 2 * B - 1;
    std::vector<cd> c, r((size_t)B * B, cd(0, 0));
    s2_forward(B, in, c);
    std::vector<double> d;
    wigner_d_all(B, beta, d);
    for (int l = 0; l < B; l++)
        for (int m = -l; m <= l; m++) {
            cd acc = 0;
            for (int mp = -l; mp <= l; mp++) {
                cd D = std::exp(cd(0, -m * alpha)) * d[((size_t)l * W + (m + B - 1)) * W + (mp + B - 1)] * std::exp(cd(0, -mp * gamma));
                acc += D * c[s2_index(l, mp)];
            }
            r[s2_index(l, m)] = acc;
        }
    s2_inverse(B, r, out);
}

void so3_correlate_scaled(int B, const double* F1, const double* F2, std::vector<double>& C) {
    const int n = 2 * B;
    so3_correlate(B, F1, F2, C);
    // first extrema in row-major order (ProjRot.h:1562-1576)
    double maxv = -DBL_MAX, minv = DBL_MAX;
    size_t maxp = 0, minp = 0;
    for (size_t o = 0; o < C.size(); o++) {
        if (C[o] > maxv) { maxv = C[o]; maxp = o; }
        if (C[o] < minv) { minv = C[o]; minp = o; }
    }
    auto rot = [&](size_t pos, std::vector<double>& R) {
        int a = (int)(pos / ((size_t)n * n)), b = (int)((pos / n) % n), g = (int)(pos % n);
        float alpha = (float)(a / (double)B * M_PI);
        float beta = (float)((2 * b + 1.0) / (4. * B) * M_PI);
        float gamma = (float)(g / (double)B * M_PI);
        R.resize((size_t)n * n);
        s2_rotate(B, F2, R.data(), alpha, beta, gamma);
    };
    std::vector<double> Rmin, Rmax;
    rot(minp, Rmin);
    rot(maxp, Rmax);
    double wmin = 0, wmax = 0;
    for (int i = 0; i < n; i++) {
        double th = std::sin(M_PI * (2 * i + 1) / 4 / (float)B);
        for (int j = 0; j < n; j++) {
            wmin += F1[(size_t)i * n + j] * Rmin[(size_t)i * n + j] * th;
            wmax += F1[(size_t)i * n + j] * Rmax[(size_t)i * n + j] * th;
        }
    }
    for (auto& v : C) v = (v - minv) / (maxv - minv) * (wmax - wmin) + wmin;
}

}  // namespace orc
