// basics.cpp -- transforms, FFT, smoothing, masks and the tricubic reslice of the oracle.
// TEST INFRASTRUCTURE ONLY (the checker); see oracle_core.hpp.
#include "oracle_core.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

namespace orc {

// ============================================================== 4x4 transforms
M4 m4_ident() {
    M4 M;
    for (int i = 0; i < 16; i++) M.a[i] = (i % 5 == 0) ? 1.0 : 0.0;
    return M;
}
M4 m4_from(const double* p) {
    M4 M;
    std::memcpy(M.a, p, sizeof(M.a));
    return M;
}
M4 m4_mul(const M4& A, const M4& B) {  // vtkMatrix4x4::Multiply4x4
    M4 C;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            C(r, c) = A(r, 0) * B(0, c) + A(r, 1) * B(1, c) + A(r, 2) * B(2, c) + A(r, 3) * B(3, c);
    return C;
}
double m4_det(const M4& A) {
    double m[4][4];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) m[r][c] = A(r, c);
    double det = 0;
    for (int c = 0; c < 4; c++) {
        double sub[3][3];
        for (int r = 1; r < 4; r++) {
            int cc = 0;
            for (int k = 0; k < 4; k++)
                if (k != c) sub[r - 1][cc++] = m[r][k];
        }
        double d3 = sub[0][0] * (sub[1][1] * sub[2][2] - sub[1][2] * sub[2][1]) -
                    sub[0][1] * (sub[1][0] * sub[2][2] - sub[1][2] * sub[2][0]) +
                    sub[0][2] * (sub[1][0] * sub[2][1] - sub[1][1] * sub[2][0]);
        det += ((c & 1) ? -1.0 : 1.0) * m[0][c] * d3;
    }
    return det;
}
M4 m4_inv(const M4& A) {  // Gauss-Jordan with partial pivoting (FP64)
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = A(r, c);
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (piv != c)
            for (int k = 0; k < 8; k++) std::swap(a[c][k], a[piv][k]);
        double d = a[c][c];
        for (int k = 0; k < 8; k++) a[c][k] /= d;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                double f = a[r][c];
                for (int k = 0; k < 8; k++) a[r][k] -= f * a[c][k];
            }
    }
    M4 I;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) I(r, c) = a[r][c + 4];
    return I;
}
M4 m4_rot(double angle, double x, double y, double z) {
    M4 M = m4_ident();
    if (angle == 0.0 || (x == 0.0 && y == 0.0 && z == 0.0)) return M;
    angle = angle * M_PI / 180.0;
    double w = std::cos(0.5 * angle);
    double f = std::sin(0.5 * angle) / std::sqrt(x * x + y * y + z * z);
    x *= f; y *= f; z *= f;
    double ww = w * w, wx = w * x, wy = w * y, wz = w * z;
    double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
    double s = ww - xx - yy - zz;
    M(0, 0) = xx * 2 + s;    M(1, 0) = (xy + wz) * 2; M(2, 0) = (xz - wy) * 2;
    M(0, 1) = (xy - wz) * 2; M(1, 1) = yy * 2 + s;    M(2, 1) = (yz + wx) * 2;
    M(0, 2) = (xz + wy) * 2; M(1, 2) = (yz - wx) * 2; M(2, 2) = zz * 2 + s;
    return M;
}
M4 m4_translate(double x, double y, double z) {
    M4 M = m4_ident();
    M(0, 3) = x; M(1, 3) = y; M(2, 3) = z;
    return M;
}
M4 m4_scale(double x, double y, double z) {
    M4 M = m4_ident();
    M(0, 0) = x; M(1, 1) = y; M(2, 2) = z;
    return M;
}
M4 m4_strip_translation(const M4& t) {
    // rotation->SetInput(t); PostMultiply(); Translate(-position)  (FIM.h:194-215, Wedge.h:323-333)
    return m4_mul(m4_translate(-t(0, 3), -t(1, 3), -t(2, 3)), t);
}
M4 m4_zyz(double a, double b, double g) {
    M4 M = m4_ident();
    pre(M, m4_rz(-g));
    pre(M, m4_ry(-b));
    pre(M, m4_rz(-a));
    return M;
}
bool transform_valid(const M4& t, double trans_restrict, double rot_restrict) {
    double sx = t(0, 3), sy = t(1, 3), sz = t(2, 3);
    if (std::fabs(sx) > trans_restrict || std::fabs(sy) > trans_restrict || std::fabs(sz) > trans_restrict)
        return false;
    // TransformPoint((0,0,1)) minus the shift, normalised; angle to z
    double px = t(0, 2), py = t(1, 2), pz = t(2, 2);
    double nrm = std::sqrt(px * px + py * py + pz * pz);
    double c = (nrm != 0.0) ? pz / nrm : 0.0;
    double angle = (c >= 1.0) ? 0.0 : std::fabs(std::acos(c) * 180.0 / M_PI);
    return angle <= rot_restrict;
}

// ============================================================== FFT
namespace {
struct FftPlan {
    int n;
    std::vector<int> fac;
    std::vector<cd> tw;  // exp(-2 pi i k / n)
};
const FftPlan& get_plan(int n) {
    static std::map<int, FftPlan> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    FftPlan P;
    P.n = n;
    int m = n;
    while (m % 4 == 0) { P.fac.push_back(4); m /= 4; }
    while (m % 2 == 0) { P.fac.push_back(2); m /= 2; }
    for (int p = 3; m > 1; p += 2)
        while (m % p == 0) { P.fac.push_back(p); m /= p; }
    P.fac.push_back(1);
    P.tw.resize(n);
    for (int k = 0; k < n; k++) P.tw[k] = cd(std::cos(2 * M_PI * k / n), -std::sin(2 * M_PI * k / n));
    return cache.emplace(n, std::move(P)).first->second;
}
// recursive mixed-radix decimation in time
void fft_rec(const cd* in, cd* out, int n, int is, const int* fac, const FftPlan& P, int tws, int sign) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = fac[0], m = n / p;
    for (int r = 0; r < p; r++) fft_rec(in + r * is, out + r * m, m, is * p, fac + 1, P, tws * p, sign);
    const int N0 = P.n;
    cd t[16];
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            cd w = P.tw[(long)r * k * tws % N0];
            if (sign > 0) w = std::conj(w);
            t[r] = out[r * m + k] * w;
        }
        for (int q = 0; q < p; q++) {
            cd s = 0;
            for (int r = 0; r < p; r++) {
                cd w = P.tw[(long)((r * q) % p) * (N0 / p) % N0];
                if (sign > 0) w = std::conj(w);
                s += t[r] * w;
            }
            out[k + q * m] = s;
        }
    }
}
}  // namespace

// iterative Stockham autosort (radix 4, 2 and odd factors), ping-pong between x and a scratch buffer
static void stockham(cd* x, int n, int sign, const FftPlan& P) {
    thread_local std::vector<cd> scratch;
    if ((int)scratch.size() < n) scratch.resize(n);
    cd* a = x;
    cd* b = scratch.data();
    int Ns = 1;
    for (size_t f = 0; f + 1 < P.fac.size(); f++) {
        const int R = P.fac[f], NB = n / R;
        for (int j = 0; j < NB; j++) {
            const int k = j % Ns;
            cd v[16];
            for (int m = 0; m < R; m++) {
                cd z = a[j + m * NB];
                if (m > 0) {
                    cd w = P.tw[(long)m * k * (n / (Ns * R))];
                    if (sign > 0) w = std::conj(w);
                    z *= w;
                }
                v[m] = z;
            }
            const int base = (j / Ns) * Ns * R + k;
            if (R == 2) {
                b[base] = v[0] + v[1];
                b[base + Ns] = v[0] - v[1];
            } else if (R == 4) {
                cd s0 = v[0] + v[2], d0 = v[0] - v[2], s1 = v[1] + v[3], d1 = v[1] - v[3];
                cd jd1 = sign < 0 ? cd(d1.imag(), -d1.real()) : cd(-d1.imag(), d1.real());
                b[base] = s0 + s1;
                b[base + 2 * Ns] = s0 - s1;
                b[base + Ns] = d0 + jd1;
                b[base + 3 * Ns] = d0 - jd1;
            } else {
                for (int q = 0; q < R; q++) {
                    cd s = 0;
                    for (int m = 0; m < R; m++) {
                        cd w = P.tw[(long)((m * q) % R) * (n / R)];
                        if (sign > 0) w = std::conj(w);
                        s += v[m] * w;
                    }
                    b[base + q * Ns] = s;
                }
            }
        }
        std::swap(a, b);
        Ns *= R;
    }
    if (a != x) std::copy(a, a + n, x);
}

void fft1d(cd* x, int n, int sign) {
    const FftPlan& P = get_plan(n);
    stockham(x, n, sign, P);
}

void r2c3(const double* in, cd* out, int n0, int n1, int n2) {
    int h = n2 / 2 + 1;
    std::vector<cd> buf(std::max(n0, std::max(n1, n2)));
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++) {
            const double* row = in + ((size_t)i * n1 + j) * n2;
            for (int k = 0; k < n2; k++) buf[k] = cd(row[k], 0.0);
            fft1d(buf.data(), n2, -1);
            cd* o = out + ((size_t)i * n1 + j) * h;
            for (int k = 0; k < h; k++) o[k] = buf[k];
        }
    for (int i = 0; i < n0; i++)
        for (int k = 0; k < h; k++) {
            for (int j = 0; j < n1; j++) buf[j] = out[((size_t)i * n1 + j) * h + k];
            fft1d(buf.data(), n1, -1);
            for (int j = 0; j < n1; j++) out[((size_t)i * n1 + j) * h + k] = buf[j];
        }
    for (int j = 0; j < n1; j++)
        for (int k = 0; k < h; k++) {
            for (int i = 0; i < n0; i++) buf[i] = out[((size_t)i * n1 + j) * h + k];
            fft1d(buf.data(), n0, -1);
            for (int i = 0; i < n0; i++) out[((size_t)i * n1 + j) * h + k] = buf[i];
        }
}

void c2r3(const cd* in, double* out, int n0, int n1, int n2) {
    int h = n2 / 2 + 1;
    std::vector<cd> G(in, in + (size_t)n0 * n1 * h);
    std::vector<cd> buf(std::max(n0, std::max(n1, n2)));
    for (int j = 0; j < n1; j++)
        for (int k = 0; k < h; k++) {
            for (int i = 0; i < n0; i++) buf[i] = G[((size_t)i * n1 + j) * h + k];
            fft1d(buf.data(), n0, +1);
            for (int i = 0; i < n0; i++) G[((size_t)i * n1 + j) * h + k] = buf[i];
        }
    for (int i = 0; i < n0; i++)
        for (int k = 0; k < h; k++) {
            for (int j = 0; j < n1; j++) buf[j] = G[((size_t)i * n1 + j) * h + k];
            fft1d(buf.data(), n1, +1);
            for (int j = 0; j < n1; j++) G[((size_t)i * n1 + j) * h + k] = buf[j];
        }
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++) {
            const cd* g = &G[((size_t)i * n1 + j) * h];
            buf[0] = cd(g[0].real(), 0.0);
            buf[n2 / 2] = cd(g[n2 / 2].real(), 0.0);
            for (int k = 1; k < n2 / 2; k++) {
                buf[k] = g[k];
                buf[n2 - k] = std::conj(g[k]);
            }
            fft1d(buf.data(), n2, +1);
            double* o = out + ((size_t)i * n1 + j) * n2;
            for (int k = 0; k < n2; k++) o[k] = buf[k].real();
        }
}

// ============================================================== smoothing and masks
void gauss_smooth3(std::vector<float>& A, int n0, int n1, int n2, double stdv, double rf) {
    const int dims[3] = {n0, n1, n2};
    const long strides[3] = {(long)n1 * n2, (long)n2, 1};
    int radius = (int)(stdv * rf);
    std::vector<float> out(A.size());
    for (int axis = 2; axis >= 0; axis--) {  // vtkImageGaussianSmooth: axis 2, then 1, then 0
        int n = dims[axis];
        long st = strides[axis];
        for (long base = 0; base < (long)A.size(); base++) {
            int idx = (int)((base / st) % n);
            int lo = idx - radius, hi = idx + radius;
            int lclip = lo < 0 ? -lo : 0;
            int rclip = hi > n - 1 ? hi - (n - 1) : 0;
            double sum = 0, ksum = 0;
            if (stdv == 0.0) { out[base] = A[base]; continue; }
            for (int x = -radius + lclip; x <= radius - rclip; x++) ksum += std::exp(-(double)(x * x) / (stdv * stdv * 2.0));
            for (int x = -radius + lclip; x <= radius - rclip; x++) {
                double kv = std::exp(-(double)(x * x) / (stdv * stdv * 2.0)) / ksum;
                sum += kv * (double)A[base + (long)x * st];
            }
            out[base] = (float)sum;
        }
        A.swap(out);
    }
}

std::vector<float> sphere_mask(int N, float radius, float sigma) {
    std::vector<float> M((size_t)N * N * N, 1.0f);
    if (radius == 0.0f) return M;
    float c = (float)(N / 2.0);
    float r2 = std::min((float)(N - 1), (float)(radius + sigma / 2.0));
    float r1 = std::max(0.0f, (float)(radius - sigma / 2.0));
    float kappa = (float)(M_PI / (r2 - r1));
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) {
                float di = i - c, dj = j - c, dk = k - c;
                float r = std::sqrt(di * di + dj * dj + dk * dk);
                float v;
                if (r <= r1) v = 1.0f;
                else if (r >= r2) v = 0.0f;
                else v = (float)((1.0 + std::cos(kappa * (r - r1))) / 2.0);
                M[((size_t)i * N + j) * N + k] = v;
            }
    return M;
}

std::vector<float> square_fft_mask(int N, int sq) {
    std::vector<float> M((size_t)N * N * N, 0.0f);
    if (sq > 0 && sq < std::floor(N / 2.0)) {
        for (int i = sq; i <= N - 1 - sq; i++)
            for (int j = sq; j <= N - 1 - sq; j++)
                for (int k = sq; k <= N - 1 - sq; k++) M[((size_t)i * N + j) * N + k] = 1.0f;
    } else {
        std::fill(M.begin(), M.end(), 1.0f);
    }
    gauss_smooth3(M, N, N, N, 2.5, 2.5 * 0.75);  // softenMask(maskfft, 2.5) (IFilt.h:475-494)
    return M;
}

void to_internal(const float* vol, int N, std::vector<double>& out) {
    out.resize((size_t)N * N * N);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) out[((size_t)i * N + j) * N + k] = vol[((size_t)k * N + j) * N + i];
}

// ============================================================== reslice (WSub.h:603-828)
namespace {
inline void keys_weights(double f, double F[4]) {  // vtkTricubicInterpWeights (vtkImageInterpolator.cxx:313-325)
    double fm1 = f - 1, fd2 = f * 0.5, ft3 = f * 3;
    F[0] = -fd2 * fm1 * fm1;
    F[1] = ((ft3 - 2) * fd2 - 1) * fm1;
    F[2] = -((ft3 - 4) * f - 1) * fd2;
    F[3] = f * fd2 * fm1;
}
inline int vfloor(double x, double& f) {  // vtkInterpolationMath::Floor with VTK_INTERPOLATE_FLOOR_TOL
    x += 7.62939453125e-06;
    int i = (int)std::floor(x);
    f = x - i;
    return i;
}
}  // namespace

void reslice_image(const float* vol, int N, const ettk_particle& d, const M4* t, std::vector<double>& out) {
    const double pos[3] = {N / 2.0, N / 2.0, N / 2.0};
    extract_image(vol, N, N, N, pos, N, d, t, out);
}

// nbfWedgedSubImage3D::getImage (WSub.h:603-802, normalize off) from a volume of nx x ny x nz voxels (x fastest)
// whose origin is moved to `pos` (the sub-volume's position): output voxel q of the N^3 box reads the input at
// pos + fin (q - N/2 - cut z), fin = Euler(normal) matrix t; VTK tricubic, taps clamped to the input, outside
// [-0.5, n - 0.5] per axis invalid; the mean of the valid (float) samples is subtracted, invalid voxels are 0
void extract_image(const float* vol, int NX, int NY, int NZ, const double* pos3, int N, const ettk_particle& d,
                   const M4* t, std::vector<double>& out) {
    M4 euler = m4_ident();
    pre(euler, m4_rz(-d.normal[2]));
    pre(euler, m4_rx(-d.normal[0]));
    pre(euler, m4_rz(-d.normal[1]));
    M4 local = m4_mul(euler, m4_from(d.matrix));
    M4 fin = t ? m4_mul(local, *t) : local;
    if (m4_det(fin) == 0.0) fin = m4_ident();
    const double half = N / 2.0;
    const double lo = -0.5, hx = NX - 1 + 0.5, hy = NY - 1 + 0.5, hz = NZ - 1 + 0.5;
    out.assign((size_t)N * N * N, 0.0);
    std::vector<unsigned char> valid((size_t)N * N * N, 0);
    double sum = 0;
    long count = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) {
                double w[3] = {i - half, j - half, k - half - d.cut_offset};
                double p[3];
                for (int r = 0; r < 3; r++) p[r] = fin(r, 0) * w[0] + fin(r, 1) * w[1] + fin(r, 2) * w[2] + fin(r, 3) + pos3[r];
                size_t o = ((size_t)i * N + j) * N + k;
                if (p[0] < lo || p[0] > hx || p[1] < lo || p[1] > hy || p[2] < lo || p[2] > hz) continue;
                double fx, fy, fz;
                int ix = vfloor(p[0], fx), iy = vfloor(p[1], fy), iz = vfloor(p[2], fz);
                double wx[4], wy[4], wz[4];
                keys_weights(fx, wx); keys_weights(fy, wy); keys_weights(fz, wz);
                int cx[4], cy[4], cz[4];
                for (int q = 0; q < 4; q++) {
                    cx[q] = std::min(std::max(ix - 1 + q, 0), NX - 1);
                    cy[q] = std::min(std::max(iy - 1 + q, 0), NY - 1);
                    cz[q] = std::min(std::max(iz - 1 + q, 0), NZ - 1);
                }
                double v = 0;
                for (int c = 0; c < 4; c++) {
                    double vy = 0;
                    for (int b = 0; b < 4; b++) {
                        const float* row = vol + ((size_t)cz[c] * NY + cy[b]) * NX;
                        double vx = wx[0] * row[cx[0]] + wx[1] * row[cx[1]] + wx[2] * row[cx[2]] + wx[3] * row[cx[3]];
                        vy += wy[b] * vx;
                    }
                    v += wz[c] * vy;
                }
                float vf = (float)v;  // the reslice output keeps the float scalar type of the fixed image
                out[o] = vf;
                valid[o] = 1;
                sum += vf;
                count++;
            }
    float mean_inside = (float)(sum / (float)count);
    for (size_t o = 0; o < out.size(); o++) out[o] = valid[o] ? out[o] - mean_inside : 0.0;
}

}  // namespace orc
