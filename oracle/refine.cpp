// refine.cpp -- FP64 restatement of the continuous constrained-CCC score used by the batched + adjoint mode,
// with an FP64 central-difference gradient and the same lock-step gradient-ascent rule.
// TEST INFRASTRUCTURE ONLY (the checker); see oracle_core.hpp.
//
// The score has no counterpart in the reference (it replaces the Powell refinement of
// nbfFourierImageMetric::refine, FIM.h:1702-1809).  It is evaluated in the particle's own frame, where the
// particle's missing wedge does not move with the pose:
//   S(x) = <X1(x), X2> / (|X1(x)| |X2|),  <a,b> = sum_h c_h Re(a_h conj b_h), c_h = 1 on the first and last
//          planes of the half axis and 2 elsewhere;
//   X2 = H r2c((I2 - mean I2) maskfft shift)            (the particle, fixed)
//   X1 = H r2c(I1'(x) maskfft shift),  I1'(x)(p) = (I1 - mean I1)(N/2 + cut z + fin(x)^-1 (p - N/2)),
//        trilinear with clamped taps (the reference extended by its border values: I1' is continuous in x);
//        fin(x) = Euler M t(x)
//   t(x) = [R0 Rx(x0) Ry(x1) Rz(x2) | tau0 + x3..5]  (degrees, voxels)
//   H = band-pass filter x binary particle wedge (lab frame) x binary reference wedge (rotated by the seed)
// Volumes are handled in the particle's memory order (x fastest): the half axis of the spectra is x.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "oracle_core.hpp"

namespace orc {

static M4 refine_pose(const M4& seed, const double* x) {
    M4 t = m4_strip_translation(seed);
    pre(t, m4_rx(x[0]));
    pre(t, m4_ry(x[1]));
    pre(t, m4_rz(x[2]));
    t(0, 3) = seed(0, 3) + x[3];
    t(1, 3) = seed(1, 3) + x[4];
    t(2, 3) = seed(2, 3) + x[5];
    return t;
}

// Lab-frame wedge normals (nbfWedge3D::set, Wedge.h:143-184, angles clamped to [-90, 90]); false when the
// wedge spans 180 degrees or more (no missing data).
static bool raw_normals(float lw, float uw, double n1[3], double n2[3]) {
    double l = lw < -90 ? -90.0 : lw, u = uw > 90 ? 90.0 : uw;
    if (u - l >= 180) return false;
    n1[0] = std::cos(-(90.0 - u) * M_PI / 180.0);
    n1[1] = 0.0;
    n1[2] = -std::sin(-(90.0 - u) * M_PI / 180.0);
    n2[0] = std::cos(-(90.0 - l) * M_PI / 180.0);
    n2[1] = 0.0;
    n2[2] = -std::sin(-(90.0 - l) * M_PI / 180.0);
    return true;
}

// Euler(normal) * M of a sub-volume (nbfWedgedSubImage3D::getImage, WSub.h:712-726)
static M4 euler_local(const ettk_particle& d) {
    M4 euler = m4_ident();
    pre(euler, m4_rz(-d.normal[2]));
    pre(euler, m4_rx(-d.normal[0]));
    pre(euler, m4_rz(-d.normal[1]));
    return m4_mul(euler, m4_from(d.matrix));
}

// I1'(p) in the particle's memory order out[(z N + y) N + x]
static void reslice_reference(const Plan& P, const ettk_particle& d, const M4& t, std::vector<double>& out) {
    const int N = P.N;
    const size_t n3 = (size_t)N * N * N;
    const M4 G = m4_inv(m4_mul(euler_local(d), t));
    const double half = N / 2.0;
    double mref = 0;
    for (size_t o = 0; o < n3; o++) mref += P.ref[o];
    mref /= (double)n3;
    out.assign(n3, 0.0);
    auto at = [&](int x, int y, int z) {  // the reference is held as (x, y, z) with z fastest
        x = std::min(std::max(x, 0), N - 1);
        y = std::min(std::max(y, 0), N - 1);
        z = std::min(std::max(z, 0), N - 1);
        return P.ref[((size_t)x * N + y) * N + z];
    };
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                double w[3] = {x - half, y - half, z - half}, q[3];
                for (int r = 0; r < 3; r++) q[r] = G(r, 0) * w[0] + G(r, 1) * w[1] + G(r, 2) * w[2] + G(r, 3) + half;
                q[2] += d.cut_offset;
                int x0 = (int)std::floor(q[0]), y0 = (int)std::floor(q[1]), z0 = (int)std::floor(q[2]);
                double ax = q[0] - x0, ay = q[1] - y0, az = q[2] - z0;
                double v = 0;
                for (int c = 0; c < 8; c++) {
                    int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                    v += (dx ? ax : 1 - ax) * (dy ? ay : 1 - ay) * (dz ? az : 1 - az) * at(x0 + dx, y0 + dy, z0 + dz);
                }
                out[((size_t)z * N + y) * N + x] = v - mref;
            }
}

// H on the particle-frame half spectrum [z][y][x <= N/2] (DC at (N/2, N/2, N/2) after the checkerboard)
static std::vector<double> frame_weights(const Plan& P, const ettk_particle& d, const M4& seed) {
    const int N = P.N, H = P.H;
    std::vector<double> Wt((size_t)N * N * H);
    double n1[3] = {0, 0, 0}, n2[3] = {0, 0, 0}, r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0};
    bool pw = P.p.missing_wedge && raw_normals(d.lwedge, d.uwedge, n1, n2);
    bool rw = false, rlook = false;
    double A3[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (P.p.missing_wedge && P.ref_wedge_kind == 1) {
        // average: binarised accumulated wedge at the reference-frame frequency A k (nearest voxel, clamped)
        M4 A = m4_strip_translation(m4_inv(m4_mul(euler_local(d), seed)));
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) A3[r * 3 + c] = A(r, c);
        rlook = true;
    } else if (P.p.missing_wedge && P.ref_wedge_kind == 0) {
        double m1[3], m2[3];
        if (raw_normals(P.refd.lwedge, P.refd.uwedge, m1, m2)) {
            // reference wedge normals in the reference frame (v = inverse(fin_ref) n), then in the particle
            // frame at the seed pose: u = A^T v with A the rotation of fin(seed)^-1
            M4 Rr = m4_strip_translation(m4_inv(euler_local(P.refd)));
            M4 A = m4_strip_translation(m4_inv(m4_mul(euler_local(d), seed)));
            double v1[3], v2[3];
            for (int r = 0; r < 3; r++) {
                v1[r] = Rr(r, 0) * m1[0] + Rr(r, 1) * m1[1] + Rr(r, 2) * m1[2];
                v2[r] = Rr(r, 0) * m2[0] + Rr(r, 1) * m2[1] + Rr(r, 2) * m2[2];
            }
            for (int c = 0; c < 3; c++) {
                r1[c] = A(0, c) * v1[0] + A(1, c) * v1[1] + A(2, c) * v1[2];
                r2[c] = A(0, c) * v2[0] + A(1, c) * v2[1] + A(2, c) * v2[2];
            }
            rw = true;
        }
    }
    for (int a = 0; a < N; a++)
        for (int b = 0; b < N; b++)
            for (int c = 0; c < H; c++) {
                size_t h = ((size_t)a * N + b) * H + c;
                double kx = c - (H - 1), ky = b - N / 2, kz = a - N / 2;
                double w = P.filter[h];  // radial: the same table in any axis order
                if (pw && (n1[0] * kx + n1[1] * ky + n1[2] * kz) * (n2[0] * kx + n2[1] * ky + n2[2] * kz) > 0) w = 0;
                if (rw && (r1[0] * kx + r1[1] * ky + r1[2] * kz) * (r2[0] * kx + r2[1] * ky + r2[2] * kz) > 0) w = 0;
                if (rlook) {
                    int q[3];
                    for (int t = 0; t < 3; t++) {
                        double pt = A3[t * 3] * kx + A3[t * 3 + 1] * ky + A3[t * 3 + 2] * kz;
                        q[t] = std::min(std::max((int)std::nearbyint(pt) + N / 2, 0), N - 1);
                    }
                    if (P.Wbin_acc[((size_t)q[0] * N + q[1]) * N + q[2]] == 0.0f) w = 0;
                }
                Wt[h] = w;
            }
    return Wt;
}

double refine_score(Plan& P, const float* part, const ettk_particle& d, const M4& seed, const double* x) {
    const int N = P.N, H = P.H;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * H;
    std::vector<double> A(n3);
    double mp = 0;
    for (size_t o = 0; o < n3; o++) mp += part[o];
    mp /= (double)n3;
    for (size_t o = 0; o < n3; o++) A[o] = ((double)part[o] - mp) * P.maskfft[o] * P.shift[o];
    std::vector<cd> F1(nh), F2(nh);
    r2c3(A.data(), F2.data(), N, N, N);
    std::vector<double> I1;
    reslice_reference(P, d, refine_pose(seed, x), I1);
    for (size_t o = 0; o < n3; o++) A[o] = I1[o] * P.maskfft[o] * P.shift[o];
    r2c3(A.data(), F1.data(), N, N, N);
    std::vector<double> Wt = frame_weights(P, d, seed);
    double dot = 0, n1 = 0, n2 = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < H; k++) {
                size_t h = ((size_t)i * N + j) * H + k;
                cd x1 = F1[h] * Wt[h], x2 = F2[h] * Wt[h];
                double c = (k == 0 || k == H - 1) ? 1.0 : 2.0;
                dot += c * std::real(x1 * std::conj(x2));
                n1 += c * std::norm(x1);
                n2 += c * std::norm(x2);
            }
    return dot / std::sqrt(n1 * n2);
}

}  // namespace orc

using namespace orc;

extern "C" {

// score and FP64 central-difference gradient (step h in degrees / voxels)
double orc_refine_eval(void* hp, const float* part, const ettk_particle* d, const double* seed, const double* x,
                       double h, double* grad) {
    Plan& P = *(Plan*)hp;
    M4 s = m4_from(seed);
    double S = refine_score(P, part, *d, s, x);
    if (grad) {
        for (int r = 0; r < 6; r++) {
            double xp[6], xm[6];
            std::memcpy(xp, x, sizeof(xp));
            std::memcpy(xm, x, sizeof(xm));
            xp[r] += h;
            xm[r] -= h;
            grad[r] = (refine_score(P, part, *d, s, xp) - refine_score(P, part, *d, s, xm)) / (2 * h);
        }
    }
    return S;
}

// The library's ascent (refine_jobs) with the FP64 gradient: steps along the gradient in the metric where one
// degree counts as c = (N/4) pi/180 voxels of displacement, of length s voxels (start 0.5); an improving
// step is accepted and s grows by 1.25, otherwise s halves.  n_iter evaluations in all.
double orc_refine(void* hp, const float* part, const ettk_particle* d, const double* seed, int n_iter, double* pose_out) {
    Plan& P = *(Plan*)hp;
    const double c = P.N / 4.0 * M_PI / 180.0;
    double x[6] = {0, 0, 0, 0, 0, 0}, g[6], xt[6], gt[6];
    double S = orc_refine_eval(hp, part, d, seed, x, 1e-4, g);
    double s = 0.5;
    for (int it = 1; it < n_iter; it++) {
        double dir[6], nrm = 0;
        for (int r = 0; r < 6; r++) {
            dir[r] = r < 3 ? g[r] / (c * c) : g[r];
            nrm += r < 3 ? c * c * dir[r] * dir[r] : dir[r] * dir[r];
        }
        nrm = std::sqrt(nrm);
        for (int r = 0; r < 6; r++) xt[r] = x[r] + (nrm > 0 ? s * dir[r] / nrm : 0.0);
        double Stry = orc_refine_eval(hp, part, d, seed, xt, 1e-4, gt);
        if (Stry > S) {
            S = Stry;
            std::memcpy(x, xt, sizeof(x));
            std::memcpy(g, gt, sizeof(g));
            s *= 1.25;
        } else {
            s *= 0.5;
        }
    }
    M4 t = refine_pose(m4_from(seed), x);
    std::memcpy(pose_out, t.a, sizeof(t.a));
    return S;
}

}  // extern "C"
