// ccc.cpp -- the constrained FFT cross-correlation of the oracle.
// TEST INFRASTRUCTURE ONLY (the checker); see oracle_core.hpp.
//
// Restates nbfFourierImageMetric::executeFourierNewHalf (inc/em/nbfFourierImageMetric.h:1078-1323),
// reinitializeHalf (:634-694), nbfFourierFilter::updateFilterHalf/executeHalf (nbfFourierFilter.h:492-606),
// buildBandPassHalf/buildBFactorHalf (:738-797), nbfCorrelationImageMetric::normalizeFourierHalf
// (nbfCorrelationImageMetric.h:498-530) and locateMaxima (:304-323), nbfWedge3D::getImageHalf
// (nbfWedge3D.h:313-360) with nbfWedgedSubImage3D::getWedgeImageHalf (nbfWedgedSubImage3D.h:941-979).
#include "oracle_core.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace orc {

static std::vector<float> build_filter_half(const ettk_params& p, int N, bool& dummy) {
    int H = N / 2 + 1;
    std::vector<float> f((size_t)N * N * H, 1.0f);
    f[((size_t)(N / 2) * N + N / 2) * H + (H - 1)] = 0.0f;  // DC
    dummy = true;
    bool use_bp = (p.bp_lo > 0.0f) || (p.bp_hi < 1.0f);
    int c = N / 2;  // center = floor(size / 2.0) (TinyVector<int>)
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < H; k++) {
                double ti = (1.0 * i - c) / c, tj = (1.0 * j - c) / c, tk = (1.0 * k - c) / c;
                float radius = (float)std::sqrt(ti * ti + tj * tj + tk * tk);
                size_t o = ((size_t)i * N + j) * H + k;
                if (use_bp) {
                    double lo = (double)p.bp_lo, hi = (double)p.bp_hi;
                    double lv = (double)p.bp_lo_decay, hv = (double)p.bp_hi_decay;
                    float bp = 1.0f;
                    if (radius < lo) bp = (float)std::exp(-(lo - radius) * (lo - radius) / 2.0 / lv / lv);
                    if (radius > hi) bp = (float)std::exp(-(radius - hi) * (radius - hi) / 2.0 / hv / hv);
                    f[o] *= bp;
                }
                if (p.bfactor != 0.0f) f[o] *= (float)std::exp(-(double)p.bfactor * (double)radius * radius);
            }
    if (use_bp || p.bfactor != 0.0f) dummy = false;
    return f;
}

std::vector<float> wedge_half(int N, float lw, float uw, const M4& t) {
    int H = N / 2 + 1;
    std::vector<float> W((size_t)N * N * H, 1.0f);
    double l = lw < -90 ? -90.0 : lw;  // nbfWedge3D::set clamps to [-90, 90] (Wedge.h:143-184)
    double u = uw > 90 ? 90.0 : uw;
    if (u - l >= 180) return W;
    double n1[3] = {std::cos(-(90.0 - u) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - u) * M_PI / 180.0)};
    double n2[3] = {std::cos(-(90.0 - l) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - l) * M_PI / 180.0)};
    M4 R = m4_strip_translation(t);
    double v1[3], v2[3];
    for (int r = 0; r < 3; r++) {
        v1[r] = R(r, 0) * n1[0] + R(r, 1) * n1[1] + R(r, 2) * n1[2];
        v2[r] = R(r, 0) * n2[0] + R(r, 1) * n2[1] + R(r, 2) * n2[2];
    }
    int c0 = N / 2, c1 = N / 2, c2 = H - 1;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < H; k++) {
                double a = v1[0] * (i - c0) + v1[1] * (j - c1) + v1[2] * (k - c2);
                double b = v2[0] * (i - c0) + v2[1] * (j - c1) + v2[2] * (k - c2);
                W[((size_t)i * N + j) * H + k] = (a * b <= 0.0) ? 1.0f : 0.0f;
            }
    gauss_smooth3(W, N, N, H, 2.0, 0.5);  // smoothWedgeImage (Wedge.h:432-445)
    return W;
}

std::vector<float> particle_wedge_half(int N, const ettk_particle& d, const M4* t) {
    M4 euler = m4_ident();
    pre(euler, m4_rz(-d.normal[2]));
    pre(euler, m4_rx(-d.normal[0]));
    pre(euler, m4_rz(-d.normal[1]));
    M4 local = m4_mul(euler, m4_from(d.matrix));
    M4 fin = t ? m4_mul(local, *t) : local;
    return wedge_half(N, d.lwedge, d.uwedge, m4_inv(fin));
}

// nbfWedge3D::getImage (Wedge.h:195-290) without a stored wedge image: the binary wedge on the full grid,
// centre floor(shape / 2), then smoothWedgeImage; all ones when the wedge spans 180 degrees
std::vector<float> wedge_full(int N, float lw, float uw, const M4& t) {
    std::vector<float> W((size_t)N * N * N, 1.0f);
    double l = lw < -90 ? -90.0 : lw;
    double u = uw > 90 ? 90.0 : uw;
    if (u - l >= 180) return W;
    double n1[3] = {std::cos(-(90.0 - u) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - u) * M_PI / 180.0)};
    double n2[3] = {std::cos(-(90.0 - l) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - l) * M_PI / 180.0)};
    M4 R = m4_strip_translation(t);
    double v1[3], v2[3];
    for (int r = 0; r < 3; r++) {
        v1[r] = R(r, 0) * n1[0] + R(r, 1) * n1[1] + R(r, 2) * n1[2];
        v2[r] = R(r, 0) * n2[0] + R(r, 1) * n2[1] + R(r, 2) * n2[2];
    }
    const int c = N / 2;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) {
                double a = v1[0] * (i - c) + v1[1] * (j - c) + v1[2] * (k - c);
                double b = v2[0] * (i - c) + v2[1] * (j - c) + v2[2] * (k - c);
                W[((size_t)i * N + j) * N + k] = (a * b <= 0.0) ? 1.0f : 0.0f;
            }
    gauss_smooth3(W, N, N, N, 2.0, 0.5);
    return W;
}

// updateAccumulatedWedgeImage (WAvg.h:496-538): for every entry with |weight| > 0, setTransform(NULL) then
// getWedgeImage(current, store) (WSub.h:841-870: fin = Euler(normal) * store, wedge of inverse(fin)),
// accum += current; setAccumulatedWedgeImage (:49-70): binarise at max/2, effective if the count is below
// 0.95 of the size, then smoothWedgeImage.  Spherical: updateAccumulatedSphericalWedgeImageMPI (:614-651) sums
// getSphericalWedgeImage(A, store) (WSub.h:1007-1041), updateSphericalWedgeImage (:654-667) binarises at max/2.
AvgWedge avg_wedge(int N, int B, int n, const ettk_particle* d, const float* w, const double* mats) {
    AvgWedge out;
    const size_t n3 = (size_t)N * N * N, nn = (size_t)4 * B * B;
    out.accum.assign(n3, 0.0f);
    std::vector<float> sacc(nn, 0.0f);
    for (int e = 0; e < n; e++) {
        if (!(std::fabs(w[e]) > 0)) continue;
        ettk_particle q = d[e];
        std::memcpy(q.matrix, mats + (size_t)e * 16, sizeof(q.matrix));
        M4 euler = m4_ident();
        pre(euler, m4_rz(-q.normal[2]));
        pre(euler, m4_rx(-q.normal[0]));
        pre(euler, m4_rz(-q.normal[1]));
        std::vector<float> cur = wedge_full(N, q.lwedge, q.uwedge, m4_inv(m4_mul(euler, m4_from(q.matrix))));
        for (size_t o = 0; o < n3; o++) out.accum[o] += cur[o];
        std::vector<float> sw = spherical_wedge(B, q);
        for (size_t o = 0; o < nn; o++) sacc[o] += sw[o];
    }
    float mx = -FLT_MAX;
    for (float v : out.accum) mx = std::max(mx, v);
    out.bin.resize(n3);
    double cnt = 0;
    for (size_t o = 0; o < n3; o++) {
        out.bin[o] = (double)out.accum[o] > (double)mx / 2.0 ? 1.0f : 0.0f;
        cnt += out.bin[o];
    }
    out.effective = cnt < 0.95 * (double)n3;
    out.wedge = out.bin;
    if (out.effective) gauss_smooth3(out.wedge, N, N, N, 2.0, 0.5);
    else std::fill(out.wedge.begin(), out.wedge.end(), 1.0f);
    float smx = -FLT_MAX;
    for (float v : sacc) smx = std::max(smx, v);
    out.sph.resize(nn);
    for (size_t o = 0; o < nn; o++) out.sph[o] = (double)sacc[o] > (double)smx / 2.0 ? 1.0f : 0.0f;
    return out;
}

void plan_init(Plan& P, const ettk_params& p) {
    P.p = p;
    int N = p.N;
    P.N = N;
    P.H = N / 2 + 1;
    size_t n3 = (size_t)N * N * N;
    P.shift.resize(n3);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) P.shift[((size_t)i * N + j) * N + k] = ((i + j + k) & 1) ? -1.0f : 1.0f;
    P.imgmask = sphere_mask(N, p.mask_radius, p.mask_sigma);
    P.refmask = P.imgmask;
    P.maskfft = square_fft_mask(N, p.square_mask);
    P.filter = build_filter_half(p, N, P.dummy_filter);
    std::vector<double> tmp(n3);
    for (size_t o = 0; o < n3; o++) tmp[o] = (double)P.refmask[o] * P.shift[o];
    P.Mfft.resize((size_t)N * N * P.H);
    r2c3(tmp.data(), P.Mfft.data(), N, N, N);
}

// executeHalf (FFilt.h:587-606)
static void execute_half(const Plan& P, std::vector<cd>& F, const std::vector<float>* wedge) {
    if (P.p.mcf) {
        for (auto& z : F) {
            double C = std::pow(std::real(z * std::conj(z)), 0.25);
            if (C > 0) z = z / C;
        }
    }
    if (!P.dummy_filter) {
        for (size_t o = 0; o < F.size(); o++) {
            float w = wedge ? P.filter[o] * (*wedge)[o] : P.filter[o];
            F[o] *= (double)w;
        }
    }
}

// normalizeFourierHalf (CorrMetric.h:498-530)
static float normalize_half(int N, std::vector<cd>& F) {
    int H = N / 2 + 1;
    F[((size_t)(N / 2) * N + N / 2) * H + (H - 1)] = 0.0;
    double s1 = 0, s2 = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const cd* row = &F[((size_t)i * N + j) * H];
            for (int k = 0; k < H - 1; k++) s1 += std::norm(row[k]);
            s2 += std::norm(row[H - 1]);
        }
    float norm = (float)(2 * s1);
    norm = (float)std::sqrt((norm + s2) / (1.0 * N * N * 2 * (H - 1)));
    for (auto& z : F) z /= (double)norm;
    return norm;
}

void plan_set_reference(Plan& P, const float* ref, const ettk_particle& d) {
    to_internal(ref, P.N, P.ref);
    P.refd = d;
    // The reference is used as it is (an average: getImage(a, NULL) -> blitzToVtk(average), WAvg.h:249-257).
    // Its wedge W1 = getWedgeImageHalf(wedge1) at its own alignment (FIM.h:1101-1105).  The F1 path
    // (FIM.h:1123-1153) depends on W1*W2(t), so it is evaluated inside every ccc_eval call, as the reference does.
    if (P.p.missing_wedge) P.W1 = particle_wedge_half(P.N, d, nullptr);
    else P.W1.clear();
    P.ref_wedge_kind = 0;
    P.Wbin_acc.clear();
    P.W1s_acc.clear();
}

// setInput1 with an average: its wedge is the accumulated one (getWedgeImageHalf takes k <= N/2 of
// getWedgeImage, WAvg.h:555-611); the spherical window is its binarised spherical sum (ones when not effective)
void plan_set_reference_average(Plan& P, const float* ref, const AvgWedge& aw) {
    ettk_particle full{};
    full.lwedge = -90.0f;
    full.uwedge = 90.0f;
    for (int k = 0; k < 16; k++) full.matrix[k] = (k % 5 == 0) ? 1.0 : 0.0;
    plan_set_reference(P, ref, full);
    const int N = P.N, H = P.H;
    P.ref_wedge_kind = aw.effective ? 1 : 2;
    if (P.p.missing_wedge) {
        P.W1.assign((size_t)N * N * H, 1.0f);
        if (aw.effective)
            for (int i = 0; i < N; i++)
                for (int j = 0; j < N; j++)
                    for (int k = 0; k < H; k++) P.W1[((size_t)i * N + j) * H + k] = aw.wedge[((size_t)i * N + j) * N + k];
    }
    P.W1acc_half.assign((size_t)N * N * H, 1.0f);
    if (aw.effective)
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++)
                for (int k = 0; k < H; k++) P.W1acc_half[((size_t)i * N + j) * H + k] = aw.wedge[((size_t)i * N + j) * N + k];
    P.Wbin_acc = aw.bin;
    P.W1s_acc = aw.effective ? aw.sph : std::vector<float>(aw.sph.size(), 1.0f);
}

// getWedgeOverlap (CorrMetric.h:533-566): wedge images of the inputs whose wedge is effective
// (nbfWedge3D::isEffective, Wedge.h:70), sum(w1 w2 filter) / sum(filter) over the half grid, 1 when neither is
float wedge_overlap(const Plan& P, const ettk_particle& d, const M4& t) {
    auto eff = [](float lw, float uw) {
        const float l = lw < -90 ? -90.0f : lw, u = uw > 90 ? 90.0f : uw;
        return l != -90.0f || u != 90.0f;
    };
    const bool e1 = P.ref_wedge_kind == 0 ? eff(P.refd.lwedge, P.refd.uwedge) : P.ref_wedge_kind == 1;
    const bool e2 = eff(d.lwedge, d.uwedge);
    if (!e1 && !e2) return 1.0f;
    const size_t nh = (size_t)P.N * P.N * P.H;
    std::vector<float> w1(nh, 1.0f), w2(nh, 1.0f);
    if (e1) w1 = P.ref_wedge_kind == 0 ? particle_wedge_half(P.N, P.refd, nullptr) : P.W1acc_half;
    if (e2) w2 = particle_wedge_half(P.N, d, &t);
    double a = 0, b = 0;
    for (size_t h = 0; h < nh; h++) {
        a += (double)(w1[h] * w2[h] * P.filter[h]);
        b += P.filter[h];
    }
    return (float)(a / b);
}

// locateMaxima with blitz maxIndex (first maximum in row-major order)
float ccc_eval(Plan& P, const float* part, const ettk_particle& d, M4& t, CccTrace* trace) {
    const int N = P.N, H = P.H;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * H;
    P.ccc_count++;
    std::vector<float> wedge;
    const std::vector<float>* wp = nullptr;
    if (P.p.missing_wedge) {
        wedge = particle_wedge_half(N, d, &t);
        for (size_t o = 0; o < nh; o++) wedge[o] *= P.W1[o];
        wp = &wedge;
    }
    // reference side
    std::vector<double> A(n3);
    std::vector<cd> F(nh);
    for (size_t o = 0; o < n3; o++) A[o] = P.ref[o] * P.maskfft[o] * P.shift[o];
    r2c3(A.data(), F.data(), N, N, N);
    execute_half(P, F, wp);
    normalize_half(N, F);
    c2r3(F.data(), A.data(), N, N, N);
    double k1m = 0;
    for (size_t o = 0; o < n3; o++) k1m += A[o] * A[o] * P.refmask[o];
    float k1_m = (float)k1m;
    for (size_t o = 0; o < n3; o++) A[o] *= P.refmask[o];
    std::vector<cd> F1m(nh);
    r2c3(A.data(), F1m.data(), N, N, N);
    // particle side
    std::vector<double> I2;
    reslice_image(part, N, d, &t, I2);
    for (size_t o = 0; o < n3; o++) A[o] = I2[o] * P.maskfft[o] * P.shift[o];
    r2c3(A.data(), F.data(), N, N, N);
    execute_half(P, F, wp);
    normalize_half(N, F);
    double normf = (double)n3;
    std::vector<cd> FFT2re(nh);
    for (size_t o = 0; o < nh; o++) FFT2re[o] = 2.0 * F1m[o] * std::conj(F[o] * normf);
    c2r3(F.data(), A.data(), N, N, N);
    for (size_t o = 0; o < n3; o++) A[o] = A[o] * A[o] * P.shift[o];
    std::vector<cd> Q(nh);
    r2c3(A.data(), Q.data(), N, N, N);
    // C = (-M conj(Q) + FFT2re) * shiftr
    std::vector<cd> C(nh);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < H; k++) {
                size_t o = ((size_t)i * N + j) * H + k;
                C[o] = (-P.Mfft[o] * std::conj(Q[o]) + FFT2re[o]) * (double)P.shift[((size_t)i * N + j) * N + k];
            }
    // embed into the (2N) half spectrum, factor (2N)^3 / N^3
    const int N2 = 2 * N, H2 = N + 1;
    double factor = (1.0 * N2 * N2 * 2 * (H2 - 1)) / (1.0 * N * N * 2 * (H - 1));
    std::vector<cd> big((size_t)N2 * N2 * H2, cd(0, 0));
    int i0 = N - (int)std::floor(N / 2.0), j0 = N - (int)std::floor(N / 2.0), k0 = H - 1;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < H; k++)
                big[((size_t)(i + i0) * N2 + (j + j0)) * H2 + (k + k0)] = C[((size_t)i * N + j) * H + k] * factor;
    std::vector<double> Tb((size_t)N2 * N2 * N2);
    c2r3(big.data(), Tb.data(), N2, N2, N2);
    // restricted search window (FIM.h:1222-1247)
    double center = N2 / 2.0;
    double restriction = center;
    float tr = P.p.trans_restrict < 0 ? FLT_MAX : P.p.trans_restrict;
    if (tr < FLT_MAX) restriction = std::max((double)tr * 2.0, 1.0);
    int I0 = (int)std::max(0.0, center - std::floor(restriction));
    int I1 = (int)std::min(center + std::floor(restriction), (double)(N2 - 1));
    int ext = I1 - I0 + 1;
    int ub2 = (ext - 1) / 2;
    double big_n3 = (double)N2 * N2 * N2;
    if (trace) {
        trace->ext = ext;
        trace->k1_m = (double)k1_m;
        trace->T.resize((size_t)ext * ext * ext);
    }
    double best = -DBL_MAX;
    int bi = -1, bj = -1, bk = -1;
    for (int i = 0; i < ext; i++)
        for (int j = 0; j < ext; j++)
            for (int k = 0; k < ext; k++) {
                int gi = I0 + i, gj = I0 + j, gk = I0 + k;
                double v = Tb[((size_t)gi * N2 + gj) * N2 + gk] * (((gi + gj + gk) & 1) ? -1.0 : 1.0) / big_n3;
                if (trace) trace->T[((size_t)i * ext + j) * ext + k] = v;
                float rad = (float)((ub2 - i) * (ub2 - i) + (ub2 - j) * (ub2 - j) + (ub2 - k) * (ub2 - k));
                if (!(rad < restriction * restriction)) v = -DBL_MAX;
                if (bi < 0 || v > best) { best = v; bi = i; bj = j; bk = k; }
            }
    double peak;
    if (P.p.overlap_norm && wp) {
        double sw = 0, sf = 0;
        for (size_t o = 0; o < nh; o++) {
            sw += (double)((*wp)[o] * P.filter[o]);
            sf += (double)P.filter[o];
        }
        peak = std::max(0.0, (double)k1_m - best) / sw * sf;
    } else {
        peak = std::max(0.0, (double)k1_m - best);
    }
    float cp = (float)peak;
    cp = (float)(cp / big_n3 / big_n3);
    P.last_peak = cp;
    double sh[3] = {(bi - (ext - 1) / 2.0) / 2.0, (bj - (ext - 1) / 2.0) / 2.0, (bk - (ext - 1) / 2.0) / 2.0};
    // shifts are Pixel (float) in the reference (TinyVector<Pixel,Dim> cccPeakPosition)
    t = m4_mul(t, m4_translate(-(float)sh[0], -(float)sh[1], -(float)sh[2]));
    return cp;
}

}  // namespace orc
