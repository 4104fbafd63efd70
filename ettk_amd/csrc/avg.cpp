// avg.cpp -- host side of the amplitude-weighted Fourier average (device code in avg.hip).
#include "plan.hpp"

#include <cfloat>
#include <cstring>

namespace ettk {
EvalGeom make_geom_public(const ettk_plan* P, const ettk_particle& d, const Mat4& t, int particle);
void launch_avg_acc(const float2* F, const float* w, int E, long nh, double* re, double* im, double* acc, hipStream_t s);
void launch_avg_final(const double* re, const double* im, const double* acc, int N, float2* F, hipStream_t s);
void launch_avg_normalise(const float* A, int N, float* out, hipStream_t s);
void launch_var_acc(const float* R, const float* mean, const float* w, int E, double wsum, const float* avg, int N,
                    double* var, hipStream_t s);
void launch_var_out(const double* var, int N, float* out, hipStream_t s);
void launch_wedge_accum(const WedgeRef* w, int m, int N, float* acc, hipStream_t s);
void launch_wedge_binarize(const float* acc, int N, float* bin, int* count, hipStream_t s);
void launch_wedge_smooth_grid(const float* bin, int N, float* full, float* half, hipStream_t s);

// updateAccumulatedWedgeImage (WAvg.h:496-538) + setAccumulatedWedgeImage (:49-70) on the device, and the
// spherical counterpart updateAccumulatedSphericalWedgeImageMPI + updateSphericalWedgeImage (:614-667) on the
// host.  Entry e adds the wedge of descs[e] at alignment matrices[e*16]: the particle's own matrix is reset
// first (setTransform(NULL), so fin = Euler(normal) * matrices[e]); entries with weights[e] == 0 are skipped.
void avg_wedge_compute(ettk_plan* P, int n, const ettk_particle* descs, const float* weights, const double* matrices) {
    ETTK_HIP(hipSetDevice(P->ctx->device));
    const int N = P->N, B = P->p.B, ns = 2 * B;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * P->H;
    hipStream_t s = P->stream;
    std::vector<WedgeRef> w;
    std::vector<float> sphf((size_t)ns * ns, 0.0f);
    for (int i = 0; i < n; i++) {
        if (!(std::fabs(weights[i]) > 0)) continue;
        ettk_particle d = descs[i];
        std::memcpy(d.matrix, matrices + (size_t)i * 16, sizeof(d.matrix));
        WedgeRef r{};
        wedge_normals(d.lwedge, d.uwedge, mat_inverse(euler_of(d) * Mat4::from(d.matrix)), r);
        w.push_back(r);
        const std::vector<float> a = spherical_wedge_bin(B, d);
        for (size_t q = 0; q < a.size(); q++) sphf[q] += a[q];  // Array<Pixel,2> += (float)
    }
    ETTK_REQUIRE(!w.empty(), ETTK_ERR_ARG, "accumulated wedge: no entry has a nonzero weight");
    P->wacc.ensure(n3);
    P->wbin.ensure(n3);
    P->wsm.ensure(n3);
    P->wsm_half.ensure(nh);
    P->wcount.ensure(1);
    DevBuf<WedgeRef> dw;
    dw.upload(w.data(), w.size(), s);
    launch_wedge_accum(dw.p, (int)w.size(), N, P->wacc.p, s);
    launch_wedge_binarize(P->wacc.p, N, P->wbin.p, P->wcount.p, s);
    int count = 0;
    ETTK_HIP(hipMemcpyAsync(&count, P->wcount.p, sizeof(int), hipMemcpyDeviceToHost, s));
    ETTK_HIP(hipStreamSynchronize(s));
    P->wacc_effective = (double)count < 0.95 * (double)n3;
    if (P->wacc_effective) launch_wedge_smooth_grid(P->wbin.p, N, P->wsm.p, P->wsm_half.p, s);
    // spherical: binarise at max / 2 (the reference applies it whatever the 3-D wedge says; it is only read
    // when the 3-D accumulated wedge is effective, WAvg.h:670-690)
    float mx = -FLT_MAX;
    for (float v : sphf) mx = std::max(mx, v);
    P->sph_acc_bin.assign(sphf.size(), 0.0f);
    for (size_t q = 0; q < sphf.size(); q++) P->sph_acc_bin[q] = (double)sphf[q] > (double)mx / 2.0 ? 1.0f : 0.0f;
    ETTK_HIP(hipStreamSynchronize(s));
    P->wacc_ready = true;
}
}  // namespace ettk

using namespace ettk;

static DevBuf<double>& acc_buf(ettk_plan* P, int which) {
    return which == 0 ? P->avg_re : which == 1 ? P->avg_im : P->avg_acc;
}

extern "C" {

int ettk_avg_reset(ettk_plan* P) {
    try {
        ETTK_REQUIRE(P, ETTK_ERR_ARG, "ettk_avg_reset: null plan");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        const size_t nh = (size_t)P->N * P->N * P->H;
        for (int w = 0; w < 3; w++) {
            DevBuf<double>& b = acc_buf(P, w);
            b.ensure(nh);
            launch_fill_zero(b.p, nh, P->stream);
        }
        ETTK_HIP(hipStreamSynchronize(P->stream));
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return ETTK_ERR_ALLOC;
    }
    set_error("");
    return ETTK_OK;
}

// getImageMPI (WAvg.h:301-351): reslice with setTransform(NULL) and the stored alignment, r2c(T * shift),
// B = |F|, acc += |w| B, re += w B Re F, im += w B Im F
int ettk_avg_accumulate(ettk_plan* P, const float* d_particles, int n, const ettk_particle* descs,
                        const float* weights, const double* matrices) {
    try {
        ETTK_REQUIRE(P && d_particles && descs && weights && matrices && n >= 0, ETTK_ERR_ARG,
                     "ettk_avg_accumulate: bad argument");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        const int N = P->N;
        const long nh = (long)N * N * P->H;
        ETTK_REQUIRE(acc_buf(P, 0).n >= (size_t)nh, ETTK_ERR_STATE, "ettk_avg_accumulate: call ettk_avg_reset first");
        hipStream_t s = P->stream;
        ensure_ccc_workspace(P, std::min(std::max(n, 1), 128));
        std::vector<int> sel;
        for (int i = 0; i < n; i++)
            if (std::fabs(weights[i]) > 0) sel.push_back(i);
        std::vector<EvalGeom> g(P->cap);
        std::vector<float> w(P->cap);
        DevBuf<float> dw;
        dw.ensure(P->cap);
        for (size_t c0 = 0; c0 < sel.size(); c0 += P->cap) {
            int E = (int)std::min<size_t>(P->cap, sel.size() - c0);
            for (int e = 0; e < E; e++) {
                int i = sel[c0 + e];
                ettk_particle d = descs[i];
                for (int k = 0; k < 16; k++) d.matrix[k] = (k % 5 == 0) ? 1.0 : 0.0;  // setTransform(NULL)
                g[e] = make_geom_public(P, d, Mat4::from(matrices + (size_t)i * 16), i);
                w[e] = weights[i];
            }
            ETTK_HIP(hipMemcpyAsync(P->geom.p, g.data(), E * sizeof(EvalGeom), hipMemcpyHostToDevice, s));
            ETTK_HIP(hipMemcpyAsync(dw.p, w.data(), E * sizeof(float), hipMemcpyHostToDevice, s));
            launch_reslice(d_particles, P->geom.p, E, N, P->R2.p, P->p_res.p, P->nb_res, s);
            launch_reslice_mean(P->p_res.p, P->nb_res, E, P->stats.p, s);
            R2CArgs a{P->R2.p, P->S2.p, N, E, PRO_RESLICED, nullptr, P->stats.p};
            fft3_r2c(P->ctx->fft, a, nullptr, s);
            launch_avg_acc(P->S2.p, dw.p, E, nh, acc_buf(P, 0).p, acc_buf(P, 1).p, acc_buf(P, 2).p, s);
            ETTK_HIP(hipStreamSynchronize(s));
        }
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return ETTK_ERR_ALLOC;
    }
    set_error("");
    return ETTK_OK;
}

// nbfCorrelationImageMetric::getImageVariance -> nbfWedgedAverageImage3D::getImageVarianceMPI (CorrMetric.h:
// 987-1003, WAvg.h:440-478): reslice each entry with |weight| > 0 at its alignment (setTransform(NULL),
// getImage(image, store, false): valid voxels minus their mean, 0 outside), var += |w| (T - avg)^2 / sum|w|
int ettk_avg_variance(ettk_plan* P, const float* d_particles, int n, const ettk_particle* descs, const float* weights,
                      const double* matrices, const float* d_average, float* d_out) {
    try {
        ETTK_REQUIRE(P && d_particles && descs && weights && matrices && d_average && d_out && n > 0, ETTK_ERR_ARG,
                     "ettk_avg_variance: bad argument");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        const int N = P->N;
        const size_t n3 = (size_t)N * N * N;
        hipStream_t s = P->stream;
        ensure_ccc_workspace(P, std::min(std::max(n, 1), 128));
        double wsum = 0;
        std::vector<int> sel;
        for (int i = 0; i < n; i++) {
            wsum += std::fabs((double)weights[i]);
            if (std::fabs(weights[i]) > 0) sel.push_back(i);
        }
        ETTK_REQUIRE(!sel.empty(), ETTK_ERR_ARG, "ettk_avg_variance: no entry has a nonzero weight");
        DevBuf<double> var;
        var.ensure(n3);
        launch_fill_zero(var.p, n3, s);
        std::vector<EvalGeom> g(P->cap);
        std::vector<float> w(P->cap);
        DevBuf<float> dw;
        dw.ensure(P->cap);
        for (size_t c0 = 0; c0 < sel.size(); c0 += P->cap) {
            const int E = (int)std::min<size_t>(P->cap, sel.size() - c0);
            for (int e = 0; e < E; e++) {
                const int i = sel[c0 + e];
                ettk_particle d = descs[i];
                for (int k = 0; k < 16; k++) d.matrix[k] = (k % 5 == 0) ? 1.0 : 0.0;  // setTransform(NULL)
                g[e] = make_geom_public(P, d, Mat4::from(matrices + (size_t)i * 16), i);
                w[e] = weights[i];
            }
            ETTK_HIP(hipMemcpyAsync(P->geom.p, g.data(), E * sizeof(EvalGeom), hipMemcpyHostToDevice, s));
            ETTK_HIP(hipMemcpyAsync(dw.p, w.data(), E * sizeof(float), hipMemcpyHostToDevice, s));
            launch_reslice(d_particles, P->geom.p, E, N, P->R2.p, P->p_res.p, P->nb_res, s);
            launch_reslice_mean(P->p_res.p, P->nb_res, E, P->stats.p, s);
            launch_var_acc(P->R2.p, P->stats.p, dw.p, E, wsum, d_average, N, var.p, s);
            ETTK_HIP(hipStreamSynchronize(s));
        }
        launch_var_out(var.p, N, d_out, s);
        ETTK_HIP(hipStreamSynchronize(s));
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return ETTK_ERR_ALLOC;
    }
    set_error("");
    return ETTK_OK;
}

// nbfWedgedAverageImage3D::getAccumulatedWedgeImage / getWedgeImage / getSphericalWedgeImage of an average
// whose alignments are the n entries (WAvg.h:496-611, 654-690); see avg_wedge_compute
int ettk_avg_wedge(ettk_plan* P, int n, const ettk_particle* descs, const float* weights, const double* matrices,
                   float* d_accum, float* d_wedge, float* out_spherical, int* out_effective) {
    try {
        ETTK_REQUIRE(P && descs && weights && matrices && n > 0, ETTK_ERR_ARG, "ettk_avg_wedge: bad argument");
        avg_wedge_compute(P, n, descs, weights, matrices);
        const int N = P->N, ns = 2 * P->p.B;
        const size_t n3 = (size_t)N * N * N;
        hipStream_t s = P->stream;
        if (d_accum) ETTK_HIP(hipMemcpyAsync(d_accum, P->wacc.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
        if (d_wedge) {
            if (P->wacc_effective) {
                ETTK_HIP(hipMemcpyAsync(d_wedge, P->wsm.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
            } else {
                std::vector<float> ones(n3, 1.0f);  // getWedgeImage: A = 1 when the wedge is not effective
                ETTK_HIP(hipMemcpyAsync(d_wedge, ones.data(), n3 * sizeof(float), hipMemcpyHostToDevice, s));
                ETTK_HIP(hipStreamSynchronize(s));
            }
        }
        if (out_spherical) {
            for (size_t q = 0; q < (size_t)ns * ns; q++)
                out_spherical[q] = P->wacc_effective ? P->sph_acc_bin[q] : 1.0f;  // getSphericalWedgeImage
        }
        if (out_effective) *out_effective = P->wacc_effective ? 1 : 0;
        ETTK_HIP(hipStreamSynchronize(s));
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return ETTK_ERR_ALLOC;
    }
    set_error("");
    return ETTK_OK;
}

// getImage (WAvg.h:225-243): F = (re + i im)/acc where acc > 0, DC = 0, c2r / N^3 * shift, RMS = 1
int ettk_avg_finalize(ettk_plan* P, float* d_out) {
    try {
        ETTK_REQUIRE(P && d_out, ETTK_ERR_ARG, "ettk_avg_finalize: null argument");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        const int N = P->N;
        ETTK_REQUIRE(acc_buf(P, 0).n >= (size_t)N * N * P->H, ETTK_ERR_STATE, "ettk_avg_finalize: nothing accumulated");
        hipStream_t s = P->stream;
        ensure_ccc_workspace(P, 1);
        launch_avg_final(acc_buf(P, 0).p, acc_buf(P, 1).p, acc_buf(P, 2).p, N, P->S1.p, s);
        fft3_c2r(P->ctx->fft, P->S1.p, P->S1.p, P->R1.p, N, 1, s);
        launch_avg_normalise(P->R1.p, N, d_out, s);
        ETTK_HIP(hipStreamSynchronize(s));
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return ETTK_ERR_ALLOC;
    }
    set_error("");
    return ETTK_OK;
}

}  // extern "C"
