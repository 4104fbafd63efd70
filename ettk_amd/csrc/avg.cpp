// avg.cpp -- host side of the amplitude-weighted Fourier average (device code in avg.hip).
#include "plan.hpp"

namespace ettk {
EvalGeom make_geom_public(const ettk_plan* P, const ettk_particle& d, const Mat4& t, int particle);
void launch_avg_acc(const float2* F, const float* w, int E, long nh, double* re, double* im, double* acc, hipStream_t s);
void launch_avg_final(const double* re, const double* im, const double* acc, int N, float2* F, hipStream_t s);
void launch_avg_normalise(const float* A, int N, float* out, hipStream_t s);
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
            ETTK_HIP(hipMemsetAsync(b.p, 0, nh * sizeof(double), P->ctx->stream));
        }
        ETTK_HIP(hipStreamSynchronize(P->ctx->stream));
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
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
        hipStream_t s = P->ctx->stream;
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
        hipStream_t s = P->ctx->stream;
        ensure_ccc_workspace(P, 1);
        launch_avg_final(acc_buf(P, 0).p, acc_buf(P, 1).p, acc_buf(P, 2).p, N, P->S1.p, s);
        fft3_c2r(P->ctx->fft, P->S1.p, P->S1.p, P->R1.p, N, 1, s);
        launch_avg_normalise(P->R1.p, N, d_out, s);
        ETTK_HIP(hipStreamSynchronize(s));
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    }
    set_error("");
    return ETTK_OK;
}

}  // extern "C"
