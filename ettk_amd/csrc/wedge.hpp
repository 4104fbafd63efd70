// wedge.hpp -- device functions of the analytic missing wedge (shared by ccc.hip and avg.hip).
#pragma once
#include "common.hpp"

namespace ettk {

// ------------------------------------------------------------------ smoothed analytic wedge
// binary wedge (v1.d)(v2.d) <= 0, d = (i - N/2, j - N/2, k - N/2), on an N x N x nk grid: the half grid
// (nk = N/2 + 1, nbfWedge3D::getImageHalf, Wedge.h:313-360) or the full grid (nk = N, nbfWedge3D::getImage,
// Wedge.h:195-290, centre floor(N/2)); smoothed with the separable 3-tap Gaussian (sigma 2, radius factor 1/2:
// smoothWedgeImage, Wedge.h:432-445) clipped and renormalised at the borders of that grid
__device__ __forceinline__ float wedge_bin(const float* v1, const float* v2, int i, int j, int k, int N) {
    float di = (float)(i - N / 2), dj = (float)(j - N / 2), dk = (float)(k - N / 2);
    float a = v1[0] * di + v1[1] * dj + v1[2] * dk;
    float b = v2[0] * di + v2[1] * dj + v2[2] * dk;
    return (a * b <= 0.f) ? 1.f : 0.f;
}
__device__ __forceinline__ void tap3(int idx, int n, float* w) {
    const float g = 0.88249690258459540286f;  // exp(-1/8)
    float l = idx > 0 ? g : 0.f, r = idx < n - 1 ? g : 0.f;
    float inv = 1.f / (l + 1.f + r);
    w[0] = l * inv;
    w[1] = inv;
    w[2] = r * inv;
}
__device__ inline float wedge_smooth(const float* v1, const float* v2, int i, int j, int k, int N, int nk) {
    // Away from both planes the 27 binary samples all equal the centre's: the smoothed value is then 0 or the
    // kernel's sum, accumulated in the same order as the full stencil below (so the value is identical).
    const float di = (float)(i - N / 2), dj = (float)(j - N / 2), dk = (float)(k - N / 2);
    const float a0 = v1[0] * di + v1[1] * dj + v1[2] * dk, b0 = v2[0] * di + v2[1] * dj + v2[2] * dk;
    const float r1 = fabsf(v1[0]) + fabsf(v1[1]) + fabsf(v1[2]), r2 = fabsf(v2[0]) + fabsf(v2[1]) + fabsf(v2[2]);
    const bool uniform = fabsf(a0) > r1 * 1.001f + 1e-3f && fabsf(b0) > r2 * 1.001f + 1e-3f;
    if (uniform && a0 * b0 > 0.f) return 0.f;
    float wi[3], wj[3], wk[3];
    tap3(i, N, wi);
    tap3(j, N, wj);
    tap3(k, nk, wk);
    if (uniform) {
        float acc = 0.f;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (wi[a] == 0.f) continue;
            float sj = 0.f;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (wj[b] == 0.f) continue;
                float sk = 0.f;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    if (wk[c] == 0.f) continue;
                    sk += wk[c] * 1.f;
                }
                sj += wj[b] * sk;
            }
            acc += wi[a] * sj;
        }
        return acc;
    }
    float acc = 0.f;
#pragma unroll
    for (int a = -1; a <= 1; a++) {
        if (wi[a + 1] == 0.f) continue;
        float sj = 0.f;
#pragma unroll
        for (int b = -1; b <= 1; b++) {
            if (wj[b + 1] == 0.f) continue;
            float sk = 0.f;
#pragma unroll
            for (int c = -1; c <= 1; c++) {
                if (wk[c + 1] == 0.f) continue;
                sk += wk[c + 1] * wedge_bin(v1, v2, i + a, j + b, k + c, N);
            }
            sj += wj[b + 1] * sk;
        }
        acc += wi[a + 1] * sj;
    }
    return acc;
}

}  // namespace ettk
