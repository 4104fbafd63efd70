// avg.hip -- kernels of the amplitude-weighted Fourier average (gfx950).
//
// nbfWedgedAverageImage3D::getImageMPI (inc/em/nbfWedgedAverageImage3D.h:301-351) accumulates, per half
// spectrum element, acc += |w| |F|, re += w |F| Re F, im += w |F| Im F over the aligned particles;
// getImage (:225-243) divides, zeroes DC, inverts and RMS-normalises.  The accumulation visits the particles
// of a chunk in order for every element (FP64 accumulators), so the sum is deterministic.
#include <cfloat>

#include "kernels.hpp"

namespace ettk {

__global__ void k_avg_acc(const float2* __restrict__ F, const float* __restrict__ w, int E, long nh,
                          double* __restrict__ re, double* __restrict__ im, double* __restrict__ acc) {
    for (long h = blockIdx.x * (long)blockDim.x + threadIdx.x; h < nh; h += (long)gridDim.x * blockDim.x) {
        double r = re[h], i = im[h], a = acc[h];
        for (int e = 0; e < E; e++) {
            float2 z = F[(size_t)e * nh + h];
            double B = sqrt((double)z.x * z.x + (double)z.y * z.y);
            double we = w[e];
            a += fabs(we) * B;
            r += we * B * z.x;
            i += we * B * z.y;
        }
        re[h] = r;
        im[h] = i;
        acc[h] = a;
    }
}

__global__ void k_avg_final(const double* __restrict__ re, const double* __restrict__ im, const double* __restrict__ acc,
                            int N, float2* __restrict__ F) {
    const int H = N / 2 + 1;
    const long nh = (long)N * N * H, dc = ((long)(N / 2) * N + N / 2) * H + (H - 1);
    for (long h = blockIdx.x * (long)blockDim.x + threadIdx.x; h < nh; h += (long)gridDim.x * blockDim.x) {
        float2 z = make_float2(0.f, 0.f);
        if (acc[h] > 0 && h != dc) z = make_float2((float)(re[h] / acc[h]), (float)(im[h] / acc[h]));
        F[h] = z;
    }
}

// avg = c2r(F) / N^3 * shift, then divided by sqrt(mean(avg^2)); written x fastest
__global__ void k_avg_normalise(const float* __restrict__ A, int N, float* __restrict__ out) {
    __shared__ double sh[256];
    const long n3 = (long)N * N * N;
    const double inv = 1.0 / (double)n3;
    double s = 0;
    for (long o = threadIdx.x; o < n3; o += blockDim.x) {
        int i, j, k;
        split3((unsigned)o, (unsigned)N, (unsigned)N, i, j, k);
        double v = A[o] * inv * (((i + j + k) & 1) ? -1.0 : 1.0);
        s += v * v;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    const float rms = (float)sqrt(sh[0] / (double)n3);  // varianceInside is a Pixel (float)
    for (long o = threadIdx.x; o < n3; o += blockDim.x) {
        int z, y, x;
        split3((unsigned)o, (unsigned)N, (unsigned)N, z, y, x);
        long src = ((long)x * N + y) * N + z;
        double v = A[src] * inv * (((x + y + z) & 1) ? -1.0 : 1.0);
        out[o] = (float)(v / rms);
    }
}

void launch_avg_acc(const float2* F, const float* w, int E, long nh, double* re, double* im, double* acc, hipStream_t s) {
    k_avg_acc<<<1024, 256, 0, s>>>(F, w, E, nh, re, im, acc);
    ETTK_HIP(hipGetLastError());
}
void launch_avg_final(const double* re, const double* im, const double* acc, int N, float2* F, hipStream_t s) {
    k_avg_final<<<1024, 256, 0, s>>>(re, im, acc, N, F);
    ETTK_HIP(hipGetLastError());
}
void launch_avg_normalise(const float* A, int N, float* out, hipStream_t s) {
    k_avg_normalise<<<1, 256, 0, s>>>(A, N, out);
    ETTK_HIP(hipGetLastError());
}

}  // namespace ettk
