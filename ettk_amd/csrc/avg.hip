// avg.hip -- kernels of the amplitude-weighted Fourier average (gfx950).
//
// nbfWedgedAverageImage3D::getImageMPI (inc/em/nbfWedgedAverageImage3D.h:301-351) accumulates, per half
// spectrum element, acc += |w| |F|, re += w |F| Re F, im += w |F| Im F over the aligned particles;
// getImage (:225-243) divides, zeroes DC, inverts and RMS-normalises.  The accumulation visits the particles
// of a chunk in order for every element (FP64 accumulators), so the sum is deterministic.
#include <cfloat>

#include "kernels.hpp"
#include "wedge.hpp"

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

// ------------------------------------------------------------------ variance map of an average
// nbfWedgedAverageImage3D::getImageVarianceMPI (WAvg.h:440-478): var += |w| (T - average)^2 / sum|w| over the
// entries in order, T the resliced particle minus its valid-voxel mean (0 outside the particle box).  R holds E
// resliced volumes in A(i,j,k) order (k fastest), `avg` and the output are x fastest.
__global__ void __launch_bounds__(256) k_var_acc(const float* __restrict__ R, const float* __restrict__ mean,
                                                 const float* __restrict__ w, int E, double wsum, const float* __restrict__ avg,
                                                 int N, double* __restrict__ var) {
    const unsigned n3 = (unsigned)N * N * N;
    for (unsigned o = blockIdx.x * blockDim.x + threadIdx.x; o < n3; o += gridDim.x * blockDim.x) {
        int i, j, k;
        split3(o, (unsigned)N, (unsigned)N, i, j, k);
        const double a = avg[((unsigned)k * N + j) * N + i];
        double v = var[o];
        for (int e = 0; e < E; e++) {
            const float x = R[(size_t)e * n3 + o];
            const double T = (x == -FLT_MAX) ? 0.0 : (double)x - (double)mean[e];
            const double d = T - a;
            v += fabs((double)w[e]) * (d * d) / wsum;
        }
        var[o] = v;
    }
}
__global__ void __launch_bounds__(256) k_var_out(const double* __restrict__ var, int N, float* __restrict__ out) {
    const unsigned n3 = (unsigned)N * N * N;
    for (unsigned o = blockIdx.x * blockDim.x + threadIdx.x; o < n3; o += gridDim.x * blockDim.x) {
        int z, y, x;
        split3(o, (unsigned)N, (unsigned)N, z, y, x);
        out[o] = (float)var[((unsigned)x * N + y) * N + z];
    }
}
void launch_var_acc(const float* R, const float* mean, const float* w, int E, double wsum, const float* avg, int N,
                    double* var, hipStream_t s) {
    const unsigned n3 = (unsigned)N * N * N;
    k_var_acc<<<(n3 + 255) / 256, 256, 0, s>>>(R, mean, w, E, wsum, avg, N, var);
    ETTK_HIP(hipGetLastError());
}
void launch_var_out(const double* var, int N, float* out, hipStream_t s) {
    const unsigned n3 = (unsigned)N * N * N;
    k_var_out<<<(n3 + 255) / 256, 256, 0, s>>>(var, N, out);
    ETTK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ accumulated wedge of an average
// nbfWedgedAverageImage3D::updateAccumulatedWedgeImage (WAvg.h:496-538): accum += getWedgeImage(current, store)
// over the (particle, candidate) entries in order, each the smoothed full-grid wedge of nbfWedge3D::getImage
// (Wedge.h:195-290).  One thread per voxel of the A(i,j,k) grid (k fastest) runs the entries in order, so the
// float sum is accumulated in the reference's order.
__global__ void __launch_bounds__(256) k_wedge_accum(const WedgeRef* __restrict__ w, int m, int N,
                                                     float* __restrict__ acc) {
    const unsigned n3 = (unsigned)N * N * N;
    for (unsigned o = blockIdx.x * blockDim.x + threadIdx.x; o < n3; o += gridDim.x * blockDim.x) {
        int i, j, k;
        split3(o, (unsigned)N, (unsigned)N, i, j, k);
        float a = 0.f;
        for (int e = 0; e < m; e++) a += w[e].full ? 1.f : wedge_smooth(w[e].v1, w[e].v2, i, j, k, N, N);
        acc[o] = a;
    }
}

// setAccumulatedWedgeImage (WAvg.h:49-70): bin = accum > max(accum)/2, count = sum(bin).  One workgroup.
__global__ void __launch_bounds__(1024) k_wedge_binarize(const float* __restrict__ acc, unsigned n3,
                                                         float* __restrict__ bin, int* __restrict__ count) {
    __shared__ float smx[16];
    __shared__ int scn[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float mx = -FLT_MAX;
    for (unsigned o = threadIdx.x; o < n3; o += blockDim.x) mx = fmaxf(mx, acc[o]);
    for (int d = 32; d > 0; d >>= 1) mx = fmaxf(mx, __shfl_xor(mx, d));
    if (lane == 0) smx[wv] = mx;
    __syncthreads();
    mx = smx[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) mx = fmaxf(mx, smx[q]);
    const float half = mx * 0.5f;  // max / 2.0: exact halving of a float
    int c = 0;
    for (unsigned o = threadIdx.x; o < n3; o += blockDim.x) {
        const bool on = acc[o] > half;
        bin[o] = on ? 1.f : 0.f;
        c += on ? 1 : 0;
    }
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
    if (lane == 0) scn[wv] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += scn[q];
        count[0] = t;
    }
}

// smoothWedgeImage of the binarised accumulation on the full grid (3-tap clipped Gaussian per axis, k then j
// then i as vtkImageGaussianSmooth runs its axes); also writes the half grid k < N/2 + 1 (getWedgeImageHalf,
// WAvg.h:602-611)
__global__ void __launch_bounds__(256) k_wedge_smooth_grid(const float* __restrict__ bin, int N, float* __restrict__ full,
                                                           float* __restrict__ half) {
    const unsigned n3 = (unsigned)N * N * N;
    const int H = N / 2 + 1;
    for (unsigned o = blockIdx.x * blockDim.x + threadIdx.x; o < n3; o += gridDim.x * blockDim.x) {
        int i, j, k;
        split3(o, (unsigned)N, (unsigned)N, i, j, k);
        float wi[3], wj[3], wk[3];
        tap3(i, N, wi);
        tap3(j, N, wj);
        tap3(k, N, wk);
        float acc = 0.f;
#pragma unroll
        for (int a = -1; a <= 1; a++) {
            if (wi[a + 1] == 0.f) continue;
            float sj = 0.f;
#pragma unroll
            for (int b = -1; b <= 1; b++) {
                if (wj[b + 1] == 0.f) continue;
                const float* row = bin + ((size_t)(i + a) * N + (j + b)) * N + k;
                float sk = 0.f;
#pragma unroll
                for (int c = -1; c <= 1; c++) {
                    if (wk[c + 1] == 0.f) continue;
                    sk += wk[c + 1] * row[c];
                }
                sj += wj[b + 1] * sk;
            }
            acc += wi[a + 1] * sj;
        }
        if (full) full[o] = acc;
        if (k < H) half[((size_t)i * N + j) * H + k] = acc;
    }
}

void launch_wedge_accum(const WedgeRef* w, int m, int N, float* acc, hipStream_t s) {
    const unsigned n3 = (unsigned)N * N * N;
    k_wedge_accum<<<(n3 + 255) / 256, 256, 0, s>>>(w, m, N, acc);
    ETTK_HIP(hipGetLastError());
}
void launch_wedge_binarize(const float* acc, int N, float* bin, int* count, hipStream_t s) {
    k_wedge_binarize<<<1, 1024, 0, s>>>(acc, (unsigned)N * N * N, bin, count);
    ETTK_HIP(hipGetLastError());
}
void launch_wedge_smooth_grid(const float* bin, int N, float* full, float* half, hipStream_t s) {
    const unsigned n3 = (unsigned)N * N * N;
    k_wedge_smooth_grid<<<(n3 + 255) / 256, 256, 0, s>>>(bin, N, full, half);
    ETTK_HIP(hipGetLastError());
}

}  // namespace ettk
