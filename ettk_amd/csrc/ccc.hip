// ccc.hip -- kernels of the wedge-constrained FFT cross-correlation (gfx950).
//
// Device side of nbfFourierImageMetric::executeFourierNewHalf (inc/em/nbfFourierImageMetric.h:1078-1323)
// for a batch of (particle, transform) evaluations:
//   reslice  -> nbfWedgedSubImage3D::getImage (inc/em/nbfWedgedSubImage3D.h:603-828), VTK tricubic
//   spec     -> nbfFourierFilter::executeHalf (nbfFourierFilter.h:587-606) with the smoothed analytic wedge
//               of nbfWedge3D::getImageHalf (nbfWedge3D.h:313-360) generated on the fly, and the Hermitian
//               weighted sums of normalizeFourierHalf (nbfCorrelationImageMetric.h:498-530)
//   combine  -> the quadratic expression of FIM.h:1202-1203
//   dftA/BC  -> the zero-padded (2N)^3 c2r of FIM.h:1209-1214 evaluated only on the search window
//               (FIM.h:1222-1247) as three separable partial DFTs, followed by the spherical mask and the
//               first-maximum search of locateMaxima (nbfCorrelationImageMetric.h:304-323)
// Per-evaluation sums are written as per-block partials and reduced in a fixed order (deterministic).
#include "kernels.hpp"

#include <hip/hip_ext.h>
#include "wedge.hpp"

#include <algorithm>
#include <cfloat>

namespace ettk {

static constexpr int kThreads = 256;

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
// block sum (result valid in thread 0)
__device__ __forceinline__ float block_sum(float v, float* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    float r = 0.f;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x + 63) / 64; i++) r += sh[i];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------ reslice
__device__ __forceinline__ void keys_w(float f, float* w) {  // vtkTricubicInterpWeights
    float fm1 = f - 1.f, fd2 = f * 0.5f, ft3 = f * 3.f;
    w[0] = -fd2 * fm1 * fm1;
    w[1] = ((ft3 - 2.f) * fd2 - 1.f) * fm1;
    w[2] = -((ft3 - 4.f) * f - 1.f) * fd2;
    w[3] = f * fd2 * fm1;
}

// Validity of a sample is decided in FP32 unless it lies within 1e-3 of the border, where the FP64 map decides
// (the reference places samples in FP64: a sample that FP32 rounding moves across the border would otherwise
// enter or leave the valid set and change the mean the volume is centred on).
__device__ __forceinline__ bool near_border(float p, float lo, float hi) {
    return fabsf(p - lo) < 1e-3f || fabsf(p - hi) < 1e-3f;
}
__device__ __forceinline__ bool inside_fp64(const double* md, int i, int j, int k, double lo, double hx, double hy,
                                            double hz) {
    const double x = md[0] * i + md[3] * j + md[6] * k + md[9];
    const double y = md[1] * i + md[4] * j + md[7] * k + md[10];
    const double z = md[2] * i + md[5] * j + md[8] * k + md[11];
    return x >= lo && x <= hx && y >= lo && y <= hy && z >= lo && z <= hz;
}

// One workgroup reslices an 8x8x8 tile of the output.  The input voxels its 64-tap stencils can touch lie in
// the bounding box of the tile's transformed corners: at most floor(7 sqrt 3) + 1 + 4 = 17 voxels per axis for
// a rigid map.  That box (its actual extent, about 14 voxels per axis for a random rotation), with the volume's
// border replicated where it reaches outside (the taps clamp to the border), is staged in LDS at fixed strides,
// so every tap is a plain LDS read and the 4 taps of a row are consecutive words.  A non-rigid map whose box does
// not fit reads the clamped taps from global memory instead.
// Each wavefront interpolates a 4x4x4 block of the tile: the taps of its voxels then cover a compact region
// of the box and, with the row and plane strides 17 and 300, fall on fewer common LDS banks than those of an
// 8x8 plane.  A 17^3 box at these strides (20 KB) lets 8 workgroups share a CU's 160 KB of LDS.
// Measured on 256 evaluations at 64^3 (one launch): 654 us, against 775 us for an 18^3 box at strides 19/343,
// 8x8 planes per wavefront and two reductions; outputs bit-identical.
static constexpr int kTile = 8, kBox = 17, kSY = 17, kSZ = 300;
// kUse names the caller in kernel traces (0: CCC evaluations, 1: S^2 projection); the code is the same
template <int kUse>
__global__ void __launch_bounds__(kThreads) k_reslice(const float* __restrict__ parts, const EvalGeom* __restrict__ geom,
                                                       int N, float* __restrict__ out, float* __restrict__ partials) {
    __shared__ float2 sh[kThreads / 64];
    __shared__ float box[kBox * kSZ];
    int e, tb;
    xcd_eval_tile(blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, e, tb);
    const EvalGeom g = geom[e];
    const float* vol = parts + (size_t)g.particle * N * N * N;
    const long n3 = (long)N * N * N;
    const int nt = N / kTile;
    const int ti = tb / (nt * nt), tj = (tb / nt) % nt, tk = tb % nt;
    const int i0 = ti * kTile, j0 = tj * kTile, k0 = tk * kTile;
    const float tol = 7.62939453125e-06f;
    // staged box [lo, hi] per input axis (x, y, z); it may reach outside the volume
    int lo[3], hi[3];
    bool fits = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float c0 = g.m[a] * i0 + g.m[3 + a] * j0 + g.m[6 + a] * k0 + g.m[9 + a];
        float ext = (kTile - 1) * (fabsf(g.m[a]) + fabsf(g.m[3 + a]) + fabsf(g.m[6 + a]));
        float mn = c0 + (kTile - 1) * (fminf(g.m[a], 0.f) + fminf(g.m[3 + a], 0.f) + fminf(g.m[6 + a], 0.f));
        lo[a] = (int)floorf(mn + tol) - 1;
        hi[a] = (int)floorf(mn + ext + tol) + 2;
        if (!(hi[a] - lo[a] + 1 <= kBox) || !(fabsf(mn) < 1e6f)) fits = false;
    }
    if (fits) {
        // 2-D thread map over (y, x) of the box, loop over z; coordinates clamped to the volume
        constexpr int kRows = kThreads / kBox;
        const int t = threadIdx.x;
        const int wx = hi[0] - lo[0] + 1, wy = hi[1] - lo[1] + 1, wz = hi[2] - lo[2] + 1;
        const int x = t % kBox, y0 = t / kBox;
        if (t < kBox * kRows && x < wx) {
            const int gx = min(max(lo[0] + x, 0), N - 1);
            for (int y = y0; y < wy; y += kRows) {
                const int gy = min(max(lo[1] + y, 0), N - 1);
                const float* src = vol + (size_t)gy * N + gx;
#pragma unroll 6
                for (int z = 0; z < wz; z++) {
                    const int gz = min(max(lo[2] + z, 0), N - 1);
                    box[z * kSZ + y * kSY + x] = src[(size_t)gz * N * N];
                }
            }
        }
    }
    __syncthreads();
    float s = 0.f, c = 0.f;
    const float lob = -0.5f, hib = (float)N - 0.5f;
    for (int t = threadIdx.x; t < kTile * kTile * kTile; t += blockDim.x) {
        // t >> 6 numbers the tile's eight 4x4x4 blocks (two passes of four wavefronts), l = t & 63 the voxel
        // within its block
        const int sb = t >> 6, l = t & 63;
        const int i = i0 + 4 * (sb >> 2) + (l >> 4);
        const int j = j0 + 4 * ((sb >> 1) & 1) + ((l >> 2) & 3);
        const int k = k0 + 4 * (sb & 1) + (l & 3);
        float px = g.m[0] * i + g.m[3] * j + g.m[6] * k + g.m[9];
        float py = g.m[1] * i + g.m[4] * j + g.m[7] * k + g.m[10];
        float pz = g.m[2] * i + g.m[5] * j + g.m[8] * k + g.m[11];
        float v = -FLT_MAX;
        bool in = px >= lob && px <= hib && py >= lob && py <= hib && pz >= lob && pz <= hib;
        if (near_border(px, lob, hib) || near_border(py, lob, hib) || near_border(pz, lob, hib))
            in = inside_fp64(g.md, i, j, k, -0.5, N - 0.5, N - 0.5, N - 0.5);
        if (in) {
            float xx = px + tol, yy = py + tol, zz = pz + tol;
            int ix = (int)floorf(xx), iy = (int)floorf(yy), iz = (int)floorf(zz);
            if (fits) {
                // the box was sized from the tile's corners; a sample within rounding of a voxel boundary can
                // floor one voxel outside it: take the neighbouring cubic piece at f just outside [0, 1) instead
                // (the Keys kernel is C1, the value agrees to O(f^2) of a few ulps)
                ix = min(max(ix, lo[0] + 1), hi[0] - 2);
                iy = min(max(iy, lo[1] + 1), hi[1] - 2);
                iz = min(max(iz, lo[2] + 1), hi[2] - 2);
            }
            float wx[4], wy[4], wz[4];
            keys_w(xx - ix, wx);
            keys_w(yy - iy, wy);
            keys_w(zz - iz, wz);
            float acc = 0.f;
            if (fits) {
                const float* b0 = box + (iz - 1 - lo[2]) * kSZ + (iy - 1 - lo[1]) * kSY + (ix - 1 - lo[0]);
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    float ay = 0.f;
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const float* row = b0 + a * kSZ + b * kSY;
                        ay += wy[b] * (wx[0] * row[0] + wx[1] * row[1] + wx[2] * row[2] + wx[3] * row[3]);
                    }
                    acc += wz[a] * ay;
                }
            } else {
                int cx[4], cy[4], cz[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    cx[q] = min(max(ix - 1 + q, 0), N - 1);
                    cy[q] = min(max(iy - 1 + q, 0), N - 1);
                    cz[q] = min(max(iz - 1 + q, 0), N - 1);
                }
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    float ay = 0.f;
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const float* row = vol + ((size_t)cz[a] * N + cy[b]) * N;
                        ay += wy[b] * (wx[0] * row[cx[0]] + wx[1] * row[cx[1]] + wx[2] * row[cx[2]] + wx[3] * row[cx[3]]);
                    }
                    acc += wz[a] * ay;
                }
            }
            v = acc;
            s += v;
            c += 1.f;
        }
        out[(size_t)e * n3 + ((size_t)i * N + j) * N + k] = v;
    }
    // the tile's (sum, count) in one reduction: butterfly per wavefront, then the wavefronts in order
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = make_float2(s, c);
    __syncthreads();
    if (threadIdx.x == 0) {
        float ts = 0.f, tc = 0.f;
        for (int w = 0; w < kThreads / 64; w++) {
            ts += sh[w].x;
            tc += sh[w].y;
        }
        partials[((size_t)e * gridDim.x + tb) * 2 + 0] = ts;  // in tile order: the same sum order for any mapping
        partials[((size_t)e * gridDim.x + tb) * 2 + 1] = tc;
    }
}

// one wave per evaluation: lane l sums the partials b = l, l + 64, ... in FP64, then a fixed butterfly
__device__ __forceinline__ double wave_dsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__global__ void __launch_bounds__(64) k_reslice_mean(const float* __restrict__ partials, int nb, int n_eval,
                                                     float* __restrict__ stats) {
    const int e = blockIdx.x, l = threadIdx.x;
    double s = 0, c = 0;
    for (int b = l; b < nb; b += 64) {
        s += partials[((size_t)e * nb + b) * 2];
        c += partials[((size_t)e * nb + b) * 2 + 1];
    }
    s = wave_dsum(s);
    c = wave_dsum(c);
    if (l == 0) stats[e] = (float)(s / (double)(float)c);  // meanInside is a float (WSub.h:794-797)
}

int reslice_blocks(int N) { return (N / kTile) * (N / kTile) * (N / kTile); }

void launch_reslice(const float* parts, const EvalGeom* geom, int n_eval, int N, float* out, float* partials,
                    int nb, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop) {
    ETTK_REQUIRE(N % kTile == 0 && nb == reslice_blocks(N), ETTK_ERR_ARG, "launch_reslice: block count mismatch");
    dim3 grid(nb, n_eval);
    if (ev_start || ev_stop)  // the events take the kernel's own start and end timestamps
        hipExtLaunchKernelGGL(k_reslice<0>, grid, dim3(kThreads), 0, s, ev_start, ev_stop, 0, parts, geom, N, out,
                              partials);
    else
        k_reslice<0><<<grid, kThreads, 0, s>>>(parts, geom, N, out, partials);
    ETTK_HIP(hipGetLastError());
}
void launch_reslice_projection(const float* parts, const EvalGeom* geom, int n, int N, float* out, float* partials,
                               int nb, hipStream_t s) {
    ETTK_REQUIRE(N % kTile == 0 && nb == reslice_blocks(N), ETTK_ERR_ARG, "launch_reslice: block count mismatch");
    k_reslice<1><<<dim3(nb, n), kThreads, 0, s>>>(parts, geom, N, out, partials);
    ETTK_HIP(hipGetLastError());
}
void launch_reslice_mean(const float* partials, int nb, int n_eval, float* stats, hipStream_t s) {
    k_reslice_mean<<<n_eval, 64, 0, s>>>(partials, nb, n_eval, stats);
    ETTK_HIP(hipGetLastError());
}

__global__ void k_wedge_ref(WedgeRef w, int N, float* __restrict__ W) {
    const int H = N / 2 + 1;
    long nh = (long)N * N * H;
    for (long h = blockIdx.x * (long)blockDim.x + threadIdx.x; h < nh; h += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3((unsigned)h, (unsigned)N, (unsigned)H, i, j, k);
        W[h] = w.full ? 1.f : wedge_smooth(w.v1, w.v2, i, j, k, N, H);
    }
}
void launch_wedge_ref(const WedgeRef& w, int N, float* W, hipStream_t s) {
    k_wedge_ref<<<512, kThreads, 0, s>>>(w, N, W);
    ETTK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ spectral filter
// both sides of one evaluation in one pass, the smoothed wedge computed once per element:
//   X1 = F1 * filter * W1 W2(e)            (reference side, F1 shared by all evaluations; n1 and sum(W filter))
//   X2 = mcf(F2[e]) * filter * W1 W2(e)    (particle side, in place; n2)
// (nbfFourierFilter::executeHalf, FFilt.h:587-606, and the Hermitian weighted sums of normalizeFourierHalf,
// CorrMetric.h:498-530)
__global__ void __launch_bounds__(kThreads) k_spec_filter2(const float2* __restrict__ F1, float2* __restrict__ X1,
                                                            float2* __restrict__ X2, const float* __restrict__ filter,
                                                            const float* __restrict__ W1, const EvalGeom* __restrict__ geom,
                                                            int N, int mcf, int apply, float* __restrict__ p_n1,
                                                            float* __restrict__ p_wf, float* __restrict__ p_n2) {
    __shared__ float sh[8];
    const int e = blockIdx.y;
    const int H = N / 2 + 1;
    const long nh = (long)N * N * H;
    const EvalGeom& g = geom[e];
    const long chunk = (nh + gridDim.x - 1) / gridDim.x;
    const long h0 = blockIdx.x * chunk, h1 = min(nh, h0 + chunk);
    const long dc = ((long)(N / 2) * N + N / 2) * H + (H - 1);
    float s1 = 0.f, swf = 0.f, s2 = 0.f;
    for (long h = h0 + threadIdx.x; h < h1; h += blockDim.x) {
        int i, j, k;
        split3((unsigned)h, (unsigned)N, (unsigned)H, i, j, k);
        float w = W1 ? W1[h] : 1.f;
        if (!g.full) w *= wedge_smooth(g.v1, g.v2, i, j, k, N, H);
        const float fw = filter[h] * w;
        const float hw = k < H - 1 ? 2.f : 1.f;
        float2 a = F1[h];
        if (apply) a = cscale(a, fw);
        if (h == dc) a = make_float2(0.f, 0.f);
        X1[(size_t)e * nh + h] = a;
        s1 += hw * (a.x * a.x + a.y * a.y);
        swf += fw;
        float2 z = X2[(size_t)e * nh + h];
        if (mcf) {
            float m2 = z.x * z.x + z.y * z.y;
            float C = sqrtf(sqrtf(m2));
            if (C > 0.f) z = make_float2(z.x / C, z.y / C);
        }
        if (apply) z = cscale(z, fw);
        if (h == dc) z = make_float2(0.f, 0.f);
        X2[(size_t)e * nh + h] = z;
        s2 += hw * (z.x * z.x + z.y * z.y);
    }
    const float t1 = block_sum(s1, sh);
    const float tw = block_sum(swf, sh);
    const float t2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        p_n1[(size_t)e * gridDim.x + blockIdx.x] = t1;
        p_wf[(size_t)e * gridDim.x + blockIdx.x] = tw;
        p_n2[(size_t)e * gridDim.x + blockIdx.x] = t2;
    }
}
void launch_spec_filter2(const float2* F1, float2* X1, float2* X2, const float* filter, const float* W1,
                         const EvalGeom* geom, int n_eval, int N, int mcf, int apply, float* p_n1, float* p_wf,
                         float* p_n2, int nb, hipStream_t s) {
    dim3 grid(nb, n_eval);
    k_spec_filter2<<<grid, kThreads, 0, s>>>(F1, X1, X2, filter, W1, geom, N, mcf, apply, p_n1, p_wf, p_n2);
    ETTK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ wedge overlap
// nbfCorrelationImageMetric::getWedgeOverlap (CorrMetric.h:533-566): sum(W1 W2(t) filter) over the half grid,
// W2 the particle's smoothed wedge at the evaluation's transform (all ones when geom.full).  Per-block FP64
// partials, then one sum per evaluation.
__global__ void __launch_bounds__(kThreads) k_wedge_overlap(const float* __restrict__ W1, const float* __restrict__ filter,
                                                             const EvalGeom* __restrict__ geom, int N,
                                                             double* __restrict__ part) {
    __shared__ double sh[kThreads / 64];
    const int e = blockIdx.y, H = N / 2 + 1;
    const unsigned nh = (unsigned)N * N * H;
    const EvalGeom& g = geom[e];
    double acc = 0.0;
    for (unsigned h = blockIdx.x * kThreads + threadIdx.x; h < nh; h += gridDim.x * kThreads) {
        int i, j, k;
        split3(h, (unsigned)N, (unsigned)H, i, j, k);
        float w = W1[h] * (g.full ? 1.f : wedge_smooth(g.v1, g.v2, i, j, k, N, H));
        acc += (double)(w * filter[h]);
    }
    for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kThreads / 64; q++) t += sh[q];
        part[(size_t)e * gridDim.x + blockIdx.x] = t;
    }
}
__global__ void k_sum_partials(const double* __restrict__ part, int nb, int n_eval, double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_eval) return;
    double t = 0.0;
    for (int b = 0; b < nb; b++) t += part[(size_t)e * nb + b];
    out[e] = t;
}
void launch_wedge_overlap(const float* W1, const float* filter, const EvalGeom* geom, int E, int N, double* part,
                          double* out, hipStream_t s) {
    const int nb = 64;
    k_wedge_overlap<<<dim3(nb, E), kThreads, 0, s>>>(W1, filter, geom, N, part);
    k_sum_partials<<<(E + 127) / 128, 128, 0, s>>>(part, nb, E, out);
    ETTK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ reference-side real-space mask
__global__ void __launch_bounds__(kThreads) k_mask_ref(float* __restrict__ R, const float* __restrict__ m1, int N,
                                                        float* __restrict__ partials) {
    __shared__ float sh[8];
    const int e = blockIdx.y;
    const long n3 = (long)N * N * N;
    const long chunk = (n3 + gridDim.x - 1) / gridDim.x;
    const long o0 = blockIdx.x * chunk, o1 = min(n3, o0 + chunk);
    float s = 0.f;
    for (long o = o0 + threadIdx.x; o < o1; o += blockDim.x) {
        float r = R[(size_t)e * n3 + o], m = m1[o];
        s += r * r * m;
        R[(size_t)e * n3 + o] = r * m;
    }
    float t = block_sum(s, sh);
    if (threadIdx.x == 0) partials[(size_t)e * gridDim.x + blockIdx.x] = t;
}
void launch_mask_ref(float* R, const float* m1, int n_eval, int N, float* partials, int nb, hipStream_t s) {
    dim3 grid(nb, n_eval);
    k_mask_ref<<<grid, kThreads, 0, s>>>(R, m1, N, partials);
    ETTK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ per-evaluation scalars
__global__ void __launch_bounds__(64) k_ccc_scalars(const float* p_n1, const float* p_n2, const float* p_k1,
                                                    const float* p_wf, int nb, int nbr, int n_eval, int N, float* scalars) {
    const int e = blockIdx.x, l = threadIdx.x;
    double s1 = 0, s2 = 0, k1 = 0, wf = 0;
    for (int b = l; b < nb; b += 64) {
        s1 += p_n1[(size_t)e * nb + b];
        s2 += p_n2[(size_t)e * nb + b];
        if (p_wf) wf += p_wf[(size_t)e * nb + b];
    }
    for (int b = l; b < nbr; b += 64) k1 += p_k1[(size_t)e * nbr + b];
    s1 = wave_dsum(s1);
    s2 = wave_dsum(s2);
    k1 = wave_dsum(k1);
    wf = wave_dsum(wf);
    if (l != 0) return;
    double n3 = (double)N * N * N;
    float n1 = (float)sqrt(s1 / n3), n2 = (float)sqrt(s2 / n3);
    scalars[e * 4 + 0] = n1;
    scalars[e * 4 + 1] = n2;
    scalars[e * 4 + 2] = (float)(k1 / ((double)n1 * (double)n1));  // k1_m of the normalised reference
    scalars[e * 4 + 3] = (float)wf;
}
void launch_ccc_scalars(const float* p_n1, const float* p_n2, const float* p_k1, const float* p_wf, int nb,
                        int nbr, int n_eval, int N, float* scalars, hipStream_t s) {
    k_ccc_scalars<<<n_eval, 64, 0, s>>>(p_n1, p_n2, p_k1, p_wf, nb, nbr, n_eval, N, scalars);
    ETTK_HIP(hipGetLastError());
}

// acc + a b as four chained FMAs (the partial DFTs' inner product)
__device__ __forceinline__ float2 cmac(float2 acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
    return acc;
}

// ------------------------------------------------------------------ combine + first partial DFT (along i)
__global__ void __launch_bounds__(kThreads) k_combine_dftA(const float2* __restrict__ F1m, const float2* __restrict__ F2,
                                                            const float2* __restrict__ Q, const float2* __restrict__ Mfft,
                                                            const float* __restrict__ scalars,
                                                            const float2* __restrict__ tabA, CccSizes z,
                                                            float2* __restrict__ G1) {
    extern __shared__ float2 lds[];
    const int N = z.N, H = z.H, E = z.E;
    float2* C = lds;             // [i][k]
    float2* A = lds + N * H;     // [x][i]
    const int j = blockIdx.x, e = blockIdx.y;
    const long nh = (long)N * N * H;
    const float n1 = scalars[e * 4 + 0], n2 = scalars[e * 4 + 1];
    const float n3 = (float)N * N * N;
    const float a = 2.f * n3 / (n1 * n2);
    const float b = 1.f / (n2 * n2);
    struct Quad {
        float2 f1, f2, q, m;
    };
    const float2* __restrict__ f1p = F1m + (size_t)e * nh;
    const float2* __restrict__ f2p = F2 + (size_t)e * nh;
    const float2* __restrict__ qp = Q + (size_t)e * nh;
    block_copy_rt<5>(
        N * H,
        [&](int t) {
            const long h = ((long)(t / H) * N + j) * H + t % H;
            return Quad{f1p[h], f2p[h], qp[h], Mfft[h]};
        },
        [&](int t, const Quad& v) {
            const int i = t / H, k = t % H;
            float2 c = csub(cscale(cmulc(v.f1, v.f2), a), cscale(cmulc(v.m, v.q), b));
            if ((i + j + k) & 1) c = make_float2(-c.x, -c.y);
            C[t] = c;
        });
    block_copy_rt<9>(E * N, [&](int t) { return tabA[t]; }, [&](int t, float2 v) { A[t] = v; });
    __syncthreads();
    for (int t = threadIdx.x; t < E * H; t += blockDim.x) {
        int x = t / H, k = t % H;
        float2 acc = make_float2(0.f, 0.f);
        for (int i = 0; i < N; i++) acc = cmac(acc, A[x * N + i], C[i * H + k]);
        G1[(((size_t)e * E + x) * N + j) * H + k] = acc;
    }
}
void launch_combine_dftA(const float2* F1m, const float2* F2, const float2* Q, const float2* Mfft,
                         const float* scalars, const float2* tabA, const CccSizes& z, int n_eval, float2* G1,
                         hipStream_t s) {
    dim3 grid(z.N, n_eval);
    size_t shm = (size_t)(z.N * z.H + z.E * z.N) * sizeof(float2);
    ensure_dynamic_lds((const void*)k_combine_dftA, shm);
    k_combine_dftA<<<grid, kThreads, shm, s>>>(F1m, F2, Q, Mfft, scalars, tabA, z, G1);
    ETTK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ second and third partial DFTs + argmax
__global__ void __launch_bounds__(kThreads) k_dftBC_argmax(const float2* __restrict__ G1, const float2* __restrict__ tabA,
                                                            const float2* __restrict__ tabK, CccSizes z,
                                                            float* __restrict__ best_val, int* __restrict__ best_idx) {
    extern __shared__ float2 lds[];
    const int N = z.N, H = z.H, E = z.E;
    float2* g1 = lds;                 // [j][k]
    float2* A = g1 + N * H;           // [y][j]
    float2* g2 = A + E * N;           // [y][k]
    float2* K = g2 + E * H;           // [zz][k]
    __shared__ float sv[kThreads];
    __shared__ int si[kThreads];
    const int x = blockIdx.x, e = blockIdx.y;
    const float2* src = G1 + ((size_t)e * E + x) * N * H;
    block_copy_rt<9>(N * H, [&](int t) { return src[t]; }, [&](int t, float2 v) { g1[t] = v; });
    block_copy_rt<9>(E * N, [&](int t) { return tabA[t]; }, [&](int t, float2 v) { A[t] = v; });
    block_copy_rt<5>(E * H, [&](int t) { return tabK[t]; }, [&](int t, float2 v) { K[t] = v; });
    __syncthreads();
    for (int t = threadIdx.x; t < E * H; t += blockDim.x) {
        int y = t / H, k = t % H;
        float2 acc = make_float2(0.f, 0.f);
        for (int j = 0; j < N; j++) acc = cmac(acc, A[y * N + j], g1[j * H + k]);
        g2[t] = acc;
    }
    __syncthreads();
    const float scale = 8.f / ((float)(2 * N) * (float)(2 * N) * (float)(2 * N));
    float bv = -FLT_MAX;
    int bi = 0x7fffffff;
    const int dx = z.ub2 - x;
    for (int t = threadIdx.x; t < E * E; t += blockDim.x) {
        int y = t / E, zz = t % E;
        int dy = z.ub2 - y, dz = z.ub2 - zz;
        float rad = (float)(dx * dx + dy * dy + dz * dz);
        if (!(rad < z.restr2)) continue;
        float acc = 0.f;
        for (int k = 0; k < H; k++) {
            const float2 p = g2[y * H + k], q = K[zz * H + k];
            acc = fmaf(p.x, q.x, acc);
            acc = fmaf(-p.y, q.y, acc);
        }
        acc *= scale;
        if (acc > bv || (acc == bv && t < bi)) { bv = acc; bi = t; }
    }
    sv[threadIdx.x] = bv;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            float v2 = sv[threadIdx.x + st];
            int i2 = si[threadIdx.x + st];
            if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
                sv[threadIdx.x] = v2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        best_val[(size_t)e * E + x] = sv[0];
        best_idx[(size_t)e * E + x] = si[0];
    }
}
void launch_dftBC_argmax(const float2* G1, const float2* tabA, const float2* tabK, const CccSizes& z, int n_eval,
                         float* best_val, int* best_idx, hipStream_t s) {
    dim3 grid(z.E, n_eval);
    size_t shm = (size_t)(z.N * z.H + z.E * z.N + 2 * z.E * z.H) * sizeof(float2);
    ensure_dynamic_lds((const void*)k_dftBC_argmax, shm);
    k_dftBC_argmax<<<grid, kThreads, shm, s>>>(G1, tabA, tabK, z, best_val, best_idx);
    ETTK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ final reduction per evaluation
__global__ void k_peak_final(const float* __restrict__ best_val, const int* __restrict__ best_idx,
                             const float* __restrict__ scalars, float sum_filter, int overlap, CccSizes z, int n_eval,
                             float* __restrict__ out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_eval) return;
    const int E = z.E;
    float bv = -FLT_MAX;
    int bx = -1, byz = 0;
    for (int x = 0; x < E; x++) {
        float v = best_val[(size_t)e * E + x];
        int id = best_idx[(size_t)e * E + x];
        if (id == 0x7fffffff) continue;
        if (bx < 0 || v > bv) { bv = v; bx = x; byz = id; }
    }
    float px = 0.f, py = 0.f, pz = 0.f;
    double maxT = -(double)FLT_MAX;
    if (bx >= 0) {
        px = (float)bx;
        py = (float)(byz / E);
        pz = (float)(byz % E);
        maxT = bv;
    }
    const float k1m = scalars[e * 4 + 2], wf = scalars[e * 4 + 3];
    double peak = fmax(0.0, (double)k1m - maxT);
    if (overlap) peak = peak / (double)wf * (double)sum_filter;
    float cp = (float)peak;
    double big = (double)(2 * z.N) * (2 * z.N) * (2 * z.N);
    cp = (float)((double)cp / big / big);
    float c = (E - 1) / 2.0f;
    out[e * 4 + 0] = cp;
    out[e * 4 + 1] = (px - c) / 2.0f;
    out[e * 4 + 2] = (py - c) / 2.0f;
    out[e * 4 + 3] = (pz - c) / 2.0f;
}
void launch_peak_final(const float* best_val, const int* best_idx, const float* scalars, float sum_filter,
                       int overlap, const CccSizes& z, int n_eval, float* out, hipStream_t s) {
    k_peak_final<<<(n_eval + 127) / 128, 128, 0, s>>>(best_val, best_idx, scalars, sum_filter, overlap, z, n_eval, out);
    ETTK_HIP(hipGetLastError());
}

template <class T>
__global__ void k_fill_zero(T* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = T(0);
}
template <class T>
static void fill_zero(T* p, size_t n, hipStream_t s) {
    if (n == 0) return;
    const unsigned blocks = (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 4096);
    k_fill_zero<T><<<blocks, kThreads, 0, s>>>(p, n);
    ETTK_HIP(hipGetLastError());
}
void launch_fill_zero(int* p, size_t n, hipStream_t s) { fill_zero(p, n, s); }
void launch_fill_zero(double* p, size_t n, hipStream_t s) { fill_zero(p, n, s); }

// ------------------------------------------------------------------ helpers used by the plan set-up
__global__ void k_xfast_to_internal(const float* __restrict__ in, float* __restrict__ out, int N) {
    long n3 = (long)N * N * N;
    for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n3; o += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3((unsigned)o, (unsigned)N, (unsigned)N, i, j, k);
        out[o] = in[((size_t)k * N + j) * N + i];
    }
}
void launch_xfast_to_internal(const float* in, float* out, int N, hipStream_t s) {
    k_xfast_to_internal<<<1024, kThreads, 0, s>>>(in, out, N);
    ETTK_HIP(hipGetLastError());
}
__global__ void k_internal_to_xfast(const float* __restrict__ in, float* __restrict__ out, int N) {
    long n3 = (long)N * N * N;
    for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n3; o += (long)gridDim.x * blockDim.x) {
        int zz, y, x;
        split3((unsigned)o, (unsigned)N, (unsigned)N, zz, y, x);
        out[o] = in[((size_t)x * N + y) * N + zz];
    }
}
void launch_internal_to_xfast(const float* in, float* out, int N, hipStream_t s) {
    k_internal_to_xfast<<<1024, kThreads, 0, s>>>(in, out, N);
    ETTK_HIP(hipGetLastError());
}
__global__ void k_mcf_inplace(float2* __restrict__ F, long n) {
    for (long h = blockIdx.x * (long)blockDim.x + threadIdx.x; h < n; h += (long)gridDim.x * blockDim.x) {
        float2 z = F[h];
        float C = sqrtf(sqrtf(z.x * z.x + z.y * z.y));
        if (C > 0.f) F[h] = make_float2(z.x / C, z.y / C);
    }
}
void launch_mcf_inplace(float2* F, long n, hipStream_t s) {
    k_mcf_inplace<<<1024, kThreads, 0, s>>>(F, n);
    ETTK_HIP(hipGetLastError());
}

}  // namespace ettk
