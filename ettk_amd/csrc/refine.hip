// refine.hip -- continuous constrained CCC and its adjoint for local pose refinement (gfx950).
//
// No reference counterpart: this replaces the Powell line searches of nbfFourierImageMetric::refine
// (inc/em/nbfFourierImageMetric.h:1702-1809), whose every function value is a full constrained CCC on a
// half-voxel shift grid, by gradient ascent on a differentiable score of the six pose parameters.
//
// The score lives in the particle's own frame, where the particle's missing wedge is fixed:
//   X2 = H r2c((I2 - mean I2) maskfft (-1)^(x+y+z))               particle, fixed: spectra computed once
//   X1 = H r2c(I1'(theta) maskfft (-1)^(x+y+z))                    reference resliced into the particle frame
//   I1'(p) = (I1 - mean I1)(q(p)),  q = G' (p - N/2, 1),  G' = fin(theta)^-1 plus (N/2, N/2, N/2 + cut)
//            (trilinear with clamped taps: the reference is extended by its border values, so that I1' and S
//            are continuous in theta)
//   S  = <X1, X2> / (|X1| |X2|),  <a, b> = sum_h c_h Re(a_h conj(b_h))  (c_h: exact Hermitian weights)
// with H = band-pass x binary particle wedge (lab frame) x binary reference wedge (rotated by the seed).
// Volumes stay in the particle's memory order [z][y][x]; the half axis of every spectrum is x.
// Adjoint: dS = <Z, dX1> with Z = X2/(n1 n2) - S X1/n1^2, so dS/dI1'(p) = maskfft (-1)^(x+y+z) c2r(H Z)(p),
// and through the interpolator dS/dtheta_r =
// sum_ab dG'_r[a][b] Q[a][b], Q[a][b] = sum_p gamma(p) dI1/dq_a(q(p)) w_b(p), w = (x-N/2, y-N/2, z-N/2, 1):
// twelve moments per evaluation, contracted with dG'_r = -G dfin_r G on the host in FP64.
#include <cfloat>

#include "kernels.hpp"

namespace ettk {

static constexpr int kT = 256;

__device__ __forceinline__ float wsum64(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double dsum64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// trilinear sample of an x-fastest volume with clamped taps, and its gradient in (x, y, z)
__device__ __forceinline__ float trilin(const float* __restrict__ vol, int N, float px, float py, float pz, float* gr) {
    float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    float ax = px - fx, ay = py - fy, az = pz - fz;
    int x0 = min(max((int)fx, 0), N - 1), x1 = min(max((int)fx + 1, 0), N - 1);
    int y0 = min(max((int)fy, 0), N - 1), y1 = min(max((int)fy + 1, 0), N - 1);
    int z0 = min(max((int)fz, 0), N - 1), z1 = min(max((int)fz + 1, 0), N - 1);
    const int r00 = (z0 * N + y0) * N, r01 = (z0 * N + y1) * N, r10 = (z1 * N + y0) * N, r11 = (z1 * N + y1) * N;
    float v000 = vol[r00 + x0], v100 = vol[r00 + x1];
    float v010 = vol[r01 + x0], v110 = vol[r01 + x1];
    float v001 = vol[r10 + x0], v101 = vol[r10 + x1];
    float v011 = vol[r11 + x0], v111 = vol[r11 + x1];
    float c00 = v000 + ax * (v100 - v000), c10 = v010 + ax * (v110 - v010);
    float c01 = v001 + ax * (v101 - v001), c11 = v011 + ax * (v111 - v011);
    float c0 = c00 + ay * (c10 - c00), c1 = c01 + ay * (c11 - c01);
    if (gr) {
        float dx00 = v100 - v000, dx10 = v110 - v010, dx01 = v101 - v001, dx11 = v111 - v011;
        float dx0 = dx00 + ay * (dx10 - dx00), dx1 = dx01 + ay * (dx11 - dx01);
        gr[0] = dx0 + az * (dx1 - dx0);
        float dy0 = c10 - c00, dy1 = c11 - c01;
        gr[1] = dy0 + az * (dy1 - dy0);
        gr[2] = c1 - c0;
    }
    return c0 + az * (c1 - c0);
}

// sampling position in the reference of voxel (x, y, z) of the particle frame
__device__ __forceinline__ void ref_pos(const RefineGeom& g, int N, int x, int y, int z, float* q) {
    const float h = 0.5f * (float)N;
    const float wx = (float)x - h, wy = (float)y - h, wz = (float)z - h;
    q[0] = g.G[0] * wx + g.G[1] * wy + g.G[2] * wz + g.G[3];
    q[1] = g.G[4] * wx + g.G[5] * wy + g.G[6] * wz + g.G[7];
    q[2] = g.G[8] * wx + g.G[9] * wy + g.G[10] * wz + g.G[11];
}

// ---------------------------------------------------------------- LDS-staged trilinear sampling
// One workgroup handles an 8x8x8 tile of the particle frame.  The reference voxels its trilinear taps can
// touch lie in the bounding box [floor(min), floor(max) + 1] of the tile's mapped corners, at most
// floor(7 sqrt 3) + 1 + 2 = 15 voxels per axis for a rigid map (about 12 for a random rotation).  That box's
// actual extent, with the reference's border replicated where it reaches outside (the taps clamp to the
// border), is staged in LDS at fixed strides, so the divergent gathers become LDS reads and every sample equals
// the clamped-tap global one; a sample that rounds one voxel past the box keeps the box's edge cell with its
// fraction just outside [0, 1) (the same value to a few ulps).  A 15^3 box (13.5 KB) lets 10 workgroups share a
// CU, which hides the staging loads' latency: on 128 evaluations at 64^3 the reference reslice takes 155 us per
// launch and the adjoint moments 180 us, against 185 and 221 us with an 18^3 box staged whole.  A map whose box
// does not fit samples from global memory instead.
static constexpr int kRT = 8, kRB = 15;

__device__ __forceinline__ bool stage_ref_box(const float* __restrict__ vol, int N, const RefineGeom& g, int x0,
                                              int y0, int z0, float* box, int lo[3], int hi[3]) {
    const float h = 0.5f * (float)N;
    bool fits = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float* r = g.G + a * 4;
        const float c0 = r[0] * ((float)x0 - h) + r[1] * ((float)y0 - h) + r[2] * ((float)z0 - h) + r[3];
        const float mn = c0 + (kRT - 1) * (fminf(r[0], 0.f) + fminf(r[1], 0.f) + fminf(r[2], 0.f));
        const float ext = (kRT - 1) * (fabsf(r[0]) + fabsf(r[1]) + fabsf(r[2]));
        lo[a] = (int)floorf(mn);
        hi[a] = (int)floorf(mn + ext) + 1;
        if (!(hi[a] - lo[a] + 1 <= kRB) || !(fabsf(mn) < 1e6f)) fits = false;
    }
    if (fits) {
        // 2-D thread map over (y, x) of the box, loop over z; coordinates clamped to the volume
        constexpr int kRows = kT / kRB;
        const int t = threadIdx.x;
        const int wx = hi[0] - lo[0] + 1, wy = hi[1] - lo[1] + 1, wz = hi[2] - lo[2] + 1;
        const int x = t % kRB, yb = t / kRB;
        if (t < kRB * kRows && x < wx) {
            const int gx = min(max(lo[0] + x, 0), N - 1);
            for (int y = yb; y < wy; y += kRows) {
                const int gy = min(max(lo[1] + y, 0), N - 1);
                const float* src = vol + (size_t)gy * N + gx;
#pragma unroll 5
                for (int z = 0; z < wz; z++) {
                    const int gz = min(max(lo[2] + z, 0), N - 1);
                    box[(z * kRB + y) * kRB + x] = src[(size_t)gz * N * N];
                }
            }
        }
    }
    __syncthreads();
    return fits;
}

// trilin() on the staged box: the same arithmetic, taps read from LDS
__device__ __forceinline__ float trilin_box(const float* box, const int lo[3], const int hi[3], float px, float py,
                                            float pz, float* gr) {
    float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    float ax = px - fx, ay = py - fy, az = pz - fz;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    if (ix < lo[0] || ix > hi[0] - 1) {
        ix = min(max(ix, lo[0]), hi[0] - 1);
        ax = px - (float)ix;
    }
    if (iy < lo[1] || iy > hi[1] - 1) {
        iy = min(max(iy, lo[1]), hi[1] - 1);
        ay = py - (float)iy;
    }
    if (iz < lo[2] || iz > hi[2] - 1) {
        iz = min(max(iz, lo[2]), hi[2] - 1);
        az = pz - (float)iz;
    }
    const float* p = box + ((iz - lo[2]) * kRB + (iy - lo[1])) * kRB + (ix - lo[0]);
    float v000 = p[0], v100 = p[1];
    float v010 = p[kRB], v110 = p[kRB + 1];
    float v001 = p[kRB * kRB], v101 = p[kRB * kRB + 1];
    float v011 = p[kRB * kRB + kRB], v111 = p[kRB * kRB + kRB + 1];
    float c00 = v000 + ax * (v100 - v000), c10 = v010 + ax * (v110 - v010);
    float c01 = v001 + ax * (v101 - v001), c11 = v011 + ax * (v111 - v011);
    float c0 = c00 + ay * (c10 - c00), c1 = c01 + ay * (c11 - c01);
    if (gr) {
        float dx00 = v100 - v000, dx10 = v110 - v010, dx01 = v101 - v001, dx11 = v111 - v011;
        float dx0 = dx00 + ay * (dx10 - dx00), dx1 = dx01 + ay * (dx11 - dx01);
        gr[0] = dx0 + az * (dx1 - dx0);
        float dy0 = c10 - c00, dy1 = c11 - c01;
        gr[1] = dy0 + az * (dy1 - dy0);
        gr[2] = c1 - c0;
    }
    return c0 + az * (c1 - c0);
}


// mean of each particle volume (FP64 accumulation): one workgroup per particle
__global__ void __launch_bounds__(kT) k_vol_mean(const float* __restrict__ vols, long n3, float* __restrict__ mean) {
    __shared__ double sh[kT / 64];
    const float* v = vols + (size_t)blockIdx.x * n3;
    double s = 0.0;
    for (long o = threadIdx.x; o < n3; o += kT) s += v[o];
    s = dsum64(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kT / 64; w++) t += sh[w];
        mean[blockIdx.x] = (float)(t / (double)n3);
    }
}

// I1'(p) for every evaluation: out[e][z][y][x]; `ref` holds I1 - mean I1, x fastest
__global__ void __launch_bounds__(kT) k_reslice_ref(const float* __restrict__ ref, const RefineGeom* __restrict__ geom,
                                                     int N, float* __restrict__ out) {
    __shared__ float box[kRB * kRB * kRB];
    const int e = blockIdx.y;
    const RefineGeom g = geom[e];
    const size_t n3 = (size_t)N * N * N;
    const int nt = N / kRT;
    const int tb = xcd_tile(blockIdx.x, gridDim.x);
    const int z0 = (tb / (nt * nt)) * kRT, y0 = ((tb / nt) % nt) * kRT, x0 = (tb % nt) * kRT;
    int lo[3], hi[3];
    const bool fits = stage_ref_box(ref, N, g, x0, y0, z0, box, lo, hi);
    for (int t = threadIdx.x; t < kRT * kRT * kRT; t += kT) {
        const int x = x0 + (t & 7), y = y0 + ((t >> 3) & 7), z = z0 + (t >> 6);
        float q[3];
        ref_pos(g, N, x, y, z, q);
        const float v = fits ? trilin_box(box, lo, hi, q[0], q[1], q[2], nullptr) : trilin(ref, N, q[0], q[1], q[2], nullptr);
        out[(size_t)e * n3 + ((size_t)z * N + y) * N + x] = v;
    }
}

// per-block FP64 partials of <H X1, H X2>, |H X1|^2, |H X2|^2 (Hermitian weights): X1 = S1[e], X2 = FP[spec]
__global__ void __launch_bounds__(kT) k_refine_dot(const float2* __restrict__ X1, const float2* __restrict__ FP,
                                                    const float* __restrict__ filter, const float* __restrict__ wbin,
                                                    const RefineGeom* __restrict__ geom, int N,
                                                    double* __restrict__ part) {
    __shared__ double sh[3][kT / 64];
    const int e = blockIdx.y, H = N / 2 + 1;
    const long nh = (long)N * N * H;
    const RefineGeom g = geom[e];
    const float2* x1p = X1 + (size_t)e * nh;
    const float2* x2p = FP + (size_t)g.spec * nh;
    double sd = 0.0, s1 = 0.0, s2 = 0.0;
    for (long h = (long)blockIdx.x * kT + threadIdx.x; h < nh; h += (long)gridDim.x * kT) {
        int a, b, c;
        split3((unsigned)h, (unsigned)N, (unsigned)H, a, b, c);
        const float w = refine_hweight(filter, wbin, g, h, a, b, c, N, H);
        const float2 u = x1p[h], v = x2p[h];
        const double ww = (c == 0 || c == H - 1) ? (double)w * w : 2.0 * (double)w * w;
        sd += ww * ((double)u.x * v.x + (double)u.y * v.y);
        s1 += ww * ((double)u.x * u.x + (double)u.y * u.y);
        s2 += ww * ((double)v.x * v.x + (double)v.y * v.y);
    }
    sd = dsum64(sd);
    s1 = dsum64(s1);
    s2 = dsum64(s2);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = sd;
        sh[1][w] = s1;
        sh[2][w] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int i = 0; i < kT / 64; i++) t += sh[threadIdx.x][i];
        part[((size_t)e * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = t;
    }
}

// sc[e] = {S, n1, n2} in FP64
__global__ void __launch_bounds__(64) k_refine_scalars(const double* __restrict__ part, int nb, int n_eval,
                                                       double* __restrict__ sc) {
    const int e = blockIdx.x, l = threadIdx.x;
    double d = 0, q1 = 0, q2 = 0;
    for (int b = l; b < nb; b += 64) {
        d += part[((size_t)e * nb + b) * 3];
        q1 += part[((size_t)e * nb + b) * 3 + 1];
        q2 += part[((size_t)e * nb + b) * 3 + 2];
    }
    for (int o = 32; o > 0; o >>= 1) {
        d += __shfl_xor(d, o, 64);
        q1 += __shfl_xor(q1, o, 64);
        q2 += __shfl_xor(q2, o, 64);
    }
    if (l != 0) return;
    const double n1 = sqrt(q1), n2 = sqrt(q2);
    sc[e * 4 + 0] = (n1 > 0 && n2 > 0) ? d / (n1 * n2) : 0.0;
    sc[e * 4 + 1] = n1;
    sc[e * 4 + 2] = n2;
}

// adjoint spectrum in place: X1 <- H (H X2 / (n1 n2) - S H X1 / n1^2)
__global__ void __launch_bounds__(kT) k_refine_adj(float2* __restrict__ X1, const float2* __restrict__ FP,
                                                    const float* __restrict__ filter, const float* __restrict__ wbin,
                                                    const RefineGeom* __restrict__ geom, int N,
                                                    const double* __restrict__ sc) {
    const int e = blockIdx.y, H = N / 2 + 1;
    const long nh = (long)N * N * H;
    const RefineGeom g = geom[e];
    const double S = sc[e * 4], n1 = sc[e * 4 + 1], n2 = sc[e * 4 + 2];
    const float ca = (n1 > 0 && n2 > 0) ? (float)(1.0 / (n1 * n2)) : 0.f;
    const float cb = n1 > 0 ? (float)(S / (n1 * n1)) : 0.f;
    float2* x1p = X1 + (size_t)e * nh;
    const float2* x2p = FP + (size_t)g.spec * nh;
    for (long h = (long)blockIdx.x * kT + threadIdx.x; h < nh; h += (long)gridDim.x * kT) {
        int a, b, c;
        split3((unsigned)h, (unsigned)N, (unsigned)H, a, b, c);
        const float w = refine_hweight(filter, wbin, g, h, a, b, c, N, H);
        const float2 u = x1p[h], v = x2p[h];
        const float w2 = w * w;
        x1p[h] = make_float2(w2 * (ca * v.x - cb * u.x), w2 * (ca * v.y - cb * u.y));
    }
}

// twelve moments per evaluation: per-workgroup partial sums of gamma(p) dI1/dq_a w_b(p) over kMT consecutive
// 8x8x8 tiles (one staged box per tile, one 12-value reduction per workgroup)
static constexpr int kMT = 4;
__global__ void __launch_bounds__(kT) k_refine_moments(const float* __restrict__ ref, const RefineGeom* __restrict__ geom,
                                                        const float* __restrict__ G, const float* __restrict__ maskfft,
                                                        int N, float* __restrict__ part) {
    __shared__ float box[kRB * kRB * kRB];
    __shared__ float sh[12][kT / 64];
    const int e = blockIdx.y;
    const RefineGeom g = geom[e];
    const size_t n3 = (size_t)N * N * N;
    const float* Ge = G + (size_t)e * n3;
    const float half = 0.5f * (float)N;
    const int nt = N / kRT, ntile = nt * nt * nt;
    const int wb = xcd_tile(blockIdx.x, gridDim.x);
    float q12[12];
#pragma unroll
    for (int r = 0; r < 12; r++) q12[r] = 0.f;
    for (int u = 0; u < kMT; u++) {
        const int tb = wb * kMT + u;
        if (tb >= ntile) break;
        const int z0 = (tb / (nt * nt)) * kRT, y0 = ((tb / nt) % nt) * kRT, x0 = (tb % nt) * kRT;
        if (u > 0) __syncthreads();  // the previous tile's box is no longer read
        int lo[3], hi[3];
        const bool fits = stage_ref_box(ref, N, g, x0, y0, z0, box, lo, hi);
        for (int t = threadIdx.x; t < kRT * kRT * kRT; t += kT) {
            const int x = x0 + (t & 7), y = y0 + ((t >> 3) & 7), z = z0 + (t >> 6);
            const size_t o = ((size_t)z * N + y) * N + x;
            float q[3];
            ref_pos(g, N, x, y, z, q);
            const float gam = maskfft[o] * (((x + y + z) & 1) ? -Ge[o] : Ge[o]);
            float gr[3];
            if (fits) trilin_box(box, lo, hi, q[0], q[1], q[2], gr);
            else trilin(ref, N, q[0], q[1], q[2], gr);
            const float w[4] = {(float)x - half, (float)y - half, (float)z - half, 1.f};
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float ga = gam * gr[a];
#pragma unroll
                for (int b = 0; b < 4; b++) q12[a * 4 + b] += ga * w[b];
            }
        }
    }
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < 12; r++) {
        const float t = wsum64(q12[r]);
        if ((threadIdx.x & 63) == 0) sh[r][wv] = t;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        float t = 0.f;
        for (int i = 0; i < kT / 64; i++) t += sh[threadIdx.x][i];
        part[((size_t)e * gridDim.x + wb) * 12 + threadIdx.x] = t;
    }
}

__global__ void __launch_bounds__(64) k_refine_reduce_moments(const float* __restrict__ part, int nb, int n_eval,
                                                              double* __restrict__ out) {
    const int e = blockIdx.x, l = threadIdx.x;
    double acc[12];
#pragma unroll
    for (int r = 0; r < 12; r++) acc[r] = 0.0;
    for (int b = l; b < nb; b += 64) {
        const float* p = part + ((size_t)e * nb + b) * 12;
#pragma unroll
        for (int r = 0; r < 12; r++) acc[r] += p[r];
    }
#pragma unroll
    for (int r = 0; r < 12; r++) {
        double v = acc[r];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (l == 0) out[e * 12 + r] = v;
    }
}

int refine_moment_blocks(int N) {
    const int nt = N / kRT;
    return (nt * nt * nt + kMT - 1) / kMT;
}
int refine_dot_blocks(int N) {
    const long nh = (long)N * N * (N / 2 + 1);
    return (int)std::min<long>(64, (nh + 4 * kT - 1) / (4 * kT));
}

void launch_vol_mean(const float* vols, int count, int N, float* mean, hipStream_t s) {
    k_vol_mean<<<count, kT, 0, s>>>(vols, (long)N * N * N, mean);
    ETTK_HIP(hipGetLastError());
}
void launch_reslice_ref(const float* ref, const RefineGeom* geom, int E, int N, float* out, hipStream_t s) {
    ETTK_REQUIRE(N % kRT == 0, ETTK_ERR_ARG, "launch_reslice_ref: box size must be a multiple of 8");
    const int nt = N / kRT;
    k_reslice_ref<<<dim3(nt * nt * nt, E), kT, 0, s>>>(ref, geom, N, out);
    ETTK_HIP(hipGetLastError());
}
void launch_refine_dot(const float2* X1, const float2* FP, const float* filter, const float* wbin, const RefineGeom* geom,
                       int E, int N,
                       double* part, double* sc, hipStream_t s) {
    const int nb = refine_dot_blocks(N);
    k_refine_dot<<<dim3(nb, E), kT, 0, s>>>(X1, FP, filter, wbin, geom, N, part);
    k_refine_scalars<<<E, 64, 0, s>>>(part, nb, E, sc);
    ETTK_HIP(hipGetLastError());
}
void launch_refine_scalars(const double* part, int nb, int E, double* sc, hipStream_t s) {
    k_refine_scalars<<<E, 64, 0, s>>>(part, nb, E, sc);
    ETTK_HIP(hipGetLastError());
}
void launch_refine_adj(float2* X1, const float2* FP, const float* filter, const float* wbin, const RefineGeom* geom,
                       int E, int N,
                       const double* sc, hipStream_t s) {
    const long nh = (long)N * N * (N / 2 + 1);
    const int nb = (int)std::min<long>(512, (nh + 4 * kT - 1) / (4 * kT));
    k_refine_adj<<<dim3(nb, E), kT, 0, s>>>(X1, FP, filter, wbin, geom, N, sc);
    ETTK_HIP(hipGetLastError());
}
void launch_refine_moments(const float* ref, const RefineGeom* geom, const float* G, const float* maskfft, int E, int N,
                           float* mpart, double* mom, hipStream_t s) {
    ETTK_REQUIRE(N % kRT == 0, ETTK_ERR_ARG, "launch_refine_moments: box size must be a multiple of 8");
    const int nb = refine_moment_blocks(N);
    k_refine_moments<<<dim3(nb, E), kT, 0, s>>>(ref, geom, G, maskfft, N, mpart);
    k_refine_reduce_moments<<<E, 64, 0, s>>>(mpart, nb, E, mom);
    ETTK_HIP(hipGetLastError());
}

}  // namespace ettk
