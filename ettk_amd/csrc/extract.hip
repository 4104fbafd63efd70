// extract.hip -- sub-volumes cut from a tomogram: nbfWedgedSubImage3D::getImage with the tomogram loaded
// (inc/em/nbfWedgedSubImage3D.h:603-802, normalize off), the step before the alignment path
// (src/CutVolumes3DFromPositions.cpp).
//
// Output voxel q of particle p reads the tomogram at  position_p + fin_p (q - N/2 - cut_p z),
// fin = Euler(normal) matrix, with VTK's tricubic (Keys a = -1/2, taps clamped to the tomogram); samples outside
// [-0.5, n - 0.5] on any axis are invalid.  The mean of the valid samples is subtracted and invalid voxels are 0,
// as getImage does before handing the sub-volume to the metric.  Outputs are N^3 floats, x fastest.
#include "kernels.hpp"
#include "plan.hpp"

#include <cfloat>

namespace ettk {

namespace {

constexpr int kExT = 256;

struct ExtractGeom {
    float m[12];   // input index = m[0..2] x + m[3..5] y + m[6..8] z + m[9..11]  (rows x, y, z)
    double md[12];  // the same map in FP64: decides validity for samples within FP32 rounding of the border
};

__device__ __forceinline__ void keys_w4(float f, float* w) {
    const float fm1 = f - 1.f, fd2 = f * 0.5f, ft3 = f * 3.f;
    w[0] = -fd2 * fm1 * fm1;
    w[1] = ((ft3 - 2.f) * fd2 - 1.f) * fm1;
    w[2] = -((ft3 - 4.f) * f - 1.f) * fd2;
    w[3] = f * fd2 * fm1;
}

// one thread per output voxel; per-block (sum, count) of the valid samples in block order
__global__ void __launch_bounds__(kExT) k_extract(const float* __restrict__ tomo, int nx, int ny, int nz,
                                                  const ExtractGeom* __restrict__ geom, int N, float* __restrict__ out,
                                                  float* __restrict__ partials) {
    __shared__ float sh[2][kExT / 64];
    const int p = blockIdx.y;
    const ExtractGeom g = geom[p];
    const long n3 = (long)N * N * N;
    const long o = (long)blockIdx.x * kExT + threadIdx.x;
    const float tol = 7.62939453125e-06f;  // VTK's floor tolerance (2^-17)
    float s = 0.f, c = 0.f;
    if (o < n3) {
        int z, y, x;
        split3((unsigned)o, (unsigned)N, (unsigned)N, z, y, x);
        const float px = g.m[0] * x + g.m[3] * y + g.m[6] * z + g.m[9];
        const float py = g.m[1] * x + g.m[4] * y + g.m[7] * z + g.m[10];
        const float pz = g.m[2] * x + g.m[5] * y + g.m[8] * z + g.m[11];
        float v = -FLT_MAX;
        const float hx = nx - 0.5f, hy = ny - 0.5f, hz = nz - 0.5f;
        bool in = px >= -0.5f && px <= hx && py >= -0.5f && py <= hy && pz >= -0.5f && pz <= hz;
        auto near = [](float p, float hi) { return fabsf(p + 0.5f) < 1e-3f || fabsf(p - hi) < 1e-3f; };
        if (near(px, hx) || near(py, hy) || near(pz, hz)) {
            const double* md = geom[p].md;
            const double qx = md[0] * x + md[3] * y + md[6] * z + md[9];
            const double qy = md[1] * x + md[4] * y + md[7] * z + md[10];
            const double qz = md[2] * x + md[5] * y + md[8] * z + md[11];
            in = qx >= -0.5 && qx <= nx - 0.5 && qy >= -0.5 && qy <= ny - 0.5 && qz >= -0.5 && qz <= nz - 0.5;
        }
        if (in) {
            const float xx = px + tol, yy = py + tol, zz = pz + tol;
            const int ix = (int)floorf(xx), iy = (int)floorf(yy), iz = (int)floorf(zz);
            float wx[4], wy[4], wz[4];
            keys_w4(xx - ix, wx);
            keys_w4(yy - iy, wy);
            keys_w4(zz - iz, wz);
            int cx[4], cy[4], cz[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                cx[q] = min(max(ix - 1 + q, 0), nx - 1);
                cy[q] = min(max(iy - 1 + q, 0), ny - 1);
                cz[q] = min(max(iz - 1 + q, 0), nz - 1);
            }
            float acc = 0.f;
#pragma unroll
            for (int a = 0; a < 4; a++) {
                float ay = 0.f;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const float* row = tomo + ((size_t)cz[a] * ny + cy[b]) * nx;
                    ay += wy[b] * (wx[0] * row[cx[0]] + wx[1] * row[cx[1]] + wx[2] * row[cx[2]] + wx[3] * row[cx[3]]);
                }
                acc += wz[a] * ay;
            }
            v = acc;
            s = v;
            c = 1.f;
        }
        out[(size_t)p * n3 + o] = v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off, 64);
        c += __shfl_down(c, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = s;
        sh[1][w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float ts = 0.f, tc = 0.f;
        for (int q = 0; q < kExT / 64; q++) {
            ts += sh[0][q];
            tc += sh[1][q];
        }
        partials[((size_t)p * gridDim.x + blockIdx.x) * 2] = ts;
        partials[((size_t)p * gridDim.x + blockIdx.x) * 2 + 1] = tc;
    }
}

// valid samples minus their mean, invalid (-FLT_MAX) to 0
__global__ void __launch_bounds__(kExT) k_extract_finish(float* __restrict__ out, const float* __restrict__ mean, long n3) {
    const int p = blockIdx.y;
    const long o = (long)blockIdx.x * kExT + threadIdx.x;
    if (o >= n3) return;
    float* v = out + (size_t)p * n3 + o;
    const float a = *v;
    *v = a > -FLT_MAX ? a - mean[p] : 0.f;
}

}  // namespace

void extract_batch(ettk_ctx* ctx, const float* d_tomo, int nx, int ny, int nz, int n, const double* positions,
                   const ettk_particle* descs, int N, float* d_out) {
    hipStream_t s = ctx->stream;
    std::vector<ExtractGeom> g(n);
    const double half = N / 2.0;
    for (int i = 0; i < n; i++) {
        Mat4 fin = euler_of(descs[i]) * Mat4::from(descs[i].matrix);
        if (mat_det(fin) == 0.0) fin = Mat4::ident();
        const double w0[3] = {-half, -half, -half - descs[i].cut_offset};
        for (int r = 0; r < 3; r++) {
            g[i].md[0 + r] = fin(r, 0);
            g[i].md[3 + r] = fin(r, 1);
            g[i].md[6 + r] = fin(r, 2);
            g[i].md[9 + r] = fin(r, 0) * w0[0] + fin(r, 1) * w0[1] + fin(r, 2) * w0[2] + fin(r, 3) + positions[3 * i + r];
            for (int c = 0; c < 4; c++) g[i].m[3 * c + r] = (float)g[i].md[3 * c + r];
        }
    }
    const long n3 = (long)N * N * N;
    const int nb = (int)((n3 + kExT - 1) / kExT);
    DevBuf<ExtractGeom> dg;
    DevBuf<float> part, mean;
    dg.upload(g.data(), n, s);
    part.ensure((size_t)2 * nb * n);
    mean.ensure(n);
    k_extract<<<dim3(nb, n), kExT, 0, s>>>(d_tomo, nx, ny, nz, dg.p, N, d_out, part.p);
    ETTK_HIP(hipGetLastError());
    launch_reslice_mean(part.p, nb, n, mean.p, s);
    k_extract_finish<<<dim3(nb, n), kExT, 0, s>>>(d_out, mean.p, n3);
    ETTK_HIP(hipGetLastError());
    ETTK_HIP(hipStreamSynchronize(s));
}

}  // namespace ettk

using namespace ettk;

extern "C" int ettk_extract_batch(ettk_ctx* ctx, const float* d_tomogram, int nx, int ny, int nz, int n,
                                  const double* positions, const ettk_particle* descs, int N, float* d_out) {
    try {
        ETTK_REQUIRE(ctx && d_tomogram && positions && descs && d_out, ETTK_ERR_ARG, "ettk_extract_batch: null argument");
        ETTK_REQUIRE(nx > 0 && ny > 0 && nz > 0 && n >= 0 && N > 0 && N <= 1024, ETTK_ERR_ARG,
                     "ettk_extract_batch: bad sizes");
        ETTK_REQUIRE((long long)nx * ny * nz < (1LL << 40), ETTK_ERR_ARG, "ettk_extract_batch: tomogram too large");
        if (n > 0) {
            ETTK_HIP(hipSetDevice(ctx->device));
            extract_batch(ctx, d_tomogram, nx, ny, nz, n, positions, descs, N, d_out);
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
