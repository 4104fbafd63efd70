// cand.hip -- candidate rotations of the SO(3) search from the local maxima of each correlation volume, on the
// device: getTrueLocalMaximaPositions and eliminateEquivalentAngles (ProjRot.h:1593-1702), up to the priority
// queue, which the host fills from the survivors in index order.
//
// Per volume (one workgroup): the maxima are sorted by index (blitz iteration order, alpha slowest); each is
// decoded to zyz(alpha, beta, gamma) = (Rg Rb) Ra from the host's rotation tables with the host's operation order
// and rounding (no fused multiply-adds), and kept when isTransformValid accepts it; then the equivalent-angle
// elimination: the reference visits every ordered pair (i, j) of kept maxima in index order and, when both are
// still alive and their 4x4 matrices are within `tol` in L1 norm, drops the worse of the two (j on a tie).  A pair
// visited as (i, j) with j < i can no longer have both ends alive (the pair was decided, or skipped because one
// end had died, when row j was visited), so only the partners j > i of each row decide anything: the partners are
// found in parallel, and the rows run in order on one wavefront, each row's partners on its lanes.  The survivors
// are written over the volume's list, in index order.
// A volume is left to the host (its raw list untouched, mode 1) when it holds more maxima than fit in LDS, when a
// maximum has more close partners than a wavefront holds, or when a rotation angle lies within 1e-9 degree of the
// rotation restriction (the host's acos decides there).
#include <cfloat>

#include "so3.hpp"

namespace ettk {

namespace {

constexpr int kCT = 256;  // threads per workgroup

__device__ __forceinline__ double dm(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double da(double a, double b) { return __dadd_rn(a, b); }

// C = A B for row-major 4x4 matrices, each element ((a0 b0 + a1 b1) + a2 b2) + a3 b3 as the host computes it
__device__ void mat4_mul_rn(const double* A, const double* B, double* C) {
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
            C[r * 4 + c] = da(da(da(dm(A[r * 4 + 0], B[0 * 4 + c]), dm(A[r * 4 + 1], B[1 * 4 + c])),
                                 dm(A[r * 4 + 2], B[2 * 4 + c])),
                              dm(A[r * 4 + 3], B[3 * 4 + c]));
}

// the rotation of grid point t = (a n + b) n + g: (Rg[g] Rb[b]) Ra[a]; tabs = [Ra; Rb; Rg], n matrices each
__device__ void grid_rotation(const double* __restrict__ tabs, int n, int t, double* m) {
    const int a = t / (n * n), b = (t / n) % n, g = t % n;
    double gb[16];
    mat4_mul_rn(tabs + ((size_t)2 * n + g) * 16, tabs + ((size_t)n + b) * 16, gb);
    mat4_mul_rn(gb, tabs + (size_t)a * 16, m);
}

__device__ __forceinline__ int block_scan_excl(int x, int* sh, int& total) {
    // exclusive prefix sum over the workgroup (kCT threads), sh holds kCT / 64 + 1 ints
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    int base = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < kCT / 64; k++) {
        if (k < w) base += sh[k];
        total += sh[k];
    }
    __syncthreads();
    return base + incl - x;
}

}  // namespace

// active[v]: 0 = volume unused, 1 = extract here, 2 = leave to the host; mode[v]: 0 = survivors written (count in
// kept[v]), 1 = the host extracts from the raw list (count[v])
__global__ void __launch_bounds__(kCT) k_lm_candidates(const int* __restrict__ count, int* __restrict__ idx,
                                                       double* __restrict__ val, int cap, const int* __restrict__ active,
                                                       const double* __restrict__ tabs, int n, double trans_restrict,
                                                       double rot_restrict, float tol, double* __restrict__ mats,
                                                       int* __restrict__ adj, int* __restrict__ kept,
                                                       int* __restrict__ mode) {
    __shared__ int key[kCandMax];
    __shared__ double vv[kCandMax];
    __shared__ float pf[3][kCandMax];  // the first matrix row in float: a conservative pre-test of closeness
    __shared__ int deg[kCandMax];
    __shared__ unsigned char alive[kCandMax];
    __shared__ int sh[kCT / 64 + 1];
    __shared__ int s_flag;
    const int v = blockIdx.x, tid = threadIdx.x;
    int* vidx = idx + (size_t)v * cap;
    double* vval = val + (size_t)v * cap;
    if (!active[v]) {
        if (tid == 0) {
            kept[v] = 0;
            mode[v] = 0;
        }
        return;
    }
    const int c = count[(size_t)v * kLmCountStride];
    if (active[v] == 2 || c > cap || c > kCandMax) {
        if (tid == 0) {
            kept[v] = 0;
            mode[v] = 1;
        }
        return;
    }
    if (tid == 0) s_flag = 0;
    // load and sort by index (bitonic over the next power of two; padding sorts last)
    int P = 1;
    while (P < c) P <<= 1;
    for (int k = tid; k < P; k += kCT) {
        key[k] = k < c ? vidx[k] : 0x7fffffff;
        vv[k] = k < c ? vval[k] : 0.0;
    }
    __syncthreads();
    for (int len = 2; len <= P; len <<= 1)
        for (int st = len >> 1; st > 0; st >>= 1) {
            for (int k = tid; k < P; k += kCT) {
                const int o = k ^ st;
                if (o > k) {
                    const bool up = (k & len) == 0;
                    const int a = key[k], b = key[o];
                    if ((a > b) == up) {
                        key[k] = b;
                        key[o] = a;
                        const double t = vv[k];
                        vv[k] = vv[o];
                        vv[o] = t;
                    }
                }
            }
            __syncthreads();
        }
    // isTransformValid of each maximum (in registers: at most kCandMax / kCT per thread)
    constexpr int NPT = kCandMax / kCT;
    int okm = 0;
#pragma unroll
    for (int u = 0; u < NPT; u++) {
        const int k = tid + u * kCT;
        if (k >= c) continue;
        double m[16];
        grid_rotation(tabs, n, key[k], m);
        bool ok = !(fabs(m[3]) > trans_restrict || fabs(m[7]) > trans_restrict || fabs(m[11]) > trans_restrict);
        const double px = m[2], py = m[6], pz = m[10];
        const double nr = sqrt(da(da(dm(px, px), dm(py, py)), dm(pz, pz)));
        const double cc = nr != 0.0 ? pz / nr : 0.0;
        const double ang = cc >= 1.0 ? 0.0 : fabs(dm(acos(cc), 180.0) / M_PI);
        if (fabs(ang - rot_restrict) <= 1e-9 * fmax(1.0, rot_restrict)) s_flag = 1;  // the host decides
        if (ok && ang <= rot_restrict) okm |= 1 << u;
    }
    // keep the valid maxima in order: compacted in place (read before the barrier inside the scan)
    int myk[NPT];
    double myv[NPT];
#pragma unroll
    for (int u = 0; u < NPT; u++) {
        const int k = tid + u * kCT;
        myk[u] = k < c ? key[k] : 0;
        myv[u] = k < c ? vv[k] : 0.0;
    }
    int nvalid = 0;
#pragma unroll
    for (int u = 0; u < NPT; u++) {
        // thread-major order must follow index order: process the elements of stride kCT block by block
        int tot = 0;
        const int pos = block_scan_excl((okm >> u) & 1, sh, tot);
        if ((okm >> u) & 1) {
            key[nvalid + pos] = myk[u];
            vv[nvalid + pos] = myv[u];
        }
        nvalid += tot;
        __syncthreads();
    }
    if (s_flag) {
        if (tid == 0) {
            kept[v] = 0;
            mode[v] = 1;
        }
        return;
    }
    const int nv = nvalid;
    double* vm = mats + (size_t)v * kCandMax * 16;
    for (int q = tid; q < nv; q += kCT) {
        double m[16];
        grid_rotation(tabs, n, key[q], m);
#pragma unroll
        for (int e = 0; e < 16; e++) vm[(size_t)q * 16 + e] = m[e];
        pf[0][q] = (float)m[0];
        pf[1][q] = (float)m[1];
        pf[2][q] = (float)m[2];
        alive[q] = 1;
    }
    __syncthreads();
    // close partners j > i of every kept maximum, in increasing j; the float pre-test rejects a pair only when the
    // first three terms of the L1 distance already exceed tol by far more than their rounding
    int* va = adj + (size_t)v * kCandMax * 64;
    const float ftol = tol + 1e-5f;
    for (int i = tid; i < nv; i += kCT) {
        const float a0 = pf[0][i], a1 = pf[1][i], a2 = pf[2][i];
        const double* mi = vm + (size_t)i * 16;
        int d = 0;
        for (int j = i + 1; j < nv; j++) {
            if (!(fabsf(a0 - pf[0][j]) + fabsf(a1 - pf[1][j]) + fabsf(a2 - pf[2][j]) < ftol)) continue;
            const double* mj = vm + (size_t)j * 16;
            double err = 0.0;
            int e = 0;
            for (; e < 16; e++) {
                err = da(err, fabs(mi[e] - mj[e]));
                if (!(err < (double)tol)) break;
            }
            if (e == 16) {
                if (d < 64) va[(size_t)i * 64 + d] = j;
                d++;
            }
        }
        deg[i] = d;
        if (d > 64) s_flag = 1;
    }
    __syncthreads();
    if (s_flag) {
        if (tid == 0) {
            kept[v] = 0;
            mode[v] = 1;
        }
        return;
    }
    // the rows that have partners, in order (over the pre-test array, free now)
    int* rows = reinterpret_cast<int*>(&pf[0][0]);
    int nrows = 0;
    for (int q0 = 0; q0 < nv; q0 += kCT) {
        const int q = q0 + tid;
        const int has = q < nv && deg[q] > 0;
        int tot = 0;
        const int pos = block_scan_excl(has, sh, tot);
        if (has) rows[nrows + pos] = q;
        nrows += tot;
    }
    __syncthreads();
    // the rows in order on wavefront 0, each row's partners on its lanes (the next row's partners are loaded while
    // a row is decided): the alive partners before the first one that is better than row i die, and if there is
    // such a partner, row i dies
    if (tid < 64) {
        const int lane = tid;
        int jn = nrows > 0 && lane < deg[rows[0]] ? va[(size_t)rows[0] * 64 + lane] : 0;
        for (int r = 0; r < nrows; r++) {
            const int i = rows[r], di = deg[i];
            const int j = jn;
            if (r + 1 < nrows) {
                const int i2 = rows[r + 1];
                jn = lane < deg[i2] ? va[(size_t)i2 * 64 + lane] : 0;
            }
            if (!alive[i]) continue;
            const float fi = (float)(-vv[i]);
            const bool al = lane < di && alive[j];
            const bool better = al && fi > (float)(-vv[j]);
            const unsigned long long bal = __ballot(better);
            const int first = bal ? __ffsll((long long)bal) - 1 : 64;
            if (al && lane < first) alive[j] = 0;
            if (lane == 0 && bal) alive[i] = 0;
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    // survivors in index order over the volume's list
    int written = 0;
    for (int q0 = 0; q0 < nv; q0 += kCT) {
        const int q = q0 + tid;
        const int a = q < nv ? alive[q] : 0;
        const int k = q < nv ? key[q] : 0;
        const double x = q < nv ? vv[q] : 0.0;
        int tot = 0;
        const int pos = block_scan_excl(a, sh, tot);
        if (a) {
            vidx[written + pos] = k;
            vval[written + pos] = x;
        }
        written += tot;
    }
    if (tid == 0) {
        kept[v] = written;
        mode[v] = 0;
    }
}

void launch_lm_candidates(const int* count, int* idx, double* val, int cap, const int* active, int np,
                          const double* tabs, int n, double trans_restrict, double rot_restrict, float tol,
                          double* mats, int* adj, int* kept, int* mode, hipStream_t s) {
    k_lm_candidates<<<np, kCT, 0, s>>>(count, idx, val, cap, active, tabs, n, trans_restrict, rot_restrict, tol, mats,
                                       adj, kept, mode);
    ETTK_HIP(hipGetLastError());
}

}  // namespace ettk
