// cand.hip -- the data-parallel part of the candidate extraction of the SO(3) search on the device
// (getTrueLocalMaximaPositions and eliminateEquivalentAngles, ProjRot.h:1593-1702).
//
// Per volume (one workgroup): the local maxima are sorted by index (blitz iteration order, alpha slowest); each is
// decoded to zyz(alpha, beta, gamma) = (Rg Rb) Ra from the host's rotation tables with the host's operation order
// and rounding (no fused multiply-adds) and kept when isTransformValid accepts it; then, for every kept maximum i,
// the kept maxima j > i whose 4x4 matrices lie within `tol` of its own in L1 norm (its partners, in increasing j).
// The sorted kept maxima are written over the volume's list and the partners as a compressed row list; the host
// then runs the elimination itself, which is sequential and short: the reference visits every ordered pair (i, j)
// in index order and, when both are still alive and close, drops the worse (j on a tie), and a pair visited as
// (i, j) with j < i can no longer have both ends alive, so the partners j > i of the rows, taken in order, decide
// everything.
// A volume is left to the host (its raw list untouched, mode 1) when it holds more maxima than fit in LDS, when a
// maximum has more partners than its row holds, or when a rotation angle lies within 1e-9 degree of the rotation
// restriction (the host's acos decides there).
#include <cfloat>

#include "so3.hpp"

namespace ettk {

namespace {

constexpr int kCT = 1024;  // threads per workgroup (four wavefronts per SIMD hide the LDS latency of the pair tests)

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

// active[v]: 0 = volume unused, 1 = extract here, 2 = leave to the host.  Results: mode[v] = 0 with kept[v] sorted
// valid maxima over the volume's list, their partner counts in deg[v kCandMax + q] and the partners of row q in
// adj[(v kCandMax + q) kCandSlots + .] (nadj[v] in all); mode[v] = 1: the host extracts from the raw list.
__global__ void __launch_bounds__(kCT) k_lm_candidates(const int* __restrict__ count, int* __restrict__ idx,
                                                       double* __restrict__ val, int cap, const int* __restrict__ active,
                                                       const double* __restrict__ tabs, int n, double trans_restrict,
                                                       double rot_restrict, float tol, double* __restrict__ mats,
                                                       int* __restrict__ deg, int* __restrict__ adj,
                                                       int* __restrict__ kept, int* __restrict__ nadj,
                                                       int* __restrict__ mode) {
    __shared__ int key[kCandMax];
    __shared__ double vv[kCandMax];
    __shared__ float pf[6][kCandMax];  // the first two matrix rows in float: a conservative pre-test of closeness
    __shared__ int sh[kCT / 64 + 1];
    __shared__ int s_flag;
    const int v = blockIdx.x, tid = threadIdx.x;
    int* vidx = idx + (size_t)v * cap;
    double* vval = val + (size_t)v * cap;
    auto leave = [&](int m) {
        if (tid == 0) {
            kept[v] = 0;
            nadj[v] = 0;
            mode[v] = m;
        }
    };
    if (!active[v]) {
        leave(0);
        return;
    }
    const int c = count[(size_t)v * kLmCountStride];
    if (active[v] == 2 || c > cap || c > kCandMax) {
        leave(1);
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
    int myk[NPT];
    double myv[NPT];
#pragma unroll
    for (int u = 0; u < NPT; u++) {
        const int k = tid + u * kCT;
        myk[u] = k < c ? key[k] : 0;
        myv[u] = k < c ? vv[k] : 0.0;
        if (k >= c) continue;
        double m[16];
        grid_rotation(tabs, n, myk[u], m);
        const bool ok = !(fabs(m[3]) > trans_restrict || fabs(m[7]) > trans_restrict || fabs(m[11]) > trans_restrict);
        const double px = m[2], py = m[6], pz = m[10];
        const double nr = sqrt(da(da(dm(px, px), dm(py, py)), dm(pz, pz)));
        const double cc = nr != 0.0 ? pz / nr : 0.0;
        const double ang = cc >= 1.0 ? 0.0 : fabs(dm(acos(cc), 180.0) / M_PI);
        if (fabs(ang - rot_restrict) <= 1e-9 * fmax(1.0, rot_restrict)) s_flag = 1;  // the host decides
        if (ok && ang <= rot_restrict) okm |= 1 << u;
    }
    // keep the valid maxima in index order, compacted in place (all reads above precede the scan's barriers)
    int nv = 0;
#pragma unroll
    for (int u = 0; u < NPT; u++) {
        int tot = 0;
        const int pos = block_scan_excl((okm >> u) & 1, sh, tot);
        if ((okm >> u) & 1) {
            key[nv + pos] = myk[u];
            vv[nv + pos] = myv[u];
        }
        nv += tot;
        __syncthreads();
    }
    if (s_flag) {
        leave(1);
        return;
    }
    double* vm = mats + (size_t)v * kCandMax * 16;
    for (int q = tid; q < nv; q += kCT) {
        double m[16];
        grid_rotation(tabs, n, key[q], m);
#pragma unroll
        for (int e = 0; e < 16; e++) vm[(size_t)q * 16 + e] = m[e];
        pf[0][q] = (float)m[0];
        pf[1][q] = (float)m[1];
        pf[2][q] = (float)m[2];
        pf[3][q] = (float)m[4];
        pf[4][q] = (float)m[5];
        pf[5][q] = (float)m[6];
    }
    __syncthreads();
    // partners j > i of every kept maximum, in increasing j, into the row's kCandSlots slots (more: the host
    // extracts).  The float pre-test rejects a pair only when six of the sixteen terms of the L1 distance already
    // exceed tol by far more than their rounding (each float element within 2^-24 of the double one).
    const float ftol = tol + 2e-5f;
    int* va = adj + (size_t)v * kCandMax * kCandSlots;
    int dsum = 0;
    for (int i = tid; i < nv; i += kCT) {
        const double* mi = vm + (size_t)i * 16;
        const float a0 = pf[0][i], a1 = pf[1][i], a2 = pf[2][i], a3 = pf[3][i], a4 = pf[4][i], a5 = pf[5][i];
        int d = 0;
#pragma unroll 4
        for (int j = i + 1; j < nv; j++) {
            const float e3 = fabsf(a0 - pf[0][j]) + fabsf(a1 - pf[1][j]) + fabsf(a2 - pf[2][j]);
            if (!(e3 + fabsf(a3 - pf[3][j]) + fabsf(a4 - pf[4][j]) + fabsf(a5 - pf[5][j]) < ftol)) continue;
            const double* mj = vm + (size_t)j * 16;
            double err = 0.0;
            int e = 0;
            for (; e < 16; e++) {
                err = da(err, fabs(mi[e] - mj[e]));
                if (!(err < (double)tol)) break;
            }
            if (e == 16) {
                if (d < kCandSlots) va[(size_t)i * kCandSlots + d] = j;
                d++;
            }
        }
        deg[(size_t)v * kCandMax + i] = d;
        dsum += d;
        if (d > kCandSlots) s_flag = 1;
    }
    int total = 0;
    block_scan_excl(dsum, sh, total);
    if (s_flag || total > kCandAdj) {
        leave(1);
        return;
    }
    // the sorted kept maxima over the volume's list
    for (int q = tid; q < nv; q += kCT) {
        vidx[q] = key[q];
        vval[q] = vv[q];
    }
    if (tid == 0) {
        kept[v] = nv;
        nadj[v] = total;
        mode[v] = 0;
    }
}

void launch_lm_candidates(const int* count, int* idx, double* val, int cap, const int* active, int np,
                          const double* tabs, int n, double trans_restrict, double rot_restrict, float tol,
                          double* mats, int* deg, int* adj, int* kept, int* nadj, int* mode, hipStream_t s) {
    k_lm_candidates<<<np, kCT, 0, s>>>(count, idx, val, cap, active, tabs, n, trans_restrict, rot_restrict, tol, mats,
                                       deg, adj, kept, nadj, mode);
    ETTK_HIP(hipGetLastError());
}

}  // namespace ettk
