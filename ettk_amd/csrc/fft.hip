// fft.hip -- batched 3-D FFTs for gfx950: LDS-resident Stockham pencils (radix 4/3/2), one pass per axis.
//
// Replaces the FFTW double-precision plans of nbfFourierFilter::initializeFFTWhalf
// (inc/em/nbfFourierFilter.h:270-358): fftw_plan_dft_r2c_3d / fftw_plan_dft_c2r_3d on [i][j][k] arrays
// with the half spectrum along k.  Each workgroup stages P pencils of length L in LDS, runs the log_r(L)
// Stockham stages there and writes them back, so every pass reads and writes each element of HBM once.
//   r2c: pass k (pairs of real rows packed as one complex row, split after the transform), pass j, pass i
//   c2r: pass i, pass j (complex, backward), then pass k (pairs of half rows -> two real rows)
#include "kernels.hpp"
#include "wedge.hpp"

#include <cfloat>
#include <cmath>

namespace ettk {

static const int kLens[9] = {16, 32, 48, 64, 96, 128, 192, 256, 0};

int fft_len_index(int L) {
    for (int i = 0; kLens[i]; i++)
        if (kLens[i] == L) return i;
    return -1;
}

void fft_tables_init(FftTables& T, hipStream_t s) {
    std::vector<float2> h;
    for (int i = 0; kLens[i]; i++) {
        T.offset[i] = (int)h.size();
        int L = kLens[i];
        for (int t = 0; t < L; t++) {
            double a = -2.0 * M_PI * (double)t / (double)L;
            h.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
        }
    }
    T.tw.upload(h.data(), h.size(), s);
    ETTK_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ compile-time radix plans
template <int L> struct Radices;
template <> struct Radices<16>  { static constexpr int n = 2; static constexpr int r[4] = {4, 4, 1, 1}; };
template <> struct Radices<32>  { static constexpr int n = 3; static constexpr int r[4] = {4, 4, 2, 1}; };
template <> struct Radices<48>  { static constexpr int n = 3; static constexpr int r[4] = {4, 4, 3, 1}; };
template <> struct Radices<64>  { static constexpr int n = 3; static constexpr int r[4] = {4, 4, 4, 1}; };
template <> struct Radices<96>  { static constexpr int n = 4; static constexpr int r[4] = {4, 4, 3, 2}; };
template <> struct Radices<128> { static constexpr int n = 4; static constexpr int r[4] = {4, 4, 4, 2}; };
template <> struct Radices<192> { static constexpr int n = 4; static constexpr int r[4] = {4, 4, 4, 3}; };
template <> struct Radices<256> { static constexpr int n = 4; static constexpr int r[4] = {4, 4, 4, 4}; };

template <int L> struct Shape {
    static constexpr int TPP = L / 4;                       // threads per pencil
    static constexpr int P = (1024 / L) > 0 ? 1024 / L : 1; // pencils per workgroup
    static constexpr int THREADS = P * TPP;
    static constexpr int LD = L + 1;                        // padded LDS row
    static_assert(THREADS >= L, "one twiddle per thread");
};

__device__ __forceinline__ float2 mul_i(float2 z, bool inv) {  // forward: -i z, inverse: +i z
    return inv ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

template <int R>
__device__ __forceinline__ void butterfly(float2* x, bool inv) {
    if constexpr (R == 2) {
        float2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const float s = inv ? 0.86602540378443864676f : -0.86602540378443864676f;
        float2 t1 = cadd(x[1], x[2]);
        float2 t2 = make_float2(x[0].x - 0.5f * t1.x, x[0].y - 0.5f * t1.y);
        float2 d = csub(x[1], x[2]);
        float2 id = make_float2(-s * d.y, s * d.x);  // i*s*d
        x[0] = cadd(x[0], t1);
        x[1] = cadd(t2, id);
        x[2] = csub(t2, id);
    } else if constexpr (R == 4) {
        float2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
        float2 c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
        float2 md = mul_i(d, inv);  // forward: -i d
        x[0] = cadd(a, c);
        x[2] = csub(a, c);
        x[1] = cadd(b, md);
        x[3] = csub(b, md);
    }
}

template <int L, int R>
__device__ __forceinline__ void stockham_stage(float2* pen, int lane, int Ns, const float2* tw, bool inv) {
    constexpr int NB = L / R;
    constexpr int TPP = Shape<L>::TPP;
    constexpr int MAXB = (NB + TPP - 1) / TPP;
    float2 v[MAXB][R];
    int base[MAXB];
    const int step = L / (Ns * R);
#pragma unroll
    for (int c = 0; c < MAXB; c++) {
        int bf = lane + c * TPP;
        base[c] = -1;
        if (bf < NB) {
            int k = bf % Ns;
#pragma unroll
            for (int m = 0; m < R; m++) {
                float2 x = pen[bf + m * NB];
                if (m > 0) {
                    float2 w = tw[m * k * step];
                    if (inv) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[c][m] = x;
            }
            butterfly<R>(v[c], inv);
            base[c] = (bf / Ns) * Ns * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXB; c++)
        if (base[c] >= 0)
#pragma unroll
            for (int m = 0; m < R; m++) pen[base[c] + m * Ns] = v[c][m];
    __syncthreads();
}

template <int L, int S>
__device__ __forceinline__ void run_stage(float2* pen, int lane, int& Ns, const float2* tw, bool inv) {
    using RD = Radices<L>;
    if constexpr (S < RD::n) {
        constexpr int R = RD::r[S];
        stockham_stage<L, R>(pen, lane, Ns, tw, inv);
        Ns *= R;
    }
}

// in-LDS transform of one pencil (every thread of the block must call it: it synchronises)
template <int L>
__device__ __forceinline__ void lds_fft(float2* pen, int lane, const float2* tw, bool inv) {
    int Ns = 1;
    run_stage<L, 0>(pen, lane, Ns, tw, inv);
    run_stage<L, 1>(pen, lane, Ns, tw, inv);
    run_stage<L, 2>(pen, lane, Ns, tw, inv);
    run_stage<L, 3>(pen, lane, Ns, tw, inv);
}

// ------------------------------------------------------------------ length 64 in registers (8 x 8)
// 64 = 8 x 8 Cooley-Tukey: each of 8 threads holds x[t + 8m] (m = 0..7), runs an 8-point DFT in registers,
// applies W64^(t k1), and the 8 x 8 block is transposed through the pencil's own LDS row; an 8-point DFT per
// thread finishes X[k1 + 8 k2].  One LDS exchange instead of three radix-4 stages.
__device__ __forceinline__ float2 mul_w8(float2 z, int n, bool inv) {  // z * exp(-+2 pi i n / 8), n = 1..3
    const float r = 0.70710678118654752440f;
    switch (n) {
        case 1: return inv ? make_float2(r * (z.x - z.y), r * (z.x + z.y)) : make_float2(r * (z.x + z.y), r * (z.y - z.x));
        case 2: return mul_i(z, inv);
        default: return inv ? make_float2(-r * (z.x + z.y), r * (z.x - z.y)) : make_float2(r * (z.y - z.x), -r * (z.x + z.y));
    }
}
// in place, natural order in and out: a[k] = sum_n a[n] exp(-+2 pi i n k / 8)
__device__ __forceinline__ void dft8(float2* a, bool inv) {
    float2 b0 = cadd(a[0], a[4]), b4 = csub(a[0], a[4]);
    float2 b1 = cadd(a[1], a[5]), b5 = mul_w8(csub(a[1], a[5]), 1, inv);
    float2 b2 = cadd(a[2], a[6]), b6 = mul_i(csub(a[2], a[6]), inv);
    float2 b3 = cadd(a[3], a[7]), b7 = mul_w8(csub(a[3], a[7]), 3, inv);
    float2 c0 = cadd(b0, b2), c2 = csub(b0, b2), c1 = cadd(b1, b3), c3 = mul_i(csub(b1, b3), inv);
    float2 c4 = cadd(b4, b6), c6 = csub(b4, b6), c5 = cadd(b5, b7), c7 = mul_i(csub(b5, b7), inv);
    a[0] = cadd(c0, c1);
    a[4] = csub(c0, c1);
    a[2] = cadd(c2, c3);
    a[6] = csub(c2, c3);
    a[1] = cadd(c4, c5);
    a[5] = csub(c4, c5);
    a[3] = cadd(c6, c7);
    a[7] = csub(c6, c7);
}
// the twiddles a thread of a length-64 pencil applies, w[k] = exp(-2 pi i lane k / 64), held in registers
// for all the pencil transforms of a kernel (length 64 only; other lengths read the LDS table)
template <int L> struct TwCache {
    float2 w[L == 64 ? 8 : 1];
};
template <int L>
__device__ __forceinline__ TwCache<L> tw_cache(const float2* tw, int lane) {
    TwCache<L> c;
    if constexpr (L == 64) {
#pragma unroll
        for (int k = 0; k < 8; k++) c.w[k] = tw[(lane * k) & 63];
    } else {
        c.w[0] = make_float2(1.f, 0.f);
    }
    return c;
}

// Orders the LDS accesses of one wavefront: the LDS instructions of a wavefront execute in order, so threads of
// one wavefront exchanging data through LDS only need the compiler kept from reordering those accesses.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// every thread of the workgroup must call it (it ends with a workgroup barrier, after which the whole pencil is
// visible to every thread).  The 8 threads of a pencil (lanes 8q .. 8q+7 of the workgroup) share a wavefront,
// so the exchange inside needs only wavefront ordering.  The 8 x 8 exchange goes through rows of 9
// (pen[k * 9 + t] = element k of thread t), so neither its writes nor its reads collide in LDS banks; the
// pencil's storage therefore spans 72 entries.
__device__ __forceinline__ void fft64_reg(float2* pen, int lane, const TwCache<64>& tc, bool inv) {
    float2 x[8];
#pragma unroll
    for (int m = 0; m < 8; m++) x[m] = pen[lane + 8 * m];
    dft8(x, inv);
#pragma unroll
    for (int k = 1; k < 8; k++) {
        float2 w = tc.w[k];
        if (inv) w.y = -w.y;
        x[k] = cmul(x[k], w);
    }
    wave_lds_order();
#pragma unroll
    for (int k = 0; k < 8; k++) pen[k * 9 + lane] = x[k];
    wave_lds_order();
#pragma unroll
    for (int t = 0; t < 8; t++) x[t] = pen[lane * 9 + t];
    dft8(x, inv);
    wave_lds_order();
#pragma unroll
    for (int k = 0; k < 8; k++) pen[lane + 8 * k] = x[k];
    __syncthreads();
}

// threads per pencil of the pencil FFT used by the fused and strided passes, and the LDS stride of a pencil
template <int L> struct PenT {
    static constexpr int TPP = (L == 64) ? 8 : Shape<L>::TPP;
    static constexpr int LD = (L == 64) ? 73 : L + 1;
};
template <int L>
__device__ __forceinline__ void pencil_fft(float2* pen, int lane, const float2* tw, const TwCache<L>& tc, bool inv) {
    if constexpr (L == 64) fft64_reg(pen, lane, tc, inv);
    else lds_fft<L>(pen, lane, tw, inv);
}

// ------------------------------------------------------------------ r2c, pass along k (real rows)
template <int L>
__device__ __forceinline__ float r2c_load(const R2CArgs& a, long row, int k) {
    long ij = row % ((long)L * L);
    int i = (int)(ij / L), j = (int)(ij % L);
    float sg = ((i + j + k) & 1) ? -1.0f : 1.0f;
    if (a.pro == PRO_PAD_MASKED) {
        // cartesianToSpherical (ProjRot.h:265-318): the masked image (image - mean) * mask + mean sits in the
        // corner [0, ns)^3 of the (2 ns)^3 box, the rest holds the value of voxel (0,0,0); times (-1)^(i+j+k)
        const long b = row / ((long)L * L);
        const int ns = a.ns;
        const float* st = a.stats + b * 4;
        if (i < ns && j < ns && k < ns) {
            long o = ((long)i * ns + j) * ns + k;
            float x = a.in[b * (long)ns * ns * ns + o];
            float img = (x == -3.402823466e+38f) ? 0.0f : x - st[0];
            float v = st[3] != 0.f ? (img - st[1]) * a.aux0[o] + st[1] : img;
            return v * sg;
        }
        return st[2] * sg;
    }
    float x = a.in[row * L + k];
    if (a.pro == PRO_NONE) return x;
    switch (a.pro) {
        case PRO_SQUARE_SHIFT: return x * x * sg;
        case PRO_SHIFT: return x * sg;
        case PRO_MASK_SHIFT: return x * a.aux0[ij * L + k] * sg;
        case PRO_RESLICED: {
            if (x == -3.402823466e+38f) return 0.0f;  // outside the particle box: -FLT_MAX background
            long b = row / ((long)L * L);
            return (x - a.stats[b]) * (a.aux0 ? a.aux0[ij * L + k] : 1.0f) * sg;
        }
    }
    return x;
}

template <int L>
__global__ void __launch_bounds__(Shape<L>::THREADS) k_r2c_rows(R2CArgs a, const float2* __restrict__ twg) {
    using S = Shape<L>;
    constexpr int H = L / 2 + 1;
    __shared__ float2 lds[S::P * S::LD];
    __shared__ float2 tw[L];
    const long npairs = (long)a.batch * L * L / 2;
    const long p0 = (long)blockIdx.x * S::P;
    const float2 twv = twg[threadIdx.x % L];  // stored after the staging loads are in flight
    // load P pairs of rows (2L contiguous floats each) and pack row 2q + i*row 2q+1
    for (int t = threadIdx.x; t < S::P * 2 * L; t += blockDim.x) {
        int q = t / (2 * L), r = t % (2 * L);
        long pair = p0 + q;
        float x = 0.0f;
        if (pair < npairs) {
            long row = pair * 2 + r / L;
            x = r2c_load<L>(a, row, r % L);
        }
        float* f = reinterpret_cast<float*>(&lds[q * S::LD + (r % L)]);
        f[r / L] = x;
    }
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    const int q = threadIdx.x / S::TPP, lane = threadIdx.x % S::TPP;
    lds_fft<L>(&lds[q * S::LD], lane, tw, false);
    // split Z = A + iB into the half spectra of the two rows
    for (int t = threadIdx.x; t < S::P * 2 * H; t += blockDim.x) {
        int qq = t / (2 * H), r = t % (2 * H);
        long pair = p0 + qq;
        if (pair >= npairs) continue;
        int which = r / H, k = r % H;
        float2 z = lds[qq * S::LD + k], zc = cconj(lds[qq * S::LD + ((L - k) % L)]);
        float2 o = which == 0 ? cscale(cadd(z, zc), 0.5f) : make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
        a.out[(pair * 2 + which) * H + k] = o;
    }
}

// ------------------------------------------------------------------ complex pass along a strided axis
struct StridedArgs {
    const float2* in;
    float2* out;
    long npencils;
    long inner;         // pencils with consecutive addresses
    long outer_stride;  // address step between groups of `inner` pencils
    long es;            // element stride along the transformed axis
    int inv;
};

template <int L> struct StridedShape {
    static constexpr int TPP = PenT<L>::TPP;
    static constexpr int P = Shape<L>::THREADS / TPP;  // pencils per workgroup
    static constexpr int THREADS = Shape<L>::THREADS;
    static constexpr int LD = PenT<L>::LD;
};

template <int L>
__global__ void __launch_bounds__(StridedShape<L>::THREADS) k_fft_strided(StridedArgs a, const float2* __restrict__ twg) {
    using S = StridedShape<L>;
    __shared__ float2 lds[S::P * S::LD];
    __shared__ float2 tw[L];
    // consecutive pencil groups on one XCD: at L = 128 a group's rows are half cache lines, whose other halves
    // belong to the neighbouring groups (fetched into one L2 instead of two)
    const long q0 = (long)xcd_tile(blockIdx.x, gridDim.x) * S::P;
    const float2 twv = twg[threadIdx.x % L];  // stored after the staging loads are in flight
    // THREADS is a multiple of P: a thread moves the elements l = t / P, t / P + THREADS / P, ... of the one
    // pencil t % P, so the pencil's base address is computed once
    const int qm = threadIdx.x % S::P;
    const long pen = q0 + qm;
    const bool valid = pen < a.npencils;
    const long base = valid ? (pen / a.inner) * a.outer_stride + (pen % a.inner) : 0;
    constexpr int STEP = S::THREADS / S::P, NL = L / STEP;
    static_assert(L % STEP == 0, "pencil elements per thread");
    const int l0 = threadIdx.x / S::P;
    {
        float2 v[NL];  // all loads in flight before the first LDS store (an invalid pencil re-reads pencil 0)
#pragma unroll
        for (int u = 0; u < NL; u++) v[u] = a.in[base + (long)(l0 + u * STEP) * a.es];
#pragma unroll
        for (int u = 0; u < NL; u++) lds[qm * S::LD + l0 + u * STEP] = valid ? v[u] : make_float2(0.f, 0.f);
    }
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    const int q = threadIdx.x / S::TPP, lane = threadIdx.x % S::TPP;
    pencil_fft<L>(&lds[q * S::LD], lane, tw, tw_cache<L>(tw, lane), a.inv != 0);
    if (valid)
#pragma unroll
        for (int u = 0; u < NL; u++) a.out[base + (long)(l0 + u * STEP) * a.es] = lds[qm * S::LD + l0 + u * STEP];
}

// ------------------------------------------------------------------ CCC spectra: one pass along i (L <= 64)
// deterministic sum over the NT threads of a workgroup (NT need not be a multiple of 64), valid in thread 0
template <int NT>
__device__ __forceinline__ float block_sum_nt(float v, float* sred) {
    constexpr int NFW = NT / 64, REM = NT % 64;
    const int w = threadIdx.x >> 6;
    if (w < NFW) {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if ((threadIdx.x & 63) == 0) sred[w] = v;
    } else {
        sred[NFW + (threadIdx.x & 63)] = v;
    }
    __syncthreads();
    float tot = 0.f;
    if (threadIdx.x == 0)
        for (int i = 0; i < NFW + REM; i++) tot += sred[i];
    __syncthreads();
    return tot;
}

// Between the k/j plane passes of the particle's r2c and the fused plane kernels, every step of the CCC runs on
// pencils along i, so one kernel does them all per pencil (j, k) of an evaluation:
//   F2 = forward i-transform of the plane-pass output (completes the particle's r2c)
//   fw = filter * W1 * W2(e) (the smoothed particle wedge at the evaluation's transform, computed once)
//   X2 = mcf(F2) * fw, DC term zeroed -> S2 (the combine's F2), and its inverse i-transform -> S3
//   X1 = F1mcf * fw, DC term zeroed -> inverse i-transform -> S1
//   per-block partials of sum hw |X1|^2, sum fw, sum hw |X2|^2 (hw = 2 off the last half plane)
// The element-wise arithmetic is that of k_spec_filter2 (nbfFourierFilter::executeHalf, FFilt.h:587-606, and
// normalizeFourierHalf, CorrMetric.h:498-530); the transforms are those of the separate i passes.
template <int L>
__global__ void __launch_bounds__(StridedShape<L>::THREADS) k_ccc_ipass(CccIpassArgs a, const float2* __restrict__ twg) {
    using S = StridedShape<L>;
    constexpr int H = L / 2 + 1, LH = L * H, STEP = S::THREADS / S::P, NL = L / STEP;
    static_assert(L % STEP == 0, "pencil elements per thread");
    __shared__ float2 lds[S::P * S::LD];
    __shared__ float2 tw[L];
    __shared__ float sred[S::THREADS / 64 + S::THREADS % 64 + 1];
    const int e = blockIdx.y;
    const float2 twv = twg[threadIdx.x % L];
    const int qm = threadIdx.x % S::P, l0 = threadIdx.x / S::P;
    const int p = blockIdx.x * S::P + qm;  // pencil (j, k) = (p / H, p % H) of evaluation e
    const bool valid = p < LH;
    const int pc = valid ? p : 0, pj = pc / H, pk = pc % H;
    const size_t vol = (size_t)e * L * LH;
    float2* __restrict__ pen = &lds[qm * S::LD];
    {
        float2 v[NL];
#pragma unroll
        for (int u = 0; u < NL; u++) v[u] = ld_stream(&a.S2[vol + (size_t)(l0 + u * STEP) * LH + pc]);
#pragma unroll
        for (int u = 0; u < NL; u++) pen[l0 + u * STEP] = valid ? v[u] : make_float2(0.f, 0.f);
    }
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    const int q = threadIdx.x / S::TPP, lane = threadIdx.x % S::TPP;
    const TwCache<L> twc = tw_cache<L>(tw, lane);
    pencil_fft<L>(&lds[q * S::LD], lane, tw, twc, false);
    // filter weights of this thread's elements (i = l0 + u STEP of pencil p), the particle side in place
    const EvalGeom& g = a.geom[e];
    const int dc_i = L / 2, dc_j = L / 2, dc_k = H - 1;
    const float hw = pk < H - 1 ? 2.f : 1.f;
    float fw[NL], flt[NL], w1[NL];
#pragma unroll
    for (int u = 0; u < NL; u++) {
        const size_t h = (size_t)(l0 + u * STEP) * LH + pc;
        flt[u] = a.filter[h];
        w1[u] = a.W1 ? a.W1[h] : 1.f;
    }
    float s2 = 0.f, swf = 0.f;
#pragma unroll
    for (int u = 0; u < NL; u++) {
        const int i = l0 + u * STEP;
        float w = w1[u];
        if (!g.full) w *= wedge_smooth(g.v1, g.v2, i, pj, pk, L, H);
        fw[u] = flt[u] * w;
        const bool dc = i == dc_i && pj == dc_j && pk == dc_k;
        float2 z = pen[i];
        if (a.mcf) {
            const float m2 = z.x * z.x + z.y * z.y;
            const float C = sqrtf(sqrtf(m2));
            if (C > 0.f) z = make_float2(z.x / C, z.y / C);
        }
        if (a.apply) z = cscale(z, fw[u]);
        if (dc) z = make_float2(0.f, 0.f);
        if (valid) {
            st_stream(&a.S2[vol + (size_t)i * LH + pc], z);
            s2 += hw * (z.x * z.x + z.y * z.y);
            swf += fw[u];
        }
        pen[i] = z;
    }
    __syncthreads();
    pencil_fft<L>(&lds[q * S::LD], lane, tw, twc, true);
    // reference side: its spectrum loads are issued before the particle side's inverse transform is stored
    float2 f1[NL];
#pragma unroll
    for (int u = 0; u < NL; u++) f1[u] = a.F1mcf[(size_t)(l0 + u * STEP) * LH + pc];
    if (valid)
#pragma unroll
        for (int u = 0; u < NL; u++) st_stream(&a.S3[vol + (size_t)(l0 + u * STEP) * LH + pc], pen[l0 + u * STEP]);
    float s1 = 0.f;
#pragma unroll
    for (int u = 0; u < NL; u++) {
        const int i = l0 + u * STEP;
        float2 z = f1[u];
        if (a.apply) z = cscale(z, fw[u]);
        if (i == dc_i && pj == dc_j && pk == dc_k) z = make_float2(0.f, 0.f);
        if (valid) s1 += hw * (z.x * z.x + z.y * z.y);
        pen[i] = z;
    }
    __syncthreads();
    pencil_fft<L>(&lds[q * S::LD], lane, tw, twc, true);
    if (valid)
#pragma unroll
        for (int u = 0; u < NL; u++) st_stream(&a.S1[vol + (size_t)(l0 + u * STEP) * LH + pc], pen[l0 + u * STEP]);
    const float t1 = block_sum_nt<S::THREADS>(s1, sred);
    const float tf = block_sum_nt<S::THREADS>(swf, sred);
    const float t2 = block_sum_nt<S::THREADS>(s2, sred);
    if (threadIdx.x == 0) {
        const size_t o = (size_t)e * gridDim.x + blockIdx.x;
        a.p_n1[o] = t1;
        a.p_wf[o] = tf;
        a.p_n2[o] = t2;
    }
}

// ------------------------------------------------------------------ c2r, pass along k (half rows)
template <int L>
__global__ void __launch_bounds__(Shape<L>::THREADS) k_c2r_rows(const float2* __restrict__ in, float* __restrict__ out,
                                                                  long npairs, const float2* __restrict__ twg) {
    using S = Shape<L>;
    constexpr int H = L / 2 + 1;
    __shared__ float2 lds[S::P * S::LD];
    __shared__ float2 tw[L];
    const long p0 = (long)blockIdx.x * S::P;
    const float2 twv = twg[threadIdx.x % L];  // stored after the staging loads are in flight
    // Z(k) = A(k) + i B(k) over the full length, with A(L-k) = conj(A(k)) and real A(0), A(L/2)
    for (int t = threadIdx.x; t < S::P * L; t += blockDim.x) {
        int q = t / L, k = t % L;
        long pair = p0 + q;
        float2 z = make_float2(0.f, 0.f);
        if (pair < npairs) {
            const float2* ra = in + (pair * 2) * H;
            const float2* rb = ra + H;
            float2 A, Bv;
            if (k <= L / 2) {
                A = ra[k];
                Bv = rb[k];
                if (k == 0 || k == L / 2) { A.y = 0.f; Bv.y = 0.f; }
            } else {
                A = cconj(ra[L - k]);
                Bv = cconj(rb[L - k]);
            }
            z = make_float2(A.x - Bv.y, A.y + Bv.x);
        }
        lds[q * S::LD + k] = z;
    }
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    const int q = threadIdx.x / S::TPP, lane = threadIdx.x % S::TPP;
    lds_fft<L>(&lds[q * S::LD], lane, tw, true);
    for (int t = threadIdx.x; t < S::P * 2 * L; t += blockDim.x) {
        int qq = t / (2 * L), r = t % (2 * L);
        long pair = p0 + qq;
        if (pair >= npairs) continue;
        float2 z = lds[qq * S::LD + (r % L)];
        out[pair * 2 * L + r] = (r < L) ? z.x : z.y;
    }
}

// ------------------------------------------------------------------ fused plane passes (L <= 64)
// One workgroup per (batch, i) plane of L x L reals: the k pass (pairs of real rows packed as complex rows) and
// the j pass run back to back in LDS, so an r2c moves each element through HBM twice (plane pass + i pass)
// instead of three times.  H pencil groups of TPP threads; the spare group of the k pass works on a dummy row.
// The k pass works on the L/2 row pairs.  The j pass works on L/2 pencils as well: the DC (k = 0) and Nyquist
// (k = L/2) columns of a real plane's row transforms are real, so they travel as one complex pencil
// DC + i Nyquist (pencil 0) and are separated after the transform by Hermitian symmetry; the other pencils
// are the columns k = 1 .. L/2 - 1.  Slot L/2 of the buffer holds the Nyquist column while a c2r packs it.
template <int L> struct Plane {
    static constexpr int H = L / 2 + 1;
    static constexpr int TPP = PenT<L>::TPP;
    static constexpr int THREADS = (L / 2) * TPP;
    static constexpr int LD = PenT<L>::LD;
    static_assert(THREADS >= L, "one twiddle per thread");
};

// c2r: pencil 0 <- h(column 0) + i h(column L/2), h(c)(j) = (c(j) + conj c(-j)) / 2.  The c2r keeps only the real
// part of the inverse i/j transforms on these two planes (FFTW's c2r; a spectrum that is not exactly Hermitian
// there, such as one weighted by a border-clipped smoothed wedge, has an imaginary part that is dropped).  The
// inverse transform of h(c) is exactly that real part, so the two columns share one transform: its real and
// imaginary parts.  Every thread of the workgroup must call it (it synchronises).
template <int L>
__device__ __forceinline__ void pack_dc_nyquist(float2* buf) {
    constexpr int LD = Plane<L>::LD;
    const int t = threadIdx.x, tm = (L - t) % L;
    float2 z = make_float2(0.f, 0.f);
    if (t < L) {
        const float2 c0 = buf[t], c0m = cconj(buf[tm]), cn = buf[(L / 2) * LD + t], cnm = cconj(buf[(L / 2) * LD + tm]);
        const float2 h0 = make_float2(0.5f * (c0.x + c0m.x), 0.5f * (c0.y + c0m.y));
        const float2 hn = make_float2(0.5f * (cn.x + cnm.x), 0.5f * (cn.y + cnm.y));
        z = make_float2(h0.x - hn.y, h0.y + hn.x);
    }
    __syncthreads();
    if (t < L) buf[t] = z;
}
// c2r along k: Z(k) = A(k) + i B(k) of rows 2pp, 2pp+1 over the full length from the inverse-j-transformed
// pencils, A(L-k) = conj(A(k)), A(0) and A(L/2) real (the real and imaginary parts of pencil 0)
template <int L>
__device__ __forceinline__ float2 c2r_row_pair(const float2* buf, int pp, int k) {
    constexpr int LD = Plane<L>::LD;
    float2 A, Bv;
    if (k == 0 || k == L / 2) {
        const float2 a = buf[2 * pp], b = buf[2 * pp + 1];
        A = make_float2(k == 0 ? a.x : a.y, 0.f);
        Bv = make_float2(k == 0 ? b.x : b.y, 0.f);
    } else if (k < L / 2) {
        A = buf[k * LD + 2 * pp];
        Bv = buf[k * LD + 2 * pp + 1];
    } else {
        A = cconj(buf[(L - k) * LD + 2 * pp]);
        Bv = cconj(buf[(L - k) * LD + 2 * pp + 1]);
    }
    return make_float2(A.x - Bv.y, A.y + Bv.x);
}
// r2c: value o of column k of row 2pp + which into the j-pass pencils; at k = 0 and k = L/2 the row transform of
// a real row is real (o.y is exactly 0) and goes to the real or imaginary part of pencil 0
template <int L>
__device__ __forceinline__ void r2c_split_store(float2* buf, int pp, int which, int k, float2 o) {
    constexpr int LD = Plane<L>::LD;
    if (k > 0 && k < L / 2) buf[k * LD + 2 * pp + which] = o;
    else reinterpret_cast<float*>(&buf[2 * pp + which])[k == 0 ? 0 : 1] = o.x;
}
// r2c: element (j, k) of the half plane after the forward j pass; DC and Nyquist from pencil 0 Z:
// X0(j) = (Z(j) + conj Z(-j)) / 2, XN(j) = (Z(j) - conj Z(-j)) / 2i
template <int L>
__device__ __forceinline__ float2 r2c_half_at(const float2* buf, int j, int k) {
    constexpr int LD = Plane<L>::LD;
    if (k > 0 && k < L / 2) return buf[k * LD + j];
    const float2 z = buf[j], zc = cconj(buf[(L - j) % L]);
    if (k == 0) return make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    return make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
}

// the k-pass input of plane i for one preprocessing mode, staged as packed row pairs buf[(j/2) LD + x]
// (.x row j even, .y row j odd); MSK: aux0 holds the N^3 mask
template <int L, int PRO, bool MSK>
__device__ __forceinline__ void r2c_stage_plane(const R2CArgs& a, int pi, const float* __restrict__ src, float mean,
                                                float2* buf) {
    constexpr int LD = Plane<L>::LD;
    const float* __restrict__ msk = MSK ? a.aux0 + (size_t)pi * L * L : nullptr;
    block_copy<Plane<L>::THREADS, L * L>(
        [&](int t) { return make_float2(src[t], MSK ? msk[t] : 1.0f); },
        [&](int t, float2 v) {
            const int j = t / L, x = t % L;
            const float sg = ((pi + j + x) & 1) ? -1.0f : 1.0f;
            float r;
            if constexpr (PRO == PRO_NONE) r = v.x;
            else if constexpr (PRO == PRO_SQUARE_SHIFT) r = v.x * v.x * sg;
            else if constexpr (PRO == PRO_SHIFT) r = v.x * sg;
            else if constexpr (PRO == PRO_MASK_SHIFT) r = v.x * v.y * sg;
            else r = (v.x == -3.402823466e+38f) ? 0.0f : (v.x - mean) * v.y * sg;  // PRO_RESLICED
            reinterpret_cast<float*>(&buf[(j >> 1) * LD + x])[j & 1] = r;
        });
}

template <int L>
__global__ void __launch_bounds__(Plane<L>::THREADS) k_r2c_plane(R2CArgs a, const float2* __restrict__ twg) {
    using PL = Plane<L>;
    constexpr int H = PL::H, LD = PL::LD;
    constexpr int NS = ((L / 2) * 2 * H + PL::THREADS - 1) / PL::THREADS;  // split outputs per thread
    __shared__ float2 buf[H * LD];  // k pass: L/2 packed row pairs; then j pass: L/2 column pencils
    __shared__ float2 tw[L];
    const long plane = blockIdx.x;  // b * L + i
    const int pb = (int)(plane / L), pi = (int)(plane % L);
    const float* src = a.in + plane * L * L;
    const float mean = a.pro == PRO_RESLICED ? a.stats[pb] : 0.0f;
    const float2 twv = twg[threadIdx.x % L];  // stored after the staging loads are in flight
    switch (a.pro) {
        case PRO_NONE: r2c_stage_plane<L, PRO_NONE, false>(a, pi, src, mean, buf); break;
        case PRO_SQUARE_SHIFT: r2c_stage_plane<L, PRO_SQUARE_SHIFT, false>(a, pi, src, mean, buf); break;
        case PRO_SHIFT: r2c_stage_plane<L, PRO_SHIFT, false>(a, pi, src, mean, buf); break;
        case PRO_MASK_SHIFT: r2c_stage_plane<L, PRO_MASK_SHIFT, true>(a, pi, src, mean, buf); break;
        case PRO_RESLICED:
            if (a.aux0) r2c_stage_plane<L, PRO_RESLICED, true>(a, pi, src, mean, buf);
            else r2c_stage_plane<L, PRO_RESLICED, false>(a, pi, src, mean, buf);
            break;
        default:
            for (int t = threadIdx.x; t < L * L; t += blockDim.x) {
                const int j = t / L, x = t % L;
                float* f = reinterpret_cast<float*>(&buf[(j >> 1) * LD + x]);
                f[j & 1] = r2c_load<L>(a, ((long)pb * L + pi) * L + j, x);
            }
    }
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    const int q = threadIdx.x / PL::TPP, lane = threadIdx.x % PL::TPP;
    const TwCache<L> twc = tw_cache<L>(tw, lane);
    pencil_fft<L>(&buf[q * LD], lane, tw, twc, false);
    // split Z = A + iB of each pair into the half spectra of its two rows, through registers
    float2 o[NS];
#pragma unroll
    for (int u = 0; u < NS; u++) {
        const int t = threadIdx.x + u * PL::THREADS;
        if (t < (L / 2) * 2 * H) {
            const int pp = t / (2 * H), r = t % (2 * H), which = r / H, k = r % H;
            const float2 z = buf[pp * LD + k], zc = cconj(buf[pp * LD + ((L - k) % L)]);
            o[u] = which == 0 ? cscale(cadd(z, zc), 0.5f) : make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NS; u++) {
        const int t = threadIdx.x + u * PL::THREADS;
        if (t < (L / 2) * 2 * H) {
            const int pp = t / (2 * H), r = t % (2 * H), which = r / H, k = r % H;
            r2c_split_store<L>(buf, pp, which, k, o[u]);
        }
    }
    __syncthreads();
    pencil_fft<L>(&buf[q * LD], lane, tw, twc, false);
    float2* op = a.out + plane * L * H;
    for (int t = threadIdx.x; t < L * H; t += blockDim.x) op[t] = r2c_half_at<L>(buf, t / H, t % H);
}

template <int L>
__global__ void __launch_bounds__(Plane<L>::THREADS) k_c2r_plane(const float2* __restrict__ in, float* __restrict__ out,
                                                                   const float2* __restrict__ twg) {
    using PL = Plane<L>;
    constexpr int H = PL::H, LD = PL::LD;
    constexpr int NZ = ((L / 2) * L + PL::THREADS - 1) / PL::THREADS;
    __shared__ float2 buf[H * LD];
    __shared__ float2 tw[L];
    const long plane = blockIdx.x;
    const float2 twv = twg[threadIdx.x % L];  // stored after the staging loads are in flight
    const float2* ip = in + plane * L * H;
    block_copy<PL::THREADS, L * H>([&](int t) { return ip[t]; }, [&](int t, float2 v) { buf[(t % H) * LD + t / H] = v; });
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    pack_dc_nyquist<L>(buf);
    __syncthreads();
    const int q = threadIdx.x / PL::TPP, lane = threadIdx.x % PL::TPP;
    const TwCache<L> twc = tw_cache<L>(tw, lane);
    pencil_fft<L>(&buf[q * LD], lane, tw, twc, true);
    float2 zr[NZ];
#pragma unroll
    for (int u = 0; u < NZ; u++) {
        const int t = threadIdx.x + u * PL::THREADS;
        if (t < (L / 2) * L) zr[u] = c2r_row_pair<L>(buf, t / L, t % L);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NZ; u++) {
        const int t = threadIdx.x + u * PL::THREADS;
        if (t < (L / 2) * L) buf[(t / L) * LD + (t % L)] = zr[u];
    }
    __syncthreads();
    pencil_fft<L>(&buf[q * LD], lane, tw, twc, true);
    float* op = out + plane * L * L;
    for (int t = threadIdx.x; t < L * L; t += blockDim.x) {
        const int j = t / L, x = t % L;
        const float2 z = buf[(j >> 1) * LD + x];
        op[t] = (j & 1) ? z.y : z.x;
    }
}

// ------------------------------------------------------------------ fused c2r -> pointwise op -> r2c (L <= 64)
// Between the inverse and the forward transform of one volume, a (b, i) plane is complete in LDS after the c2r
// plane passes: the pointwise operation and the r2c plane passes run there, so the real volume never goes
// through HBM.  `io` holds the half spectrum after the inverse i pass and receives the forward j/k passes.
//   PLANE_OP_MASK:          r <- r m (m = aux, N^3), partials[plane] = sum r^2 m over the plane
//   PLANE_OP_SQUARE_SHIFT:  r <- r^2 (-1)^(i+j+k)
template <int L>
__global__ void __launch_bounds__(Plane<L>::THREADS) k_c2r_r2c_plane(float2* __restrict__ io, int op,
                                                                       const float* __restrict__ aux,
                                                                       float* __restrict__ partials,
                                                                       const float2* __restrict__ twg) {
    using PL = Plane<L>;
    constexpr int H = PL::H, LD = PL::LD, T = PL::THREADS;
    constexpr int NZ = ((L / 2) * L + T - 1) / T;
    constexpr int NS = ((L / 2) * 2 * H + T - 1) / T;
    constexpr int NFW = T / 64, REM = T % 64;  // full waves, threads of the partial wave
    __shared__ float2 buf[H * LD];
    __shared__ float2 tw[L];
    __shared__ float sred[NFW + REM + 1];
    const long plane = blockIdx.x;
    const int pi = (int)(plane % L);
    const float2 twv = twg[threadIdx.x % L];  // stored after the staging loads are in flight
    float2* ip = io + plane * L * H;
    block_copy<T, L * H>([&](int t) { return ip[t]; }, [&](int t, float2 v) { buf[(t % H) * LD + t / H] = v; });
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    pack_dc_nyquist<L>(buf);
    __syncthreads();
    const int q = threadIdx.x / PL::TPP, lane = threadIdx.x % PL::TPP;
    const TwCache<L> twc = tw_cache<L>(tw, lane);
    pencil_fft<L>(&buf[q * LD], lane, tw, twc, true);
    // c2r along k (as k_c2r_plane)
    float2 zr[NZ];
#pragma unroll
    for (int u = 0; u < NZ; u++) {
        const int t = threadIdx.x + u * T;
        if (t < (L / 2) * L) zr[u] = c2r_row_pair<L>(buf, t / L, t % L);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NZ; u++) {
        const int t = threadIdx.x + u * T;
        if (t < (L / 2) * L) buf[(t / L) * LD + (t % L)] = zr[u];
    }
    __syncthreads();
    pencil_fft<L>(&buf[q * LD], lane, tw, twc, true);
    __syncthreads();
    // buf[p LD + x] = (r(row 2p, x), r(row 2p + 1, x)): the pointwise operation in place
    float sacc = 0.f;
    if (op == PLANE_OP_MASK) {
        // mask values of rows j0, j0 + 1 at x: all loads in flight before the first use
        const float* __restrict__ mp = aux + (size_t)pi * L * L;
        block_copy<T, (L / 2) * L>(
            [&](int t) { return make_float2(mp[(2 * (t / L)) * L + t % L], mp[(2 * (t / L) + 1) * L + t % L]); },
            [&](int t, float2 m) {
                const int pp = t / L, x = t % L;
                float2 z = buf[pp * LD + x];
                sacc += z.x * z.x * m.x;
                sacc += z.y * z.y * m.y;
                z.x *= m.x;
                z.y *= m.y;
                buf[pp * LD + x] = z;
            });
    } else {
        for (int t = threadIdx.x; t < (L / 2) * L; t += T) {
            const int pp = t / L, x = t % L, j0 = 2 * pp;
            float2 z = buf[pp * LD + x];
            const float sg = ((pi + j0 + x) & 1) ? -1.0f : 1.0f;
            z.x = z.x * z.x * sg;
            z.y = z.y * z.y * (-sg);
            buf[pp * LD + x] = z;
        }
    }
    if (op == PLANE_OP_MASK) {
        // deterministic block sum: shuffles inside the full waves, the partial wave's values through LDS
        const int w = threadIdx.x >> 6;
        if (w < NFW) {
            for (int o = 32; o > 0; o >>= 1) sacc += __shfl_down(sacc, o, 64);
            if ((threadIdx.x & 63) == 0) sred[w] = sacc;
        } else {
            sred[NFW + (threadIdx.x & 63)] = sacc;
        }
    }
    __syncthreads();
    if (op == PLANE_OP_MASK && threadIdx.x == 0) {
        float tot = 0.f;
        for (int i = 0; i < NFW + REM; i++) tot += sred[i];
        partials[plane] = tot;
    }
    // r2c along k then j (as k_r2c_plane after its load)
    pencil_fft<L>(&buf[q * LD], lane, tw, twc, false);
    float2 o[NS];
#pragma unroll
    for (int u = 0; u < NS; u++) {
        const int t = threadIdx.x + u * T;
        if (t < (L / 2) * 2 * H) {
            const int pp = t / (2 * H), r = t % (2 * H), which = r / H, k = r % H;
            const float2 z = buf[pp * LD + k], zc = cconj(buf[pp * LD + ((L - k) % L)]);
            o[u] = which == 0 ? cscale(cadd(z, zc), 0.5f) : make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NS; u++) {
        const int t = threadIdx.x + u * T;
        if (t < (L / 2) * 2 * H) {
            const int pp = t / (2 * H), r = t % (2 * H), which = r / H, k = r % H;
            r2c_split_store<L>(buf, pp, which, k, o[u]);
        }
    }
    __syncthreads();
    pencil_fft<L>(&buf[q * LD], lane, tw, twc, false);
    for (int t = threadIdx.x; t < L * H; t += blockDim.x) ip[t] = r2c_half_at<L>(buf, t / H, t % H);
}

// ------------------------------------------------------------------ combine + first DFT axis by FFT (N = 64)
// The first axis of the zero-padded 2N-point inverse transform on the search window (k_combine_dftA's table
// A[x][i] = (-1)^f e^{+i pi f u / N}, f = i - N/2, u = I0 + x - N) as two length-N FFTs per pencil: with
// u = 2m + b, G(u) = sum_f [(-1)^f C(f) e^{+2 pi i f b / 2N}] e^{+2 pi i f m / N}, a backward N-point transform of
// the twiddled pencil (f periodic mod N) for each parity b, read at m mod N.  One workgroup per (j, evaluation):
// 2H pencils (k, b) of 8 threads, each a length-64 register FFT.
// FWD: F1m and Q arrive before their forward i transform (the fused plane kernels' outputs); the workgroup runs
// those transforms on its 2H pencils first.
static constexpr int kCfP = 2 * 33, kCfT = kCfP * 8;
template <bool FWD>
__global__ void __launch_bounds__(kCfT) k_combine_fftA(const float2* __restrict__ F1m, const float2* __restrict__ F2,
                                                       const float2* __restrict__ Q, const float2* __restrict__ Mfft,
                                                       const float* __restrict__ scalars, CccSizes z,
                                                       float2* __restrict__ G1, const float2* __restrict__ tw64,
                                                       const float2* __restrict__ tw128) {
    constexpr int L = 64, H = L / 2 + 1, LD = PenT<64>::LD, NE = (L * H + kCfT - 1) / kCfT;
    static_assert(L * H == NE * kCfT, "four elements per thread");
    __shared__ float2 pens[kCfP * LD];
    __shared__ float2 tw[L];
    __shared__ float2 wb[L];  // e^{+2 pi i (i - L/2) / 2L} for i = 0 .. L-1
    // the j planes of one evaluation on one XCD: a plane's rows of 33 complex values are not aligned to cache
    // lines, and the lines they share with the neighbouring planes are then fetched into one L2 only
    int j, e;
    xcd_eval_tile(blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, e, j);
    const int E = z.E;
    const long nh = (long)L * L * H;
    if (threadIdx.x < L) {
        tw[threadIdx.x] = tw64[threadIdx.x];
        wb[threadIdx.x] = cconj(tw128[(threadIdx.x - L / 2 + 2 * L) % (2 * L)]);
    }
    const float n1 = scalars[e * 4 + 0], n2 = scalars[e * 4 + 1];
    const float n3 = (float)L * L * L;
    const float a = 2.f * n3 / (n1 * n2);
    const float b = 1.f / (n2 * n2);
    const float2* __restrict__ f1p = F1m + (size_t)e * nh;
    const float2* __restrict__ f2p = F2 + (size_t)e * nh;
    const float2* __restrict__ qp = Q + (size_t)e * nh;
    const int q = threadIdx.x >> 3, lane = threadIdx.x & 7;
    // this thread's elements t = threadIdx.x + u kCfT of the (i, k) plane
    float2 c0[NE], c1[NE];
    if constexpr (FWD) {
        {
            float2 vf1[NE], vq[NE];
#pragma unroll
            for (int u = 0; u < NE; u++) {
                const int t = threadIdx.x + u * kCfT;
                const long h = ((long)(t / H) * L + j) * H + t % H;
                vf1[u] = ld_stream(&f1p[h]);
                vq[u] = ld_stream(&qp[h]);
            }
#pragma unroll
            for (int u = 0; u < NE; u++) {
                const int t = threadIdx.x + u * kCfT, i = t / H, k = t % H;
                pens[k * LD + i] = vf1[u];
                pens[(H + k) * LD + i] = vq[u];
            }
        }
        __syncthreads();
        fft64_reg(&pens[q * LD], lane, tw_cache<64>(tw, lane), false);  // F1m and Q along i
        float2 vf2[NE], vm[NE];  // loaded after the transforms: fewer registers held across them
#pragma unroll
        for (int u = 0; u < NE; u++) {
            const int t = threadIdx.x + u * kCfT;
            const long h = ((long)(t / H) * L + j) * H + t % H;
            vf2[u] = ld_stream(&f2p[h]);
            vm[u] = Mfft[h];
        }
#pragma unroll
        for (int u = 0; u < NE; u++) {
            const int t = threadIdx.x + u * kCfT, i = t / H, k = t % H;
            float2 c = csub(cscale(cmulc(pens[k * LD + i], vf2[u]), a), cscale(cmulc(vm[u], pens[(H + k) * LD + i]), b));
            if ((i + j + k) & 1) c = make_float2(-c.x, -c.y);
            if ((i - L / 2) & 1) c = make_float2(-c.x, -c.y);  // (-1)^f of the table
            c0[u] = c;
            c1[u] = cmul(c, wb[i]);
        }
        __syncthreads();
    } else {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NE; u++) {
            const int t = threadIdx.x + u * kCfT;
            const long h = ((long)(t / H) * L + j) * H + t % H;
            const float2 f1 = ld_stream(&f1p[h]), f2 = ld_stream(&f2p[h]), qv = ld_stream(&qp[h]), m = Mfft[h];
            const int i = t / H, k = t % H;
            float2 c = csub(cscale(cmulc(f1, f2), a), cscale(cmulc(m, qv), b));
            if ((i + j + k) & 1) c = make_float2(-c.x, -c.y);
            if ((i - L / 2) & 1) c = make_float2(-c.x, -c.y);
            c0[u] = c;
            c1[u] = cmul(c, wb[i]);
        }
    }
#pragma unroll
    for (int u = 0; u < NE; u++) {
        const int t = threadIdx.x + u * kCfT, i = t / H, k = t % H;
        const int n = (i + L / 2) & (L - 1);  // f mod L
        pens[k * LD + n] = c0[u];
        pens[(H + k) * LD + n] = c1[u];
    }
    __syncthreads();
    fft64_reg(&pens[q * LD], lane, tw_cache<64>(tw, lane), true);
    const int u0 = z.I0 - L;
    for (int t = threadIdx.x; t < E * H; t += kCfT) {
        const int x = t / H, k = t % H, u = u0 + x, par = u & 1, m = ((u - par) / 2) & (L - 1);
        G1[(((size_t)e * E + x) * L + j) * H + k] = pens[(par * H + k) * LD + m];
    }
}

// The second axis the same way (k_dftBC_argmax's A along j), then the third axis on the crop, the spherical mask
// and the first maximum as k_dftBC_argmax computes them.  One workgroup per (x, evaluation).
__global__ void __launch_bounds__(kCfT) k_dftBC_fft(const float2* __restrict__ G1, const float2* __restrict__ tabK,
                                                    CccSizes z, float* __restrict__ best_val, int* __restrict__ best_idx,
                                                    const float2* __restrict__ tw64, const float2* __restrict__ tw128) {
    constexpr int L = 64, H = L / 2 + 1, LD = PenT<64>::LD;
    extern __shared__ float2 K[];  // [zz][k], E x H
    __shared__ float2 pens[kCfP * LD];
    __shared__ float2 tw[L];
    __shared__ float2 wb[L];
    __shared__ float sv[kCfT];
    __shared__ int si[kCfT];
    const int x = blockIdx.x, e = blockIdx.y, E = z.E;
    if (threadIdx.x < L) {
        tw[threadIdx.x] = tw64[threadIdx.x];
        wb[threadIdx.x] = cconj(tw128[(threadIdx.x - L / 2 + 2 * L) % (2 * L)]);
    }
    __syncthreads();
    const float2* src = G1 + ((size_t)e * E + x) * L * H;  // [j][k]
    block_copy_rt<4>(
        L * H, [&](int t) { return src[t]; },
        [&](int t, float2 c) {
            const int jj = t / H, k = t % H;
            if ((jj - L / 2) & 1) c = make_float2(-c.x, -c.y);
            const int n = (jj + L / 2) & (L - 1);
            pens[k * LD + n] = c;
            pens[(H + k) * LD + n] = cmul(c, wb[jj]);
        });
    block_copy_rt<3>(E * H, [&](int t) { return tabK[t]; }, [&](int t, float2 v) { K[t] = v; });
    __syncthreads();
    const int q = threadIdx.x >> 3, lane = threadIdx.x & 7;
    fft64_reg(&pens[q * LD], lane, tw_cache<64>(tw, lane), true);
    // g2(y, k) = pens[(par H + k) LD + m] with u = I0 + y - L = 2m + par
    const float scale = 8.f / ((float)(2 * L) * (float)(2 * L) * (float)(2 * L));
    float bv = -FLT_MAX;
    int bi = 0x7fffffff;
    const int dx = z.ub2 - x, u0 = z.I0 - L;
    for (int t = threadIdx.x; t < E * E; t += kCfT) {
        const int y = t / E, zz = t % E;
        const int dy = z.ub2 - y, dz = z.ub2 - zz;
        const float rad = (float)(dx * dx + dy * dy + dz * dz);
        if (!(rad < z.restr2)) continue;
        const int u = u0 + y, par = u & 1, m = ((u - par) / 2) & (L - 1);
        const float2* gy = &pens[par * H * LD + m];
        float acc = 0.f;
        for (int k = 0; k < H; k++) {
            const float2 p = gy[k * LD], qv = K[zz * H + k];
            acc = fmaf(p.x, qv.x, acc);
            acc = fmaf(-p.y, qv.y, acc);
        }
        acc *= scale;
        if (acc > bv || (acc == bv && t < bi)) {
            bv = acc;
            bi = t;
        }
    }
    // first maximum over the workgroup (value descending, then index ascending: a total order, so the halving
    // tree over any count gives the same result as a scan)
    sv[threadIdx.x] = bv;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int cnt = kCfT; cnt > 1;) {
        const int st = (cnt + 1) / 2;
        if ((int)threadIdx.x < cnt - st) {
            const int r = threadIdx.x;
            const float v2 = sv[r + st];
            const int i2 = si[r + st];
            if (v2 > sv[r] || (v2 == sv[r] && i2 < si[r])) {
                sv[r] = v2;
                si[r] = i2;
            }
        }
        __syncthreads();
        cnt = st;
    }
    if (threadIdx.x == 0) {
        best_val[(size_t)e * E + x] = sv[0];
        best_idx[(size_t)e * E + x] = si[0];
    }
}

// ------------------------------------------------------------------ host launchers
// profile record names: k_r2c_rows<L>, k_fft_strided<L>, k_c2r_rows<L>
template <int L>
static const char* plane_name(int which) {
    static const std::string n[2] = {"k_r2c_plane<" + std::to_string(L) + ">", "k_c2r_plane<" + std::to_string(L) + ">"};
    return n[which].c_str();
}
template <int L>
static const char* prof_name(int which) {
    static const std::string n[3] = {"k_r2c_rows<" + std::to_string(L) + ">", "k_fft_strided<" + std::to_string(L) + ">",
                                     "k_c2r_rows<" + std::to_string(L) + ">"};
    return n[which].c_str();
}

template <int L>
static void launch_r2c(const FftTables& T, const R2CArgs& a, float2* scratch, hipStream_t s) {
    using S = Shape<L>;
    constexpr int H = L / 2 + 1;
    const float2* tw = T.tw.p + T.offset[fft_len_index(L)];
    long npairs = (long)a.batch * L * L / 2;
    const double vol_c = (double)a.batch * L * L * H * sizeof(float2);  // one pass over the half spectrum
    const double vol_r = a.pro == PRO_PAD_MASKED ? (double)a.batch * a.ns * a.ns * a.ns * sizeof(float)
                                                 : (double)a.batch * L * L * L * sizeof(float);
    if constexpr (L <= 64) {
        // passes k and j fused per (b, i) plane
        ProfScope ps(T.prof, plane_name<L>(0), vol_r + vol_c, s);
        k_r2c_plane<L><<<(unsigned)((long)a.batch * L), Plane<L>::THREADS, 0, s>>>(a, tw);
    } else {
        // pass k: real rows -> a.out (half rows)
        {
            ProfScope ps(T.prof, prof_name<L>(0), vol_r + vol_c, s);
            k_r2c_rows<L><<<(unsigned)((npairs + S::P - 1) / S::P), S::THREADS, 0, s>>>(a, tw);
        }
        // pass j: pencils over (b, i, k), element stride H
        StridedArgs sj{a.out, a.out, (long)a.batch * L * H, (long)H, (long)L * H, (long)H, 0};
        {
            ProfScope ps(T.prof, prof_name<L>(1), 2 * vol_c, s);
            k_fft_strided<L><<<(unsigned)((sj.npencils + StridedShape<L>::P - 1) / StridedShape<L>::P), StridedShape<L>::THREADS, 0, s>>>(sj, tw);
        }
    }
    // pass i: pencils over (b, j, k), element stride L*H
    StridedArgs si{a.out, a.out, (long)a.batch * L * H, (long)L * H, (long)L * L * H, (long)L * H, 0};
    {
        ProfScope ps(T.prof, prof_name<L>(1), 2 * vol_c, s);
        k_fft_strided<L><<<(unsigned)((si.npencils + StridedShape<L>::P - 1) / StridedShape<L>::P), StridedShape<L>::THREADS, 0, s>>>(si, tw);
    }
    (void)scratch;
}

template <int L>
static void launch_c2r(const FftTables& T, const float2* in, float2* tmp, float* out, int batch, hipStream_t s) {
    using S = Shape<L>;
    constexpr int H = L / 2 + 1;
    const float2* tw = T.tw.p + T.offset[fft_len_index(L)];
    const double vol_c = (double)batch * L * L * H * sizeof(float2), vol_r = (double)batch * L * L * L * sizeof(float);
    StridedArgs si{in, tmp, (long)batch * L * H, (long)L * H, (long)L * L * H, (long)L * H, 1};
    {
        ProfScope ps(T.prof, prof_name<L>(1), 2 * vol_c, s);
        k_fft_strided<L><<<(unsigned)((si.npencils + StridedShape<L>::P - 1) / StridedShape<L>::P), StridedShape<L>::THREADS, 0, s>>>(si, tw);
    }
    if constexpr (L <= 64) {
        ProfScope ps(T.prof, plane_name<L>(1), vol_c + vol_r, s);
        k_c2r_plane<L><<<(unsigned)((long)batch * L), Plane<L>::THREADS, 0, s>>>(tmp, out, tw);
    } else {
        StridedArgs sj{tmp, tmp, (long)batch * L * H, (long)H, (long)L * H, (long)H, 1};
        {
            ProfScope ps(T.prof, prof_name<L>(1), 2 * vol_c, s);
            k_fft_strided<L><<<(unsigned)((sj.npencils + StridedShape<L>::P - 1) / StridedShape<L>::P), StridedShape<L>::THREADS, 0, s>>>(sj, tw);
        }
        long npairs = (long)batch * L * L / 2;
        {
            ProfScope ps(T.prof, prof_name<L>(2), vol_c + vol_r, s);
            k_c2r_rows<L><<<(unsigned)((npairs + S::P - 1) / S::P), S::THREADS, 0, s>>>(tmp, out, npairs, tw);
        }
    }
}

template <int L>
static void launch_c2r_op_r2c(const FftTables& T, const float2* in, float2* io, int batch, int op, const float* aux,
                              float* partials, hipStream_t s, bool fwd_i = true) {
    constexpr int H = L / 2 + 1;
    const float2* tw = T.tw.p + T.offset[fft_len_index(L)];
    const double vol_c = (double)batch * L * L * H * sizeof(float2);
    StridedArgs si{in, io, (long)batch * L * H, (long)L * H, (long)L * L * H, (long)L * H, 1};
    if (in) {
        ProfScope ps(T.prof, prof_name<L>(1), 2 * vol_c, s);
        k_fft_strided<L><<<(unsigned)((si.npencils + StridedShape<L>::P - 1) / StridedShape<L>::P), StridedShape<L>::THREADS, 0, s>>>(si, tw);
    }
    {
        static const std::string name = "k_c2r_r2c_plane<" + std::to_string(L) + ">";
        ProfScope ps(T.prof, name.c_str(), 2 * vol_c, s);
        k_c2r_r2c_plane<L><<<(unsigned)((long)batch * L), Plane<L>::THREADS, 0, s>>>(io, op, aux, partials, tw);
    }
    StridedArgs sf{io, io, (long)batch * L * H, (long)L * H, (long)L * L * H, (long)L * H, 0};
    if (fwd_i) {
        ProfScope ps(T.prof, prof_name<L>(1), 2 * vol_c, s);
        k_fft_strided<L><<<(unsigned)((sf.npencils + StridedShape<L>::P - 1) / StridedShape<L>::P), StridedShape<L>::THREADS, 0, s>>>(sf, tw);
    }
}

bool fft3_fused_supported(int n) { return n == 16 || n == 32 || n == 48 || n == 64; }

bool ccc_combine_fft_supported(const CccSizes& z) { return z.N == 64; }
void launch_dftBC_fft(const FftTables& T, const float2* G1, const float2* tabK, const CccSizes& z, int n_eval,
                      float* best_val, int* best_idx, hipStream_t s) {
    ETTK_REQUIRE(ccc_combine_fft_supported(z), ETTK_ERR_UNSUPPORTED, "launch_dftBC_fft: box size must be 64");
    const float2* tw64 = T.tw.p + T.offset[fft_len_index(64)];
    const float2* tw128 = T.tw.p + T.offset[fft_len_index(128)];
    const size_t shm = (size_t)z.E * z.H * sizeof(float2);
    ensure_dynamic_lds((const void*)k_dftBC_fft, shm);
    k_dftBC_fft<<<dim3(z.E, n_eval), kCfT, shm, s>>>(G1, tabK, z, best_val, best_idx, tw64, tw128);
    ETTK_HIP(hipGetLastError());
}
void launch_combine_fftA(const FftTables& T, const float2* F1m, const float2* F2, const float2* Q, const float2* Mfft,
                         const float* scalars, const CccSizes& z, int n_eval, float2* G1, bool forward_i, hipStream_t s) {
    ETTK_REQUIRE(ccc_combine_fft_supported(z), ETTK_ERR_UNSUPPORTED, "launch_combine_fftA: box size must be 64");
    const float2* tw64 = T.tw.p + T.offset[fft_len_index(64)];
    const float2* tw128 = T.tw.p + T.offset[fft_len_index(128)];
    if (forward_i)
        k_combine_fftA<true><<<dim3(z.N, n_eval), kCfT, 0, s>>>(F1m, F2, Q, Mfft, scalars, z, G1, tw64, tw128);
    else
        k_combine_fftA<false><<<dim3(z.N, n_eval), kCfT, 0, s>>>(F1m, F2, Q, Mfft, scalars, z, G1, tw64, tw128);
    ETTK_HIP(hipGetLastError());
}

template <int L>
static int ccc_ipass_blocks_t() {
    return (L * (L / 2 + 1) + StridedShape<L>::P - 1) / StridedShape<L>::P;
}
int fft3_ccc_ipass_blocks(int n) {
    switch (n) {
        case 16: return ccc_ipass_blocks_t<16>();
        case 32: return ccc_ipass_blocks_t<32>();
        case 48: return ccc_ipass_blocks_t<48>();
        case 64: return ccc_ipass_blocks_t<64>();
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_ccc_ipass: unsupported size " + std::to_string(n)};
    }
}

template <int L>
static void launch_ccc_ipass(const FftTables& T, const CccIpassArgs& a, int batch, hipStream_t s) {
    constexpr int H = L / 2 + 1;
    const float2* tw = T.tw.p + T.offset[fft_len_index(L)];
    const double vol_c = (double)batch * L * L * H * sizeof(float2);
    static const std::string name = "k_ccc_ipass<" + std::to_string(L) + ">";
    ProfScope ps(T.prof, name.c_str(), 4 * vol_c, s);  // S2 read and written, S3 and S1 written
    k_ccc_ipass<L><<<dim3((unsigned)ccc_ipass_blocks_t<L>(), (unsigned)batch), StridedShape<L>::THREADS, 0, s>>>(a, tw);
}
void fft3_ccc_ipass(const FftTables& T, const CccIpassArgs& a, int n, int batch, hipStream_t s) {
    switch (n) {
        case 16: launch_ccc_ipass<16>(T, a, batch, s); break;
        case 32: launch_ccc_ipass<32>(T, a, batch, s); break;
        case 48: launch_ccc_ipass<48>(T, a, batch, s); break;
        case 64: launch_ccc_ipass<64>(T, a, batch, s); break;
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_ccc_ipass: unsupported size " + std::to_string(n)};
    }
    ETTK_HIP(hipGetLastError());
}

template <int L>
static void launch_r2c_planes(const FftTables& T, const R2CArgs& a, hipStream_t s) {
    constexpr int H = L / 2 + 1;
    const float2* tw = T.tw.p + T.offset[fft_len_index(L)];
    const double vol_c = (double)a.batch * L * L * H * sizeof(float2), vol_r = (double)a.batch * L * L * L * sizeof(float);
    ProfScope ps(T.prof, plane_name<L>(0), vol_r + vol_c, s);
    k_r2c_plane<L><<<(unsigned)((long)a.batch * L), Plane<L>::THREADS, 0, s>>>(a, tw);
}
// ------------------------------------------------------------------ refinement: score and adjoint in the i passes
// FP64 sums of the three score terms over a workgroup (halving tree, any thread count), valid in thread 0
template <int NT>
__device__ __forceinline__ void block_sum3_f64(double v0, double v1, double v2, double* sred) {
    sred[threadIdx.x] = v0;
    sred[NT + threadIdx.x] = v1;
    sred[2 * NT + threadIdx.x] = v2;
    __syncthreads();
    for (int cnt = NT; cnt > 1;) {
        const int st = (cnt + 1) / 2;
        if ((int)threadIdx.x < cnt - st)
#pragma unroll
            for (int r = 0; r < 3; r++) sred[r * NT + threadIdx.x] += sred[r * NT + threadIdx.x + st];
        __syncthreads();
        cnt = st;
    }
}

// The forward i pass of X1 (in place, after the plane passes of its r2c) and k_refine_dot's terms on the
// transformed elements: per workgroup (32 pencils (j, k) of one evaluation) the FP64 partials of <H X1, H X2>,
// |H X1|^2 and |H X2|^2 with the Hermitian weights (2 off the k = 0 and k = N/2 planes).
template <int L>
__global__ void __launch_bounds__(StridedShape<L>::THREADS) k_refine_ipass_dot(
    float2* __restrict__ X1, const float2* __restrict__ FP, const float* __restrict__ filter,
    const float* __restrict__ wbin, const RefineGeom* __restrict__ geom, double* __restrict__ part,
    const float2* __restrict__ twg) {
    using S = StridedShape<L>;
    constexpr int H = L / 2 + 1, LH = L * H, STEP = S::THREADS / S::P, NL = L / STEP;
    static_assert(L % STEP == 0, "pencil elements per thread");
    __shared__ float2 lds[S::P * S::LD];
    __shared__ float2 tw[L];
    __shared__ double sred[3 * S::THREADS];
    const int e = blockIdx.y;
    const float2 twv = twg[threadIdx.x % L];
    const int qm = threadIdx.x % S::P, l0 = threadIdx.x / S::P;
    const int p = blockIdx.x * S::P + qm;  // pencil (j, k) = (p / H, p % H) of evaluation e
    const bool valid = p < LH;
    const int pc = valid ? p : 0, pj = pc / H, pk = pc % H;
    const RefineGeom g = geom[e];
    float2* x1p = X1 + (size_t)e * L * LH;
    const float2* x2p = FP + (size_t)g.spec * L * LH;
    float2* __restrict__ pen = &lds[qm * S::LD];
    float2 v2[NL];
    {
        float2 v[NL];
#pragma unroll
        for (int u = 0; u < NL; u++) {
            v[u] = x1p[(size_t)(l0 + u * STEP) * LH + pc];
            v2[u] = x2p[(size_t)(l0 + u * STEP) * LH + pc];
        }
#pragma unroll
        for (int u = 0; u < NL; u++) pen[l0 + u * STEP] = valid ? v[u] : make_float2(0.f, 0.f);
    }
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    const int q = threadIdx.x / S::TPP, lane = threadIdx.x % S::TPP;
    pencil_fft<L>(&lds[q * S::LD], lane, tw, tw_cache<L>(tw, lane), false);
    __syncthreads();
    double sd = 0.0, s1 = 0.0, s2 = 0.0;
    if (valid) {
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int i = l0 + u * STEP;
            const size_t h = (size_t)i * LH + pc;
            const float2 z = pen[i];
            x1p[h] = z;
            const float w = refine_hweight(filter, wbin, g, (long)h, i, pj, pk, L, H);
            const double ww = (pk == 0 || pk == H - 1) ? (double)w * w : 2.0 * (double)w * w;
            const float2 v = v2[u];
            sd += ww * ((double)z.x * v.x + (double)z.y * v.y);
            s1 += ww * ((double)z.x * z.x + (double)z.y * z.y);
            s2 += ww * ((double)v.x * v.x + (double)v.y * v.y);
        }
    }
    block_sum3_f64<S::THREADS>(sd, s1, s2, sred);
    if (threadIdx.x < 3) part[((size_t)e * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = sred[threadIdx.x * S::THREADS];
}

// k_refine_adj's adjoint spectrum H (H X2 / (n1 n2) - S H X1 / n1^2), formed as the pencils are loaded, then
// the inverse i pass into `out` (X1 is left unchanged)
template <int L>
__global__ void __launch_bounds__(StridedShape<L>::THREADS) k_refine_ipass_adj(
    const float2* __restrict__ X1, const float2* __restrict__ FP, const float* __restrict__ filter,
    const float* __restrict__ wbin, const RefineGeom* __restrict__ geom, const double* __restrict__ sc,
    float2* __restrict__ out, const float2* __restrict__ twg) {
    using S = StridedShape<L>;
    constexpr int H = L / 2 + 1, LH = L * H, STEP = S::THREADS / S::P, NL = L / STEP;
    static_assert(L % STEP == 0, "pencil elements per thread");
    __shared__ float2 lds[S::P * S::LD];
    __shared__ float2 tw[L];
    const int e = blockIdx.y;
    const float2 twv = twg[threadIdx.x % L];
    const int qm = threadIdx.x % S::P, l0 = threadIdx.x / S::P;
    const int p = blockIdx.x * S::P + qm;
    const bool valid = p < LH;
    const int pc = valid ? p : 0, pj = pc / H, pk = pc % H;
    const RefineGeom g = geom[e];
    const double Sv = sc[e * 4], n1 = sc[e * 4 + 1], n2 = sc[e * 4 + 2];
    const float ca = (n1 > 0 && n2 > 0) ? (float)(1.0 / (n1 * n2)) : 0.f;
    const float cb = n1 > 0 ? (float)(Sv / (n1 * n1)) : 0.f;
    const float2* x1p = X1 + (size_t)e * L * LH;
    const float2* x2p = FP + (size_t)g.spec * L * LH;
    float2* __restrict__ pen = &lds[qm * S::LD];
    {
        float2 x[NL], v[NL];
#pragma unroll
        for (int u = 0; u < NL; u++) {
            x[u] = x1p[(size_t)(l0 + u * STEP) * LH + pc];
            v[u] = x2p[(size_t)(l0 + u * STEP) * LH + pc];
        }
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int i = l0 + u * STEP;
            const float w = refine_hweight(filter, wbin, g, (long)((size_t)i * LH + pc), i, pj, pk, L, H);
            const float w2 = w * w;
            pen[i] = valid ? make_float2(w2 * (ca * v[u].x - cb * x[u].x), w2 * (ca * v[u].y - cb * x[u].y))
                           : make_float2(0.f, 0.f);
        }
    }
    if (threadIdx.x < L) tw[threadIdx.x] = twv;
    __syncthreads();
    const int q = threadIdx.x / S::TPP, lane = threadIdx.x % S::TPP;
    pencil_fft<L>(&lds[q * S::LD], lane, tw, tw_cache<L>(tw, lane), true);
    __syncthreads();
    if (valid) {
        float2* op = out + (size_t)e * L * LH;
#pragma unroll
        for (int u = 0; u < NL; u++) op[(size_t)(l0 + u * STEP) * LH + pc] = pen[l0 + u * STEP];
    }
}

template <int L>
static void launch_refine_dot_t(const FftTables& T, float2* X1, const float2* FP, const float* filter,
                                const float* wbin, const RefineGeom* geom, int batch, double* part, hipStream_t s) {
    const float2* tw = T.tw.p + T.offset[fft_len_index(L)];
    ProfScope ps(T.prof, "k_refine_ipass_dot", 3.0 * batch * L * L * (L / 2 + 1) * sizeof(float2), s);
    k_refine_ipass_dot<L><<<dim3((unsigned)ccc_ipass_blocks_t<L>(), (unsigned)batch), StridedShape<L>::THREADS, 0, s>>>(
        X1, FP, filter, wbin, geom, part, tw);
}
template <int L>
static void launch_refine_adj_c2r_t(const FftTables& T, const float2* X1, const float2* FP, const float* filter,
                                    const float* wbin, const RefineGeom* geom, const double* sc, int batch,
                                    float2* tmp, float* out, hipStream_t s) {
    constexpr int H = L / 2 + 1;
    const float2* tw = T.tw.p + T.offset[fft_len_index(L)];
    {
        ProfScope ps(T.prof, "k_refine_ipass_adj", 3.0 * batch * L * L * H * sizeof(float2), s);
        k_refine_ipass_adj<L><<<dim3((unsigned)ccc_ipass_blocks_t<L>(), (unsigned)batch), StridedShape<L>::THREADS, 0,
                                s>>>(X1, FP, filter, wbin, geom, sc, tmp, tw);
    }
    ETTK_HIP(hipGetLastError());
    const double vol_c = (double)batch * L * L * H * sizeof(float2), vol_r = (double)batch * L * L * L * sizeof(float);
    ProfScope ps(T.prof, plane_name<L>(1), vol_c + vol_r, s);
    k_c2r_plane<L><<<(unsigned)((long)batch * L), Plane<L>::THREADS, 0, s>>>(tmp, out, tw);
}
int fft3_refine_ipass_blocks(int n) { return fft3_ccc_ipass_blocks(n); }
void fft3_refine_dot(const FftTables& T, float2* X1, const float2* FP, const float* filter, const float* wbin,
                     const RefineGeom* geom, int n, int batch, double* part, hipStream_t s) {
    switch (n) {
        case 16: launch_refine_dot_t<16>(T, X1, FP, filter, wbin, geom, batch, part, s); break;
        case 32: launch_refine_dot_t<32>(T, X1, FP, filter, wbin, geom, batch, part, s); break;
        case 48: launch_refine_dot_t<48>(T, X1, FP, filter, wbin, geom, batch, part, s); break;
        case 64: launch_refine_dot_t<64>(T, X1, FP, filter, wbin, geom, batch, part, s); break;
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_refine_dot: unsupported size " + std::to_string(n)};
    }
    ETTK_HIP(hipGetLastError());
}
void fft3_refine_adj_c2r(const FftTables& T, const float2* X1, const float2* FP, const float* filter,
                         const float* wbin, const RefineGeom* geom, const double* sc, int n, int batch, float2* tmp,
                         float* out, hipStream_t s) {
    switch (n) {
        case 16: launch_refine_adj_c2r_t<16>(T, X1, FP, filter, wbin, geom, sc, batch, tmp, out, s); break;
        case 32: launch_refine_adj_c2r_t<32>(T, X1, FP, filter, wbin, geom, sc, batch, tmp, out, s); break;
        case 48: launch_refine_adj_c2r_t<48>(T, X1, FP, filter, wbin, geom, sc, batch, tmp, out, s); break;
        case 64: launch_refine_adj_c2r_t<64>(T, X1, FP, filter, wbin, geom, sc, batch, tmp, out, s); break;
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_refine_adj_c2r: unsupported size " + std::to_string(n)};
    }
    ETTK_HIP(hipGetLastError());
}

void fft3_r2c_planes(const FftTables& T, const R2CArgs& a, hipStream_t s) {
    ETTK_REQUIRE(a.pro != PRO_PAD_MASKED, ETTK_ERR_ARG, "fft3_r2c_planes: padded prologue not supported");
    switch (a.n) {
        case 16: launch_r2c_planes<16>(T, a, s); break;
        case 32: launch_r2c_planes<32>(T, a, s); break;
        case 48: launch_r2c_planes<48>(T, a, s); break;
        case 64: launch_r2c_planes<64>(T, a, s); break;
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_r2c_planes: unsupported size " + std::to_string(a.n)};
    }
    ETTK_HIP(hipGetLastError());
}

void fft3_op_r2c_after_ipass(const FftTables& T, float2* io, int n, int batch, int op, const float* aux, float* partials,
                             bool forward_i, hipStream_t s) {
    switch (n) {
        case 16: launch_c2r_op_r2c<16>(T, nullptr, io, batch, op, aux, partials, s, forward_i); break;
        case 32: launch_c2r_op_r2c<32>(T, nullptr, io, batch, op, aux, partials, s, forward_i); break;
        case 48: launch_c2r_op_r2c<48>(T, nullptr, io, batch, op, aux, partials, s, forward_i); break;
        case 64: launch_c2r_op_r2c<64>(T, nullptr, io, batch, op, aux, partials, s, forward_i); break;
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_op_r2c_after_ipass: unsupported size " + std::to_string(n)};
    }
    ETTK_HIP(hipGetLastError());
}

void fft3_c2r_op_r2c(const FftTables& T, const float2* in, float2* io, int n, int batch, int op, const float* aux,
                     float* partials, hipStream_t s) {
    ETTK_REQUIRE(in != nullptr, ETTK_ERR_ARG, "fft3_c2r_op_r2c: null input");
    switch (n) {
        case 16: launch_c2r_op_r2c<16>(T, in, io, batch, op, aux, partials, s); break;
        case 32: launch_c2r_op_r2c<32>(T, in, io, batch, op, aux, partials, s); break;
        case 48: launch_c2r_op_r2c<48>(T, in, io, batch, op, aux, partials, s); break;
        case 64: launch_c2r_op_r2c<64>(T, in, io, batch, op, aux, partials, s); break;
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_c2r_op_r2c: unsupported size " + std::to_string(n)};
    }
    ETTK_HIP(hipGetLastError());
}

void fft3_r2c(const FftTables& T, const R2CArgs& a, float2* scratch, hipStream_t s) {
    switch (a.n) {
        case 16: launch_r2c<16>(T, a, scratch, s); break;
        case 32: launch_r2c<32>(T, a, scratch, s); break;
        case 48: launch_r2c<48>(T, a, scratch, s); break;
        case 64: launch_r2c<64>(T, a, scratch, s); break;
        case 96: launch_r2c<96>(T, a, scratch, s); break;
        case 128: launch_r2c<128>(T, a, scratch, s); break;
        case 192: launch_r2c<192>(T, a, scratch, s); break;
        case 256: launch_r2c<256>(T, a, scratch, s); break;
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_r2c: unsupported size " + std::to_string(a.n)};
    }
    ETTK_HIP(hipGetLastError());
}

void fft3_c2r(const FftTables& T, const float2* in, float2* tmp, float* out, int n, int batch, hipStream_t s) {
    switch (n) {
        case 16: launch_c2r<16>(T, in, tmp, out, batch, s); break;
        case 32: launch_c2r<32>(T, in, tmp, out, batch, s); break;
        case 48: launch_c2r<48>(T, in, tmp, out, batch, s); break;
        case 64: launch_c2r<64>(T, in, tmp, out, batch, s); break;
        case 96: launch_c2r<96>(T, in, tmp, out, batch, s); break;
        case 128: launch_c2r<128>(T, in, tmp, out, batch, s); break;
        case 192: launch_c2r<192>(T, in, tmp, out, batch, s); break;
        case 256: launch_c2r<256>(T, in, tmp, out, batch, s); break;
        default: throw Error{ETTK_ERR_UNSUPPORTED, "fft3_c2r: unsupported size " + std::to_string(n)};
    }
    ETTK_HIP(hipGetLastError());
}

}  // namespace ettk
