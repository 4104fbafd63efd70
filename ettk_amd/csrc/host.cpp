// host.cpp -- plan set-up, the constrained-CCC batch driver and the C ABI entry points of libettk_align.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <mutex>
#include <cstdlib>
#include <thread>

#include "plan.hpp"

namespace ettk {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

void ensure_dynamic_lds(const void* fn, size_t bytes) {
    static std::mutex mu;
    static std::vector<std::pair<const void*, size_t>> done;
    std::lock_guard<std::mutex> g(mu);
    for (auto& d : done)
        if (d.first == fn) {
            if (bytes <= d.second) return;
            ETTK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
            d.second = bytes;
            return;
        }
    ETTK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done.emplace_back(fn, bytes);
}

// ============================================================== FP64 transforms (vtkTransform semantics)
Mat4 operator*(const Mat4& A, const Mat4& B) {
    Mat4 C;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            C(r, c) = A(r, 0) * B(0, c) + A(r, 1) * B(1, c) + A(r, 2) * B(2, c) + A(r, 3) * B(3, c);
    return C;
}
Mat4 mat_rot(double angle, double x, double y, double z) {
    Mat4 M = Mat4::ident();
    if (angle == 0.0 || (x == 0.0 && y == 0.0 && z == 0.0)) return M;
    angle = angle * M_PI / 180.0;
    double w = std::cos(0.5 * angle), f = std::sin(0.5 * angle) / std::sqrt(x * x + y * y + z * z);
    x *= f; y *= f; z *= f;
    double ww = w * w, wx = w * x, wy = w * y, wz = w * z;
    double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
    double s = ww - xx - yy - zz;
    M(0, 0) = xx * 2 + s;    M(1, 0) = (xy + wz) * 2; M(2, 0) = (xz - wy) * 2;
    M(0, 1) = (xy - wz) * 2; M(1, 1) = yy * 2 + s;    M(2, 1) = (yz + wx) * 2;
    M(0, 2) = (xz + wy) * 2; M(1, 2) = (yz - wx) * 2; M(2, 2) = zz * 2 + s;
    return M;
}
Mat4 mat_translate(double x, double y, double z) {
    Mat4 M = Mat4::ident();
    M(0, 3) = x; M(1, 3) = y; M(2, 3) = z;
    return M;
}
Mat4 mat_scale(double x, double y, double z) {
    Mat4 M = Mat4::ident();
    M(0, 0) = x; M(1, 1) = y; M(2, 2) = z;
    return M;
}
double mat_det(const Mat4& A) {
    double d = 0;
    for (int c = 0; c < 4; c++) {
        double s[3][3];
        for (int r = 1; r < 4; r++)
            for (int k = 0, cc = 0; k < 4; k++)
                if (k != c) s[r - 1][cc++] = A(r, k);
        double m = s[0][0] * (s[1][1] * s[2][2] - s[1][2] * s[2][1]) - s[0][1] * (s[1][0] * s[2][2] - s[1][2] * s[2][0]) +
                   s[0][2] * (s[1][0] * s[2][1] - s[1][1] * s[2][0]);
        d += ((c & 1) ? -1.0 : 1.0) * A(0, c) * m;
    }
    return d;
}
Mat4 mat_inverse(const Mat4& A) {
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = A(r, c);
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        for (int k = 0; k < 8 && piv != c; k++) std::swap(a[c][k], a[piv][k]);
        double d = a[c][c];
        for (int k = 0; k < 8; k++) a[c][k] /= d;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                double f = a[r][c];
                for (int k = 0; k < 8; k++) a[r][k] -= f * a[c][k];
            }
    }
    Mat4 I;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) I(r, c) = a[r][c + 4];
    return I;
}
Mat4 strip_translation(const Mat4& t) { return mat_translate(-t(0, 3), -t(1, 3), -t(2, 3)) * t; }
Mat4 zyz(double a, double b, double g) { return mat_rot(-g, 0, 0, 1) * mat_rot(-b, 0, 1, 0) * mat_rot(-a, 0, 0, 1); }
bool transform_ok(const Mat4& t, double tr, double rot) {
    if (std::fabs(t(0, 3)) > tr || std::fabs(t(1, 3)) > tr || std::fabs(t(2, 3)) > tr) return false;
    double px = t(0, 2), py = t(1, 2), pz = t(2, 2), n = std::sqrt(px * px + py * py + pz * pz);
    double c = n != 0.0 ? pz / n : 0.0;
    double ang = c >= 1.0 ? 0.0 : std::fabs(std::acos(c) * 180.0 / M_PI);
    return ang <= rot;
}
Mat4 euler_of(const ettk_particle& d) {
    return mat_rot(-d.normal[2], 0, 0, 1) * mat_rot(-d.normal[0], 1, 0, 0) * mat_rot(-d.normal[1], 0, 0, 1);
}

// ============================================================== constant arrays of a plan
// separable vtkImageGaussianSmooth on an A(i,j,k) float array: axes 2, 1, 0, kernel clipped and renormalised
int host_threads() {
    static const int n = [] {
        int t = (int)std::thread::hardware_concurrency();
        if (const char* e = std::getenv("OMP_NUM_THREADS")) {
            const int v = std::atoi(e);
            if (v > 0) t = v;
        }
        return std::max(1, std::min(t, 16));
    }();
    return n;
}

void smooth3(std::vector<float>& A, int n0, int n1, int n2, double sd, double rf) {
    int dims[3] = {n0, n1, n2};
    long st[3] = {(long)n1 * n2, n2, 1};
    int rad = (int)(sd * rf);
    std::vector<double> g(2 * rad + 1);  // kernel taps, the same values for every voxel
    for (int x = -rad; x <= rad; x++) g[x + rad] = std::exp(-(double)(x * x) / (sd * sd * 2.0));
    std::vector<float> o(A.size());
    const int nt = 2 * rad + 1;
    for (int ax = 2; ax >= 0; ax--) {
        const int n = dims[ax];
        const long sa = st[ax];
        // the clipped, renormalised taps g[x] / sum_{lo..hi} g of each position along the axis (per position, not
        // per voxel: the same doubles the voxel loop multiplies by)
        std::vector<double> w((size_t)n * nt, 0.0);
        std::vector<int> wlo(n), whi(n);
        for (int idx = 0; idx < n; idx++) {
            const int lo = std::max(-rad, -idx), hi = std::min(rad, n - 1 - idx);
            double ks = 0;
            for (int x = lo; x <= hi; x++) ks += g[x + rad];
            for (int x = lo; x <= hi; x++) w[(size_t)idx * nt + x + rad] = g[x + rad] / ks;
            wlo[idx] = lo;
            whi[idx] = hi;
        }
        const long outer = (long)A.size() / ((long)n * sa);
        for (long ob = 0; ob < outer; ob++)
            for (int idx = 0; idx < n; idx++) {
                const double* wi = &w[(size_t)idx * nt + rad];
                const int lo = wlo[idx], hi = whi[idx];
                for (long in = 0; in < sa; in++) {
                    const long b = (ob * n + idx) * sa + in;
                    double s = 0;
                    for (int x = lo; x <= hi; x++) s += wi[x] * A[b + x * sa];
                    o[b] = (float)s;
                }
            }
        A.swap(o);
    }
}

void wedge_normals(float lw, float uw, const Mat4& t, WedgeRef& w) {
    double l = lw < -90 ? -90.0 : lw, u = uw > 90 ? 90.0 : uw;
    w.full = (u - l >= 180) ? 1 : 0;
    double n1[3] = {std::cos(-(90.0 - u) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - u) * M_PI / 180.0)};
    double n2[3] = {std::cos(-(90.0 - l) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - l) * M_PI / 180.0)};
    Mat4 R = strip_translation(t);
    for (int r = 0; r < 3; r++) {
        w.v1[r] = (float)(R(r, 0) * n1[0] + R(r, 1) * n1[1] + R(r, 2) * n1[2]);
        w.v2[r] = (float)(R(r, 0) * n2[0] + R(r, 1) * n2[1] + R(r, 2) * n2[2]);
    }
}

// search window of the upsampled correlation and its partial-DFT tables (FIM.h:1222-1247); rebuilt when
// the translation restriction changes
static void search_window(ettk_plan* P) {
    const int N = P->N, H = P->H;
    hipStream_t s = P->stream;
    double center = N;  // (2N)/2
    double restr = center;
    if (P->trans_restrict < (double)FLT_MAX) restr = std::max(P->trans_restrict * 2.0, 1.0);
    int I0 = (int)std::max(0.0, center - std::floor(restr));
    int I1 = (int)std::min(center + std::floor(restr), (double)(2 * N - 1));
    int E = I1 - I0 + 1;
    P->z.N = N;
    P->z.H = H;
    P->z.E = E;
    P->z.I0 = I0;
    P->z.ub2 = (E - 1) / 2;
    P->z.restr2 = (float)restr * (float)restr;
    size_t lds_a = (size_t)(N * H + E * N) * sizeof(float2);
    size_t lds_b = (size_t)(N * H + E * N + 2 * E * H) * sizeof(float2) + 2048;
    ETTK_REQUIRE(lds_a <= 160 * 1024 && lds_b <= 160 * 1024, ETTK_ERR_UNSUPPORTED,
                 "translation search window too large for the LDS-resident upsampled DFT (reduce trans_restrict)");
    std::vector<float2> tA((size_t)E * N), tK((size_t)E * H);
    for (int x = 0; x < E; x++) {
        int u = I0 + x - N;
        for (int i = 0; i < N; i++) {
            int fr = i - N / 2;
            double ang = M_PI * (double)fr * (double)u / (double)N;
            double sg = (fr & 1) ? -1.0 : 1.0;
            tA[(size_t)x * N + i] = make_float2((float)(sg * std::cos(ang)), (float)(sg * std::sin(ang)));
        }
        for (int k = 0; k < H; k++) {
            int fr = k - N / 2;
            double ang = M_PI * (double)fr * (double)u / (double)N;
            double sg = (fr & 1) ? -1.0 : 1.0;
            double w = (k == H - 1) ? 1.0 : 2.0;
            tK[(size_t)x * H + k] = make_float2((float)(w * sg * std::cos(ang)), (float)(w * sg * std::sin(ang)));
        }
    }
    P->tabA.upload(tA.data(), tA.size(), s);
    P->tabK.upload(tK.data(), tK.size(), s);
    ETTK_HIP(hipStreamSynchronize(s));
}

static void plan_constants(ettk_plan* P) {
    const ettk_params& p = P->p;
    const int N = P->N, H = P->H;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * H;
    hipStream_t s = P->stream;
    // spherical raised-cosine image mask (nbfImageFilter::buildMask, IFilt.h:384-472)
    std::vector<float> m1(n3, 1.0f);
    if (p.mask_radius != 0.0f) {
        float c = (float)(N / 2.0);
        float r2 = std::min((float)(N - 1), (float)(p.mask_radius + p.mask_sigma / 2.0));
        float r1 = std::max(0.0f, (float)(p.mask_radius - p.mask_sigma / 2.0));
        float kappa = (float)(M_PI / (r2 - r1));
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++)
                for (int k = 0; k < N; k++) {
                    float di = i - c, dj = j - c, dk = k - c;
                    float r = std::sqrt(di * di + dj * dj + dk * dk);
                    float v = r <= r1 ? 1.0f : (r >= r2 ? 0.0f : (float)((1.0 + std::cos(kappa * (r - r1))) / 2.0));
                    m1[((size_t)i * N + j) * N + k] = v;
                }
    }
    // square FFT window softened by a Gaussian of sigma 2.5 (FIM.h:642-666)
    std::vector<float> mf(n3, 0.0f);
    int sq = p.square_mask;
    if (sq > 0 && sq < std::floor(N / 2.0)) {
        for (int i = sq; i <= N - 1 - sq; i++)
            for (int j = sq; j <= N - 1 - sq; j++)
                for (int k = sq; k <= N - 1 - sq; k++) mf[((size_t)i * N + j) * N + k] = 1.0f;
    } else {
        std::fill(mf.begin(), mf.end(), 1.0f);
    }
    smooth3(mf, N, N, N, 2.5, 2.5 * 0.75);
    // band-pass x B-factor filter on the half spectrum, DC = 0 (FFilt.h:492-547, 738-797)
    std::vector<float> f(nh, 1.0f);
    f[((size_t)(N / 2) * N + N / 2) * H + (H - 1)] = 0.0f;
    bool bp = (p.bp_lo > 0.0f) || (p.bp_hi < 1.0f);
    int c = N / 2;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < H; k++) {
                double ti = (1.0 * i - c) / c, tj = (1.0 * j - c) / c, tk = (1.0 * k - c) / c;
                float rad = (float)std::sqrt(ti * ti + tj * tj + tk * tk);
                size_t o = ((size_t)i * N + j) * H + k;
                if (bp) {
                    double lo = p.bp_lo, hi = p.bp_hi, lv = p.bp_lo_decay, hv = p.bp_hi_decay;
                    float b = 1.0f;
                    if (rad < lo) b = (float)std::exp(-(lo - rad) * (lo - rad) / 2.0 / lv / lv);
                    if (rad > hi) b = (float)std::exp(-(rad - hi) * (rad - hi) / 2.0 / hv / hv);
                    f[o] *= b;
                }
                if (p.bfactor != 0.0f) f[o] *= (float)std::exp(-(double)p.bfactor * (double)rad * rad);
            }
    P->dummy_filter = !(bp || p.bfactor != 0.0f);
    P->sum_filter = 0;
    for (float v : f) P->sum_filter += v;
    P->nb_real = (int)((n3 + 1023) / 1024);
    P->nb_spec = (int)((nh + 1023) / 1024);
    P->nb_res = reslice_blocks(N);
    P->h_m1 = m1;
    P->m1.upload(m1.data(), n3, s);
    P->maskfft.upload(mf.data(), n3, s);
    P->filter.upload(f.data(), nh, s);
    // M = r2c(m1 * shift)
    P->Mfft.ensure(nh);
    R2CArgs a{P->m1.p, P->Mfft.p, N, 1, PRO_SHIFT, nullptr, nullptr};
    fft3_r2c(P->ctx->fft, a, nullptr, s);
    search_window(P);
}

// spectral partials per evaluation: per k_spec_filter2 block, or per k_ccc_ipass block (fused transforms)
static size_t spectral_partials(const ettk_plan* P) {
    return (size_t)std::max(P->nb_spec, fft3_fused_supported(P->N) ? fft3_ccc_ipass_blocks(P->N) : 0);
}

// one region holding a CccWork for E evaluations: every piece starts on a 256-byte boundary
template <class F>
static void ccc_work_pieces(const ettk_plan* P, int E, F&& piece) {
    const int N = P->N, H = P->H;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * H, nb_sp = spectral_partials(P);
    piece(0, n3 * E * sizeof(float));                                   // R1
    piece(1, n3 * E * sizeof(float));                                   // R2
    piece(2, (size_t)2 * P->nb_res * E * sizeof(float));                // p_res
    piece(3, nb_sp * E * sizeof(float));                                // p_n1
    piece(4, nb_sp * E * sizeof(float));                                // p_n2
    piece(5, nb_sp * E * sizeof(float));                                // p_wf
    piece(6, (size_t)std::max(P->nb_real, N) * E * sizeof(float));      // p_k1
    piece(7, (size_t)E * sizeof(float));                                // stats
    piece(8, (size_t)4 * E * sizeof(float));                            // scalars
    piece(9, (size_t)P->z.E * E * sizeof(float));                       // bval
    piece(10, (size_t)P->z.E * E * sizeof(int));                        // bidx
    piece(11, nh * E * sizeof(float2));                                 // S1
    piece(12, nh * E * sizeof(float2));                                 // S2
    piece(13, nh * E * sizeof(float2));                                 // S3
    piece(14, (size_t)P->z.E * N * H * E * sizeof(float2));             // G1
}

size_t ccc_work_bytes(const ettk_plan* P, int E) {
    size_t tot = 0;
    ccc_work_pieces(P, E, [&](int, size_t b) { tot += (b + 255) & ~(size_t)255; });
    return tot;
}

CccWork ccc_work_carve(const ettk_plan* P, int E, void* base) {
    void* ptr[15];
    size_t off = 0;
    ccc_work_pieces(P, E, [&](int i, size_t b) {
        ptr[i] = static_cast<char*>(base) + off;
        off += (b + 255) & ~(size_t)255;
    });
    CccWork W;
    W.R1 = (float*)ptr[0];
    W.R2 = (float*)ptr[1];
    W.p_res = (float*)ptr[2];
    W.p_n1 = (float*)ptr[3];
    W.p_n2 = (float*)ptr[4];
    W.p_wf = (float*)ptr[5];
    W.p_k1 = (float*)ptr[6];
    W.stats = (float*)ptr[7];
    W.scalars = (float*)ptr[8];
    W.bval = (float*)ptr[9];
    W.bidx = (int*)ptr[10];
    W.S1 = (float2*)ptr[11];
    W.S2 = (float2*)ptr[12];
    W.S3 = (float2*)ptr[13];
    W.G1 = (float2*)ptr[14];
    return W;
}

void ensure_ccc_workspace(ettk_plan* P, int cap) {
    if (cap <= P->cap && P->z.E <= P->ws_E) return;
    cap = std::max(cap, P->cap);
    P->ws_E = P->z.E;
    const int N = P->N, H = P->H;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * H;
    P->R1.ensure(n3 * cap);
    P->R2.ensure(n3 * cap);
    P->S1.ensure(nh * cap);
    P->S2.ensure(nh * cap);
    P->S3.ensure(nh * cap);
    P->G1.ensure((size_t)P->z.E * N * H * cap);
    P->p_res.ensure((size_t)2 * P->nb_res * cap);
    const size_t nb_sp = spectral_partials(P);
    P->p_n1.ensure(nb_sp * cap);
    P->p_n2.ensure(nb_sp * cap);
    P->p_wf.ensure(nb_sp * cap);
    P->p_k1.ensure((size_t)std::max(P->nb_real, N) * cap);  // mask partials: per block, or per plane (fused)
    P->stats.ensure(cap);
    P->scalars.ensure((size_t)4 * cap);
    P->bval.ensure((size_t)P->z.E * cap);
    P->bidx.ensure((size_t)P->z.E * cap);
    P->out4.ensure((size_t)4 * cap);
    P->geom.ensure(cap);
    P->cap = cap;
    ccc_graphs_reset(P);
}

void upload_before(ettk_plan* P, void* dst, const void* src, size_t bytes) {
    ETTK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, P->up_stream));
    ETTK_HIP(hipEventRecord(P->ev_up, P->up_stream));
    ETTK_HIP(hipStreamWaitEvent(P->stream, P->ev_up, 0));
}

void ccc_graphs_reset(ettk_plan* P) {
    for (auto& g : P->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    P->graphs.clear();
    P->graph_warm.clear();
    P->ws_gen++;
}

// The CCC batch with its chunks replayed from graphs captured once per (chunk size, particle array): every chunk
// reads its geometry from, and writes its results to, the plan's fixed per-chunk buffers, which are filled and
// drained by copies enqueued around each replay.  A chunk size runs eagerly once before it is captured (the
// kernels' attributes are set by then).
static void ccc_batch_graph(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n, const int* pidx,
                            Mat4* t, float* dist, const ettk_particle* eval_descs) {
    hipStream_t s = P->stream;
    const int chunk = std::max(1, std::min(P->ccc_chunk, n));
    ensure_ccc_workspace(P, chunk);
    P->geom.ensure(P->cap);
    P->out4.ensure((size_t)4 * P->cap);
    P->h_geom.ensure(n);
    P->h_out_pinned.ensure((size_t)4 * n);
    EvalGeom* g = P->h_geom.p;
    for (int e = 0; e < n; e++) g[e] = make_geom_public(P, eval_descs ? eval_descs[e] : descs[pidx[e]], t[e], pidx[e]);
    CccWork W{P->R1.p, P->R2.p, P->p_res.p, P->p_n1.p, P->p_n2.p, P->p_wf.p, P->p_k1.p, P->stats.p, P->scalars.p,
              P->bval.p, P->bidx.p, P->S1.p, P->S2.p, P->S3.p, P->G1.p};
    ETTK_HIP(hipEventRecord(P->ev0, s));
    for (int e0 = 0; e0 < n; e0 += chunk) {
        const int E = std::min(chunk, n - e0);
        ETTK_HIP(hipMemcpyAsync(P->geom.p, g + e0, (size_t)E * sizeof(EvalGeom), hipMemcpyHostToDevice, s));
        CccGraph* cg = nullptr;
        for (auto& x : P->graphs)
            if (x.E == E && x.parts == d_parts && x.gen == P->ws_gen) cg = &x;
        const bool warm = std::find(P->graph_warm.begin(), P->graph_warm.end(), E) != P->graph_warm.end();
        if (!cg && warm) {
            hipGraph_t graph = nullptr;
            ETTK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                ccc_enqueue_chunk(P, W, d_parts, P->geom.p, E, P->out4.p, s);
            } catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            ETTK_HIP(hipStreamEndCapture(s, &graph));
            CccGraph ng;
            ng.E = E;
            ng.parts = d_parts;
            ng.gen = P->ws_gen;
            const hipError_t ie = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            ETTK_HIP(ie);
            P->graphs.push_back(ng);
            cg = &P->graphs.back();
        }
        if (cg) {
            ETTK_HIP(hipGraphLaunch(cg->exec, s));
        } else {
            ccc_enqueue_chunk(P, W, d_parts, P->geom.p, E, P->out4.p, s);
            if (!warm) P->graph_warm.push_back(E);
        }
        ETTK_HIP(hipMemcpyAsync(P->h_out_pinned.p + (size_t)4 * e0, P->out4.p, (size_t)4 * E * sizeof(float),
                                hipMemcpyDeviceToHost, s));
    }
    ETTK_HIP(hipEventRecord(P->ev1, s));
    ETTK_HIP(hipStreamSynchronize(s));
    float ms = 0;
    ETTK_HIP(hipEventElapsedTime(&ms, P->ev0, P->ev1));
    P->timing.ccc_ms += ms;
    P->timing.n_ccc += n;
    const float* o = P->h_out_pinned.p;
    for (int e = 0; e < n; e++) {
        dist[e] = o[(size_t)e * 4];
        t[e] = t[e] * mat_translate(-(double)o[(size_t)e * 4 + 1], -(double)o[(size_t)e * 4 + 2], -(double)o[(size_t)e * 4 + 3]);
    }
}

// ============================================================== constrained CCC batch
EvalGeom make_geom_public(const ettk_plan* P, const ettk_particle& d, const Mat4& t, int particle) {
    EvalGeom g{};
    const int N = P->N;
    Mat4 fin = euler_of(d) * Mat4::from(d.matrix) * t;
    if (mat_det(fin) == 0.0) fin = Mat4::ident();
    const double half = N / 2.0, pos = N / 2.0;
    // input index = fin3 * (q - half - cut z) + fin_t + pos
    double w0[3] = {-half, -half, -half - d.cut_offset};
    for (int r = 0; r < 3; r++) {
        g.md[0 + r] = fin(r, 0);
        g.md[3 + r] = fin(r, 1);
        g.md[6 + r] = fin(r, 2);
        g.md[9 + r] = fin(r, 0) * w0[0] + fin(r, 1) * w0[1] + fin(r, 2) * w0[2] + fin(r, 3) + pos;
        for (int c = 0; c < 4; c++) g.m[3 * c + r] = (float)g.md[3 * c + r];
    }
    WedgeRef w{};
    wedge_normals(d.lwedge, d.uwedge, mat_inverse(fin), w);
    for (int r = 0; r < 3; r++) {
        g.v1[r] = w.v1[r];
        g.v2[r] = w.v2[r];
    }
    g.full = w.full || !P->p.missing_wedge;
    g.particle = particle;
    g.cut = d.cut_offset;
    return g;
}

// getWedgeOverlap (CorrMetric.h:533-566) of the reference against particle descs[i] at transform t[i]: the
// wedges are used whatever the missing-wedge setting; 1 when neither wedge is effective (nbfWedge3D::isEffective,
// Wedge.h:70: limits other than exactly -90 / 90 after the clamp of set())
void wedge_overlaps(ettk_plan* P, const ettk_particle* descs, const Mat4* t, int n, float* out) {
    if (n <= 0) return;
    hipStream_t s = P->stream;
    const int N = P->N;
    auto effective = [](float lw, float uw) {
        const float l = lw < -90 ? -90.0f : lw, u = uw > 90 ? 90.0f : uw;
        return l != -90.0f || u != 90.0f;
    };
    const bool eff1 = P->ref_wedge_kind == 0 ? effective(P->refd.lwedge, P->refd.uwedge) : P->ref_wedge_kind == 1;
    std::vector<EvalGeom> g(n);
    std::vector<int> eff2(n);
    parallel_for(n, [&](int i) {
        const ettk_particle& d = descs[i];
        Mat4 fin = euler_of(d) * Mat4::from(d.matrix) * t[i];
        if (mat_det(fin) == 0.0) fin = Mat4::ident();
        WedgeRef w{};
        wedge_normals(d.lwedge, d.uwedge, mat_inverse(fin), w);
        EvalGeom& e = g[i];
        e = EvalGeom{};
        for (int r = 0; r < 3; r++) {
            e.v1[r] = w.v1[r];
            e.v2[r] = w.v2[r];
        }
        eff2[i] = effective(d.lwedge, d.uwedge);
        e.full = (w.full || !eff2[i]) ? 1 : 0;
    });
    // the plan's buffers (no allocation per call: a free would wait for the whole device)
    P->ov_geom.ensure(n);
    P->ov_part.ensure((size_t)n * 64);
    P->ov_sums.ensure(n);
    ETTK_HIP(hipMemcpyAsync(P->ov_geom.p, g.data(), (size_t)n * sizeof(EvalGeom), hipMemcpyHostToDevice, s));
    launch_wedge_overlap(P->W1.p, P->filter.p, P->ov_geom.p, n, N, P->ov_part.p, P->ov_sums.p, s);
    std::vector<double> h(n);
    ETTK_HIP(hipMemcpyAsync(h.data(), P->ov_sums.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    ETTK_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) out[i] = (eff1 || eff2[i]) ? (float)(h[i] / P->sum_filter) : 1.0f;
}

void ccc_enqueue_chunk(ettk_plan* P, const CccWork& W, const float* d_parts, const EvalGeom* geo, int E, float* out4,
                       hipStream_t s) {
    const int N = P->N, H = P->H;
    const long nh = (long)N * N * H;
    KernelProf* kp = P->ctx->fft.prof;
    const float* W1 = P->p.missing_wedge ? P->W1.p : nullptr;
    const int apply = P->dummy_filter ? 0 : 1;
    const double Rb = (double)E * N * N * N * sizeof(float), Hb = (double)E * nh * sizeof(float2);
    const double Gb = (double)E * P->z.E * N * H * sizeof(float2);
    // particle side: F2 = r2c((reslice - mean) maskfft shift)
    {
        ProfScope ps(kp, "k_reslice", 2 * Rb, s, true);
        launch_reslice(d_parts, geo, E, N, W.R2, W.p_res, P->nb_res, s, ps.start(), ps.stop());
        ps.done();
    }
    launch_reslice_mean(W.p_res, P->nb_res, E, W.stats, s);
    R2CArgs a2{W.R2, W.S2, N, E, PRO_RESLICED, P->maskfft.p, W.stats};
    const bool fused = fft3_fused_supported(N);
    int nb_sp = P->nb_spec;
    if (fused) {
        // F2 = r2c((reslice - mean) maskfft shift) up to its i pass; then on pencils along i: the i pass, both
        // spectra weighted by filter x W1 W2 (X1 = F1mcf, the reference with MCF applied once; X2 = mcf(F2)),
        // X2 kept for the combine, and both inverse i passes of the transforms below
        fft3_r2c_planes(P->ctx->fft, a2, s);
        CccIpassArgs ia{W.S2, W.S3, W.S1, P->F1mcf.p, P->filter.p, W1, geo, P->p.mcf, apply, W.p_n1, W.p_wf, W.p_n2};
        fft3_ccc_ipass(P->ctx->fft, ia, N, E, s);
        nb_sp = fft3_ccc_ipass_blocks(N);
        // reference side: c2r -> r1 m1 (with the k1_m partials) -> r2c = F1m, one partial per (evaluation, plane)
        // (at 64 the forward i passes of both run inside the combine kernel)
        const bool fwd = !ccc_combine_fft_supported(P->z);
        fft3_op_r2c_after_ipass(P->ctx->fft, W.S1, N, E, PLANE_OP_MASK, P->m1.p, W.p_k1, fwd, s);
        // particle side: Q = r2c(c2r(X2)^2 shift)
        fft3_op_r2c_after_ipass(P->ctx->fft, W.S3, N, E, PLANE_OP_SQUARE_SHIFT, nullptr, nullptr, fwd, s);
    } else {
        fft3_r2c(P->ctx->fft, a2, nullptr, s);
        // both spectra weighted by filter x W1 W2 in one pass: X1 = F1mcf (reference, MCF applied once), X2 = mcf(F2)
        {
            ProfScope ps(kp, "k_spec_filter", 4 * Hb, s);
            launch_spec_filter2(P->F1mcf.p, W.S1, W.S2, P->filter.p, W1, geo, E, N, P->p.mcf, apply, W.p_n1, W.p_wf,
                                W.p_n2, P->nb_spec, s);
        }
        fft3_c2r(P->ctx->fft, W.S1, W.S1, W.R1, N, E, s);
        {
            ProfScope ps(kp, "k_mask_ref", 2 * Rb, s);
            launch_mask_ref(W.R1, P->m1.p, E, N, W.p_k1, P->nb_real, s);
        }
        R2CArgs a1{W.R1, W.S1, N, E, PRO_NONE, nullptr, nullptr};
        fft3_r2c(P->ctx->fft, a1, nullptr, s);
        fft3_c2r(P->ctx->fft, W.S2, W.S3, W.R2, N, E, s);
        R2CArgs a3{W.R2, W.S3, N, E, PRO_SQUARE_SHIFT, nullptr, nullptr};
        fft3_r2c(P->ctx->fft, a3, nullptr, s);
    }
    // combine, upsampled correlation on the search window, peak
    launch_ccc_scalars(W.p_n1, W.p_n2, W.p_k1, W.p_wf, nb_sp, fused ? N : P->nb_real, E, N, W.scalars, s);
    const bool cfft = ccc_combine_fft_supported(P->z);
    {
        ProfScope ps(kp, cfft ? "k_combine_fftA" : "k_combine_dftA", 3 * Hb + Gb, s);
        if (cfft)
            launch_combine_fftA(P->ctx->fft, W.S1, W.S2, W.S3, P->Mfft.p, W.scalars, P->z, E, W.G1, fused, s);
        else
            launch_combine_dftA(W.S1, W.S2, W.S3, P->Mfft.p, W.scalars, P->tabA.p, P->z, E, W.G1, s);
    }
    {
        ProfScope ps(kp, cfft ? "k_dftBC_fft" : "k_dftBC_argmax", Gb, s);
        if (cfft)
            launch_dftBC_fft(P->ctx->fft, W.G1, P->tabK.p, P->z, E, W.bval, W.bidx, s);
        else
            launch_dftBC_argmax(W.G1, P->tabA.p, P->tabK.p, P->z, E, W.bval, W.bidx, s);
    }
    launch_peak_final(W.bval, W.bidx, W.scalars, (float)P->sum_filter, (P->p.overlap_norm && P->p.missing_wedge) ? 1 : 0,
                      P->z, E, out4, s);
}

bool ccc_lanes_available(const ettk_plan* P) {
    const KernelProf* prof = P->ctx->fft.prof;
    return !(P->use_graphs && !(prof && prof->on));
}

void ccc_batch_begin(ettk_plan* P, int li, const float* d_parts, const ettk_particle* descs, int n, const int* pidx,
                     Mat4* t, float* dist, const ettk_particle* eval_descs) {
    ETTK_REQUIRE(P->have_ref, ETTK_ERR_STATE, "reference not set");
    ETTK_REQUIRE(li == 0 || li == 1, ETTK_ERR_ARG, "ccc_batch_begin: lane must be 0 or 1");
    CccLane& L = P->lane[li];
    ETTK_REQUIRE(!L.pending, ETTK_ERR_STATE, "ccc_batch_begin: the lane's previous batch was not ended");
    hipStream_t s = P->stream;
    const int CH = std::max(1, P->ccc_chunk);  // evaluations per chunk
    ensure_ccc_workspace(P, std::min(n, CH));
    const int chunk = std::min(P->cap, CH);
    if (!L.ev0) ETTK_HIP(hipEventCreate(&L.ev0));
    if (!L.ev1) ETTK_HIP(hipEventCreate(&L.ev1));
    if (!L.done) ETTK_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    L.t_begin = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
    // geometry of each chunk computed on the host just before its kernels are enqueued, in pinned memory (so the
    // upload does not wait for the earlier chunks' kernels, which run meanwhile); one download of the results at
    // the end
    L.geom_all.ensure(n);
    L.out_all.ensure((size_t)4 * n);
    L.h_geom.ensure(n);
    L.h_out.ensure((size_t)4 * n);
    EvalGeom* g = L.h_geom.p;
    auto geom_chunk = [&](int e0, int E) {
        const int nthr = std::max(1, std::min<int>(E / 64, 4));
        auto geom_of = [&](int e) {
            g[e] = make_geom_public(P, eval_descs ? eval_descs[e] : descs[pidx[e]], t[e], pidx[e]);
        };
        if (nthr <= 1) {
            for (int e = e0; e < e0 + E; e++) geom_of(e);
        } else {
            std::vector<std::thread> th;
            for (int w = 0; w < nthr; w++)
                th.emplace_back([&, w] {
                    for (int e = e0 + w; e < e0 + E; e += nthr) geom_of(e);
                });
            for (auto& x : th) x.join();
        }
        upload_before(P, L.geom_all.p + e0, g + e0, (size_t)E * sizeof(EvalGeom));
    };
    ETTK_HIP(hipEventRecord(L.ev0, s));
    CccWork W{P->R1.p, P->R2.p, P->p_res.p, P->p_n1.p, P->p_n2.p, P->p_wf.p, P->p_k1.p, P->stats.p, P->scalars.p,
              P->bval.p, P->bidx.p, P->S1.p, P->S2.p, P->S3.p, P->G1.p};
    for (int e0 = 0; e0 < n; e0 += chunk) {
        const int E = std::min(chunk, n - e0);
        geom_chunk(e0, E);
        ccc_enqueue_chunk(P, W, d_parts, L.geom_all.p + e0, E, L.out_all.p + (size_t)4 * e0, s);
    }
    ETTK_HIP(hipEventRecord(L.ev1, s));
    ETTK_HIP(hipMemcpyAsync(L.h_out.p, L.out_all.p, (size_t)4 * n * sizeof(float), hipMemcpyDeviceToHost, s));
    ETTK_HIP(hipEventRecord(L.done, s));
    L.pending = true;
    L.n = n;
    L.t = t;
    L.dist = dist;
    P->timing.n_ccc += n;
}

void ccc_batch_end(ettk_plan* P, int li) {
    CccLane& L = P->lane[li];
    if (!L.pending) return;
    L.pending = false;
    ETTK_HIP(hipEventSynchronize(L.done));
    float ms = 0;
    ETTK_HIP(hipEventElapsedTime(&ms, L.ev0, L.ev1));
    P->timing.ccc_ms += ms;
    if (KernelProf* kp = P->ctx->fft.prof) {
        const double now =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
        kp->add_host("ccc_batch_wall", now - L.t_begin);
    }
    const float* o = L.h_out.p;
    for (int e = 0; e < L.n; e++) {
        L.dist[e] = o[(size_t)e * 4];
        const double sx = o[(size_t)e * 4 + 1], sy = o[(size_t)e * 4 + 2], sz = o[(size_t)e * 4 + 3];
        L.t[e] = L.t[e] * mat_translate(-sx, -sy, -sz);
    }
}

void ccc_batch(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n, const int* pidx, Mat4* t,
               float* dist, const ettk_particle* eval_descs) {
    ETTK_REQUIRE(P->have_ref, ETTK_ERR_STATE, "reference not set");
    if (!ccc_lanes_available(P)) {  // chunks replayed from graphs (per-kernel timing needs the eager launches)
        ccc_batch_graph(P, d_parts, descs, n, pidx, t, dist, eval_descs);
        return;
    }
    ccc_batch_end(P, 0);  // a batch still in flight on lane 0 is finished first
    ccc_batch_begin(P, 0, d_parts, descs, n, pidx, t, dist, eval_descs);
    ccc_batch_end(P, 0);
}

}  // namespace ettk

// ============================================================== C ABI
using namespace ettk;

#define ETTK_API_BEGIN try {
#define ETTK_API_END                                     \
    }                                                    \
    catch (const ettk::Error& e) {                       \
        set_error(e.msg);                                \
        return e.code;                                   \
    }                                                    \
    catch (const std::exception& e) {                    \
        set_error(e.what());                             \
        return ETTK_ERR_ALLOC;                           \
    }                                                    \
    set_error("");                                       \
    return ETTK_OK;

// frees a plan and everything it holds (also a partly built one)
static void plan_release(ettk_plan* P) {
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    if (P->copy_stream) (void)hipStreamSynchronize(P->copy_stream);
    if (P->aux_stream) (void)hipStreamSynchronize(P->aux_stream);
    if (P->up_stream) (void)hipStreamSynchronize(P->up_stream);
    for (auto& g : P->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    P->graphs.clear();
    so3_state_destroy(P);
    if (P->ev0) (void)hipEventDestroy(P->ev0);
    if (P->ev1) (void)hipEventDestroy(P->ev1);
    if (P->ev_up) (void)hipEventDestroy(P->ev_up);
    hipStream_t s = P->stream, cs = P->copy_stream, as = P->aux_stream, us = P->up_stream;
    delete P;  // device buffers are released before their streams
    if (s) (void)hipStreamDestroy(s);
    if (cs) (void)hipStreamDestroy(cs);
    if (as) (void)hipStreamDestroy(as);
    if (us) (void)hipStreamDestroy(us);
}

extern "C" {

const char* ettk_last_error(void) { return g_err.c_str(); }

int ettk_fft3(ettk_ctx* ctx, int n, int batch, int kind, const void* in, void* out, const float* aux) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(ctx && in && out && batch > 0, ETTK_ERR_ARG, "ettk_fft3: bad argument");
    ETTK_REQUIRE(fft_len_index(n) >= 0, ETTK_ERR_UNSUPPORTED, "ettk_fft3: unsupported size " + std::to_string(n));
    ETTK_REQUIRE(kind >= ETTK_FFT_R2C && kind <= ETTK_FFT_SQUARE, ETTK_ERR_ARG, "ettk_fft3: unknown kind");
    ETTK_REQUIRE(kind < ETTK_FFT_MASK || fft3_fused_supported(n), ETTK_ERR_UNSUPPORTED,
                 "ettk_fft3: the fused transforms need n <= 64");
    ETTK_REQUIRE(kind != ETTK_FFT_MASK || aux, ETTK_ERR_ARG, "ettk_fft3: ETTK_FFT_MASK needs the mask");
    ETTK_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const size_t nh = (size_t)n * n * (n / 2 + 1);
    DevBuf<float2> tmp;
    DevBuf<float> part;
    switch (kind) {
        case ETTK_FFT_R2C: {
            R2CArgs a{static_cast<const float*>(in), static_cast<float2*>(out), n, batch, PRO_NONE, nullptr, nullptr};
            fft3_r2c(ctx->fft, a, nullptr, s);
            break;
        }
        case ETTK_FFT_C2R:
            tmp.ensure(nh * batch);
            fft3_c2r(ctx->fft, static_cast<const float2*>(in), tmp.p, static_cast<float*>(out), n, batch, s);
            break;
        case ETTK_FFT_MASK:
            part.ensure((size_t)n * batch);
            fft3_c2r_op_r2c(ctx->fft, static_cast<const float2*>(in), static_cast<float2*>(out), n, batch, PLANE_OP_MASK,
                            aux, part.p, s);
            break;
        case ETTK_FFT_SQUARE:
            fft3_c2r_op_r2c(ctx->fft, static_cast<const float2*>(in), static_cast<float2*>(out), n, batch,
                            PLANE_OP_SQUARE_SHIFT, nullptr, nullptr, s);
            break;
    }
    ETTK_HIP(hipStreamSynchronize(s));
    ETTK_API_END
}

int ettk_ctx_create(int device, ettk_ctx** out) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(out != nullptr, ETTK_ERR_ARG, "ettk_ctx_create: null output");
    int n = 0;
    ETTK_HIP(hipGetDeviceCount(&n));
    ETTK_REQUIRE(device >= 0 && device < n, ETTK_ERR_ARG, "ettk_ctx_create: no such HIP device");
    ETTK_HIP(hipSetDevice(device));
    auto* c = new ettk_ctx();
    c->device = device;
    ETTK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    ETTK_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    try {
        fft_tables_init(c->fft, c->stream);
        ETTK_HIP(hipStreamSynchronize(c->stream));  // the tables are read from every plan's stream
        noalloc_scratch_init(c);
    } catch (...) {
        (void)hipStreamDestroy(c->stream);
        (void)hipStreamDestroy(c->copy_stream);
        delete c;
        throw;
    }
    *out = c;
    ETTK_API_END
}

int ettk_ctx_destroy(ettk_ctx* ctx) {
    ETTK_API_BEGIN
    if (ctx) {
        ctx_plan_cache_destroy(ctx);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->fft.tw.release();
        if (ctx->fft.prof) {
            ctx->fft.prof->reset();
            for (hipEvent_t e : ctx->fft.prof->pool) (void)hipEventDestroy(e);
            delete ctx->fft.prof;
        }
        (void)hipStreamDestroy(ctx->stream);
        if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
        delete ctx;
    }
    ETTK_API_END
}

int ettk_plan_create(ettk_ctx* ctx, const ettk_params* params, ettk_plan** out) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(ctx && params && out, ETTK_ERR_ARG, "ettk_plan_create: null argument");
    const ettk_params& p = *params;
    ETTK_REQUIRE(p.N >= 16 && p.N % 2 == 0 && fft_len_index(p.N) >= 0 && fft_len_index(2 * p.N) >= 0, ETTK_ERR_UNSUPPORTED,
                 "ettk_plan_create: box size must be one of 16, 32, 48, 64, 96, 128");
    ETTK_REQUIRE(p.B >= 4 && fft_len_index(2 * p.B) >= 0, ETTK_ERR_UNSUPPORTED,
                 "ettk_plan_create: bandwidth must be one of 8, 16, 24, 32, 48, 64, 96, 128");
    ETTK_REQUIRE(p.n_candidates >= 0 && p.n_candidates <= ETTK_MAX_CANDIDATES, ETTK_ERR_ARG,
                 "ettk_plan_create: n_candidates out of range");
    ETTK_HIP(hipSetDevice(ctx->device));
    auto* P = new ettk_plan();
    try {
        P->ctx = ctx;
        P->p = p;
        P->N = p.N;
        P->H = p.N / 2 + 1;
        P->trans_restrict = p.trans_restrict < 0 ? (double)FLT_MAX : (double)p.trans_restrict;
        P->rot_restrict = std::max((double)p.rot_restrict, 1.0);
        ETTK_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
        ETTK_HIP(hipStreamCreateWithFlags(&P->copy_stream, hipStreamNonBlocking));
        ETTK_HIP(hipStreamCreateWithFlags(&P->aux_stream, hipStreamNonBlocking));
        ETTK_HIP(hipStreamCreateWithFlags(&P->up_stream, hipStreamNonBlocking));
        ETTK_HIP(hipEventCreateWithFlags(&P->ev_up, hipEventDisableTiming));
        ETTK_HIP(hipEventCreate(&P->ev0));
        ETTK_HIP(hipEventCreate(&P->ev1));
        plan_constants(P);
        // evaluations per CCC chunk: 256, or as many as about 4 GB of workspace holds (multiples of 8), unless the
        // tuning cache holds a measured choice for this shape
        const size_t per_eval = ccc_work_bytes(P, 1);
        P->ccc_chunk = (int)std::max<size_t>(8, std::min<size_t>(256, (size_t)4e9 / per_eval)) & ~7;
        tune_lookup(P);
    } catch (...) {
        plan_release(P);
        throw;
    }
    *out = P;
    ETTK_API_END
}

int ettk_plan_destroy(ettk_plan* P) {
    ETTK_API_BEGIN
    if (P) plan_release(P);
    ETTK_API_END
}

int ettk_plan_set_restrictions(ettk_plan* P, float rot_restrict, float trans_restrict) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(P, ETTK_ERR_ARG, "ettk_plan_set_restrictions: null plan");
    ETTK_HIP(hipSetDevice(P->ctx->device));
    ETTK_HIP(hipStreamSynchronize(P->stream));
    const ettk_params saved = P->p;
    const double tr0 = P->trans_restrict, rot0 = P->rot_restrict;
    const CccSizes z0 = P->z;
    P->p.rot_restrict = rot_restrict;
    P->p.trans_restrict = trans_restrict;
    P->trans_restrict = trans_restrict < 0 ? (double)FLT_MAX : (double)trans_restrict;
    P->rot_restrict = std::max((double)rot_restrict, 1.0);
    ccc_graphs_reset(P);
    try {
        search_window(P);
    } catch (...) {
        P->p = saved;
        P->trans_restrict = tr0;
        P->rot_restrict = rot0;
        P->z = z0;
        search_window(P);
        throw;
    }
    ETTK_API_END
}

}  // extern "C"

// input1 of the metric (setInput1, nbfImageMetric.h:184-189): the volume, its wedge W1, the cached reference
// side of the CCC (FIM.h:1123-1153) and of the search (ProjRot.h:445-453).  wedge_kind 0: W1 is the analytic
// wedge of `desc` at its own alignment (FIM.h:1101-1105); 1: the accumulated wedge of an average
// (P->wsm_half, WAvg.h:555-611); 2: an average whose accumulated wedge is not effective (W1 = 1).
static void reference_load(ettk_plan* P, const float* d_ref, const ettk_particle& desc, int wedge_kind) {
    ETTK_HIP(hipSetDevice(P->ctx->device));
    const int N = P->N, H = P->H;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * H;
    hipStream_t s = P->stream;
    P->have_ref = false;
    ccc_graphs_reset(P);
    P->refd = desc;
    P->ref_wedge_kind = wedge_kind;
    P->ref.ensure(n3);
    launch_xfast_to_internal(d_ref, P->ref.p, N, s);
    P->W1.ensure(nh);
    WedgeRef w{};
    if (wedge_kind == 0) {
        Mat4 fin = euler_of(desc) * Mat4::from(desc.matrix);
        wedge_normals(desc.lwedge, desc.uwedge, mat_inverse(fin), w);
        launch_wedge_ref(w, N, P->W1.p, s);
    } else {
        w.full = 1;
        if (wedge_kind == 1) {
            ETTK_HIP(hipMemcpyAsync(P->W1.p, P->wsm_half.p, nh * sizeof(float), hipMemcpyDeviceToDevice, s));
            P->ref_wbin.ensure(n3);
            ETTK_HIP(hipMemcpyAsync(P->ref_wbin.p, P->wbin.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
        } else {
            launch_wedge_ref(w, N, P->W1.p, s);
        }
    }
    P->wref = w;
    // F1 = r2c(I1 * maskfft * shift), MCF applied once (executeHalf order: MCF, then filter x wedge)
    P->F1mcf.ensure(nh);
    R2CArgs a{P->ref.p, P->F1mcf.p, N, 1, PRO_MASK_SHIFT, P->maskfft.p, nullptr};
    fft3_r2c(P->ctx->fft, a, nullptr, s);
    if (P->p.mcf) launch_mcf_inplace(P->F1mcf.p, (long)nh, s);
    // continuous refinement: I1 - mean(I1), x fastest (set-up path: one host round trip)
    std::vector<float> h(n3);
    ETTK_HIP(hipMemcpyAsync(h.data(), d_ref, n3 * sizeof(float), hipMemcpyDeviceToHost, s));
    ETTK_HIP(hipStreamSynchronize(s));
    double mean = 0;
    for (float v : h) mean += v;
    mean /= (double)n3;
    for (float& v : h) v = (float)((double)v - mean);
    P->ref_x.upload(h.data(), n3, s);
    ETTK_HIP(hipStreamSynchronize(s));
    P->have_ref = true;
    so3_reference_set(P);
}

extern "C" {

int ettk_reference_set(ettk_plan* P, const float* d_ref, const ettk_particle* ref_desc) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(P && d_ref && ref_desc, ETTK_ERR_ARG, "ettk_reference_set: null argument");
    reference_load(P, d_ref, *ref_desc, 0);
    ETTK_API_END
}

// setInput1 with an nbfWedgedAverageImage3D: its wedge is the accumulated wedge of the particles it averages
int ettk_reference_set_average(ettk_plan* P, const float* d_ref, int n, const ettk_particle* descs,
                               const float* weights, const double* matrices) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(P && d_ref && descs && weights && matrices && n > 0, ETTK_ERR_ARG,
                 "ettk_reference_set_average: bad argument");
    avg_wedge_compute(P, n, descs, weights, matrices);
    ettk_particle full{};
    full.lwedge = -90.0f;
    full.uwedge = 90.0f;
    for (int k = 0; k < 16; k++) full.matrix[k] = (k % 5 == 0) ? 1.0 : 0.0;
    reference_load(P, d_ref, full, P->wacc_effective ? 1 : 2);
    ETTK_API_END
}

int ettk_ccc_eval_batch(ettk_plan* P, const float* d_particles, int n_particles, const ettk_particle* descs,
                        int n_evals, const int* particle_index, const double* transforms, float* out_distance,
                        double* out_transforms) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(P && d_particles && descs && particle_index && transforms && out_distance && out_transforms,
                 ETTK_ERR_ARG, "ettk_ccc_eval_batch: null argument");
    ETTK_REQUIRE(n_evals >= 0 && n_particles > 0, ETTK_ERR_ARG, "ettk_ccc_eval_batch: bad sizes");
    for (int e = 0; e < n_evals; e++)
        ETTK_REQUIRE(particle_index[e] >= 0 && particle_index[e] < n_particles, ETTK_ERR_ARG,
                     "ettk_ccc_eval_batch: particle index out of range");
    ETTK_HIP(hipSetDevice(P->ctx->device));
    std::vector<Mat4> t(n_evals);
    for (int e = 0; e < n_evals; e++) t[e] = Mat4::from(transforms + (size_t)e * 16);
    P->timing = Timing{};
    ccc_batch(P, d_particles, descs, n_evals, particle_index, t.data(), out_distance);
    for (int e = 0; e < n_evals; e++) std::memcpy(out_transforms + (size_t)e * 16, t[e].a, sizeof(t[e].a));
    ETTK_API_END
}

int ettk_profile(ettk_ctx* ctx, int enable) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(ctx, ETTK_ERR_ARG, "ettk_profile: null context");
    ETTK_HIP(hipStreamSynchronize(ctx->stream));
    if (!ctx->fft.prof) ctx->fft.prof = new KernelProf();
    ctx->fft.prof->reset();
    ctx->fft.prof->on = enable != 0;
    ETTK_API_END
}

int ettk_profile_only(ettk_ctx* ctx, const char* kernel) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(ctx, ETTK_ERR_ARG, "ettk_profile_only: null context");
    if (!ctx->fft.prof) ctx->fft.prof = new KernelProf();
    ctx->fft.prof->only = kernel ? kernel : "";
    ETTK_API_END
}

int ettk_profile_query(ettk_ctx* ctx, const char* kernel, double* total_ms, long long* launches, double* bytes) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(ctx && kernel, ETTK_ERR_ARG, "ettk_profile_query: null argument");
    ETTK_HIP(hipStreamSynchronize(ctx->stream));
    double ms = 0, by = 0;
    long long nl = 0;
    if (ctx->fft.prof) ctx->fft.prof->query(kernel, ms, nl, by);
    if (total_ms) *total_ms = ms;
    if (launches) *launches = nl;
    if (bytes) *bytes = by;
    ETTK_API_END
}

int ettk_last_timing(ettk_plan* P, double* ccc_ms, long long* n_ccc, double* total_ms) {
    ETTK_API_BEGIN
    ETTK_REQUIRE(P, ETTK_ERR_ARG, "ettk_last_timing: null plan");
    if (ccc_ms) *ccc_ms = P->timing.ccc_ms;
    if (n_ccc) *n_ccc = P->timing.n_ccc;
    if (total_ms) *total_ms = P->timing.total_ms;
    ETTK_API_END
}

}  // extern "C"
