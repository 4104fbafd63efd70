// align.cpp -- host driver of the pairwise alignment (nbfProjectionRotationMetric3D::execute) on top of the
// GPU stages: S^2 projections and SO(3) correlations in batches, candidate extraction, the direct/mirror
// CCC test of the candidate peaks and the Powell refinement of the best candidates.
//
// Modes: 0 = plain (particle after particle, one CCC evaluation per launch, the reference's control flow);
//        1 = batched (every stage over all particles; the Powell refinements of all particles run as coroutines
//            on the calling thread, and every round of their pending CCC requests is served by one batched launch);
//        2 = batched + adjoint refinement (refine.cpp).
#include <algorithm>
#include <cstring>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <coroutine>
#include <exception>
#include <functional>
#include <map>
#include <mutex>
#include <queue>
#include <future>
#include <memory>
#include <thread>

#include "so3.hpp"

namespace ettk {

// ------------------------------------------------------------------ spherical wedge (host, FP64)
// nbfWedgedSubImage3D::getSphericalWedgeImage -> nbfWedge3D::getSphericalImage (Wedge.h:364-429): the binary
// wedge on the grid theta_i = i pi/(2B-1), phi_j = j 2pi/(2B-1) (Pixel arithmetic), all ones without a wedge
// grid directions (cos theta_j sin phi_i, sin theta_j sin phi_i, cos phi_i), phi_i = i pi/(2B-1),
// theta_j = j 2pi/(2B-1), in the Pixel (float) arithmetic of nbfWedge3D::getSphericalImage; one table per B
static const std::vector<double>& sph_dirs(int B) {
    static std::mutex mu;
    static std::map<int, std::vector<double>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(B);
    if (it != cache.end()) return it->second;
    const int n = 2 * B;
    std::vector<double> t((size_t)3 * n * n);
    float phiK = (float)(M_PI / (n - 1)), thetaK = (float)(2.0 * M_PI / (n - 1));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            float si = std::sin((float)(i * phiK)), ci = std::cos((float)(i * phiK));
            float sj = std::sin((float)(j * thetaK)), cj = std::cos((float)(j * thetaK));
            double* d3 = &t[3 * ((size_t)i * n + j)];
            d3[0] = (double)(cj * si);
            d3[1] = (double)(sj * si);
            d3[2] = (double)ci;
        }
    return cache.emplace(B, std::move(t)).first->second;
}

std::vector<float> spherical_wedge_bin(int B, const ettk_particle& d) {
    const int n = 2 * B;
    std::vector<float> W((size_t)n * n, 1.0f);
    double l = d.lwedge < -90 ? -90.0 : d.lwedge, u = d.uwedge > 90 ? 90.0 : d.uwedge;
    if (u - l < 180) {
        Mat4 R = strip_translation(mat_inverse(euler_of(d) * Mat4::from(d.matrix)));
        double n1[3] = {std::cos(-(90.0 - u) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - u) * M_PI / 180.0)};
        double n2[3] = {std::cos(-(90.0 - l) * M_PI / 180.0), 0.0, -std::sin(-(90.0 - l) * M_PI / 180.0)};
        double v1[3], v2[3];
        for (int r = 0; r < 3; r++) {
            v1[r] = R(r, 0) * n1[0] + R(r, 1) * n1[1] + R(r, 2) * n1[2];
            v2[r] = R(r, 0) * n2[0] + R(r, 1) * n2[1] + R(r, 2) * n2[2];
        }
        const std::vector<double>& dir = sph_dirs(B);
        for (size_t o = 0; o < (size_t)n * n; o++) {
            const double* d3 = &dir[3 * o];
            double a = v1[0] * d3[0] + v1[1] * d3[1] + v1[2] * d3[2];
            double b = v2[0] * d3[0] + v2[1] * d3[1] + v2[2] * d3[2];
            W[o] = (a * b <= 0.0) ? 1.0f : 0.0f;
        }
    }
    return W;
}

// the window of the search: the binary wedge conditioned by conditionSphericalWindow (ProjRot.h:1322-1345:
// Gaussian sigma 2, radius factor 1) when missing-wedge compensation is on, all ones otherwise
std::vector<float> spherical_wedge_host(int B, const ettk_particle& d, bool wedge_on) {
    if (!wedge_on) return std::vector<float>((size_t)4 * B * B, 1.0f);
    std::vector<float> W = spherical_wedge_bin(B, d);
    smooth3(W, 2 * B, 2 * B, 1, 2.0, 1.0);
    return W;
}

// ------------------------------------------------------------------ candidates (ProjRot.h:1593-1702)
struct Cand {
    float a, b, g, v;
    int k = -1;  // peak_rounds: index of this candidate's direct evaluation among the particle's evaluations
};
constexpr float kEquivTol = 0.25f;  // L1 distance of equivalent rotation matrices in the search (ProjRot.h:142, 1633)
struct CandCmp {
    bool operator()(const Cand& x, const Cand& y) const { return x.v > y.v; }  // greaterPD: min-heap on [3]
};
using CandQueue = std::priority_queue<Cand, std::vector<Cand>, CandCmp>;

// The reference compares every pair (i, j) in index order and drops the worse of two rotations whose 4x4
// matrices are within `tol` in L1 norm.  A pair within tol differs by less than tol in every matrix element, so
// its third columns fall in the same or adjacent cells of a grid of pitch tol: only those pairs are visited,
// in the same (i, then increasing j) order, which leaves every decision of the exhaustive loop unchanged.
static void eliminate_equivalent(std::vector<Cand>& pos, const std::vector<Mat4>& mats, CandQueue& pq, float tol) {
    const float NEG = -FLT_MAX;
    const size_t n = pos.size();
    auto close = [&](size_t i, size_t j) {  // the partial sums never decrease: stop once they reach tol
        double err = 0;
        for (int m = 0; m < 16; m++) {
            err += std::fabs(mats[i].a[m] - mats[j].a[m]);
            if (!(err < tol)) return false;
        }
        return true;
    };
    auto visit = [&](size_t i, size_t j) {
        if (i != j && pos[j].v > NEG && close(i, j)) {
            if (pos[i].v > pos[j].v) pos[i].v = NEG;
            else pos[j].v = NEG;
        }
    };
    if (n < 64) {
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < n; j++) {
                if (pos[i].v == NEG) break;
                visit(i, j);
            }
    } else {
        // matrix elements of a rotation lie in [-1, 1]: a dense grid of side G covers every cell (+1 margin)
        const int G = (int)std::ceil(2.0 / tol) + 3;
        auto cell = [&](size_t i, int c) {
            long v = (long)std::floor(mats[i].a[c] / (double)tol) + G / 2;
            return (int)std::min<long>(std::max<long>(v, 1), G - 2);
        };
        // the points sorted by cell (counting sort, increasing index within a cell), their matrices packed in that
        // order: the three cells adjacent along the last key axis are one contiguous range
        const size_t G3 = (size_t)G * G * G;
        std::vector<int> key(n), start(G3 + 1, 0), order(n);
        for (size_t i = 0; i < n; i++) {
            key[i] = (cell(i, 2) * G + cell(i, 6)) * G + cell(i, 10);
            start[key[i] + 1]++;
        }
        for (size_t k = 0; k < G3; k++) start[k + 1] += start[k];
        std::vector<double> pm(16 * n);
        {
            std::vector<int> fill(start.begin(), start.end() - 1);
            for (size_t i = 0; i < n; i++) {
                const int s = fill[key[i]]++;
                order[s] = (int)i;
                std::memcpy(&pm[(size_t)s * 16], mats[i].a, 16 * sizeof(double));
            }
        }
        // closeness does not depend on the visiting order: the close partners of i are found first, then visited
        // in increasing index order as the exhaustive loop would (a partner already dropped is skipped there)
        std::vector<size_t> js;
        for (size_t i = 0; i < n; i++) {
            if (pos[i].v == NEG) continue;
            js.clear();
            const double* mi = mats[i].a;
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++) {
                    const int k0 = key[i] + (a * G + b) * G - 1;
                    for (int sj = start[k0]; sj < start[k0 + 3]; sj++) {
                        const int j = order[sj];
                        if ((size_t)j == i || !(pos[j].v > NEG)) continue;
                        const double* mj = &pm[(size_t)sj * 16];
                        // the partial sum over the first three elements (the same additions in the same order)
                        // already rejects most neighbours, with a branch that is nearly always taken
                        const double d3 = std::fabs(mi[0] - mj[0]) + std::fabs(mi[1] - mj[1]) + std::fabs(mi[2] - mj[2]);
                        if (!(d3 < tol)) continue;
                        double err = 0;
                        int m = 0;
                        for (; m < 16; m++) {
                            err += std::fabs(mi[m] - mj[m]);
                            if (!(err < tol)) break;
                        }
                        if (m == 16) js.push_back((size_t)j);
                    }
                }
            std::sort(js.begin(), js.end());
            for (size_t j : js) {
                if (pos[i].v == NEG) break;
                if (pos[j].v > NEG) {
                    if (pos[i].v > pos[j].v) pos[i].v = NEG;
                    else pos[j].v = NEG;
                }
            }
        }
    }
    for (auto& p : pos)
        if (p.v > NEG) pq.push(p);
}

static void eliminate_equivalent(std::vector<Cand>& pos, CandQueue& pq, float tol) {
    std::vector<Mat4> mats(pos.size());
    for (size_t i = 0; i < pos.size(); i++) mats[i] = zyz(pos[i].a, pos[i].b, pos[i].g);
    eliminate_equivalent(pos, mats, pq, tol);
}

// the Euler angles of the (2B)^3 correlation grid (ProjRot.h:1620-1640) and the three rotations zyz multiplies for
// each of their 2B values: zyz(alpha, beta, gamma) = (rg[g] rb[b]) ra[a], the same products as computed directly
struct ZyzTables {
    std::vector<float> alpha, beta, gamma;
    std::vector<Mat4> ra, rb, rg;
};
static const ZyzTables& zyz_tables(int B) {
    static std::mutex mu;
    static std::map<int, ZyzTables> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(B);
    if (it != cache.end()) return it->second;
    const int n = 2 * B;
    ZyzTables z;
    for (int i = 0; i < n; i++) {
        const float alpha = (float)(i / (double)B * 180.0);
        float beta = (float)((2 * i + 1.0) / (4. * B) * 180.0);
        if ((float)B <= 32) beta = (float)((2 * i + 0.0) / (4. * B) * 180.0);
        const float gamma = (float)(i / (double)B * 180.0);
        z.alpha.push_back(alpha);
        z.beta.push_back(beta);
        z.gamma.push_back(gamma);
        z.ra.push_back(mat_rot(-(double)alpha, 0, 0, 1));
        z.rb.push_back(mat_rot(-(double)beta, 0, 1, 0));
        z.rg.push_back(mat_rot(-(double)gamma, 0, 0, 1));
    }
    return cache.emplace(B, std::move(z)).first->second;
}

// The elimination of equivalent angles on the device's output (k_lm_candidates: the valid maxima in index order,
// deg[i] partners j > i of each, their indices row after row in adj): the rows in order, each row's partners in
// increasing j, the worse of two alive partners dropped (j on a tie) -- the decisions of eliminate_equivalent, whose
// visits (i, j) with j < i find one end dead -- then the survivors pushed in index order as there.
static void candidates_select(const ettk_plan* P, const std::vector<int>& idx, const std::vector<double>& val,
                              const std::vector<int>& deg, const std::vector<int>& adj, CandQueue& pq) {
    const int B = P->p.B, n = 2 * B;
    const ZyzTables& Z = zyz_tables(B);
    const size_t nv = idx.size();
    std::vector<float> v(nv);
    std::vector<char> alive(nv, 1);
    for (size_t k = 0; k < nv; k++) v[k] = (float)(-val[k]);
    size_t o = 0;
    for (size_t i = 0; i < nv; i++) {
        const size_t d = (size_t)deg[i];
        for (size_t k = 0; k < d && alive[i]; k++) {
            const int j = adj[o + k];
            if (alive[j]) {
                if (v[i] > v[j]) alive[i] = 0;
                else alive[j] = 0;
            }
        }
        o += d;
    }
    for (size_t k = 0; k < nv; k++) {
        if (!alive[k]) continue;
        const int t = idx[k], a = t / (n * n), b = (t / n) % n, g = t % n;
        pq.push(Cand{Z.alpha[a], Z.beta[b], Z.gamma[g], v[k]});
    }
}

static void candidates_from_maxima(const ettk_plan* P, std::vector<std::pair<int, double>>& lm, CandQueue& pq) {
    const int B = P->p.B, n = 2 * B;
    const ZyzTables& Z = zyz_tables(B);
    std::sort(lm.begin(), lm.end());  // blitz iteration order (alpha slowest)
    std::vector<Cand> pos;
    std::vector<Mat4> mats;
    for (auto& e : lm) {
        int t = e.first;
        int a = t / (n * n), b = (t / n) % n, g = t % n;
        const Mat4 m = Z.rg[g] * Z.rb[b] * Z.ra[a];
        if (transform_ok(m, P->trans_restrict, P->rot_restrict)) {
            pos.push_back(Cand{Z.alpha[a], Z.beta[b], Z.gamma[g], (float)(-e.second)});
            mats.push_back(m);
        }
    }
    eliminate_equivalent(pos, mats, pq, kEquivTol);
}

struct SearchOut {
    CandQueue pq1, pq2;
    bool two = false;
    int use_window = 0;
};

// Host-side inputs of the search, per chunk of particles (computed on the host threads while the previous
// chunk runs on the GPU): smoothed spherical wedge W2 (nbfWedge3D::getSphericalImage +
// conditionSphericalWindow), whether it is "effective" (sum <= 0.95 size, ProjRot.h:526-540), and the reslice
// geometry of the projection.  The arrays are sized for the whole call (W2 and geom in the plan's pinned buffers,
// uploaded asynchronously) and stay unchanged until it ends.
struct SearchInputs {
    float* W2 = nullptr;
    EvalGeom* geom = nullptr;
    std::vector<int> uw2;
};

// sizes the inputs of a call with n particles and every device buffer of its chunks (up to `chunk` particles)
static void search_inputs_init(ettk_plan* P, int n, int chunk, SearchInputs& in) {
    So3State& S = so3_state(P);
    const size_t nn = (size_t)4 * P->p.B * P->p.B;
    S.hW2.ensure((size_t)n * nn);
    S.hGeom.ensure(n);
    in.W2 = S.hW2.p;
    in.geom = S.hGeom.p;
    in.uw2.assign(n, 0);
    so3_search_reserve(P, chunk);
}

static void search_inputs(const ettk_plan* P, const ettk_particle* descs, int n, int p0, int cn, SearchInputs& in) {
    const int B = P->p.B, nn = 4 * B * B;
    (void)n;
    auto work = [&](int i) {
        std::vector<float> w = spherical_wedge_host(B, descs[i], P->p.missing_wedge != 0);
        double sw = 0;
        for (float v : w) sw += v;
        in.uw2[i] = P->p.missing_wedge && (sw <= 0.95f * (float)nn);
        std::copy(w.begin(), w.end(), in.W2 + (size_t)i * nn);
        in.geom[i] = make_geom_public(P, descs[i], Mat4::ident(), i);
    };
    const int nthr = std::max(1, std::min<int>(cn / 4, host_threads() / 2));
    if (nthr <= 1) {
        for (int i = p0; i < p0 + cn; i++) work(i);
        return;
    }
    std::vector<std::thread> th;
    for (int w = 0; w < nthr; w++)
        th.emplace_back([&, w] {
            for (int i = p0 + w; i < p0 + cn; i += nthr) work(i);
        });
    for (auto& t : th) t.join();
}

// One chunk of the search in flight: the host sources of its uploads (kept alive until the chunk is done),
// its local-maximum buffers on the device (a slot of the plan's pool) and, on the host, the downloaded lists.
struct SearchLists {
    int p0 = 0, cn = 0;
    bool uw1 = false;
    std::vector<int> uw2;
    std::vector<SoftInst> inst;
    std::vector<ComboOp> ops;
    LmBuffers* lm = nullptr;
    hipEvent_t maxima = nullptr, done = nullptr;  // local maxima found (main stream); lists packed (aux stream)
    ~SearchLists() {
        if (maxima) (void)hipEventDestroy(maxima);
        if (done) (void)hipEventDestroy(done);
    }
    std::chrono::steady_clock::time_point t_ready;  // when its kernels had completed
};
// the first beta bin the rotation restriction removes from the correlation volumes (they hold -FLT_MAX from it on,
// ProjRot.h:601-602, 631-632)
static int restrict_b0(const ettk_plan* P) {
    const int B = P->p.B, n = 2 * B;
    const double f = std::floor(P->rot_restrict * B / 90.0 - 1.0);
    return (int)std::max(1.0, std::min(f, (double)(n - 1)));
}
// local maxima kept per correlation volume ((2B)^3 voxels): 16384 up to B = 32, four times that per doubling, plus
// the restricted beta bins, every voxel of which is a maximum of the >= stencil (a plateau at -FLT_MAX)
static int lm_cap(const ettk_plan* P) {
    const int B = P->p.B, n = 2 * B;
    const long base = B <= 32 ? 16384 : (B <= 64 ? 65536 : 262144);
    return (int)std::min((long)n * n * n, base + (long)n * n * (n - restrict_b0(P)));
}
static constexpr int kSearchSlots = 4;  // search chunks in flight (device list buffers and host threads)

// GPU part of the search for particles [p0, p0+cn): projections, SO(3) correlations, local maxima.  Nothing
// waits here: L.done marks the end of the chunk's kernels, the next chunk is enqueued right behind it.
static void search_gpu(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int p0, int cn,
                       const SearchInputs& in, int slot, SearchLists& L) {
    So3State& S = so3_state(P);
    const int B = P->p.B, n = 2 * B, nn = n * n;
    const long n3 = (long)n * n * n;
    hipStream_t s = P->stream;
    KernelProf* kp = P->ctx->fft.prof;
    L.p0 = p0;
    L.cn = cn;
    L.uw2.assign(in.uw2.begin() + p0, in.uw2.begin() + p0 + cn);
    L.uw1 = S.use_window1;
    L.lm = &lm_slot(S, slot, 2 * cn, lm_cap(P));
    S.Pp2.ensure((size_t)cn * nn);
    S.Mx.ensure(cn);
    project_images(P, d_parts, descs, nullptr, cn, false, S.Pp2.p, in.geom + p0);
    launch_ppmax(S.Pp1.p, S.Pp2.p, cn, B, S.Mx.p, s);
    S.W2.ensure((size_t)cn * nn);
    ETTK_HIP(hipMemcpyAsync(S.W2.p, in.W2 + (size_t)p0 * nn, (size_t)cn * nn * sizeof(float),
                            hipMemcpyHostToDevice, s));
    const bool uw1 = L.uw1;
    std::vector<SoftInst>& inst = L.inst;
    std::vector<ComboOp>& ops = L.ops;
    inst.clear();
    ops.assign(2 * cn, ComboOp{CO_COPY, {0, 0, 0, 0}});
    for (int q = 0; q < cn; q++) {
        int base = (int)inst.size();
        if (uw1 && L.uw2[q]) {
            inst.push_back({q, RC_P1SQ_W1, RC_W2, 1});
            inst.push_back({q, RC_W1, RC_P2SQ_W2, 1});
            inst.push_back({q, RC_P1_W1, RC_P2_W2, 1});
            inst.push_back({q, RC_W1, RC_W2, 1});
            inst.push_back({q, RC_P1, RC_P2, 0});
            ops[2 * q] = ComboOp{CO_BOTH, {base, base + 1, base + 2, base + 3}};
            ops[2 * q + 1] = ComboOp{CO_COPY, {base + 4, 0, 0, 0}};
        } else if (uw1) {
            inst.push_back({q, RC_W1, RC_P2SQ, 1});
            inst.push_back({q, RC_P1_W1, RC_P2, 1});
            inst.push_back({q, RC_P1, RC_P2, 0});
            ops[2 * q] = ComboOp{CO_DIV_SQRT, {base + 1, base, 0, 0}};
            ops[2 * q + 1] = ComboOp{CO_COPY, {base + 2, 0, 0, 0}};
        } else if (L.uw2[q]) {
            inst.push_back({q, RC_P1SQ, RC_W2, 1});
            inst.push_back({q, RC_P1, RC_P2_W2, 1});
            inst.push_back({q, RC_P1, RC_P2, 0});
            ops[2 * q] = ComboOp{CO_DIV_SQRT, {base + 1, base, 0, 0}};
            ops[2 * q + 1] = ComboOp{CO_COPY, {base + 2, 0, 0, 0}};
        } else {
            inst.push_back({q, RC_P1, RC_P2, 0});
            ops[2 * q] = ComboOp{CO_COPY, {base, 0, 0, 0}};
            ops[2 * q + 1] = ComboOp{CO_COPY, {base, 0, 0, 0}};
        }
    }
    const int nq = (int)inst.size();
    S.C.ensure((size_t)nq * n3);
    S.Cc.ensure((size_t)2 * cn * n3);
    L.lm->hinst.ensure(inst.size());
    L.lm->hops.ensure(ops.size());
    std::copy(inst.begin(), inst.end(), L.lm->hinst.p);
    std::copy(ops.begin(), ops.end(), L.lm->hops.p);
    run_soft(P, inst, S.Pp2.p, S.W2.p, S.Mx.p, S.C.p, L.lm->hinst.p);
    S.ops.ensure(ops.size());
    ETTK_HIP(hipMemcpyAsync(S.ops.p, L.lm->hops.p, ops.size() * sizeof(ComboOp), hipMemcpyHostToDevice, s));
    const int b0 = restrict_b0(P);
    {
        ProfScope ps(kp, "k_combine", (double)(nq + 2 * cn) * n3 * 8, s);
        launch_combine(S.C.p, S.ops.p, 2 * cn, B, b0, S.Cc.p, s);
    }
    launch_fill_zero(L.lm->count.p, (size_t)2 * cn * kLmCountStride, s);
    {
        ProfScope ps(kp, "k_localmax", 2.0 * cn * n3 * 8, s);
        launch_localmax(S.Cc.p, 2 * cn, B, lm_cap(P), L.lm->count.p, L.lm->idx.p, L.lm->val.p, s);
    }
    // the candidate extraction and the packing run on the plan's auxiliary stream, beside the next chunk's kernels
    ETTK_HIP(hipEventCreateWithFlags(&L.maxima, hipEventDisableTiming));
    ETTK_HIP(hipEventRecord(L.maxima, s));
    hipStream_t as = P->aux_stream;
    ETTK_HIP(hipStreamWaitEvent(as, L.maxima, 0));
    {
        // candidate extraction of the active volumes on the device (the second volume of a particle is used only
        // with a spherical window)
        // ETTK_HOST_CANDIDATES=1 leaves every volume to the host's extraction (the same candidates; for checks)
        LmBuffers& lb = *L.lm;
        const char* hc = std::getenv("ETTK_HOST_CANDIDATES");
        const int act = hc && hc[0] == '1' ? 2 : 1;
        for (int v = 0; v < 2 * cn; v++) lb.hact.p[v] = ((v % 2 == 0) || uw1 || L.uw2[v / 2]) ? act : 0;
        ETTK_HIP(hipMemcpyAsync(lb.act.p, lb.hact.p, (size_t)2 * cn * sizeof(int), hipMemcpyHostToDevice, as));
        if (S.zyz_tabs.n == 0) {
            const ZyzTables& Z = zyz_tables(B);
            std::vector<double> t((size_t)3 * n * 16);
            for (int i = 0; i < n; i++) {
                std::memcpy(&t[(size_t)i * 16], Z.ra[i].a, sizeof(Z.ra[i].a));
                std::memcpy(&t[((size_t)n + i) * 16], Z.rb[i].a, sizeof(Z.rb[i].a));
                std::memcpy(&t[((size_t)2 * n + i) * 16], Z.rg[i].a, sizeof(Z.rg[i].a));
            }
            S.zyz_tabs.ensure(t.size());
            ETTK_HIP(hipMemcpy(S.zyz_tabs.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        S.cand_mats.ensure((size_t)2 * cn * kCandMax * 16);
        ProfScope ps(kp, "k_lm_candidates", 0.0, as);
        launch_lm_candidates(lb.count.p, lb.idx.p, lb.val.p, lm_cap(P), lb.act.p, 2 * cn, S.zyz_tabs.p, n,
                             (double)P->trans_restrict, (double)P->rot_restrict, kEquivTol, S.cand_mats.p, lb.deg.p,
                             lb.adj.p, lb.kept.p, lb.nadj.p, lb.mode.p, as);
    }
    launch_lm_compact(L.lm->count.p, L.lm->idx.p, L.lm->val.p, L.lm->kept.p, L.lm->mode.p, L.lm->deg.p, L.lm->nadj.p,
                      L.lm->adj.p, 2 * cn, lm_cap(P), L.lm->head.p, L.lm->pidx.p, L.lm->pval.p, L.lm->pdeg.p,
                      L.lm->padj.p, as);
    ETTK_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    ETTK_HIP(hipEventRecord(L.done, as));
}

// Host part: download the chunk's compacted maxima (after its kernels, on the context's copy stream: the plan's
// stream does not wait for it) and extract the candidate lists of its volumes on several host threads.
static void search_host(const ettk_plan* P, SearchLists& L, std::vector<SearchOut>& out, bool last = false) {
    KernelProf* kp = P->ctx->fft.prof;
    ETTK_HIP(hipEventSynchronize(L.done));
    auto th0 = std::chrono::steady_clock::now();
    L.t_ready = th0;
    const int p0 = L.p0, cn = L.cn;
    const bool uw1 = L.uw1;
    const int nv = 2 * cn;
    hipStream_t cs = P->copy_stream;
    LmBuffers& lb = *L.lm;
    const int* head = lb.hhead.p;
    ETTK_HIP(hipMemcpyAsync(lb.hhead.p, lb.head.p, (size_t)(4 * nv + 2) * sizeof(int), hipMemcpyDeviceToHost, cs));
    ETTK_HIP(hipStreamSynchronize(cs));
    const int total = head[nv], atotal = head[4 * nv + 1];
    if (total > 0) {
        lb.hpidx.ensure(total);
        lb.hpval.ensure(total);
        lb.hpdeg.ensure(total);
        ETTK_HIP(hipMemcpyAsync(lb.hpidx.p, lb.pidx.p, (size_t)total * sizeof(int), hipMemcpyDeviceToHost, cs));
        ETTK_HIP(hipMemcpyAsync(lb.hpval.p, lb.pval.p, (size_t)total * sizeof(double), hipMemcpyDeviceToHost, cs));
        ETTK_HIP(hipMemcpyAsync(lb.hpdeg.p, lb.pdeg.p, (size_t)total * sizeof(int), hipMemcpyDeviceToHost, cs));
    }
    if (atotal > 0) {
        lb.hpadj.ensure(atotal);
        ETTK_HIP(hipMemcpyAsync(lb.hpadj.p, lb.padj.p, (size_t)atotal * sizeof(int), hipMemcpyDeviceToHost, cs));
    }
    ETTK_HIP(hipStreamSynchronize(cs));
    const int* pidx = lb.hpidx.p;
    const double* pval = lb.hpval.p;
    const int* pdeg = lb.hpdeg.p;
    const int* padj = lb.hpadj.p;
    if (kp) kp->add_host("search_download", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th0).count());
    std::vector<std::vector<int>> idx(nv), deg(nv), adj(nv);
    std::vector<std::vector<double>> val(nv);
    std::vector<int> mode(nv, 0);
    for (int v = 0; v < nv; v++) {
        int q = v / 2;
        bool two = uw1 || L.uw2[q];
        if (v % 2 == 1 && !two) continue;
        const int raw = head[nv + 1 + v];
        ETTK_REQUIRE(raw <= lm_cap(P), ETTK_ERR_UNSUPPORTED,
                     "more SO(3) local maxima than the candidate list holds (" + std::to_string(raw) + ")");
        mode[v] = head[2 * nv + 1 + v];
        const int c = head[v + 1] - head[v];  // sorted valid maxima (mode 0) or the raw maxima (mode 1)
        idx[v].assign(pidx + head[v], pidx + head[v] + c);
        val[v].assign(pval + head[v], pval + head[v] + c);
        if (mode[v] == 0) {
            deg[v].assign(pdeg + head[v], pdeg + head[v] + c);
            const int a0 = head[3 * nv + 1 + v], a1 = head[3 * nv + 2 + v];
            adj[v].assign(padj + a0, padj + a1);
        }
        if (kp && kp->on) {
            kp->add_host("lm_count", (double)raw);
            if (mode[v]) kp->add_host("lm_host_volumes", 1.0);
        }
    }
    for (int q = 0; q < cn; q++) {
        SearchOut& o = out[p0 + q];
        o.two = uw1 || L.uw2[q];
        o.use_window = (uw1 ? 1 : 0) | (L.uw2[q] ? 2 : 0);
    }
    auto work = [&](int v) {
        int q = v / 2;
        if (v % 2 == 1 && !out[p0 + q].two) return;
        SearchOut& o = out[p0 + q];
        CandQueue& pq = v % 2 == 0 ? o.pq1 : o.pq2;
        if (mode[v] == 0) {
            candidates_select(P, idx[v], val[v], deg[v], adj[v], pq);  // sorted, checked and paired on the device
            return;
        }
        std::vector<std::pair<int, double>> lm(idx[v].size());
        for (size_t k = 0; k < lm.size(); k++) lm[k] = {idx[v][k], val[v][k]};
        candidates_from_maxima(P, lm, pq);
    };
    // the next chunk's extraction and input preparation run at the same time: half of the threads each, all of
    // them for the last chunk of a call
    int nthr = std::max(1, std::min<int>(2 * cn, last ? host_threads() : host_threads() / 2));
    if (nthr <= 1) {
        for (int v = 0; v < 2 * cn; v++) work(v);
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < nthr; w++)
            th.emplace_back([&, w] {
                for (int v = w; v < 2 * cn; v += nthr) work(v);
            });
        for (auto& t : th) t.join();
    }
    if (kp) kp->add_host("search_host", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th0).count());
}

// The search of particles [0, n): chunks of `chunk` particles enqueued back to back; each chunk's lists are
// processed on a host thread as soon as its kernels are done, while later chunks run on the GPU.
static void search_all(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n, int chunk,
                       std::vector<SearchOut>& out) {
    KernelProf* kp = P->ctx->fft.prof;
    auto ts0 = std::chrono::steady_clock::now();
    SearchInputs in;
    std::vector<std::unique_ptr<SearchLists>> lists;
    std::vector<std::future<void>> pending;
    std::exception_ptr err;
    // the inputs of chunk k + 1 are prepared on a second host thread while chunk k is enqueued (disjoint ranges of
    // arrays sized for the whole call by the first, synchronous preparation)
    std::future<void> next;
    try {
        search_inputs_init(P, n, chunk, in);
        search_inputs(P, descs, n, 0, std::min(chunk, n), in);
        for (int p0 = 0, k = 0; p0 < n; p0 += chunk, k++) {
            const int cn = std::min(chunk, n - p0);
            auto ti = std::chrono::steady_clock::now();
            if (next.valid()) next.get();
            auto tg = std::chrono::steady_clock::now();
            if (p0 + chunk < n) {
                const int q0 = p0 + chunk, qn = std::min(chunk, n - q0);
                next = std::async(std::launch::async, [P, descs, n, q0, qn, &in] { search_inputs(P, descs, n, q0, qn, in); });
            }
            // chunk k reuses the device list buffers of chunk k - kSearchSlots: its host part must be done
            if (k >= kSearchSlots) {
                std::future<void>& old = pending[k - kSearchSlots];
                try {
                    old.get();
                } catch (...) {
                    if (!err) err = std::current_exception();
                }
                if (err) break;
            }
            lists.emplace_back(new SearchLists());
            SearchLists* L = lists.back().get();
            search_gpu(P, d_parts, descs, p0, cn, in, k % kSearchSlots, *L);
            if (kp) {
                auto te = std::chrono::steady_clock::now();
                kp->add_host("search_inputs", std::chrono::duration<double, std::milli>(tg - ti).count());
                kp->add_host("search_enqueue", std::chrono::duration<double, std::milli>(te - tg).count());
            }
            const bool last = p0 + chunk >= n;
            pending.push_back(std::async(std::launch::async, [P, L, &out, last] { search_host(P, *L, out, last); }));
        }
    } catch (...) {
        err = std::current_exception();
    }
    if (next.valid()) {
        try {
            next.get();
        } catch (...) {
            if (!err) err = std::current_exception();
        }
    }
    for (auto& f : pending) {
        if (!f.valid()) continue;  // already collected when its slot was reused
        try {
            f.get();
        } catch (...) {
            if (!err) err = std::current_exception();
        }
    }
    if (err) std::rethrow_exception(err);
    if (kp) {
        const auto te = std::chrono::steady_clock::now();
        kp->add_host("search_stage_wall", std::chrono::duration<double, std::milli>(te - ts0).count());
        if (!lists.empty())  // host time after the last chunk's kernels
            kp->add_host("search_tail", std::chrono::duration<double, std::milli>(te - lists.back()->t_ready).count());
    }
}

// ------------------------------------------------------------------ evaluations awaited by coroutines
// The refinement of a particle is a coroutine that suspends at every CCC evaluation it needs.  An evaluation is
// served either at once (a sink with `direct` set: the plain form, or results computed beforehand) or queued:
// the driver then evaluates every queued request of all particles in one batched launch and resumes them, so
// the particles advance in lock-step rounds without a host thread each.
namespace {
struct EvalReq {
    int pi;
    Mat4* t;  // the transform; receives the peak's shift as ccc_batch returns it
    float out;
    std::coroutine_handle<> h;
};
struct EvalSink {
    std::function<float(int, Mat4&)> direct;  // set: evaluate at once
    std::vector<EvalReq*> pending;
};
struct EvalAwait {
    EvalSink* sink;
    EvalReq r;
    bool await_ready() {
        if (!sink->direct) return false;
        r.out = sink->direct(r.pi, *r.t);
        return true;
    }
    void await_suspend(std::coroutine_handle<> h) {
        r.h = h;
        sink->pending.push_back(&r);
    }
    float await_resume() const noexcept { return r.out; }
};
struct Evaluator {
    EvalSink* sink;
    int* count;  // evaluations made, or null
    EvalAwait operator()(int pi, Mat4& t) const {
        if (count) (*count)++;
        return EvalAwait{sink, EvalReq{pi, &t, 0.0f, {}}};
    }
};

// A lazily started coroutine returning T; awaiting it runs it and resumes the awaiting coroutine when it is done
// (symmetric transfer), and rethrows its exception there.
template <class T>
struct TaskPromiseBase {
    std::exception_ptr err;
    std::coroutine_handle<> cont;
    std::suspend_always initial_suspend() noexcept { return {}; }
    template <class P>
    struct Final {
        bool await_ready() noexcept { return false; }
        std::coroutine_handle<> await_suspend(std::coroutine_handle<P> h) noexcept {
            const std::coroutine_handle<> c = h.promise().cont;
            return c ? c : std::noop_coroutine();
        }
        void await_resume() noexcept {}
    };
    void unhandled_exception() { err = std::current_exception(); }
};
template <class T>
class Task {
   public:
    struct promise_type : TaskPromiseBase<T> {
        T value{};
        Task get_return_object() { return Task(std::coroutine_handle<promise_type>::from_promise(*this)); }
        typename TaskPromiseBase<T>::template Final<promise_type> final_suspend() noexcept { return {}; }
        void return_value(T v) { value = std::move(v); }
    };
    explicit Task(std::coroutine_handle<promise_type> h) : h_(h) {}
    Task(Task&& o) noexcept : h_(std::exchange(o.h_, {})) {}
    Task(const Task&) = delete;
    ~Task() {
        if (h_) h_.destroy();
    }
    bool await_ready() const noexcept { return false; }
    std::coroutine_handle<> await_suspend(std::coroutine_handle<> c) noexcept {
        h_.promise().cont = c;
        return h_;
    }
    T await_resume() {
        if (h_.promise().err) std::rethrow_exception(h_.promise().err);
        return std::move(h_.promise().value);
    }
    void start() { h_.resume(); }
    bool done() const { return h_.done(); }
    void rethrow() const {
        if (h_.promise().err) std::rethrow_exception(h_.promise().err);
    }

   private:
    std::coroutine_handle<promise_type> h_;
};
template <>
class Task<void> {
   public:
    struct promise_type : TaskPromiseBase<void> {
        Task get_return_object() { return Task(std::coroutine_handle<promise_type>::from_promise(*this)); }
        TaskPromiseBase<void>::Final<promise_type> final_suspend() noexcept { return {}; }
        void return_void() {}
    };
    explicit Task(std::coroutine_handle<promise_type> h) : h_(h) {}
    Task(Task&& o) noexcept : h_(std::exchange(o.h_, {})) {}
    Task(const Task&) = delete;
    ~Task() {
        if (h_) h_.destroy();
    }
    bool await_ready() const noexcept { return false; }
    std::coroutine_handle<> await_suspend(std::coroutine_handle<> c) noexcept {
        h_.promise().cont = c;
        return h_;
    }
    void await_resume() {
        if (h_.promise().err) std::rethrow_exception(h_.promise().err);
    }
    void start() { h_.resume(); }
    bool done() const { return h_.done(); }
    void rethrow() const {
        if (h_.promise().err) std::rethrow_exception(h_.promise().err);
    }

   private:
    std::coroutine_handle<promise_type> h_;
};
}  // namespace

// ------------------------------------------------------------------ Powell (vnl_powell, maxfev = 1)
// Restatement of vnl_powell::minimize (ITK 4.2.1 vnl/algo/vnl_powell.cxx:60-175) with
// vnl_bracket_minimum and vnl_brent_minimizer::minimize_given_bounds, as nbfFourierImageMetric::refine
// configures it (FIM.h:1742-1753): initial step 5, line-search x tolerance 1e-2, f tolerance 0.1.
namespace {
struct PowellState {
    std::function<Task<double>(const double*)> f;
    double end_error = 0;
    long nevals = 0;
    void report(double e) {
        if (nevals == 0) end_error = e;
        if (e < end_error) end_error = e;
        nevals++;
    }
};
const double GOLD = 1.618033988749894848, CGOLD = 0.381966011250105152;
struct Line {
    PowellState* S;
    double x0[3], dx[3];
    Task<double> operator()(double lam) {
        double x[3] = {x0[0] + lam * dx[0], x0[1] + lam * dx[1], x0[2] + lam * dx[2]};
        double e = co_await S->f(x);
        S->report(e);
        co_return e;
    }
};
Task<void> bracket(Line& f, double& a, double& b, double& c) {
    const double EPS = 1e-7, EPSqr = 1e-14;
    if (b == a) b = a + 1.0;
    double fa = co_await f(a), fb = co_await f(b);
    if (fb > fa) { std::swap(a, b); std::swap(fa, fb); }
    c = b + GOLD * (b - a);
    double fc = co_await f(c);
    while (fc < fb) {
        double x1 = a - b, f1 = fa - fb, x2 = c - b, f2 = fc - fb;
        double p = x2 * x2 * f1 - x1 * x1 * f2, q = 2 * (x2 * f1 - x1 * f2);
        if (q >= 0 && q < EPSqr) q = EPSqr;
        else if (q < 0 && q + EPSqr > 0) q = -1.0 * EPSqr;
        double du = p / q, tol = EPS * (1.0 + std::max(std::fabs(b), std::fabs(c)));
        if (du >= 0 && du < tol) du = tol;
        else if (du < 0 && du + tol > 0) du = -1.0 * tol;
        double u = b + du;
        if ((u - c) < tol && (u - c) >= 0) u += tol;
        else if ((c - u) < tol && (c - u) >= 0) u -= tol;
        double ulim = b + 100 * (c - b), fu = 0.0;
        if ((u - b) * (c - u) > 0.0) {
            fu = co_await f(u);
            if (fu < fc) {
                a = b; fa = fb; b = u; fb = fu;
                if (a > c) { std::swap(a, c); std::swap(fa, fc); }
                co_return;
            } else if (fu > fb) {
                c = u; fc = fu;
                if (a > c) { std::swap(a, c); std::swap(fa, fc); }
                co_return;
            }
            u = c + GOLD * (c - b);
            fu = co_await f(u);
        } else if ((u - c) * (ulim - u) > 0.0) {
            fu = co_await f(u);
            if (fu > fc) {
                a = b; fa = fb; b = c; fb = fc; c = u; fc = fu;
                if (a > c) { std::swap(a, c); std::swap(fa, fc); }
                co_return;
            }
        } else if ((ulim - c) * (u - ulim) >= 0) {
            u = ulim;
            fu = co_await f(u);
        } else {
            u = c + GOLD * (c - b);
            fu = co_await f(u);
        }
        a = b; b = c; c = u;
        fa = fb; fb = fc; fc = fu;
    }
    if (a > c) { std::swap(a, c); std::swap(fa, fc); }
}
Task<double> brent(Line& f, double ax, double bx, double cx, double xtol, double& fmin) {
    const double EPS = 1e-8;
    double fb = co_await f(bx);
    double x = bx, w = x, v = w, u, fx = fb, fw = fx, fv = fw, fu;
    double m = 0.5 * (ax + cx), tol = EPS * std::fabs(x) + xtol, tol2 = 2 * tol, d = 0.0, e = 0.0;
    while (std::fabs(x - m) > (tol2 - 0.5 * (cx - ax))) {
        double p = 0.0, q = 0.0, r = 0.0;
        if (std::fabs(e) > tol) {
            r = (x - w) * (fx - fv);
            q = (x - v) * (fx - fw);
            p = (x - v) * q - (x - w) * r;
            q = 2 * (q - r);
            if (q > 0) p *= -1.0; else q *= -1.0;
            r = e; e = d;
        }
        if (std::fabs(p) < std::fabs(0.5 * q * r) && p > (q * (ax - x)) && p < (q * (cx - x))) {
            d = p / q;
            u = x + d;
            if (u - ax < tol2 || cx - u < tol2) d = (x < m ? tol : -tol);
        } else {
            e = (x < m ? cx : ax) - x;
            d = CGOLD * e;
        }
        if (std::fabs(d) >= tol) u = x + d;
        else u = (d > 0) ? x + tol : x - tol;
        fu = co_await f(u);
        if (fu <= fx) {
            if (u < x) cx = x; else ax = x;
            v = w; fv = fw; w = x; fw = fx; x = u; fx = fu;
        } else {
            if (u < x) ax = u; else cx = u;
            if (fu <= fw || w == x) { v = w; fv = fw; w = u; fw = fu; }
            else if (fu <= fv || v == x || v == w) { v = u; fv = fu; }
        }
        m = 0.5 * (ax + cx);
        tol = EPS * std::fabs(x) + xtol;
        tol2 = 2 * tol;
    }
    fmin = fx;
    co_return x;
}
Task<void> powell(PowellState& S, double p[3], double step, double lxtol, double ftol) {
    const int n = 3;
    double xi[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, ptt[3], xit[3], pt[3];
    double fret = co_await S.f(p);
    S.report(fret);
    for (int j = 0; j < n; j++) pt[j] = p[j];
    for (int iter = 0; iter < 1; iter++) {
        double fp = fret, del = 0.0;
        int ibig = 0;
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) xit[j] = xi[j][i];
            double fptt = fret;
            Line L{&S, {p[0], p[1], p[2]}, {xit[0], xit[1], xit[2]}};
            double ax = 0.0, xx = step, bx, fm;
            co_await bracket(L, ax, xx, bx);
            xx = co_await brent(L, ax, xx, bx, lxtol, fm);
            fret = fm;
            for (int j = 0; j < n; j++) p[j] = L.x0[j] + xx * L.dx[j];
            if (std::fabs(fptt - fret) > del) { del = std::fabs(fptt - fret); ibig = i; }
        }
        if (2.0 * std::fabs(fp - fret) <= ftol * (std::fabs(fp) + std::fabs(fret))) co_return;
        for (int j = 0; j < n; j++) {
            ptt[j] = 2.0 * p[j] - pt[j];
            xit[j] = p[j] - pt[j];
            pt[j] = p[j];
        }
        double fptt = co_await S.f(ptt);
        S.report(fret);
        if (fptt < fp) {
            double t = 2.0 * (fp - 2.0 * fret + fptt) * (fp - fret - del) * (fp - fret - del) - del * (fp - fptt) * (fp - fptt);
            if (t < 0.0) {
                Line L{&S, {p[0], p[1], p[2]}, {xit[0], xit[1], xit[2]}};
                double ax = 0.0, xx = 1.0, bx, fm;
                co_await bracket(L, ax, xx, bx);
                xx = co_await brent(L, ax, xx, bx, lxtol, fm);
                fret = fm;
                for (int j = 0; j < n; j++) p[j] = L.x0[j] + xx * L.dx[j];
                for (int j = 0; j < n; j++) {
                    xi[j][ibig] = xi[j][n - 1];
                    xi[j][n - 1] = xit[j];
                }
            }
        }
    }
}
}  // namespace

// ------------------------------------------------------------------ per-particle control flow
using EvalFn = std::function<float(int particle, Mat4& t)>;

// nbfFourierImageMetric::refine (FIM.h:1702-1809) with f (:129-190)
static Task<float> refine_powell(const ettk_plan* P, int pi, Mat4& t, bool run, Evaluator eval) {
    Mat4 seed = strip_translation(t);
    float last = 0.0f;
    auto trial = [&](const double* x) {
        return seed * mat_rot(x[0], 1, 0, 0) * mat_rot(x[1], 0, 1, 0) * mat_rot(x[2], 0, 0, 1);
    };
    PowellState S;
    S.f = [&](const double* x) -> Task<double> {
        Mat4 tt = trial(x);
        if (!transform_ok(tt, P->trans_restrict, P->rot_restrict)) co_return (double)FLT_MAX;
        float peak = co_await eval(pi, tt);
        last = peak;
        if (!transform_ok(tt, P->trans_restrict, P->rot_restrict)) co_return (double)FLT_MAX;
        co_return (double)peak;
    };
    double x[3] = {0, 0, 0};
    if (run) co_await powell(S, x, 5.0, 1e-2, 1e-1);
    if (S.end_error != (double)FLT_MAX) {
        Mat4 vt = trial(x);
        last = co_await eval(pi, vt);
        t = vt;
    } else {
        last = co_await eval(pi, t);
    }
    co_return last;
}

// candidate loop of nbfProjectionRotationMetric3D::execute (ProjRot.h:678-901), split in two parts:
// peak_loop evaluates the direct and mirrored transform of the SO(3) peaks of each queue until Kp (Kp/3 for
// the second queue) of them beat their mirror, then eliminates equivalent angles and returns the Kc best;
// final_candidates refines them (Powell) or, with `seeds` set (batched + adjoint mode), gives each the
// refine(t, 0) evaluation only (one CCC at the candidate rotation, FIM.h:1758-1766) and hands the transform
// back as a seed of the continuous refinement.
static void peak_loop(const ettk_plan* P, int pi, SearchOut& so, const EvalFn& eval, std::vector<Cand>& selected,
                      int& nccc) {
    std::vector<Cand> combined;
    int nq = so.two ? 2 : 1;
    for (int q = 0; q < nq; q++) {
        CandQueue& pq = q == 0 ? so.pq1 : so.pq2;
        int peaks = P->p.n_peaks;
        if (q == 1) peaks = (int)std::max(1.0, std::floor(P->p.n_peaks / 3.0));
        int index = 0;
        while (index < peaks && !pq.empty()) {
            Cand p = pq.top();
            if (p.v == FLT_MAX) break;
            Mat4 t1 = zyz(p.a, p.b, p.g), td = t1;
            nccc++;
            float D = eval(pi, td);
            if (transform_ok(td, P->trans_restrict, P->rot_restrict)) {
                Mat4 mir = mat_scale(-1, -1, -1) * t1;
                nccc++;
                float Dm = eval(pi, mir);
                if (Dm >= D) index++;
                combined.push_back(Cand{p.a, p.b, p.g, D});
            }
            pq.pop();
        }
    }
    if (combined.empty()) combined.push_back(Cand{0, 0, 0, 1.0f});
    CandQueue pqc;
    eliminate_equivalent(combined, pqc, 2.5f);
    selected.clear();
    while (!pqc.empty() && (int)selected.size() < P->p.n_candidates) {
        selected.push_back(pqc.top());
        pqc.pop();
    }
}

static Task<void> final_candidates(const ettk_plan* P, int pi, const std::vector<Cand>& selected, SearchOut& so,
                                   EvalSink* sink, ettk_result& out, int nccc, std::vector<Mat4>* seeds = nullptr) {
    const Evaluator counted{sink, &nccc};
    float best = FLT_MAX;
    Mat4 bestT = Mat4::ident();
    int index = 0;
    for (const Cand& p : selected) {
        Mat4 t = zyz(p.a, p.b, p.g);
        float dist = co_await refine_powell(P, pi, t, seeds == nullptr && P->p.refine != 0, counted);
        if (seeds) seeds->push_back(t);
        if (dist < best) {
            best = dist;
            bestT = t;
        }
        out.cand_distance[index] = dist;
        std::memcpy(out.cand_matrix[index], t.a, sizeof(t.a));
        out.so3_angles[index][0] = p.a;
        out.so3_angles[index][1] = p.b;
        out.so3_angles[index][2] = p.g;
        index++;
    }
    out.n_cand = index;
    out.distance = best;
    std::memcpy(out.matrix, bestT.a, sizeof(bestT.a));
    out.n_ccc = nccc;
    out.use_window = so.use_window;
}

// runs a task whose evaluations are all served at once (a sink with `direct` set): it completes on its first resume
static void run_direct(Task<void>& task) {
    task.start();
    ETTK_REQUIRE(task.done(), ETTK_ERR_STATE, "an evaluation with a direct sink suspended");
    task.rethrow();
}

// Runs the tasks in lock-step rounds: each round evaluates every request the tasks have queued in `sink` in one
// ccc_batch (at most `chunk` evaluations per launch, as ccc_batch splits them) and resumes the requesting tasks.
static void run_batched(ettk_plan* P, const float* d_parts, const ettk_particle* descs, std::vector<Task<void>>& tasks,
                        EvalSink& sink) {
    for (auto& t : tasks) t.start();
    std::vector<EvalReq*> batch;
    std::vector<int> pidx;
    std::vector<Mat4> tr;
    std::vector<float> d;
    while (!sink.pending.empty()) {
        batch.swap(sink.pending);
        sink.pending.clear();
        const size_t m = batch.size();
        pidx.resize(m);
        tr.resize(m);
        d.resize(m);
        for (size_t k = 0; k < m; k++) {
            pidx[k] = batch[k]->pi;
            tr[k] = *batch[k]->t;
        }
        ccc_batch(P, d_parts, descs, (int)m, pidx.data(), tr.data(), d.data());
        for (size_t k = 0; k < m; k++) {
            *batch[k]->t = tr[k];
            batch[k]->out = d[k];
        }
        for (size_t k = 0; k < m; k++) batch[k]->h.resume();  // each runs to its next evaluation or its end
        batch.clear();
    }
    for (auto& t : tasks) {
        ETTK_REQUIRE(t.done(), ETTK_ERR_STATE, "a refinement task neither finished nor asked for an evaluation");
        t.rethrow();
    }
}

static void candidate_stage(const ettk_plan* P, int pi, SearchOut& so, const EvalFn& eval, ettk_result& out) {
    std::vector<Cand> selected;
    int nccc = 0;
    peak_loop(P, pi, so, eval, selected, nccc);
    EvalSink sink;
    sink.direct = eval;
    Task<void> task = final_candidates(P, pi, selected, so, &sink, out, nccc);
    run_direct(task);
}

// The peak loop of every particle in lock-step rounds (batched modes).  A queue that still needs `need`
// peaks beating their mirror will evaluate at least its next `need` peaks (each raises the count by at most
// one), so each round evaluates those peaks, direct and mirrored, for all particles in one batch; the results
// are then consumed in the order of the sequential loop.  A mirror whose direct transform turns out invalid
// is discarded (and not counted), so every particle ends with the same candidates, distances and CCC count as
// peak_loop.
// evaluated: when set, receives per particle the (transform with its peak shift, distance) of every direct
// evaluation a combined candidate came from (Cand::k indexes it)
// after_rounds, when set, is called once the last round's results are in, before the host's elimination of
// equivalent candidates (work that does not depend on the candidates can be enqueued there)
static void peak_rounds(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n,
                        std::vector<SearchOut*>& so, std::vector<std::vector<Cand>>& selected, std::vector<int>& nccc,
                        std::vector<std::vector<std::pair<Mat4, float>>>* evaluated = nullptr,
                        const std::function<void()>* after_rounds = nullptr) {
    struct State {
        int q = 0, index = 0, peaks = 0;
        bool done = false;
        std::vector<Cand> combined, popped;
        std::vector<std::pair<Mat4, float>> evald;
    };
    std::vector<State> st(n);
    for (int i = 0; i < n; i++) st[i].peaks = P->p.n_peaks;
    nccc.assign(n, 0);
    selected.assign(n, {});
    auto queue = [&](int i) -> CandQueue& { return st[i].q == 0 ? so[i]->pq1 : so[i]->pq2; };
    auto advance = [&](int i) {  // current queue finished: next queue or done
        State& s = st[i];
        if (s.q == 0 && so[i]->two) {
            s.q = 1;
            s.index = 0;
            s.peaks = (int)std::max(1.0, std::floor(P->p.n_peaks / 3.0));
        } else {
            s.done = true;
        }
    };
    // the particles in one or two groups, each with its batch of the current round: with two, the host consumes
    // one group's results and builds its next round while the GPU runs the other group's batch
    struct Group {
        int i0, i1;
        std::vector<int> pidx;
        std::vector<Mat4> t;
        std::vector<float> d;
        bool active = false;
    };
    auto build = [&](Group& G) {  // this round's evaluations of the group; false when it has none left
        G.pidx.clear();
        G.t.clear();
        for (int i = G.i0; i < G.i1; i++) {
            State& s = st[i];
            s.popped.clear();
            while (!s.done) {
                CandQueue& pq = queue(i);
                const int need = s.peaks - s.index;
                while ((int)s.popped.size() < need && !pq.empty() && pq.top().v != FLT_MAX) {
                    s.popped.push_back(pq.top());
                    pq.pop();
                }
                if (!s.popped.empty()) break;
                advance(i);  // nothing left to evaluate in this queue
            }
            for (const Cand& p : s.popped) {
                Mat4 t1 = zyz(p.a, p.b, p.g);
                G.pidx.push_back(i);
                G.t.push_back(t1);
                G.pidx.push_back(i);
                G.t.push_back(mat_scale(-1, -1, -1) * t1);
            }
        }
        G.d.resize(G.pidx.size());
        return !G.pidx.empty();
    };
    auto consume = [&](Group& G) {
        size_t r = 0;
        for (int i = G.i0; i < G.i1; i++) {
            State& s = st[i];
            if (s.popped.empty()) continue;
            for (const Cand& p : s.popped) {
                const float D = G.d[r];
                const Mat4& td = G.t[r];
                const float Dm = G.d[r + 1];
                r += 2;
                nccc[i]++;
                if (transform_ok(td, P->trans_restrict, P->rot_restrict)) {
                    nccc[i]++;
                    if (Dm >= D) s.index++;
                    if (evaluated) s.evald.emplace_back(td, D);
                    s.combined.push_back(Cand{p.a, p.b, p.g, D, evaluated ? (int)s.evald.size() - 1 : -1});
                }
            }
            CandQueue& pq = queue(i);
            if (s.index >= s.peaks || pq.empty() || pq.top().v == FLT_MAX) advance(i);
        }
    };
    if (n >= 2 && ccc_lanes_available(P)) {
        ensure_ccc_workspace(P, P->ccc_chunk);  // no workspace growth while a batch is in flight
        Group G[2];
        G[0].i0 = 0;
        G[0].i1 = n / 2;
        G[1].i0 = n / 2;
        G[1].i1 = n;
        try {
            for (int h = 0; h < 2; h++)
                if ((G[h].active = build(G[h])))
                    ccc_batch_begin(P, h, d_parts, descs, (int)G[h].pidx.size(), G[h].pidx.data(), G[h].t.data(),
                                    G[h].d.data());
            while (G[0].active || G[1].active) {
                for (int h = 0; h < 2; h++) {
                    if (!G[h].active) continue;
                    ccc_batch_end(P, h);
                    consume(G[h]);
                    if ((G[h].active = build(G[h])))
                        ccc_batch_begin(P, h, d_parts, descs, (int)G[h].pidx.size(), G[h].pidx.data(), G[h].t.data(),
                                        G[h].d.data());
                }
            }
        } catch (...) {
            // the other lane may still be in flight: let it finish before its arrays go away
            (void)hipStreamSynchronize(P->stream);
            P->lane[0].pending = P->lane[1].pending = false;
            throw;
        }
    } else {
        Group all;
        all.i0 = 0;
        all.i1 = n;
        while (build(all)) {
            ccc_batch(P, d_parts, descs, (int)all.pidx.size(), all.pidx.data(), all.t.data(), all.d.data());
            consume(all);
        }
    }
    if (after_rounds) (*after_rounds)();
    parallel_for(n, [&](int i) {  // particles are independent
        std::vector<Cand>& combined = st[i].combined;
        if (combined.empty()) combined.push_back(Cand{0, 0, 0, 1.0f});
        CandQueue pqc;
        eliminate_equivalent(combined, pqc, 2.5f);
        while (!pqc.empty() && (int)selected[i].size() < P->p.n_candidates) {
            selected[i].push_back(pqc.top());
            pqc.pop();
        }
    });
    if (evaluated) {
        evaluated->resize(n);
        for (int i = 0; i < n; i++) (*evaluated)[i] = std::move(st[i].evald);
    }
}

void refine_adjoint_batch(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n,
                          std::vector<SearchOut*>& so, ettk_result* out);

// result.overlap: getWedgeOverlap at the reported transform of each result (particle descriptors as the run used)
static void set_overlaps(ettk_plan* P, const ettk_particle* descs, int n, ettk_result* out) {
    std::vector<Mat4> tf(n);
    std::vector<float> ov(n);
    for (int i = 0; i < n; i++) tf[i] = Mat4::from(out[i].matrix);
    wedge_overlaps(P, descs, tf.data(), n, ov.data());
    for (int i = 0; i < n; i++) out[i].overlap = ov[i];
}

void align_run(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n, int mode, ettk_result* out) {
    ETTK_REQUIRE(P->have_ref, ETTK_ERR_STATE, "reference not set");
    ETTK_REQUIRE(mode >= 0 && mode <= 2, ETTK_ERR_ARG, "ettk_align_batch: mode must be 0, 1 or 2");
    std::memset(out, 0, sizeof(ettk_result) * n);
    P->timing = Timing{};
    ScopedEvent a, b;
    ETTK_HIP(hipEventRecord(a, P->stream));
    std::vector<SearchOut> so(n);
    if (mode == 0) {
        EvalFn ev = [&](int pi, Mat4& t) {
            float d;
            ccc_batch(P, d_parts, descs, 1, &pi, &t, &d);
            return d;
        };
        SearchInputs in;
        search_inputs_init(P, n, 1, in);
        for (int i = 0; i < n; i++) {
            search_inputs(P, descs, n, i, 1, in);
            SearchLists L;
            search_gpu(P, d_parts, descs, i, 1, in, 0, L);
            search_host(P, L, so);
            candidate_stage(P, i, so[i], ev, out[i]);
        }
    } else {
        search_all(P, d_parts, descs, n, 32, so);
        if (mode == 1) {
            std::vector<SearchOut*> sp(n);
            for (int i = 0; i < n; i++) sp[i] = &so[i];
            std::vector<std::vector<Cand>> sel;
            std::vector<int> nccc;
            peak_rounds(P, d_parts, descs, n, sp, sel, nccc);
            EvalSink sink;
            std::vector<Task<void>> tasks;
            tasks.reserve(n);
            for (int i = 0; i < n; i++) tasks.push_back(final_candidates(P, i, sel[i], so[i], &sink, out[i], nccc[i]));
            run_batched(P, d_parts, descs, tasks, sink);
        } else {
            std::vector<SearchOut*> sp(n);
            for (int i = 0; i < n; i++) sp[i] = &so[i];
            refine_adjoint_batch(P, d_parts, descs, n, sp, out);
        }
    }
    set_overlaps(P, descs, n, out);
    ETTK_HIP(hipEventRecord(b, P->stream));
    ETTK_HIP(hipEventSynchronize(b));
    float ms = 0;
    ETTK_HIP(hipEventElapsedTime(&ms, a, b));
    P->timing.total_ms = ms;
}


// batched + adjoint mode: candidate stage with batched CCCs, then lock-step continuous refinement of every
// seed, then one constrained CCC per refined pose for the reported distance (the continuous pose is kept)
void refine_adjoint_batch(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n,
                          std::vector<SearchOut*>& so, ettk_result* out) {
    std::vector<std::vector<Mat4>> seeds(n);
    KernelProf* kp = P->ctx->fft.prof;
    auto tc0 = std::chrono::steady_clock::now();
    {
        std::vector<std::vector<Cand>> sel;
        std::vector<int> nccc;
        std::vector<std::vector<std::pair<Mat4, float>>> evald;
        // the refinement's particle spectra depend on the particles only: enqueued while the host eliminates the
        // equivalent candidates
        const std::function<void()> spectra = [&] { refine_spectra_ahead(P, d_parts, 0, n); };
        peak_rounds(P, d_parts, descs, n, so, sel, nccc, &evald, &spectra);
        // refine(t, 0) of every selected candidate: the constrained CCC at the candidate's rotation, which
        // peak_rounds has already evaluated for a candidate that came from an SO(3) peak (the same transform, the
        // same kernels: the same distance and shift); the others (the identity placeholder) in one batch.  The
        // evaluation counts stay those of the reference method (final_candidates counts this one).
        std::vector<int> pidx, slot;
        std::vector<Mat4> t, tb;
        std::vector<float> d;
        for (int i = 0; i < n; i++)
            for (const Cand& p : sel[i]) {
                if (p.k >= 0) {
                    t.push_back(evald[i][p.k].first);
                    d.push_back(evald[i][p.k].second);
                } else {
                    slot.push_back((int)t.size());
                    pidx.push_back(i);
                    tb.push_back(strip_translation(zyz(p.a, p.b, p.g)));
                    t.push_back(Mat4::ident());
                    d.push_back(0.0f);
                }
            }
        if (!pidx.empty()) {
            std::vector<float> db(pidx.size());
            ccc_batch(P, d_parts, descs, (int)pidx.size(), pidx.data(), tb.data(), db.data());
            for (size_t q = 0; q < pidx.size(); q++) {
                t[slot[q]] = tb[q];
                d[slot[q]] = db[q];
            }
        }
        std::vector<size_t> first(n + 1, 0);  // the particle's first entry in t and d
        for (int i = 0; i < n; i++) first[i + 1] = first[i] + sel[i].size();
        parallel_for(n, [&](int i) {  // particles are independent
            size_t k = first[i];
            EvalSink sink;
            sink.direct = [&](int, Mat4& tt) {
                tt = t[k];
                return d[k++];
            };
            Task<void> task = final_candidates(P, i, sel[i], *so[i], &sink, out[i], nccc[i], &seeds[i]);
            run_direct(task);
        });
    }
    std::vector<int> pidx, cand;
    std::vector<Mat4> sd;
    for (int i = 0; i < n; i++)
        for (size_t c = 0; c < seeds[i].size(); c++) {
            pidx.push_back(i);
            cand.push_back((int)c);
            sd.push_back(seeds[i][c]);
        }
    const int nj = (int)sd.size();
    if (nj == 0) return;
    std::vector<Mat4> pose(nj), fin(nj);
    std::vector<double> score(nj);
    auto tr0 = std::chrono::steady_clock::now();
    if (kp) kp->add_host("candidate_stage_wall", std::chrono::duration<double, std::milli>(tr0 - tc0).count());
    // the reported distance: one CCC per refined pose, begun for each refinement chunk as soon as its poses are
    // final (alternating lanes), so that it runs while the other chunks finish their last steps
    std::vector<float> dist(nj);
    const bool lanes = ccc_lanes_available(P);
    int next_lane = 0;
    const std::function<void(int, int)> final_ccc = [&](int q0, int q1) {
        for (int q = q0; q < q1; q++) fin[q] = pose[q];
        const int h = next_lane;
        next_lane ^= 1;
        ccc_batch_end(P, h);
        ccc_batch_begin(P, h, d_parts, descs, q1 - q0, pidx.data() + q0, fin.data() + q0, dist.data() + q0);
    };
    try {
        refine_with_seeds(P, d_parts, descs, nj, pidx.data(), sd.data(), 20, pose.data(), score.data(),
                          lanes ? &final_ccc : nullptr);
    } catch (...) {
        (void)hipStreamSynchronize(P->stream);
        P->lane[0].pending = P->lane[1].pending = false;
        throw;
    }
    if (kp)
        kp->add_host("refine_wall",
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr0).count());
    if (lanes) {
        ccc_batch_end(P, 0);
        ccc_batch_end(P, 1);
    } else {
        fin = pose;
        ccc_batch(P, d_parts, descs, nj, pidx.data(), fin.data(), dist.data());
    }
    std::vector<float> best(n, FLT_MAX);
    for (int j = 0; j < nj; j++) {
        int i = pidx[j], c = cand[j];
        out[i].cand_distance[c] = dist[j];
        std::memcpy(out[i].cand_matrix[c], pose[j].a, sizeof(pose[j].a));
        out[i].n_ccc++;
        if (dist[j] < best[i]) {
            best[i] = dist[j];
            out[i].distance = dist[j];
            std::memcpy(out[i].matrix, pose[j].a, sizeof(pose[j].a));
        }
    }
}

// nbfCorrelationImageMetric::getDistances worker semantics (CorrMetric.h:1911-1990) for one reference:
//   type 0: S^2/SO(3) search + refinement (refinementOn), type 3: search without refinement (refinementOff);
//   type 1: nbfFourierImageMetric::execute (FIM.h:1812-1840): refine from the volume's stored transform, with
//           the volume's own transform set to NULL while it runs;
//   type 2: nbfFourierImageMetricCore::execute (Core.h:35-66): one constrained CCC at the identity (the
//           volume's stored transform applies), the shifted transform kept only when it is valid.
// `form` selects plain (0), batched (1) or batched + adjoint refinement (2) execution.
void distances_run(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n, int type, int form,
                   ettk_result* out) {
    ETTK_REQUIRE(P->have_ref, ETTK_ERR_STATE, "reference not set");
    ETTK_REQUIRE(type >= 0 && type <= 3, ETTK_ERR_ARG, "ettk_distances_batch: type must be 0..3");
    ETTK_REQUIRE(form >= 0 && form <= 2, ETTK_ERR_ARG, "ettk_distances_batch: form must be 0, 1 or 2");
    if (type == 0 || type == 3) {
        const int saved = P->p.refine;
        P->p.refine = type == 0 ? 1 : 0;
        try {
            align_run(P, d_parts, descs, n, (type == 3 && form == 2) ? 1 : form, out);
        } catch (...) {
            P->p.refine = saved;
            throw;
        }
        P->p.refine = saved;
        return;
    }
    std::memset(out, 0, sizeof(ettk_result) * n);
    std::vector<int> pidx(n);
    for (int i = 0; i < n; i++) pidx[i] = i;
    if (type == 2) {
        std::vector<Mat4> t(n, Mat4::ident());
        std::vector<float> d(n);
        if (form == 0) {
            for (int i = 0; i < n; i++) ccc_batch(P, d_parts, descs, 1, &pidx[i], &t[i], &d[i]);
        } else {
            ccc_batch(P, d_parts, descs, n, pidx.data(), t.data(), d.data());
        }
        for (int i = 0; i < n; i++) {
            Mat4 r = transform_ok(t[i], P->trans_restrict, P->rot_restrict) ? t[i] : Mat4::ident();
            out[i].distance = d[i];
            out[i].cand_distance[0] = d[i];
            std::memcpy(out[i].matrix, r.a, sizeof(r.a));
            std::memcpy(out[i].cand_matrix[0], r.a, sizeof(r.a));
            out[i].n_cand = 1;
            out[i].n_ccc = 1;
        }
        set_overlaps(P, descs, n, out);
        return;
    }
    // type 1: the stored transform becomes the starting point, the volumes are resliced without it
    std::vector<ettk_particle> dz(descs, descs + n);
    std::vector<Mat4> t(n);
    for (int i = 0; i < n; i++) {
        t[i] = Mat4::from(descs[i].matrix);
        const Mat4 I = Mat4::ident();
        std::memcpy(dz[i].matrix, I.a, sizeof(I.a));
    }
    std::vector<float> dist(n);
    std::vector<int> nccc(n, 0);
    if (form == 2) {
        std::vector<Mat4> pose(n), fin(n);
        std::vector<double> score(n);
        refine_with_seeds(P, d_parts, dz.data(), n, pidx.data(), t.data(), 20, pose.data(), score.data());
        fin = pose;
        ccc_batch(P, d_parts, dz.data(), n, pidx.data(), fin.data(), dist.data());
        t = pose;
        for (int i = 0; i < n; i++) nccc[i] = 1;
    } else {
        EvalSink sink;
        auto run_one = [&](int i) -> Task<void> {
            dist[i] = co_await refine_powell(P, i, t[i], true, Evaluator{&sink, &nccc[i]});
        };
        if (form == 0) {
            sink.direct = [&](int pi, Mat4& tt) {
                float d;
                ccc_batch(P, d_parts, dz.data(), 1, &pi, &tt, &d);
                return d;
            };
            for (int i = 0; i < n; i++) {
                Task<void> task = run_one(i);
                run_direct(task);
            }
        } else {
            std::vector<Task<void>> tasks;
            tasks.reserve(n);
            for (int i = 0; i < n; i++) tasks.push_back(run_one(i));
            run_batched(P, d_parts, dz.data(), tasks, sink);
        }
    }
    for (int i = 0; i < n; i++) {
        out[i].distance = dist[i];
        out[i].cand_distance[0] = dist[i];
        std::memcpy(out[i].matrix, t[i].a, sizeof(t[i].a));
        std::memcpy(out[i].cand_matrix[0], t[i].a, sizeof(t[i].a));
        out[i].n_cand = 1;
        out[i].n_ccc = nccc[i];
    }
    set_overlaps(P, dz.data(), n, out);
}

}  // namespace ettk


// ============================================================== C ABI
using namespace ettk;

extern "C" {

int ettk_project_batch(ettk_plan* P, const float* d_particles, int n, const ettk_particle* descs, float* out_pp) {
    try {
        ETTK_REQUIRE(P && d_particles && descs && out_pp && n > 0, ETTK_ERR_ARG, "ettk_project_batch: bad argument");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        So3State& S = so3_state(P);
        const int B = P->p.B;
        S.Pp2.ensure((size_t)n * 4 * B * B);
        project_images(P, d_particles, descs, nullptr, n, false, S.Pp2.p);
        ETTK_HIP(hipMemcpyAsync(out_pp, S.Pp2.p, (size_t)n * 4 * B * B * sizeof(float), hipMemcpyDeviceToHost, P->stream));
        ETTK_HIP(hipStreamSynchronize(P->stream));
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

static int so3_correlate_entry(ettk_plan* P, int n, const double* sig, const double* pat, double* out_c, int scaled) {
    try {
        ETTK_REQUIRE(P && sig && pat && out_c && n > 0, ETTK_ERR_ARG, "ettk_so3_correlate_batch: bad argument");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        So3State& S = so3_state(P);
        const int B = P->p.B;
        const size_t n3 = (size_t)8 * B * B * B;
        S.C.ensure(n3 * n);
        soft_given(P, n, sig, pat, S.C.p, scaled);
        ETTK_HIP(hipMemcpyAsync(out_c, S.C.p, n3 * n * sizeof(double), hipMemcpyDeviceToHost, P->stream));
        ETTK_HIP(hipStreamSynchronize(P->stream));
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

int ettk_so3_correlate_batch(ettk_plan* P, int n, const double* sig, const double* pat, double* out_c) {
    return so3_correlate_entry(P, n, sig, pat, out_c, 0);
}

int ettk_so3_correlate_scaled_batch(ettk_plan* P, int n, const double* sig, const double* pat, double* out_c) {
    return so3_correlate_entry(P, n, sig, pat, out_c, 1);
}

int ettk_sph_rotate_batch(ettk_plan* P, int n, const float* images, const double* transforms, float* out) {
    try {
        ETTK_REQUIRE(P && images && transforms && out && n > 0, ETTK_ERR_ARG, "ettk_sph_rotate_batch: bad argument");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        so3_state(P);
        sph_rotate(P, n, images, transforms, out);
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

int ettk_distances_batch(ettk_plan* P, const float* d_particles, int n, const ettk_particle* descs, int type,
                         int form, ettk_result* out) {
    try {
        ETTK_REQUIRE(P && d_particles && descs && out && n > 0, ETTK_ERR_ARG, "ettk_distances_batch: bad argument");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        distances_run(P, d_particles, descs, n, type, form, out);
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

int ettk_align_batch(ettk_plan* P, const float* d_particles, int n, const ettk_particle* descs, int mode,
                     ettk_result* out) {
    try {
        ETTK_REQUIRE(P && d_particles && descs && out && n > 0, ETTK_ERR_ARG, "ettk_align_batch: bad argument");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        align_run(P, d_particles, descs, n, mode, out);
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

}  // extern "C"
