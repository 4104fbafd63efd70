// plan.hpp -- host-side state of libettk_align: device context, per-size plan, FP64 transform helpers.
#pragma once
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>

#include "kernels.hpp"

namespace ettk {

// ------------------------------------------------------------------ 4x4 (vtkTransform semantics, FP64)
struct Mat4 {
    double a[16];
    double& operator()(int r, int c) { return a[r * 4 + c]; }
    double operator()(int r, int c) const { return a[r * 4 + c]; }
    static Mat4 ident() {
        Mat4 M;
        for (int i = 0; i < 16; i++) M.a[i] = (i % 5 == 0) ? 1.0 : 0.0;
        return M;
    }
    static Mat4 from(const double* p) {
        Mat4 M;
        std::memcpy(M.a, p, sizeof(M.a));
        return M;
    }
};
Mat4 operator*(const Mat4& A, const Mat4& B);
Mat4 mat_rot(double deg, double x, double y, double z);  // vtkTransformConcatenation::Rotate (quaternion)
Mat4 mat_translate(double x, double y, double z);
Mat4 mat_scale(double x, double y, double z);
Mat4 mat_inverse(const Mat4& A);
double mat_det(const Mat4& A);
Mat4 strip_translation(const Mat4& t);
Mat4 zyz(double a, double b, double g);  // RotateZ(-g) RotateY(-b) RotateZ(-a), pre-multiplied
bool transform_ok(const Mat4& t, double trans_restrict, double rot_restrict);  // nbfImageMetric::isTransformValid
Mat4 euler_of(const ettk_particle& d);   // RotateZ(-n2) RotateX(-n0) RotateZ(-n1)
// separable vtkImageGaussianSmooth on an A(i,j,k) float array (axes 2, 1, 0; clipped, renormalised kernel)
void smooth3(std::vector<float>& A, int n0, int n1, int n2, double sd, double rf);

// ------------------------------------------------------------------ state
struct So3State;
// one eagerly launched CCC batch: its geometry (host, pinned, and device), its results (device, and pinned host),
// and the caller's arrays that ccc_batch_end fills
struct CccLane {
    DevBuf<EvalGeom> geom_all;
    PinnedBuf<EvalGeom> h_geom;
    DevBuf<float> out_all;
    PinnedBuf<float> h_out;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, done = nullptr;
    bool pending = false;
    int n = 0;
    Mat4* t = nullptr;
    float* dist = nullptr;
    double t_begin = 0;  // host clock at ccc_batch_begin (ms)
    CccLane() = default;
    CccLane(const CccLane&) = delete;
    CccLane& operator=(const CccLane&) = delete;
    ~CccLane() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (done) (void)hipEventDestroy(done);
    }
};

// One chunk of evaluations (at most 128, the CCC workspace) of the score and gradient: its geometry and results in
// pinned host memory (asynchronous copies), the host matrices its gradient needs, and the event that marks its
// kernels and downloads done.  Chunks share the device workspaces in stream order; their geometry, score and
// moment regions of the device arrays are their own.
struct RefineChunk {
    int e0 = 0, E = 0;
    PinnedBuf<RefineGeom> hg;
    PinnedBuf<double> hsc, hmom;  // 4 and 12 per evaluation
    std::vector<Mat4> Ginv, EMs;
    hipEvent_t done = nullptr;   // scores and moments downloaded
    hipEvent_t kdone = nullptr;  // the chunk's kernels finished (its downloads run on the copy stream after it)
    RefineChunk() = default;
    RefineChunk(const RefineChunk&) = delete;
    RefineChunk& operator=(const RefineChunk&) = delete;
    ~RefineChunk() {
        if (done) (void)hipEventDestroy(done);
        if (kdone) (void)hipEventDestroy(kdone);
    }
};
// a captured CCC chunk pipeline (E evaluations, particle base pointer, workspace generation)
struct CccGraph {
    int E = 0;
    const float* parts = nullptr;
    int gen = -1;
    hipGraphExec_t exec = nullptr;
};
struct Timing {
    double ccc_ms = 0, total_ms = 0;
    long long n_ccc = 0;
};

}  // namespace ettk

struct ettk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // host downloads issued from worker threads (non-blocking)
    ettk::FftTables fft;
    void* scratch = nullptr;  // device address of this device's fixed CCC scratch region (noalloc.hip)
    struct ettk_plan_cache* plans = nullptr;  // plans kept for ettk_align_buckets (tune.cpp)
};

struct ettk_plan {
    ettk_ctx* ctx = nullptr;
    // the plan's own streams (default priority): plans of one context run concurrently from different host threads
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // host downloads issued from worker threads (non-blocking)
    hipStream_t aux_stream = nullptr;   // the candidate extraction of a search chunk, beside the next chunk's kernels
    // uploads of per-chunk geometry, so that they run beside the previous chunk's kernels instead of between them
    hipStream_t up_stream = nullptr;
    hipEvent_t ev_up = nullptr;
    ettk_params p{};
    int N = 0, H = 0;
    bool dummy_filter = true;
    double sum_filter = 0;
    double trans_restrict = 0, rot_restrict = 180;
    ettk::CccSizes z{};
    // constant device data
    ettk::DevBuf<float> maskfft, m1, filter, W1;
    ettk::DevBuf<float2> Mfft, F1mcf, tabA, tabK;
    std::vector<float> h_m1;
    // reference
    bool have_ref = false;
    ettk_particle refd{};
    ettk::DevBuf<float> ref;
    // CCC workspace
    int cap = 0, ws_E = 0;
    int nb_real = 0, nb_spec = 0, nb_res = 0;
    ettk::DevBuf<float> R1, R2, p_res, p_n1, p_n2, p_k1, p_wf, stats, scalars, bval, out4;
    ettk::DevBuf<int> bidx;
    ettk::DevBuf<float2> S1, S2, S3, G1;
    ettk::DevBuf<ettk::EvalGeom> geom;
    ettk::PinnedBuf<ettk::EvalGeom> h_geom;  // host geometry of a graph-replayed CCC batch
    ettk::CccLane lane[2];                   // eagerly launched CCC batches in flight (ccc_batch_begin / _end)
    std::vector<std::unique_ptr<ettk::RefineChunk>> refine_pool;  // refinement chunks (refine.cpp)
    // wedge overlaps (wedge_overlaps): geometry and per-block sums, kept between calls
    ettk::DevBuf<ettk::EvalGeom> ov_geom;
    ettk::DevBuf<double> ov_part, ov_sums;
    // execution choices (ettk_plan_tune): evaluations per CCC chunk, chunks replayed from captured graphs
    int ccc_chunk = 256;
    bool use_graphs = false;
    int ws_gen = 0;  // bumped whenever a buffer or constant a captured chunk reads is replaced
    std::vector<ettk::CccGraph> graphs;
    std::vector<int> graph_warm;  // chunk sizes already run eagerly on this workspace generation
    ettk::PinnedBuf<float> h_out_pinned;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ettk::Timing timing;
    // S^2 / SO(3) state (so3.hip)
    ettk::So3State* so3 = nullptr;
    // averaging accumulators
    ettk::DevBuf<double> avg_re, avg_im, avg_acc;
    // continuous refinement (refine.cpp): reference minus its mean (x fastest), its wedge normals, particle
    // spectra and per-evaluation workspace
    ettk::DevBuf<float> ref_x;
    ettk::WedgeRef wref{};
    // accumulated wedge of an average (WAvg.h:49-70, 496-538, 654-667), full grid A(i,j,k) with k fastest:
    // sum, binarised sum, smoothed binarised sum and its half grid; the binarised spherical sum (2B x 2B).
    // ref_wedge_kind: 0 = the reference's wedge is analytic (refd), 1 = accumulated and effective,
    // 2 = accumulated and not effective (all ones)
    int ref_wedge_kind = 0;
    bool wacc_ready = false, wacc_effective = false;
    ettk::DevBuf<float> wacc, wbin, wsm, wsm_half;
    ettk::DevBuf<int> wcount;
    std::vector<float> sph_acc_bin;
    ettk::DevBuf<float> ref_wbin;  // the reference's copy of wbin (kind 1), read by the refinement
    ettk::DevBuf<float2> r_fp;
    const float* r_fp_parts = nullptr;  // r_fp holds the spectra of particles [r_fp_p0, r_fp_p1) of this array,
    int r_fp_p0 = 0, r_fp_p1 = 0;       // enqueued ahead of the refinement that uses them (refine_spectra_ahead)
    ettk::DevBuf<float> r_mean, r_mpart;
    ettk::DevBuf<double> r_pdot, r_sc, r_mom;
    ettk::DevBuf<ettk::RefineGeom> r_geom;
};

namespace ettk {
// device workspace of one chunk of E constrained-CCC evaluations: the plan's own buffers (ccc_batch) or a
// caller-independent fixed region (ettk_ccc_eval_batch_noalloc)
struct CccWork {
    float *R1, *R2, *p_res, *p_n1, *p_n2, *p_wf, *p_k1, *stats, *scalars, *bval;
    int* bidx;
    float2 *S1, *S2, *S3, *G1;
};
// device bytes of a CccWork for E evaluations, and its carving out of one region (256-byte aligned pieces)
size_t ccc_work_bytes(const ettk_plan* P, int E);
CccWork ccc_work_carve(const ettk_plan* P, int E, void* base);
// per-evaluation geometry of the reslice and the particle wedge (FP64 on the host)
EvalGeom make_geom_public(const ettk_plan* P, const ettk_particle& d, const Mat4& t, int particle);
// the kernels of one chunk of E evaluations: geometry geo[E] on the device, results out4[E][4] (distance, shift)
void ccc_enqueue_chunk(ettk_plan* P, const CccWork& W, const float* d_parts, const EvalGeom* geo, int E, float* out4,
                       hipStream_t s);
// device address of the fixed CCC scratch region of the context's device into ctx->scratch (noalloc.hip)
void noalloc_scratch_init(ettk_ctx* ctx);
// the context's plans for ettk_align_buckets (created on first use) and their release (tune.cpp)
struct ettk_plan_cache* ctx_plan_cache(ettk_ctx* ctx);
void ctx_plan_cache_destroy(ettk_ctx* ctx);
// applies the recorded tuner choice for the plan's shape, when the tuning cache holds one (tune.cpp)
bool tune_lookup(ettk_plan* P);
// drops the captured chunk pipelines (after a workspace, reference or restriction change)
void ccc_graphs_reset(ettk_plan* P);
// constrained CCC for a batch of evaluations: on return t[e] <- t[e] * Translate(-shift), dist[e].  Evaluation e
// reads volume pidx[e] with descriptor descs[pidx[e]], or eval_descs[e] when eval_descs is given.
void ccc_batch(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n, const int* pidx,
               Mat4* t, float* dist, const ettk_particle* eval_descs = nullptr);
// the same batch split in two, on lane 0 or 1 of the plan (eager launches): _begin enqueues it and returns at
// once; _end waits for it and fills t and dist.  Two lanes in flight let the host finish one batch's results while
// the GPU runs the other's kernels; t, dist, pidx and the descriptors must stay valid until _end.
bool ccc_lanes_available(const ettk_plan* P);
void ccc_batch_begin(ettk_plan* P, int lane, const float* d_parts, const ettk_particle* descs, int n, const int* pidx,
                     Mat4* t, float* dist, const ettk_particle* eval_descs = nullptr);
void ccc_batch_end(ettk_plan* P, int lane);
// continuous refinement of n (particle, seed) jobs, n_iter ascent steps each (refine.cpp); on_done, when given, is
// called with [q0, q1) as soon as those jobs' poses and scores are final (before the call returns)
void refine_with_seeds(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n, const int* pidx,
                       const Mat4* seeds, int n_iter, Mat4* out_pose, double* out_score,
                       const std::function<void(int, int)>* on_done = nullptr);
// the particle spectra the refinement of particles [p0, p1) needs, enqueued ahead of it (it then skips them)
void refine_spectra_ahead(ettk_plan* P, const float* d_parts, int p0, int p1);
void ensure_ccc_workspace(ettk_plan* P, int cap);
// host-to-device copy on the plan's upload stream, ordered before the work enqueued next on the plan's stream
// (src must be pinned and stay untouched until that work has started)
void upload_before(ettk_plan* P, void* dst, const void* src, size_t bytes);
// getWedgeOverlap of the reference against descs[i] at t[i] (CorrMetric.h:533-566)
void wedge_overlaps(ettk_plan* P, const ettk_particle* descs, const Mat4* t, int n, float* out);
void so3_state_destroy(ettk_plan* P);
void so3_reference_set(ettk_plan* P);
void wedge_normals(float lw, float uw, const Mat4& t, WedgeRef& w);
std::vector<float> spherical_wedge_bin(int B, const ettk_particle& d);
// accumulated wedge of the entries with nonzero weight into P->wacc/wbin/wsm/wsm_half/sph_acc_bin
void avg_wedge_compute(ettk_plan* P, int n, const ettk_particle* descs, const float* weights, const double* matrices);
}  // namespace ettk
