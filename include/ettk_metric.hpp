/* ettk_metric.hpp -- C++ drop-in for the alignment path of nbfCorrelationImageMetric<float,3>
 * (inc/em/nbfCorrelationImageMetric.h) on top of the C ABI of libettk_align (ettk_align.h).
 *
 * The class keeps the reference's method names, argument meaning and result layout so that a caller such as
 * MPI_Classification.cpp:954/1087 or nbfLoopClustering.h:1560-1563 switches by changing the type it
 * instantiates:
 *   - volumes are device arrays (N^3 float, x fastest, the VTK/MRC order) plus the sub-volume descriptor
 *     (wedge, Euler normal, cut offset, stored alignment) that nbfWedgedSubImage3D carries;
 *   - getDistances(refs, vols, D, type) fills D(i, j, 0..16, k) = {distance, M00..M33} of candidate k,
 *     computing only the entries whose D(i, j, 0, 0) == -1 (resume semantics, CorrMetric.h:1950);
 *   - getImage(average) builds the amplitude-weighted Fourier average of the average's volumes under their
 *     alignments [weight, M16] (WAvg.h:182-214, 261-356);
 *   - mraAssignment(...) is the best-reference assignment of nbfLoopClustering::doMra (:1523-1592).
 * Errors: the reference prints to cerr and returns sentinels; here every failure of the C ABI throws
 * ettk::MetricError carrying the code and ettk_last_error() text (no sentinel distances are invented).
 * Device memory: the metric never takes ownership of caller volumes; when the volumes of one call are not
 * contiguous in one allocation it gathers them into a scratch buffer it owns.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "ettk_align.h"

namespace ettk {

struct MetricError : std::runtime_error {
    int code;
    MetricError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char* what) {
    if (rc != ETTK_OK) throw MetricError(rc, std::string(what) + ": " + ettk_last_error());
}
inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw MetricError(ETTK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

/* nbfWedgedSubImage3D<float>: one particle held as a fixed image. */
struct WedgedSubImage3D {
    const float* d_volume = nullptr;  // device, N^3 floats, x fastest
    ettk_particle desc{-90.0f, 90.0f, {0, 0, 0}, 0.0f, {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};

    void setWedge(float lwedge, float uwedge) {  // nbfWedge3D::set (Wedge.h:143-184)
        desc.lwedge = lwedge;
        desc.uwedge = uwedge;
    }
    void setTransform(const double* m16) {  // NULL resets to the identity
        static const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        std::memcpy(desc.matrix, m16 ? m16 : I, sizeof(desc.matrix));
    }
    void getTransform(double* m16) const { std::memcpy(m16, desc.matrix, sizeof(desc.matrix)); }
    void setCutOffset(float c) { desc.cut_offset = c; }
    void setNormal(float n0, float n1, float n2) {
        desc.normal[0] = n0;
        desc.normal[1] = n1;
        desc.normal[2] = n2;
    }
};

/* nbfWedgedAverageImage3D<float>: a reference.  `d_average` holds the current average (input of the
   alignments, output of getImage); `volumes` and `alignments` ([n][17][K]: weight, M16 per candidate,
   setAlignments WAvg.h:182-214) define the average getImage builds. */
struct WedgedAverageImage3D {
    float* d_average = nullptr;  // device, N^3 floats, x fastest
    ettk_particle desc{-90.0f, 90.0f, {0, 0, 0}, 0.0f, {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
    std::vector<WedgedSubImage3D>* volumes = nullptr;
    std::vector<float> alignments;  // n * 17 * K
    int K = 1;

    void setVolumes(std::vector<WedgedSubImage3D>* v) { volumes = v; }
    void setAlignments(const std::vector<float>& a, int k) {
        alignments = a;
        K = k;
    }
    float& alignment(int i, int c, int k) { return alignments[((size_t)i * 17 + c) * K + k]; }
};

/* blitz::Array<float, 4> D(nRef, nVol, depth, K) of getDistances; depth 17 = {distance, M00..M33}. */
class DistanceArray {
   public:
    DistanceArray(int rows, int cols, int depth = 17, int K = 1)
        : rows_(rows), cols_(cols), depth_(depth), K_(K), v_((size_t)rows * cols * depth * K, -1.0f) {}
    float& operator()(int i, int j, int c, int k) { return v_[(((size_t)i * cols_ + j) * depth_ + c) * K_ + k]; }
    float operator()(int i, int j, int c, int k) const { return v_[(((size_t)i * cols_ + j) * depth_ + c) * K_ + k]; }
    int rows() const { return rows_; }
    int cols() const { return cols_; }
    int depth() const { return depth_; }
    int extent3() const { return K_; }
    const float* data() const { return v_.data(); }  // [i][j][c][k], k fastest

   private:
    int rows_, cols_, depth_, K_;
    std::vector<float> v_;
};

class CorrelationImageMetric {
   public:
    /* The plan is built for box size N; B is the spherical-harmonic bandwidth (the reference's public B). */
    CorrelationImageMetric(int N, int B = 32, int device = 0) : B(B) {
        check(ettk_ctx_create(device, &ctx_), "ettk_ctx_create");
        p_.N = N;
        p_.B = B;
        p_.bp_lo = 0.0f;
        p_.bp_hi = 1.0f;
        p_.bp_lo_decay = p_.bp_hi_decay = 0.01f;
        p_.bfactor = 0.0f;
        p_.mask_radius = 0.0f;
        p_.mask_sigma = 0.0f;
        p_.square_mask = 0;
        p_.mcf = 0;
        p_.overlap_norm = 0;
        p_.missing_wedge = 0;
        p_.n_peaks = 5;  // ProjRot.h:209
        p_.n_candidates = 1;  // CorrMetric.h:185
        p_.rot_restrict = 180.0f;
        p_.trans_restrict = -1.0f;
        p_.refine = 1;
    }
    ~CorrelationImageMetric() {
        if (plan_) ettk_plan_destroy(plan_);
        if (scratch_) (void)hipFree(scratch_);
        if (ctx_) ettk_ctx_destroy(ctx_);
    }
    CorrelationImageMetric(const CorrelationImageMetric&) = delete;
    CorrelationImageMetric& operator=(const CorrelationImageMetric&) = delete;

    int B;

    // ---------------------------------------------------------------- configuration (reference names)
    void setRotationSearchRestriction(float deg) { set_(p_.rot_restrict, std::max(deg, 1.0f), true); }
    void setTranslationSearchRestriction(float vox) { set_(p_.trans_restrict, vox, true); }
    float getTranslationSearchRestriction() const { return p_.trans_restrict; }
    float getRotationSearchRestriction() const { return p_.rot_restrict; }
    void setToUseMutualCorrelation(bool on) { set_(p_.mcf, on ? 1 : 0); }
    void setToComputeOverlapNormalizedDistances(bool on) { set_(p_.overlap_norm, on ? 1 : 0); }
    void setMissingWedgeCompensation(bool on) { set_(p_.missing_wedge, on ? 1 : 0); }
    void setNumberOfCandidatePeaksToSearch(int k) { set_(p_.n_peaks, k); }
    void setNumberOfCandidates(int k) { set_(p_.n_candidates, k); }
    int getNumberOfCandidates() const { return p_.n_candidates; }
    void refinementOn() { set_(p_.refine, 1); }
    void refinementOff() { set_(p_.refine, 0); }
    // nbfFourierFilter::bandPassOn (FFilt.h:700-712) and bfactorOn
    void bandPassOn(float lo, float hi, float lo_decay, float hi_decay) {
        set_(p_.bp_lo, lo);
        set_(p_.bp_hi, hi);
        set_(p_.bp_lo_decay, lo_decay);
        set_(p_.bp_hi_decay, hi_decay);
    }
    void bfactorOn(float b) { set_(p_.bfactor, b); }
    // nbfImageFilter::maskOn with an analytic spherical mask (IFilt.h:384-472) and the square FFT window size
    void maskOn(float radius, float sigma) {
        set_(p_.mask_radius, radius);
        set_(p_.mask_sigma, sigma);
    }
    void setSquareMaskSize(int sq) { set_(p_.square_mask, sq); }
    /* Execution form, no reference counterpart: 0 plain (one evaluation per launch, the reference's control
       flow), 1 batched, 2 batched with adjoint-gradient refinement. */
    void setExecutionForm(int form) { form_ = form; }

    // ---------------------------------------------------------------- batch alignment
    /* getDistances(refs, vols, D, type) (CorrMetric.h:659-704): type 0 search + refine, 1 refine only,
       2 translation only, 3 search without refinement. */
    void getDistances(std::vector<WedgedAverageImage3D>& refs, std::vector<WedgedSubImage3D>& vols,
                      DistanceArray& D, int type) {
        if (D.rows() != (int)refs.size() || D.cols() != (int)vols.size())
            throw MetricError(ETTK_ERR_ARG, "getDistances: D must be refs x vols");
        if (D.depth() < 17 || D.depth() > 20)
            throw MetricError(ETTK_ERR_ARG, "getDistances: D rows hold 17 to 20 columns");
        // column layout of CorrMetric.h:1966-2000: distance, [correlation scale], [wedge overlap], M16
        const int off = 1 + (D.depth() > 17) + (D.depth() > 18);
        const int K = D.extent3();
        if (K < 1 || K > ETTK_MAX_CANDIDATES) throw MetricError(ETTK_ERR_ARG, "getDistances: bad candidate count");
        if (type != 1 && type != 2 && K != p_.n_candidates)
            throw MetricError(ETTK_ERR_ARG, "getDistances: D's last extent must equal the number of candidates");
        ensure_plan_();
        for (int i = 0; i < (int)refs.size(); i++) {
            std::vector<int> todo;
            for (int j = 0; j < (int)vols.size(); j++)
                if (D(i, j, 0, 0) == -1.0f) todo.push_back(j);
            if (todo.empty()) continue;
            set_reference_(refs[i]);
            std::vector<ettk_particle> descs;
            const float* base = gather_(vols, todo, descs);
            std::vector<ettk_result> res(todo.size());
            check(ettk_distances_batch(plan_, base, (int)todo.size(), descs.data(), type, form_, res.data()),
                  "ettk_distances_batch");
            for (size_t q = 0; q < todo.size(); q++) store_row_(D, i, todo[q], res[q], K, off);
        }
    }

    /* getDistances(list, D, type) (CorrMetric.h:570-612): the symmetric matrix of a list of sub-volumes.  The
       entries (i, j), i < j, whose D(i, j, 0, 0) == -1 are computed with list[i] as the reference (input1) and
       list[j] as the moving volume (input2); makeDistanceMatrix then completes the matrix.  The reference side is
       list[i]'s volume as stored with its wedge at its alignment; the reference reslices input1 through its
       alignment, so the entries used as references must hold the identity (as sub-volumes do before any
       alignment), otherwise MetricError(ETTK_ERR_UNSUPPORTED).  D is n x n x depth x K with depth 19 or 20
       (distance, correlation scale, wedge overlap, M16: the layout makeDistanceMatrix reads). */
    void getDistances(std::vector<WedgedSubImage3D>& list, DistanceArray& D, int type) {
        const int n = (int)list.size();
        if (D.rows() != n || D.cols() != n) throw MetricError(ETTK_ERR_ARG, "getDistances: D must be n x n");
        if (D.depth() < 19 || D.depth() > 20)
            throw MetricError(ETTK_ERR_ARG, "getDistances: the symmetric matrix holds rows of 19 or 20 columns");
        const int K = D.extent3();
        if (K < 1 || K > ETTK_MAX_CANDIDATES) throw MetricError(ETTK_ERR_ARG, "getDistances: bad candidate count");
        if (type != 1 && type != 2 && K != p_.n_candidates)
            throw MetricError(ETTK_ERR_ARG, "getDistances: D's last extent must equal the number of candidates");
        ensure_plan_();
        for (int i = 0; i < n; i++) {
            std::vector<int> todo;
            for (int j = i + 1; j < n; j++)
                if (D(i, j, 0, 0) == -1.0f) todo.push_back(j);
            if (todo.empty()) continue;
            static const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
            if (std::memcmp(list[i].desc.matrix, I, sizeof(I)) != 0)
                throw MetricError(ETTK_ERR_UNSUPPORTED,
                                  "getDistances: a sub-volume used as the reference must hold the identity alignment");
            check(ettk_reference_set(plan_, list[i].d_volume, &list[i].desc), "ettk_reference_set");
            std::vector<ettk_particle> descs;
            const float* base = gather_(list, todo, descs);
            std::vector<ettk_result> res(todo.size());
            check(ettk_distances_batch(plan_, base, (int)todo.size(), descs.data(), type, form_, res.data()),
                  "ettk_distances_batch");
            for (size_t q = 0; q < todo.size(); q++) store_row_(D, i, todo[q], res[q], K, 3);
        }
        makeDistanceMatrix(D);
    }

    /* makeDistanceMatrix (CorrMetric.h:2291-2328): (j, i) receives the distance, correlation scale and overlap
       of (i, j) and the inverse of each of its candidate transforms; the diagonal is 0 with the identity. */
    static void makeDistanceMatrix(DistanceArray& D) {
        if (D.rows() != D.cols() || D.depth() < 19)
            throw MetricError(ETTK_ERR_ARG, "makeDistanceMatrix: D must be n x n with at least 19 columns");
        const int n = D.rows(), K = D.extent3();
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++)
                for (int k = 0; k < K; k++) {
                    for (int c = 0; c < 3; c++) D(j, i, c, k) = D(i, j, c, k);
                    double m[16], inv[16];
                    for (int c = 0; c < 16; c++) m[c] = D(i, j, 3 + c, k);
                    invert4(m, inv);
                    for (int c = 0; c < 16; c++) D(j, i, 3 + c, k) = (float)inv[c];
                }
        for (int i = 0; i < n; i++)
            for (int k = 0; k < K; k++) {
                for (int c = 0; c < D.depth(); c++) D(i, i, c, k) = 0.0f;
                for (int c = 0; c < 16; c++) D(i, i, 3 + c, k) = (c % 5 == 0) ? 1.0f : 0.0f;
            }
    }

    // ---------------------------------------------------------------- averaging
    /* getImage(average) (CorrMetric.h:803-984 with WAvg.h:261-356): sum over volumes and candidates with
       non-zero weight of the amplitude-weighted spectra, then the RMS-normalised average into d_average. */
    void getImage(WedgedAverageImage3D& avg) {
        if (!avg.volumes || avg.volumes->empty()) throw MetricError(ETTK_ERR_ARG, "getImage: no volumes");
        const int n = (int)avg.volumes->size();
        if (avg.alignments.size() != (size_t)n * 17 * avg.K) throw MetricError(ETTK_ERR_ARG, "getImage: alignments");
        ensure_plan_();
        std::vector<int> all(n);
        for (int j = 0; j < n; j++) all[j] = j;
        std::vector<ettk_particle> descs;
        const float* base = gather_(*avg.volumes, all, descs);
        check(ettk_avg_reset(plan_), "ettk_avg_reset");
        std::vector<float> w(n);
        std::vector<double> m((size_t)n * 16);
        for (int k = 0; k < avg.K; k++) {
            bool any = false;
            for (int j = 0; j < n; j++) {
                w[j] = avg.alignment(j, 0, k);
                any = any || w[j] != 0.0f;
                for (int c = 0; c < 16; c++) m[(size_t)j * 16 + c] = avg.alignment(j, 1 + c, k);
            }
            if (any) check(ettk_avg_accumulate(plan_, base, n, descs.data(), w.data(), m.data()), "ettk_avg_accumulate");
        }
        check(ettk_avg_finalize(plan_, avg.d_average), "ettk_avg_finalize");
    }

    /* getImageVariance(average) (CorrMetric.h:987-1003, WAvg.h:440-478): the weighted variance of the aligned
       volumes around avg.d_average (computed by getImage first), into d_variance (N^3 floats, x fastest). */
    void getImageVariance(WedgedAverageImage3D& avg, float* d_variance) {
        if (!avg.volumes || avg.volumes->empty()) throw MetricError(ETTK_ERR_ARG, "getImageVariance: no volumes");
        const int n = (int)avg.volumes->size();
        if (avg.alignments.size() != (size_t)n * 17 * avg.K)
            throw MetricError(ETTK_ERR_ARG, "getImageVariance: alignments");
        ensure_plan_();
        std::vector<int> all(n);
        for (int j = 0; j < n; j++) all[j] = j;
        std::vector<ettk_particle> d0;
        const float* base = gather_(*avg.volumes, all, d0);
        // entries in the reference's (volume, candidate) order; the volumes are repeated per candidate
        std::vector<int> idx;
        std::vector<ettk_particle> descs;
        std::vector<float> w;
        std::vector<double> m;
        for (int j = 0; j < n; j++)
            for (int k = 0; k < avg.K; k++) {
                idx.push_back(j);
                descs.push_back(d0[j]);
                w.push_back(avg.alignment(j, 0, k));
                for (int c = 0; c < 16; c++) m.push_back(avg.alignment(j, 1 + c, k));
            }
        const float* vols = base;
        if (avg.K > 1) {
            std::vector<WedgedSubImage3D> rep;
            for (int j : idx) rep.push_back((*avg.volumes)[j]);
            std::vector<int> all2(rep.size());
            for (size_t q = 0; q < rep.size(); q++) all2[q] = (int)q;
            std::vector<ettk_particle> dd;
            vols = gather_(rep, all2, dd);
        }
        check(ettk_avg_variance(plan_, vols, (int)descs.size(), descs.data(), w.data(), m.data(), avg.d_average,
                                d_variance),
              "ettk_avg_variance");
    }

    // ---------------------------------------------------------------- multi-reference assignment
    /* nbfLoopClustering::doMra best-reference selection (nbfLoopClustering.h:1523-1592): each volume is
       evaluated under each of its alignments to the references (A(j, i, 1..16, 0)) against every reference
       at zero shift (type 2, translation restriction 0); the alignment whose distances sum lowest over the
       references is marked by negating A(j, i, 0, 0).  Returns the chosen alignment per volume. */
    std::vector<int> mraAssignment(std::vector<WedgedAverageImage3D>& refs, std::vector<WedgedSubImage3D>& vols,
                                   DistanceArray& A) {
        const int R = (int)refs.size(), n = (int)vols.size();
        if (A.rows() != R || A.cols() != n || A.depth() != 17)
            throw MetricError(ETTK_ERR_ARG, "mraAssignment: A must be refs x vols x 17");
        std::vector<int> best(n, 0);
        if (R < 2) return best;
        std::vector<WedgedSubImage3D> dup((size_t)n * R);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < R; j++) {
                double m[16];
                for (int c = 0; c < 16; c++) m[c] = A(j, i, 1 + c, 0);
                dup[(size_t)i * R + j] = vols[i];
                dup[(size_t)i * R + j].setTransform(m);
            }
        const float saved = p_.trans_restrict;
        setTranslationSearchRestriction(0.0f);
        DistanceArray E(R, n * R, 17, 1);
        try {
            getDistances(refs, dup, E, 2);
        } catch (...) {
            setTranslationSearchRestriction(saved);
            throw;
        }
        setTranslationSearchRestriction(saved);
        for (int i = 0; i < n; i++) {
            double bestv = 0;
            for (int j = 0; j < R; j++) {
                double s = 0;
                for (int r = 0; r < R; r++) s += E(r, i * R + j, 0, 0);
                if (j == 0 || s < bestv) {  // first minimum (blitz minIndex)
                    bestv = s;
                    best[i] = j;
                }
            }
            A(best[i], i, 0, 0) *= -1.0f;
        }
        return best;
    }

    ettk_plan* plan() {
        ensure_plan_();
        return plan_;
    }

   private:
    // one result into D(i, j, ., .): the layout of CorrMetric.h:1966-2000 (distance, [correlation scale],
    // [wedge overlap], M16 from column `off`)
    static void store_row_(DistanceArray& D, int i, int j, const ettk_result& r, int K, int off) {
        if (K == 1) {
            D(i, j, 0, 0) = r.distance;
            if (off > 1) D(i, j, 1, 0) = 1.0f;        // getCorrelationScale: 1 for these metrics
            if (off > 2) D(i, j, 2, 0) = r.overlap;   // getWedgeOverlap(getTransform())
            for (int c = 0; c < 16; c++) D(i, j, off + c, 0) = (float)r.matrix[c];
        } else {
            for (int k = 0; k < K; k++) D(i, j, 0, k) = std::numeric_limits<float>::max();
            for (int k = 0; k < std::min(K, r.n_cand); k++) {
                D(i, j, 0, k) = r.cand_distance[k];
                if (off > 1) D(i, j, 1, k) = 1.0f;    // candidate scales and overlaps are stored
                if (off > 2) D(i, j, 2, k) = 1.0f;    //   as 1 (ProjRot.h:879-880, FIM.h:1872-1876)
                for (int c = 0; c < 16; c++) D(i, j, off + c, k) = (float)r.cand_matrix[k][c];
            }
        }
    }
    // inverse of a row-major 4x4 matrix by cofactors (a singular matrix gives zeros, as vtkMatrix4x4::Invert
    // leaves its output unchanged at determinant 0)
    static void invert4(const double* a, double* out) {
        double c[16];
        c[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
        c[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
        c[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
        c[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
        c[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
        c[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
        c[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
        c[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
        c[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
        c[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
        c[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
        c[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
        c[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
        c[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
        c[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
        c[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
        const double det = a[0] * c[0] + a[1] * c[4] + a[2] * c[8] + a[3] * c[12];
        for (int q = 0; q < 16; q++) out[q] = det == 0.0 ? 0.0 : c[q] / det;
    }
    template <class T>
    void set_(T& field, T v, bool restriction = false) {
        if (field == v) return;
        field = v;
        if (!plan_) return;
        if (restriction) {
            check(ettk_plan_set_restrictions(plan_, p_.rot_restrict, p_.trans_restrict), "ettk_plan_set_restrictions");
        } else {
            ettk_plan_destroy(plan_);  // the masks and filters are plan constants: rebuilt on next use
            plan_ = nullptr;
        }
    }
    void ensure_plan_() {
        if (plan_) return;
        p_.B = B;
        check(ettk_plan_create(ctx_, &p_, &plan_), "ettk_plan_create");
    }
    // device base of the selected volumes, contiguous in the given order (gathered when they are not)
    /* setInput1(average): an average that carries its volumes and alignments uses their accumulated wedge
       (updateAccumulatedWedgeImage, WAvg.h:496-611, entries in (volume, candidate) order); one without them,
       or whose weights are all zero, uses its own descriptor. */
    void set_reference_(WedgedAverageImage3D& a) {
        const int n = a.volumes ? (int)a.volumes->size() : 0;
        if (n > 0 && a.alignments.size() == (size_t)n * 17 * a.K) {
            std::vector<ettk_particle> d;
            std::vector<float> w;
            std::vector<double> m;
            bool any = false;
            for (int j = 0; j < n; j++)
                for (int k = 0; k < a.K; k++) {
                    d.push_back((*a.volumes)[j].desc);
                    w.push_back(a.alignment(j, 0, k));
                    any = any || w.back() != 0.0f;
                    for (int c = 0; c < 16; c++) m.push_back(a.alignment(j, 1 + c, k));
                }
            if (any) {
                check(ettk_reference_set_average(plan_, a.d_average, (int)d.size(), d.data(), w.data(), m.data()),
                      "ettk_reference_set_average");
                return;
            }
        }
        check(ettk_reference_set(plan_, a.d_average, &a.desc), "ettk_reference_set");
    }

    const float* gather_(const std::vector<WedgedSubImage3D>& vols, const std::vector<int>& idx,
                         std::vector<ettk_particle>& descs) {
        const size_t n3 = (size_t)p_.N * p_.N * p_.N;
        descs.resize(idx.size());
        bool contiguous = true;
        for (size_t q = 0; q < idx.size(); q++) {
            descs[q] = vols[idx[q]].desc;
            if (vols[idx[q]].d_volume != vols[idx[0]].d_volume + q * n3) contiguous = false;
        }
        if (contiguous) return vols[idx[0]].d_volume;
        const size_t bytes = idx.size() * n3 * sizeof(float);
        if (bytes > scratch_bytes_) {
            if (scratch_) check_hip(hipFree(scratch_), "hipFree");
            scratch_ = nullptr;
            check_hip(hipMalloc(&scratch_, bytes), "hipMalloc");
            scratch_bytes_ = bytes;
        }
        for (size_t q = 0; q < idx.size(); q++)
            check_hip(hipMemcpy(scratch_ + q * n3, vols[idx[q]].d_volume, n3 * sizeof(float), hipMemcpyDeviceToDevice),
                      "hipMemcpy");
        return scratch_;
    }

    ettk_ctx* ctx_ = nullptr;
    ettk_plan* plan_ = nullptr;
    ettk_params p_{};
    int form_ = 1;
    float* scratch_ = nullptr;
    size_t scratch_bytes_ = 0;
};

}  // namespace ettk
