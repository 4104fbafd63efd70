// classify.cpp -- one iteration of multi-reference alignment and classification (ettk_mra_iteration).
//
// The iteration of nbfLoopClustering::doMra (inc/em/nbfLoopClustering.h:1218-1647) that the classification
// driver of MPI_Classification repeats (src/MPI_Classification.cpp:880-1166), on the alignment path of this
// library:
//   1. every particle is aligned to every reference (getDistances of the metric type requested,
//      nbfLoopClustering.h:2011-2030; the references are averages, their wedge the accumulated wedge of the
//      particles they were built from);
//   2. best-reference assignment (nbfLoopClustering.h:1523-1592): each particle is evaluated under each of its
//      K alignments against all K references at zero shift (type 2 with translation restriction 0); the
//      alignment whose distances sum lowest (first minimum) gives the class;
//   3. selection (MPI_Classification.cpp:924-935, 1137-1150): within each class the particles whose distance to
//      their class reference is at most the top_percentage quantile of the class's sorted distances;
//   4. the new references: the amplitude-weighted average of each class's selected particles at their
//      alignments to that class (getImage, CorrMetric.h:803-984), RMS-normalised.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "plan.hpp"

namespace ettk {

static void abi(int rc) {
    if (rc != ETTK_OK) throw Error{rc, ettk_last_error()};
}

static void mra_iteration(ettk_plan* P, int K, float* d_refs, const int* ref_members, const double* member_matrices,
                          const float* d_parts, int n, const ettk_particle* descs, int type, int form, float top,
                          int* out_class, int* out_selected, float* out_distance, double* out_matrices) {
    const int N = P->N;
    const size_t n3 = (size_t)N * N * N;
    ettk_particle full{};
    full.lwedge = -90.0f;
    full.uwedge = 90.0f;
    for (int k = 0; k < 16; k++) full.matrix[k] = (k % 5 == 0) ? 1.0 : 0.0;
    std::vector<float> w(n);
    // the reference of class k: the previous members of k (weight 1, their alignments) define its wedge
    auto set_ref = [&](int k) {
        bool any = false;
        if (ref_members && member_matrices)
            for (int i = 0; i < n; i++) {
                w[i] = ref_members[i] == k ? 1.0f : 0.0f;
                any = any || w[i] != 0.0f;
            }
        if (any) abi(ettk_reference_set_average(P, d_refs + k * n3, n, descs, w.data(), member_matrices));
        else abi(ettk_reference_set(P, d_refs + k * n3, &full));
    };
    // 1. alignments A[k][i]
    std::vector<ettk_result> A((size_t)K * n);
    for (int k = 0; k < K; k++) {
        set_ref(k);
        abi(ettk_distances_batch(P, d_parts, n, descs, type, form, A.data() + (size_t)k * n));
    }
    // 2. assignment: E[r][i K + j] = distance to reference r of particle i under its alignment to reference j
    std::vector<int> cls(n, 0);
    if (K > 1) {
        std::vector<ettk_particle> dup((size_t)n * K);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < K; j++) {
                ettk_particle d = descs[i];
                std::memcpy(d.matrix, A[(size_t)j * n + i].matrix, sizeof(d.matrix));
                dup[(size_t)i * K + j] = d;
            }
        const float rot0 = P->p.rot_restrict, tr0 = P->p.trans_restrict;
        abi(ettk_plan_set_restrictions(P, rot0, 0.0f));
        std::vector<double> S((size_t)n * K, 0.0);
        std::vector<float> dist((size_t)n * K);
        std::vector<Mat4> t((size_t)n * K);
        try {
            for (int r = 0; r < K; r++) {
                set_ref(r);
                for (auto& m : t) m = Mat4::ident();
                // type 2: one CCC at the identity under the stored alignment; evaluation i K + j reads volume i
                std::vector<int> vol_of((size_t)n * K);
                for (size_t q = 0; q < vol_of.size(); q++) vol_of[q] = (int)(q / K);
                ccc_batch(P, d_parts, nullptr, (int)dup.size(), vol_of.data(), t.data(), dist.data(), dup.data());
                for (size_t q = 0; q < S.size(); q++) S[q] += dist[q];
            }
        } catch (...) {
            (void)ettk_plan_set_restrictions(P, rot0, tr0);
            throw;
        }
        abi(ettk_plan_set_restrictions(P, rot0, tr0));
        for (int i = 0; i < n; i++) {
            int best = 0;
            for (int j = 1; j < K; j++)
                if (S[(size_t)i * K + j] < S[(size_t)i * K + best]) best = j;  // first minimum (blitz minIndex)
            cls[i] = best;
        }
    }
    // 3. selection within each class
    std::vector<int> sel(n, 1);
    for (int k = 0; k < K; k++) {
        std::vector<float> d;
        for (int i = 0; i < n; i++)
            if (cls[i] == k) d.push_back(A[(size_t)k * n + i].distance);
        if (d.empty() || !(top < 1.0f)) continue;
        std::sort(d.begin(), d.end());
        const float cut = d[(size_t)(top * (float)(d.size() - 1))];
        for (int i = 0; i < n; i++)
            if (cls[i] == k) sel[i] = A[(size_t)k * n + i].distance <= cut ? 1 : 0;
    }
    // 4. class averages of the selected members at their alignments to the class
    std::vector<double> mats((size_t)n * 16);
    for (int i = 0; i < n; i++) std::memcpy(&mats[(size_t)i * 16], A[(size_t)cls[i] * n + i].matrix, 16 * sizeof(double));
    for (int k = 0; k < K; k++) {
        bool any = false;
        for (int i = 0; i < n; i++) {
            w[i] = (cls[i] == k && sel[i]) ? 1.0f : 0.0f;
            any = any || w[i] != 0.0f;
        }
        if (!any) continue;  // an empty class keeps its reference
        abi(ettk_avg_reset(P));
        abi(ettk_avg_accumulate(P, d_parts, n, descs, w.data(), mats.data()));
        abi(ettk_avg_finalize(P, d_refs + k * n3));
    }
    for (int i = 0; i < n; i++) {
        out_class[i] = cls[i];
        if (out_selected) out_selected[i] = sel[i];
        if (out_distance) out_distance[i] = A[(size_t)cls[i] * n + i].distance;
        if (out_matrices) std::memcpy(out_matrices + (size_t)i * 16, &mats[(size_t)i * 16], 16 * sizeof(double));
    }
}

}  // namespace ettk

using namespace ettk;

extern "C" int ettk_mra_iteration(ettk_plan* P, int K, float* d_refs, const int* ref_members,
                                  const double* member_matrices, const float* d_particles, int n,
                                  const ettk_particle* descs, int type, int form, float top_percentage,
                                  int* out_class, int* out_selected, float* out_distance, double* out_matrices) {
    try {
        ETTK_REQUIRE(P && d_refs && d_particles && descs && out_class && n > 0 && K >= 1, ETTK_ERR_ARG,
                     "ettk_mra_iteration: bad argument");
        ETTK_REQUIRE(type >= 0 && type <= 3, ETTK_ERR_ARG, "ettk_mra_iteration: metric type must be 0..3");
        ETTK_REQUIRE((ref_members == nullptr) == (member_matrices == nullptr), ETTK_ERR_ARG,
                     "ettk_mra_iteration: ref_members and member_matrices go together");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        mra_iteration(P, K, d_refs, ref_members, member_matrices, d_particles, n, descs, type, form, top_percentage,
                      out_class, out_selected, out_distance, out_matrices);
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
