// noalloc.hip -- the constrained-CCC entry point that allocates no memory (ettk_ccc_eval_batch_noalloc).
//
// Callers that cannot allocate device memory while they run (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc) still need the CCC.  Its workspace here is one fixed device array of
// kScratchBytes per GPU, defined below; every call on a GPU shares that GPU's array, so calls that overlap on
// one GPU (two host threads, two plans of one device) race and must be serialised by the caller.  Evaluations
// are processed in chunks that fit the array; each chunk's geometry and results pass through fixed host
// arrays on the calling thread's stack.  Nothing in the array has to start at zero: every kernel of the chain
// writes each workspace element it later reads.
#include <algorithm>
#include <cstring>

#include "plan.hpp"

namespace ettk {

constexpr size_t kScratchBytes = (size_t)128 << 20;  // 128 MiB per GPU, held until the process ends
constexpr int kNoallocChunk = 64;                     // evaluations per chunk at most (host stack arrays)

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of this device's copy of `scratch`.  In host code the name `scratch` is only a host-side
// stand-in; kernels receive this address.  Called with the context's device current.
void noalloc_scratch_init(ettk_ctx* ctx) {
    void* ptr = nullptr;
    ETTK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    ETTK_REQUIRE(ptr != nullptr, ETTK_ERR_HIP, "noalloc: no device address for the scratch region");
    ctx->scratch = ptr;
}

}  // namespace ettk

using namespace ettk;

extern "C" int ettk_ccc_eval_batch_noalloc(ettk_plan* P, const float* d_particles, int n_particles,
                                           const ettk_particle* descs, int n_evals, const int* particle_index,
                                           const double* transforms, float* out_distance, double* out_transforms) {
    try {
        ETTK_REQUIRE(P && d_particles && descs && particle_index && transforms && out_distance && out_transforms,
                     ETTK_ERR_ARG, "ettk_ccc_eval_batch_noalloc: null argument");
        ETTK_REQUIRE(n_evals >= 0 && n_particles > 0, ETTK_ERR_ARG, "ettk_ccc_eval_batch_noalloc: bad sizes");
        ETTK_REQUIRE(P->have_ref, ETTK_ERR_STATE, "ettk_ccc_eval_batch_noalloc: reference not set");
        for (int e = 0; e < n_evals; e++)
            ETTK_REQUIRE(particle_index[e] >= 0 && particle_index[e] < n_particles, ETTK_ERR_ARG,
                         "ettk_ccc_eval_batch_noalloc: particle index out of range");
        ETTK_REQUIRE(P->ctx->scratch != nullptr, ETTK_ERR_STATE, "ettk_ccc_eval_batch_noalloc: scratch not initialised");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        // per chunk: the workspace, then the geometry and the results, each on a 256-byte boundary
        auto need = [&](int E) {
            return ccc_work_bytes(P, E) + (((size_t)E * sizeof(EvalGeom) + 255) & ~(size_t)255) +
                   (size_t)4 * E * sizeof(float);
        };
        int chunk = 0;
        for (int E = kNoallocChunk; E >= 1; E--)
            if (need(E) <= kScratchBytes) {
                chunk = E;
                break;
            }
        ETTK_REQUIRE(chunk > 0, ETTK_ERR_UNSUPPORTED,
                     "ettk_ccc_eval_batch_noalloc: one evaluation needs " + std::to_string(need(1)) +
                         " bytes of workspace, more than the " + std::to_string(kScratchBytes) +
                         "-byte scratch region (reduce the box size or the translation restriction)");
        if (chunk >= 8) chunk &= ~7;  // whole groups of eight evaluations: one per XCD in the dealt kernels
        char* base = static_cast<char*>(P->ctx->scratch);
        const CccWork W = ccc_work_carve(P, chunk, base);
        const size_t wb = ccc_work_bytes(P, chunk);
        EvalGeom* d_geo = reinterpret_cast<EvalGeom*>(base + wb);
        float* d_out4 = reinterpret_cast<float*>(base + wb + (((size_t)chunk * sizeof(EvalGeom) + 255) & ~(size_t)255));
        hipStream_t s = P->stream;
        EvalGeom h_geo[kNoallocChunk];
        float h_out4[4 * kNoallocChunk];
        for (int e0 = 0; e0 < n_evals; e0 += chunk) {
            const int E = std::min(chunk, n_evals - e0);
            for (int e = 0; e < E; e++) {
                const int q = particle_index[e0 + e];
                h_geo[e] = make_geom_public(P, descs[q], Mat4::from(transforms + (size_t)(e0 + e) * 16), q);
            }
            ETTK_HIP(hipMemcpyAsync(d_geo, h_geo, (size_t)E * sizeof(EvalGeom), hipMemcpyHostToDevice, s));
            ccc_enqueue_chunk(P, W, d_particles, d_geo, E, d_out4, s);
            ETTK_HIP(hipMemcpyAsync(h_out4, d_out4, (size_t)4 * E * sizeof(float), hipMemcpyDeviceToHost, s));
            ETTK_HIP(hipStreamSynchronize(s));  // h_geo and h_out4 are reused by the next chunk
            for (int e = 0; e < E; e++) {
                const Mat4 t = Mat4::from(transforms + (size_t)(e0 + e) * 16) *
                               mat_translate(-(double)h_out4[4 * e + 1], -(double)h_out4[4 * e + 2],
                                             -(double)h_out4[4 * e + 3]);
                out_distance[e0 + e] = h_out4[4 * e];
                std::memcpy(out_transforms + (size_t)(e0 + e) * 16, t.a, sizeof(t.a));
            }
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
