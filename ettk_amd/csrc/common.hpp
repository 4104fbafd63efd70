// common.hpp -- shared host/device definitions of libettk_align (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <list>
#include <mutex>
#include <string>
#include <vector>
#include <thread>
#include <exception>
#include <algorithm>

#include "../../include/ettk_align.h"

namespace ettk {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);

// host worker threads for one parallel section: OMP_NUM_THREADS when set (the process's share of the CPU),
// otherwise the hardware thread count, at most 16
int host_threads();
// fn(i) for i in [0, n) on up to host_threads() threads (in the calling thread when n is small)
template <class F>
void parallel_for(int n, F fn, int min_per_thread = 32);
struct Error {
    int code;
    std::string msg;
};

#define ETTK_HIP(call)                                                                                  \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::ettk::Error{ETTK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};      \
    } while (0)
#define ETTK_REQUIRE(cond, code, msg)                                                                   \
    do {                                                                                                \
        if (!(cond)) throw ::ettk::Error{(code), (msg)};                                                \
    } while (0)

template <class F>
void parallel_for(int n, F fn, int min_per_thread) {
    const int nthr = std::max(1, std::min(host_threads(), n / std::max(1, min_per_thread)));
    if (nthr <= 1) {
        for (int i = 0; i < n; i++) fn(i);
        return;
    }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err(nthr);
    try {
        for (int w = 0; w < nthr; w++)
            th.emplace_back([&, w] {
                try {
                    for (int i = w; i < n; i += nthr) fn(i);
                } catch (...) {
                    err[w] = std::current_exception();
                }
            });
    } catch (...) {  // a thread that could not be started: join the running ones before reporting it
        for (auto& t : th) t.join();
        throw;
    }
    for (auto& t : th) t.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// Raises the dynamic-LDS limit of kernel `fn` to at least `bytes` (hipFuncSetAttribute), at most once per size
// increase; safe when plans are driven from several host threads.
void ensure_dynamic_lds(const void* fn, size_t bytes);

// ------------------------------------------------------------------ complex helpers (float2)
__host__ __device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline float2 cmulc(float2 a, float2 b) { return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }  // a*conj(b)
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Streaming accesses (the nontemporal bit): data that one evaluation reads or writes once per pass, marked so
// that the L2 evicts it before the data every evaluation of a launch shares (reference spectrum, masks, filter).
typedef float nt_float2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_stream(const float2* p) {
    const nt_float2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_float2*>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(float2* p, float2 v) {
    nt_float2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<nt_float2*>(p));
}

// Workgroup-wide copies in which every global load of a round is issued before the first loaded value is used,
// so a round costs one memory latency instead of one per element.  Thread t moves elements t, t + NT, ...;
// out-of-range slots re-read the last element and are dropped.
//   block_copy<NT, NE>: NE elements, NT threads, one round
//   block_copy_rt<U>:   ne elements (runtime), blockDim.x threads, U elements per thread per round
template <int NT, int NE, class Ld, class St>
__device__ __forceinline__ void block_copy(Ld ld, St st) {
    constexpr int NI = (NE + NT - 1) / NT;
    decltype(ld(0)) v[NI];
#pragma unroll
    for (int u = 0; u < NI; u++) {
        const int t = (int)threadIdx.x + u * NT;
        v[u] = ld((NE % NT == 0 || t < NE) ? t : NE - 1);
    }
#pragma unroll
    for (int u = 0; u < NI; u++) {
        const int t = (int)threadIdx.x + u * NT;
        if (NE % NT == 0 || t < NE) st(t, v[u]);
    }
}
template <int U, class Ld, class St>
__device__ __forceinline__ void block_copy_rt(int ne, Ld ld, St st) {
    const int nt = (int)blockDim.x;
    for (int t0 = (int)threadIdx.x; t0 < ne; t0 += U * nt) {
        decltype(ld(0)) v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u * nt;
            v[u] = ld(t < ne ? t : ne - 1);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u * nt;
            if (t < ne) st(t, v[u]);
        }
    }
}

// ------------------------------------------------------------------ device buffers
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t count) {
        if (count <= n) return;
        release();
        ETTK_HIP(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        ensure(count);
        ETTK_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// a HIP event owned by a scope (destroyed on every exit, including an exception)
struct ScopedEvent {
    hipEvent_t e = nullptr;
    explicit ScopedEvent(unsigned flags = hipEventDefault) { ETTK_HIP(hipEventCreateWithFlags(&e, flags)); }
    ScopedEvent(const ScopedEvent&) = delete;
    ScopedEvent& operator=(const ScopedEvent&) = delete;
    ~ScopedEvent() {
        if (e) (void)hipEventDestroy(e);
    }
    operator hipEvent_t() const { return e; }
};

// pinned host buffer (hipHostMalloc): host-to-device copies from it are asynchronous with respect to the host
template <class T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t count) {
        if (count <= n) return;
        release();
        ETTK_HIP(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
};

// ------------------------------------------------------------------ per-kernel HIP-event timing
// When enabled, every instrumented launch is bracketed by two events on its stream; totals are read back by
// ettk_profile_query.  Algorithmic bytes per launch are accumulated beside the times.
struct KernelProf {
    bool on = false;
    std::string only;  // non-empty: only launches of this kernel are timed (host stages are always recorded)
    struct Rec {
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
        double bytes = 0;
        double ms = 0;
        long long launches = 0;
    };
    std::mutex mu;  // host stages are recorded from worker threads while launches are being recorded
    std::list<std::pair<std::string, Rec>> recs;  // stable addresses: scopes keep a pointer to their record
    std::vector<hipEvent_t> pool;
    Rec& rec_locked(const char* name) {
        for (auto& r : recs)
            if (r.first == name) return r.second;
        recs.emplace_back(name, Rec{});
        return recs.back().second;
    }
    hipEvent_t take_locked() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        ETTK_HIP(hipEventCreate(&e));
        return e;
    }
    // fold recorded event pairs into totals (call after the stream has been synchronised)
    void collect() {
        std::lock_guard<std::mutex> g(mu);
        for (auto& r : recs) {
            for (auto& p : r.second.ev) {
                float ms = 0;
                ETTK_HIP(hipEventElapsedTime(&ms, p.first, p.second));
                r.second.ms += ms;
                pool.push_back(p.first);
                pool.push_back(p.second);
            }
            r.second.ev.clear();
        }
    }
    void reset() {
        collect();
        std::lock_guard<std::mutex> g(mu);
        recs.clear();
    }
    void add_host(const char* name, double ms) {
        if (!on) return;
        std::lock_guard<std::mutex> g(mu);
        Rec& r = rec_locked(name);
        r.ms += ms;
        r.launches++;
    }
    bool query(const char* name, double& ms, long long& launches, double& bytes) {
        collect();
        std::lock_guard<std::mutex> g(mu);
        for (auto& r : recs)
            if (r.first == name) {
                ms = r.second.ms;
                launches = r.second.launches;
                bytes = r.second.bytes;
                return true;
            }
        return false;
    }
};
struct ProfScope {
    KernelProf* P;
    KernelProf::Rec* r = nullptr;
    hipEvent_t b = nullptr, e = nullptr;
    hipStream_t s;
    bool by_launch, launched = false;
    // by_launch: the launch records the two events itself (hipExtLaunchKernelGGL with start() and stop(), whose
    // timestamps are the kernel's own start and end, as a kernel trace reports them), so the scope records neither;
    // otherwise the events bracket the launch on its stream
    ProfScope(KernelProf* prof, const char* name, double bytes, hipStream_t st, bool by_launch_ = false)
        : P(prof), s(st), by_launch(by_launch_) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        // a launch captured into a graph carries no events
        if (by_launch && P && P->on && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
            return;
        if (P && P->on && (P->only.empty() || P->only == name)) {
            {
                std::lock_guard<std::mutex> g(P->mu);
                r = &P->rec_locked(name);
                r->bytes += bytes;
                r->launches++;
                b = P->take_locked();
                if (by_launch) e = P->take_locked();
            }
            if (!by_launch) ETTK_HIP(hipEventRecord(b, s));
        }
    }
    hipEvent_t start() const { return b; }
    hipEvent_t stop() const { return e; }
    void done() { launched = true; }  // by_launch: the launch that carries the events succeeded
    ~ProfScope() {
        if (!r) return;
        if (by_launch) {
            std::lock_guard<std::mutex> g(P->mu);
            if (launched) {
                r->ev.emplace_back(b, e);
            } else {
                P->pool.push_back(b);
                P->pool.push_back(e);
            }
            return;
        }
        {
            std::lock_guard<std::mutex> g(P->mu);
            e = P->take_locked();
        }
        if (hipEventRecord(e, s) == hipSuccess) {
            std::lock_guard<std::mutex> g(P->mu);
            r->ev.emplace_back(b, e);
        }
    }
};

// ------------------------------------------------------------------ FFT plans (twiddle tables)
struct FftTables {
    KernelProf* prof = nullptr;
    // twiddle[L] = exp(-2 pi i t / L), t in [0, L), for every supported length, packed
    DevBuf<float2> tw;
    int offset[9] = {};  // offsets per supported length index
};
int fft_len_index(int L);  // -1 when unsupported
void fft_tables_init(FftTables& T, hipStream_t s);

// batched 3-D transforms on [b][i][j][k] volumes (k fastest), FFTW conventions:
//   r2c: real n^3 -> half n x n x (n/2+1) complex, forward (e^{-i})
//   c2r: half -> real, backward (e^{+i}), unnormalised; transforms along i and j first, then the
//        half-complex transform along k (real part of the k=0 and k=n/2 terms)
// The r2c load applies an element-wise prologue selected by `pro`.
enum Prologue : int {
    PRO_NONE = 0,        // x
    PRO_SQUARE_SHIFT,    // x*x*(-1)^(i+j+k)
    PRO_MASK_SHIFT,      // x*aux0[i,j,k]*(-1)^(i+j+k)
    PRO_RESLICED,        // valid(x) ? (x - mean_b) * aux0 * (-1)^(i+j+k) : 0   (mean from stats[b])
    PRO_SHIFT,           // x*(-1)^(i+j+k)
    PRO_PAD_MASKED,      // zero-padded ... constant-padded masked small volume (S^2 projection), see fft.hip
};
struct R2CArgs {
    const float* in;      // [b][n][n][n]  (PRO_PAD_MASKED: [b][ns][ns][ns])
    float2* out;          // [b][n][n][n/2+1]
    int n, batch;
    int pro;
    const float* aux0;    // mask (n^3) for PRO_MASK_SHIFT / PRO_RESLICED; (ns^3) for PRO_PAD_MASKED
    const float* stats;   // per batch: mean (PRO_RESLICED); {valid mean, image mean, pad constant, mask on} (PAD)
    int ns = 0;           // small box size for PRO_PAD_MASKED
};
void fft3_r2c(const FftTables& T, const R2CArgs& a, float2* scratch, hipStream_t s);
// c2r: in is preserved when `in_scratch` != in (the i/j passes run out of place into in_scratch)
void fft3_c2r(const FftTables& T, const float2* in, float2* in_scratch, float* out, int n, int batch, hipStream_t s);
// r2c(op(c2r(in))) without the real volume in HBM (n <= 64): the inverse i pass reads `in` and writes `io`
// (in == io allowed), the fused plane kernel and the forward i pass work in place on `io`.
enum : int { PLANE_OP_MASK = 0, PLANE_OP_SQUARE_SHIFT = 1 };
bool fft3_fused_supported(int n);
void fft3_c2r_op_r2c(const FftTables& T, const float2* in, float2* io, int n, int batch, int op, const float* aux,
                     float* partials, hipStream_t s);

}  // namespace ettk
