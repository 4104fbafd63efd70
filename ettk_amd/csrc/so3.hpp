// so3.hpp -- S^2 projection / SO(3) search state and launch wrappers.
#pragma once
#include <algorithm>
#include <memory>

#include "plan.hpp"

namespace ettk {

// sample recipes of the soft() calls (ProjRot.h:526-584)
enum Recipe : int {
    RC_P1SQ_W1 = 0, RC_W1, RC_P1_W1, RC_P2SQ_W2, RC_W2, RC_P2_W2, RC_P1, RC_P2, RC_P1SQ, RC_P2SQ
};
struct SoftInst {
    int particle;  // slot of the particle inside the current chunk
    int sig, pat;  // recipes of signal (F1) and pattern (F2)
    int scaled;    // softScaled (1) or soft (0)
};
enum ComboKind : int { CO_COPY = 0, CO_DIV_SQRT, CO_BOTH };
struct ComboOp {
    int kind;
    int i[4];      // instance indices
};

// local-maximum lists of one search chunk (count per volume, then up to `cap` indices and values per volume); the
// per-volume counters sit kLmCountStride ints apart, one 128-byte line each, so that appends to different
// volumes do not serialise on one line
constexpr int kLmCountStride = 32;
constexpr int kCandMax = 4096;   // maxima per volume the device candidate extraction takes (more: the host does)
constexpr int kCandSlots = 64;   // equivalent partners per maximum it keeps (more: the host extracts)
constexpr int kCandAdj = 16384;  // packed partner entries per volume
struct LmBuffers {
    DevBuf<int> count, idx;
    DevBuf<double> val;
    // device candidate extraction: volumes to process (uploaded from hact); per volume the kept maxima, the partner
    // entries and the mode; per maximum its partner count, per volume its partner lists; their packed copies
    DevBuf<int> act, kept, nadj, mode, deg, adj, pdeg, padj;
    PinnedBuf<int> hact, hpdeg, hpadj;
    // compacted lists: head = [offsets (nvol + 1), counts (nvol)], then the packed indices and values
    DevBuf<int> head, pidx;
    DevBuf<double> pval;
    // the chunk's SOFT instances and combine operations, staged in pinned memory for asynchronous uploads
    PinnedBuf<SoftInst> hinst;
    PinnedBuf<ComboOp> hops;
    // host copies of the compacted lists, pinned: downloads from pageable memory pass through one staging buffer
    // of the runtime, which serialises the chunks' downloads issued from concurrent host threads
    PinnedBuf<int> hhead, hpidx;
    PinnedBuf<double> hpval;
};

struct So3State {
    int B = 0;
    int d_slice = 0;
    bool use_window1 = false;
    DevBuf<double> legF, legI, dtab, dtab2;
    DevBuf<double> dtabP;  // the Wigner-d table per order pair for the matrix-core synthesis: [c1 n + c2][l][beta]
    DevBuf<double> rtab;  // d^l_{m m'} at the softScaled angles beta_b = (float)((2b+1) pi / 4B), [b][l][m][m']
    DevBuf<int> doff;
    DevBuf<double2> tw;
    // reference
    DevBuf<float> Pp1, W1;
    // projection workspace
    DevBuf<float> projR, projMag, projStats, projPartials;
    DevBuf<double> projSum;  // per (particle, part) sums of the image statistics
    DevBuf<float2> projF;
    DevBuf<EvalGeom> projGeom;
    // soft workspace
    DevBuf<SoftInst> inst;
    DevBuf<double> sig, pat, vals, Rr, wv, pvals;
    DevBuf<int> ppos;
    DevBuf<double2> fm, cs, cp, T1, T2, rot, g;
    DevBuf<double> zyz_tabs;  // [Ra; Rb; Rg] of the correlation grid, 2B 4x4 matrices each (device candidates)
    DevBuf<double> cand_mats;  // scratch of the device candidates: the matrices of the kept maxima per volume
    DevBuf<double2> csT, cpT;  // coefficients instance-fastest for the matrix-core synthesis: [l l + l - m][q]
    DevBuf<int> pos;
    // search workspace (align driver)
    DevBuf<float> Pp2, W2, Mx;
    DevBuf<double> C, Cc;
    DevBuf<ComboOp> ops;
    std::vector<std::unique_ptr<LmBuffers>> lm_pool;  // one slot per chunk in flight
    // host inputs of a search call (spherical wedges, reslice geometry), pinned for asynchronous uploads
    PinnedBuf<float> hW2;
    PinnedBuf<EvalGeom> hGeom;
};

inline LmBuffers& lm_slot(So3State& S, int slot, int nvol, int cap) {
    while ((int)S.lm_pool.size() <= slot) S.lm_pool.emplace_back(new LmBuffers());
    LmBuffers& b = *S.lm_pool[slot];
    b.count.ensure((size_t)nvol * kLmCountStride);
    b.idx.ensure((size_t)nvol * cap);
    b.val.ensure((size_t)nvol * cap);
    b.head.ensure((size_t)4 * nvol + 2);
    b.act.ensure(nvol);
    b.kept.ensure(nvol);
    b.nadj.ensure(nvol);
    b.mode.ensure(nvol);
    b.deg.ensure((size_t)nvol * kCandMax);
    b.adj.ensure((size_t)nvol * kCandMax * kCandSlots);
    b.pdeg.ensure((size_t)nvol * cap);
    b.padj.ensure((size_t)nvol * kCandAdj);
    b.hact.ensure(nvol);
    b.pidx.ensure((size_t)nvol * cap);
    b.pval.ensure((size_t)nvol * cap);
    // the host lists start at 2048 entries per volume (grown by the download when a chunk holds more)
    b.hhead.ensure((size_t)4 * nvol + 2);
    b.hpidx.ensure((size_t)nvol * std::min(cap, 2048));
    b.hpval.ensure((size_t)nvol * std::min(cap, 2048));
    return b;
}

EvalGeom make_geom_public(const ettk_plan* P, const ettk_particle& d, const Mat4& t, int particle);
// smoothed analytic spherical wedge (nbfWedge3D::getSphericalImage + conditionSphericalWindow), host FP64
std::vector<float> spherical_wedge_host(int B, const ettk_particle& d, bool wedge_on);
// `h_geom` (optional): reslice geometry of the n images on the host, kept alive by the caller until the stream
// has consumed it; computed here when null
void project_images(ettk_plan* P, const float* d_parts, const ettk_particle* descs, const int* pidx, int n, bool is_ref,
                    float* d_Pp, const EvalGeom* h_geom = nullptr);
// `h_inst` (optional): a copy of `inst` in pinned memory, kept unchanged until the stream has consumed it
void run_soft(ettk_plan* P, const std::vector<SoftInst>& inst, const float* d_Pp2, const float* d_W2, const float* d_Mx,
              double* d_C, const SoftInst* h_inst = nullptr);
// sizes every device buffer of the search for chunks of up to `cn` particles (at most five SOFT instances each),
// so that none is reallocated (a free synchronises the device) while chunks are in flight
void so3_search_reserve(ettk_plan* P, int cn);
// soft (scaled = 0) or softScaled (scaled = 1) of nq host signal/pattern pairs into d_C ([q][a][b][g])
void soft_given(ettk_plan* P, int nq, const double* h_sig, const double* h_pat, double* d_C, int scaled = 0);
// harmonics(A, t, R) (ProjRot.h:1348-1454) of n (2B x 2B) host images by the rotation parts of h_t[i*16]
void sph_rotate(ettk_plan* P, int n, const float* h_img, const double* h_t, float* h_out);
So3State& so3_state(ettk_plan* P);
void launch_combine(const double* C, const ComboOp* d_ops, int np, int B, int b0, double* out, hipStream_t s);
void launch_localmax(const double* C, int np, int B, int cap, int* count, int* idx, double* val, hipStream_t s);
// packs the np lists (at most `cap` entries each) back to back: head = [offsets (np + 1), counts (np)]
void launch_lm_compact(const int* count, const int* idx, const double* val, const int* kept, const int* mode,
                       const int* deg, const int* nadj, const int* adj, int np, int cap, int* head, int* pidx,
                       double* pval, int* pdeg, int* padj, hipStream_t s);
// the data-parallel part of the candidate extraction of each active volume on the device (cand.hip): its sorted
// valid maxima over the volume's list with their equivalent partners (kept[v], deg, adj, nadj[v], mode[v] = 0), or
// mode[v] = 1 where the host extracts from the raw list (active[v]: 0 unused, 1 device, 2 host)
void launch_lm_candidates(const int* count, int* idx, double* val, int cap, const int* active, int np,
                          const double* tabs, int n, double trans_restrict, double rot_restrict, float tol,
                          double* mats, int* deg, int* adj, int* kept, int* nadj, int* mode, hipStream_t s);
void launch_ppmax(const float* Pp1, const float* Pp2, int np, int B, float* Mx, hipStream_t s);

}  // namespace ettk
