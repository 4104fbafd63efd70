// refine.cpp -- batched evaluation of the continuous CCC and its gradient in the particle frame, and the
// lock-step gradient ascent that replaces Powell in the batched + adjoint mode.
#include <algorithm>
#include <cfloat>
#include <functional>
#include <cmath>
#include <memory>

#include "plan.hpp"

namespace ettk {

// pose of the refinement parameters: t(x) = [R0 Rx(x0) Ry(x1) Rz(x2) | tau0 + (x3, x4, x5)], angles in degrees
static Mat4 pose(const Mat4& seed, const double* x) {
    Mat4 R = strip_translation(seed);
    Mat4 t = R * mat_rot(x[0], 1, 0, 0) * mat_rot(x[1], 0, 1, 0) * mat_rot(x[2], 0, 0, 1);
    t(0, 3) = seed(0, 3) + x[3];
    t(1, 3) = seed(1, 3) + x[4];
    t(2, 3) = seed(2, 3) + x[5];
    return t;
}
// derivative of the rotation about axis `ax` by `deg` degrees, per degree
static Mat4 drot(int ax, double deg) {
    double a = deg * M_PI / 180.0, c = std::cos(a), s = std::sin(a), k = M_PI / 180.0;
    Mat4 D;
    for (double& v : D.a) v = 0.0;
    int i = (ax + 1) % 3, j = (ax + 2) % 3;  // rotation in the (i, j) plane
    D(i, i) = -s * k;
    D(i, j) = -c * k;
    D(j, i) = c * k;
    D(j, j) = -s * k;
    return D;
}
// lab-frame wedge normals (nbfWedge3D::set, Wedge.h:143-184); false when nothing is missing
static bool lab_normals(float lw, float uw, double n1[3], double n2[3]) {
    double l = lw < -90 ? -90.0 : lw, u = uw > 90 ? 90.0 : uw;
    if (u - l >= 180) return false;
    n1[0] = std::cos(-(90.0 - u) * M_PI / 180.0);
    n1[1] = 0.0;
    n1[2] = -std::sin(-(90.0 - u) * M_PI / 180.0);
    n2[0] = std::cos(-(90.0 - l) * M_PI / 180.0);
    n2[1] = 0.0;
    n2[2] = -std::sin(-(90.0 - l) * M_PI / 180.0);
    return true;
}

// Spectra r2c((I2 - mean I2) maskfft shift) of particles [p0, p1) into P->r_fp (slot = particle - p0).
static void particle_spectra(ettk_plan* P, const float* d_parts, int p0, int p1) {
    const int N = P->N, H = P->H;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * H;
    const int cnt = p1 - p0;
    hipStream_t s = P->stream;
    P->r_fp.ensure((size_t)cnt * nh);
    P->r_mean.ensure(cnt);
    launch_vol_mean(d_parts + (size_t)p0 * n3, cnt, N, P->r_mean.p, s);
    const int CH = 64;
    for (int c0 = 0; c0 < cnt; c0 += CH) {
        int c = std::min(CH, cnt - c0);
        R2CArgs a{d_parts + (size_t)(p0 + c0) * n3, P->r_fp.p + (size_t)c0 * nh, N, c, PRO_RESLICED, P->maskfft.p,
                  P->r_mean.p + c0};
        fft3_r2c(P->ctx->fft, a, nullptr, s);
    }
}

// device arrays for n evaluations in chunks of `cap`, and the chunks themselves (kept by the plan between calls:
// their pinned buffers and events are allocated once)
static std::vector<RefineChunk*> refine_prepare(ettk_plan* P, int n) {
    const int N = P->N;
    ensure_ccc_workspace(P, std::min(n, 128));
    const int cap = std::min(P->cap, 128);
    const int nbd = refine_dot_blocks(N), nbm = refine_moment_blocks(N);
    P->r_pdot.ensure((size_t)3 * std::max(nbd, fft3_fused_supported(N) ? fft3_refine_ipass_blocks(N) : 0) * cap);
    P->r_mpart.ensure((size_t)12 * nbm * cap);
    P->r_sc.ensure((size_t)4 * n);
    P->r_mom.ensure((size_t)12 * n);
    P->r_geom.ensure(n);
    std::vector<RefineChunk*> chunks;
    size_t k = 0;
    for (int e0 = 0; e0 < n; e0 += cap, k++) {
        if (k == P->refine_pool.size()) P->refine_pool.emplace_back(new RefineChunk());
        RefineChunk& c = *P->refine_pool[k];
        c.e0 = e0;
        c.E = std::min(cap, n - e0);
        c.hg.ensure(c.E);
        c.hsc.ensure((size_t)4 * c.E);
        c.hmom.ensure((size_t)12 * c.E);
        c.Ginv.resize(c.E);
        c.EMs.resize(c.E);
        if (!c.done) ETTK_HIP(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
        if (!c.kdone) ETTK_HIP(hipEventCreateWithFlags(&c.kdone, hipEventDisableTiming));
        chunks.push_back(&c);
    }
    return chunks;
}

// host geometry (FP64) of the chunk's evaluations q (particle pidx[q], seed seeds[q], parameters x6[q])
static void refine_geometry(const ettk_plan* P, const ettk_particle* descs, const int* pidx, int p0, const Mat4* seeds,
                            const double* x6, RefineChunk& c) {
    const int N = P->N;
    for (int k = 0; k < c.E; k++) {
        const int q = c.e0 + k;
        const ettk_particle& d = descs[pidx[q]];
        const Mat4 EM = euler_of(d) * Mat4::from(d.matrix);
        Mat4 fin = EM * pose(seeds[q], x6 + (size_t)q * 6);
        if (mat_det(fin) == 0.0) fin = Mat4::ident();
        const Mat4 Gi = mat_inverse(fin);
        c.EMs[k] = EM;
        c.Ginv[k] = Gi;
        RefineGeom& r = c.hg.p[k];
        r = RefineGeom{};
        const double cc[3] = {N / 2.0, N / 2.0, N / 2.0 + d.cut_offset};
        for (int a = 0; a < 3; a++) {
            for (int b = 0; b < 3; b++) r.G[a * 4 + b] = (float)Gi(a, b);
            r.G[a * 4 + 3] = (float)(Gi(a, 3) + cc[a]);
        }
        double n1[3], n2[3];
        r.pw = (P->p.missing_wedge && lab_normals(d.lwedge, d.uwedge, n1, n2)) ? 1 : 0;
        if (r.pw)
            for (int a = 0; a < 3; a++) {
                r.n1[a] = (float)n1[a];
                r.n2[a] = (float)n2[a];
            }
        r.rw = !P->p.missing_wedge ? 0 : P->ref_wedge_kind == 1 ? 2 : (P->ref_wedge_kind == 0 && !P->wref.full) ? 1 : 0;
        if (r.rw) {
            // particle-frame frequency k -> reference-frame frequency A k at the seed; the analytic reference
            // wedge normals v become u = A^T v
            Mat4 fs = EM * seeds[q];
            if (mat_det(fs) == 0.0) fs = Mat4::ident();
            const Mat4 A = mat_inverse(fs);
            for (int rr = 0; rr < 3; rr++)
                for (int ci = 0; ci < 3; ci++) r.A[rr * 3 + ci] = (float)A(rr, ci);
            if (r.rw == 1)
                for (int ci = 0; ci < 3; ci++) {
                    double a1 = 0, a2 = 0;
                    for (int rr = 0; rr < 3; rr++) {
                        a1 += A(rr, ci) * P->wref.v1[rr];
                        a2 += A(rr, ci) * P->wref.v2[rr];
                    }
                    r.u1[ci] = (float)a1;
                    r.u2[ci] = (float)a2;
                }
        }
        r.spec = pidx[q] - p0;
    }
}

// upload of the chunk's geometry, its kernels and the download of its scores and moments; returns at once.  The
// particle spectra of the referenced particles must be in P->r_fp.
static void refine_enqueue(ettk_plan* P, RefineChunk& c) {
    const int N = P->N, H = P->H;
    const long nh = (long)N * N * H;
    hipStream_t s = P->stream;
    KernelProf* kp = P->ctx->fft.prof;
    const int E = c.E, e0 = c.e0;
    RefineGeom* geo = P->r_geom.p + e0;
    upload_before(P, geo, c.hg.p, (size_t)E * sizeof(RefineGeom));
    const double Rb = (double)E * N * N * N * sizeof(float), Hb = (double)E * nh * sizeof(float2);
    {
        ProfScope ps(kp, "k_reslice_ref", Rb, s);
        launch_reslice_ref(P->ref_x.p, geo, E, N, P->R2.p, s);
    }
    R2CArgs a{P->R2.p, P->S2.p, N, E, PRO_MASK_SHIFT, P->maskfft.p, nullptr};
    double* sc = P->r_sc.p + (size_t)4 * e0;
    if (fft3_fused_supported(N)) {
        // the score terms in the epilogue of the forward i pass, the adjoint spectrum in the prologue of the
        // inverse one: X1 is read once by each instead of once more by separate score and adjoint passes
        fft3_r2c_planes(P->ctx->fft, a, s);
        fft3_refine_dot(P->ctx->fft, P->S2.p, P->r_fp.p, P->filter.p, P->ref_wbin.p, geo, N, E, P->r_pdot.p, s);
        launch_refine_scalars(P->r_pdot.p, fft3_refine_ipass_blocks(N), E, sc, s);
        fft3_refine_adj_c2r(P->ctx->fft, P->S2.p, P->r_fp.p, P->filter.p, P->ref_wbin.p, geo, sc, N, E, P->S3.p, P->R1.p,
                            s);
    } else {
        fft3_r2c(P->ctx->fft, a, nullptr, s);
        {
            ProfScope ps(kp, "k_refine_dot", 2 * Hb, s);
            launch_refine_dot(P->S2.p, P->r_fp.p, P->filter.p, P->ref_wbin.p, geo, E, N, P->r_pdot.p, sc, s);
        }
        {
            ProfScope ps(kp, "k_refine_adj", 3 * Hb, s);
            launch_refine_adj(P->S2.p, P->r_fp.p, P->filter.p, P->ref_wbin.p, geo, E, N, sc, s);
        }
        fft3_c2r(P->ctx->fft, P->S2.p, P->S3.p, P->R1.p, N, E, s);
    }
    double* mom = P->r_mom.p + (size_t)12 * e0;
    {
        ProfScope ps(kp, "k_refine_moments", Rb, s);
        launch_refine_moments(P->ref_x.p, geo, P->R1.p, P->maskfft.p, E, N, P->r_mpart.p, mom, s);
    }
    // the downloads on the copy stream: the next chunk's kernels do not wait for them
    hipStream_t cs = P->copy_stream;
    ETTK_HIP(hipEventRecord(c.kdone, s));
    ETTK_HIP(hipStreamWaitEvent(cs, c.kdone, 0));
    ETTK_HIP(hipMemcpyAsync(c.hsc.p, sc, (size_t)4 * E * sizeof(double), hipMemcpyDeviceToHost, cs));
    ETTK_HIP(hipMemcpyAsync(c.hmom.p, mom, (size_t)12 * E * sizeof(double), hipMemcpyDeviceToHost, cs));
    ETTK_HIP(hipEventRecord(c.done, cs));
}

// score S[q] and gradient grad6[q * 6] of the chunk's evaluations once its kernels are done
static void refine_finish(RefineChunk& c, const Mat4* seeds, const double* x6, double* S, double* grad6) {
    ETTK_HIP(hipEventSynchronize(c.done));
    for (int k = 0; k < c.E; k++) {
        const int q = c.e0 + k;
        const double* x = x6 + (size_t)q * 6;
        S[q] = c.hsc.p[(size_t)k * 4];
        const Mat4 R0 = strip_translation(seeds[q]);
        const Mat4 Rx = mat_rot(x[0], 1, 0, 0), Ry = mat_rot(x[1], 0, 1, 0), Rz = mat_rot(x[2], 0, 0, 1);
        Mat4 dt[6];
        dt[0] = R0 * drot(0, x[0]) * Ry * Rz;
        dt[1] = R0 * Rx * drot(1, x[1]) * Rz;
        dt[2] = R0 * Rx * Ry * drot(2, x[2]);
        for (int r = 0; r < 3; r++) {
            dt[r](3, 3) = 0.0;
            dt[r](0, 3) = dt[r](1, 3) = dt[r](2, 3) = 0.0;
        }
        for (int r = 3; r < 6; r++) {
            for (double& v : dt[r].a) v = 0.0;
            dt[r](r - 3, 3) = 1.0;
        }
        for (int r = 0; r < 6; r++) {
            // d(fin^-1) = -fin^-1 dfin fin^-1
            const Mat4 dG = c.Ginv[k] * (c.EMs[k] * dt[r]) * c.Ginv[k];
            double acc = 0;
            for (int aa = 0; aa < 3; aa++)
                for (int b = 0; b < 4; b++) acc -= dG(aa, b) * c.hmom.p[(size_t)k * 12 + aa * 4 + b];
            grad6[(size_t)q * 6 + r] = acc;
        }
    }
}

// Score and gradient for n evaluations (particle pidx[q], seed seeds[q], parameters x6[q]).  The particle
// spectra of the referenced particles must be in P->r_fp (slot = particle - p0).
static void refine_eval_batch(ettk_plan* P, const ettk_particle* descs, int n, const int* pidx, int p0,
                              const Mat4* seeds, const double* x6, double* S, double* grad6) {
    const std::vector<RefineChunk*> chunks = refine_prepare(P, n);
    for (auto& c : chunks) {
        refine_geometry(P, descs, pidx, p0, seeds, x6, *c);
        refine_enqueue(P, *c);
    }
    for (auto& c : chunks) refine_finish(*c, seeds, x6, S, grad6);
}

// Lock-step gradient ascent for every (particle, seed) job, n_iter evaluations per job: a step of s voxels
// (start 0.5) along the gradient in the metric where one degree counts as c = (N/4) pi/180 voxels of
// displacement; an improving step is accepted and s grows by 1.25, otherwise s halves.  Jobs are independent:
// each chunk of jobs advances on its own, round robin, so the host finishes one chunk's step and enqueues its next
// evaluation while the GPU evaluates the other chunks.
static void refine_jobs(ettk_plan* P, const ettk_particle* descs, int n, const int* pidx, int p0, const Mat4* seeds,
                        int n_iter, Mat4* out_pose, double* out_score,
                        const std::function<void(int, int)>* on_done) {
    const double cm = P->N / 4.0 * M_PI / 180.0;
    std::vector<double> x((size_t)6 * n, 0.0), xt((size_t)6 * n, 0.0), g((size_t)6 * n), gt((size_t)6 * n);
    std::vector<double> Sc(n), St(n), step(n, 0.5);
    const std::vector<RefineChunk*> chunks = refine_prepare(P, n);
    for (auto& c : chunks) {
        refine_geometry(P, descs, pidx, p0, seeds, x.data(), *c);
        refine_enqueue(P, *c);
    }
    std::vector<int> evals(chunks.size(), 0);
    size_t active = chunks.size();
    while (active > 0) {
        for (size_t ci = 0; ci < chunks.size(); ci++) {
            if (evals[ci] >= n_iter) continue;
            RefineChunk& c = *chunks[ci];
            const int q0 = c.e0, q1 = c.e0 + c.E;
            if (evals[ci] == 0) {
                refine_finish(c, seeds, x.data(), Sc.data(), g.data());
            } else {
                refine_finish(c, seeds, xt.data(), St.data(), gt.data());
                for (int q = q0; q < q1; q++) {
                    if (St[q] > Sc[q]) {
                        Sc[q] = St[q];
                        for (int r = 0; r < 6; r++) {
                            x[q * 6 + r] = xt[q * 6 + r];
                            g[q * 6 + r] = gt[q * 6 + r];
                        }
                        step[q] *= 1.25;
                    } else {
                        step[q] *= 0.5;
                    }
                }
            }
            if (++evals[ci] >= n_iter) {
                // the chunk's jobs are final: their poses and scores, and the caller told
                for (int q = q0; q < q1; q++) {
                    out_pose[q] = pose(seeds[q], x.data() + (size_t)q * 6);
                    out_score[q] = Sc[q];
                }
                if (on_done) (*on_done)(q0, q1);
                active--;
                continue;
            }
            for (int q = q0; q < q1; q++) {
                double dir[6], nrm = 0;
                for (int r = 0; r < 6; r++) {
                    dir[r] = r < 3 ? g[q * 6 + r] / (cm * cm) : g[q * 6 + r];
                    nrm += r < 3 ? cm * cm * dir[r] * dir[r] : dir[r] * dir[r];
                }
                nrm = std::sqrt(nrm);
                for (int r = 0; r < 6; r++) xt[q * 6 + r] = x[q * 6 + r] + (nrm > 0 ? step[q] * dir[r] / nrm : 0.0);
            }
            refine_geometry(P, descs, pidx, p0, seeds, xt.data(), c);
            refine_enqueue(P, c);
        }
    }
}

// Range of particles referenced by a batch of jobs
static void particle_range(int n, const int* pidx, int& p0, int& p1) {
    p0 = 0;
    p1 = 0;
    if (n == 0) return;
    p0 = *std::min_element(pidx, pidx + n);
    p1 = *std::max_element(pidx, pidx + n) + 1;
}

void refine_spectra_ahead(ettk_plan* P, const float* d_parts, int p0, int p1) {
    particle_spectra(P, d_parts, p0, p1);
    P->r_fp_parts = d_parts;
    P->r_fp_p0 = p0;
    P->r_fp_p1 = p1;
}

void refine_with_seeds(ettk_plan* P, const float* d_parts, const ettk_particle* descs, int n, const int* pidx,
                       const Mat4* seeds, int n_iter, Mat4* out_pose, double* out_score,
                       const std::function<void(int, int)>* on_done) {
    ETTK_REQUIRE(P->have_ref, ETTK_ERR_STATE, "reference not set");
    if (n == 0) return;
    int p0, p1;
    particle_range(n, pidx, p0, p1);
    if (!(P->r_fp_parts == d_parts && P->r_fp_p0 == p0 && P->r_fp_p1 == p1)) particle_spectra(P, d_parts, p0, p1);
    P->r_fp_parts = nullptr;  // spectra enqueued ahead serve this call only
    refine_jobs(P, descs, n, pidx, p0, seeds, n_iter, out_pose, out_score, on_done);
}

}  // namespace ettk

using namespace ettk;

extern "C" int ettk_refine_batch(ettk_plan* P, const float* d_particles, int n_particles, const ettk_particle* descs,
                                 int n, const int* particle_index, const double* seeds, int n_iter, double* out_poses,
                                 double* out_scores) {
    try {
        ETTK_REQUIRE(P && d_particles && descs && particle_index && seeds && out_poses && out_scores && n >= 0 &&
                         n_iter >= 1,
                     ETTK_ERR_ARG, "ettk_refine_batch: bad argument");
        for (int e = 0; e < n; e++)
            ETTK_REQUIRE(particle_index[e] >= 0 && particle_index[e] < n_particles, ETTK_ERR_ARG,
                         "ettk_refine_batch: particle index out of range");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        std::vector<Mat4> sd(n), ps(n);
        for (int e = 0; e < n; e++) sd[e] = Mat4::from(seeds + (size_t)e * 16);
        refine_with_seeds(P, d_particles, descs, n, particle_index, sd.data(), n_iter, ps.data(), out_scores);
        for (int e = 0; e < n; e++) std::memcpy(out_poses + (size_t)e * 16, ps[e].a, sizeof(ps[e].a));
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

extern "C" int ettk_ccc_grad_batch(ettk_plan* P, const float* d_particles, int n_particles, const ettk_particle* descs,
                                   int n_evals, const int* particle_index, const double* seeds, const double* params6,
                                   double* out_value, double* out_grad) {
    try {
        ETTK_REQUIRE(P && d_particles && descs && particle_index && seeds && params6 && out_value && out_grad &&
                         n_evals >= 0,
                     ETTK_ERR_ARG, "ettk_ccc_grad_batch: bad argument");
        for (int e = 0; e < n_evals; e++)
            ETTK_REQUIRE(particle_index[e] >= 0 && particle_index[e] < n_particles, ETTK_ERR_ARG,
                         "ettk_ccc_grad_batch: particle index out of range");
        ETTK_REQUIRE(P->have_ref, ETTK_ERR_STATE, "reference not set");
        ETTK_HIP(hipSetDevice(P->ctx->device));
        if (n_evals > 0) {
            std::vector<Mat4> sd(n_evals);
            for (int e = 0; e < n_evals; e++) sd[e] = Mat4::from(seeds + (size_t)e * 16);
            int p0, p1;
            particle_range(n_evals, particle_index, p0, p1);
            particle_spectra(P, d_particles, p0, p1);
            refine_eval_batch(P, descs, n_evals, particle_index, p0, sd.data(), params6, out_value, out_grad);
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
