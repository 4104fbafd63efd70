/*
 * ettk_align.h -- C ABI of the MI355X sub-volume alignment library (libettk_align.so).
 *
 * The library replaces the hot path that ETTK runs behind nbfCorrelationImageMetric<float,3>
 * (inc/em/nbfCorrelationImageMetric.h): the per-pair metric nbfProjectionRotationMetric3D::execute
 * (inc/em/nbfProjectionRotationMetric3D.h:416-1052), the constrained CCC
 * nbfFourierImageMetric::executeFourierNewHalf (inc/em/nbfFourierImageMetric.h:1078-1323) and the
 * amplitude-weighted Fourier average nbfWedgedAverageImage3D::getImageMPI
 * (inc/em/nbfWedgedAverageImage3D.h:261-356).  The reference has no FFI of its own; these entry points
 * are what a C/C++ caller (or the ctypes binding shown in INTEGRATION.md) binds in place of the templated
 * classes.  Every function returns 0 (ETTK_OK) or a negative ETTK_ERR_* code; the text of the last error
 * of the calling thread is returned by ettk_last_error().  No exceptions cross this boundary.
 *
 * Conventions shared by every entry point
 *   - Volumes are float32 N^3 arrays in VTK/MRC order, x fastest: v[(z*N + y)*N + x], the order in which
 *     nbfMrcReader hands data to the reference (inc/io/nbfVTKInterface.cpp:77,134-135).
 *   - Arguments named d_* are device (HBM) pointers on the plan's GPU; all other pointers are host memory.
 *   - 4x4 transforms are row-major doubles, exactly vtkMatrix4x4::DeepCopy(double[16], ...) order.
 *   - Angles are degrees, as in the reference.
 *   - One plan owns one HIP stream; calls on one plan are serialised by the caller (the reference metric
 *     is not re-entrant either: inc/em/nbfWedgedSubImage3D.h:172-201).
 */
#ifndef ETTK_ALIGN_H
#define ETTK_ALIGN_H

#ifdef __cplusplus
extern "C" {
#endif

#define ETTK_OK                0
#define ETTK_ERR_ARG          -1
#define ETTK_ERR_HIP          -2
#define ETTK_ERR_ALLOC        -3
#define ETTK_ERR_STATE        -4
#define ETTK_ERR_UNSUPPORTED  -5

#define ETTK_MAX_CANDIDATES    8

/* Alignment parameters.  Each field names the reference setter it mirrors. */
typedef struct ettk_params {
    int   N;               /* cubic box size, even (nbfFourierFilter::initializeFFTWhalf asserts even)   */
    int   B;               /* S^2 bandwidth; nbfProjectionRotationMetric3D::B (public float member)     */
    float bp_lo;           /* nbfFourierFilter::bandPassOn(lowCut, highCut, lowVariance, highVariance)    */
    float bp_hi;           /*   active only when lowCut > 0 or highCut < 1 (FFilt.h:700-712)              */
    float bp_lo_decay;
    float bp_hi_decay;
    float bfactor;         /* nbfFourierFilter::bfactorOn; 0 = off                                        */
    float mask_radius;     /* nbfImageFilter::setMaskSize(x,x,x,sigma); 0 = mask off (mask == 1)          */
    float mask_sigma;      /*   raised-cosine width of the spherical mask (IFilt.h:426-446)               */
    int   square_mask;     /* nbfImageFilter::squareMaskSize (edge of the FFT window, FIM.h:642-664)      */
    int   mcf;             /* setToUseMutualCorrelation: F <- F/|F|^(1/2)                                 */
    int   overlap_norm;    /* setToComputeOverlapNormalizedDistances                                      */
    int   missing_wedge;   /* setMissingWedgeCompensation                                                 */
    int   n_peaks;         /* setNumberOfCandidatePeaksToSearch (Kp, default 5)                           */
    int   n_candidates;    /* setNumberOfCandidates (Kc, default 1), <= ETTK_MAX_CANDIDATES              */
    float rot_restrict;    /* setRotationSearchRestriction, degrees (clamped to >= 1 as in the reference) */
    float trans_restrict;  /* setTranslationSearchRestriction, voxels; < 0 means unrestricted             */
    int   refine;          /* refinementOn()/Off(): Powell refinement of the Kc best candidates          */
} ettk_params;

/* One particle (nbfWedgedSubImage3D held as a fixed image, WSub.h:124-148) or the reference. */
typedef struct ettk_particle {
    float  lwedge;         /* nbfWedge3D::set(lwedge, uwedge): missing-wedge limits about Y, degrees     */
    float  uwedge;         /*   uwedge - lwedge >= 180 means no missing wedge                             */
    float  normal[3];      /* Euler "normal" angles of the sub-volume (WSub.h:741-749); 0 for fixed images */
    float  cut_offset;     /* nbfWedgedSubImage3D::setCutOffset                                            */
    double matrix[16];     /* current alignment (nbfWedgedSubImage3D::setTransform), row-major            */
} ettk_particle;

/* Result of one pairwise alignment (reference x particle), the row of D in getDistances
   (CorrMetric.h:2509-2530): distance, then the 4x4 of each refined candidate. */
typedef struct ettk_result {
    float  distance;                                 /* correlationPeak of the best candidate          */
    double matrix[16];                               /* best transform (metric->getTransform())        */
    int    n_cand;                                   /* candidates refined (<= Kc)                     */
    float  cand_distance[ETTK_MAX_CANDIDATES];       /* candidateCorrelationPeaks                      */
    double cand_matrix[ETTK_MAX_CANDIDATES][16];     /* candidateTransforms                            */
    float  so3_angles[ETTK_MAX_CANDIDATES][3];       /* (alpha, beta, gamma) of each refined candidate */
    int    n_ccc;                                    /* constrained-CCC evaluations the reference's
                                                        method spends on the pair (an evaluation the
                                                        device reuses still counts once here)          */
    int    use_window;                               /* bit0: reference window, bit1: particle window  */
    float  overlap;                                  /* getWedgeOverlap(getTransform()), CorrMetric.h:
                                                        533-566: sum(W1 W2 filter) / sum(filter) at the
                                                        best transform, 1 when neither wedge is effective;
                                                        the correlation scale of these metrics is 1      */
} ettk_result;

typedef struct ettk_ctx  ettk_ctx;
typedef struct ettk_plan ettk_plan;

/* Text of the last error raised on the calling thread ("" when none). */
const char* ettk_last_error(void);

/* Device context: selects the HIP device, creates the stream. */
int ettk_ctx_create(int device, ettk_ctx** out);
int ettk_ctx_destroy(ettk_ctx* ctx);

/* Plan for one box size and bandwidth: masks, filters, FFT twiddles, Legendre and Wigner-d tables.
   Replaces the lazily built state of nbfFourierImageMetric::reinitializeHalf (FIM.h:634-694),
   nbfFourierFilter::updateFilterHalf (FFilt.h:492-547) and
   nbfProjectionRotationMetric3D::initializeHarmonics (ProjRot.h:1055-1134). */
int ettk_plan_create(ettk_ctx* ctx, const ettk_params* params, ettk_plan** out);
int ettk_plan_destroy(ettk_plan* plan);

/* setRotationSearchRestriction / setTranslationSearchRestriction (nbfImageMetric.h:81-95): the rotation
   restriction is clamped to >= 1 degree, a negative translation restriction means unrestricted; the
   upsampled search window and its DFT tables are rebuilt.  Restriction 0 evaluates the CCC at zero shift
   (as the multi-reference assignment does, nbfLoopClustering.h:1560-1563). */
int ettk_plan_set_restrictions(ettk_plan* plan, float rot_restrict, float trans_restrict);

/* Reference volume (input1 of the metric): setInput1 (nbfImageMetric.h:184-189).  The volume is used as
   is (an average is already RMS-normalised, WAvg.h:241-243); its wedge is analytic from desc->l/uwedge.
   Caches the reference side of the CCC (F1m, k1_m: FIM.h:1123-1153) and its S^2 projection Pp1
   (ProjRot.h:445-453). */
int ettk_reference_set(ettk_plan* plan, const float* d_ref, const ettk_particle* ref_desc);

/* Constrained CCC at given transforms: nbfFourierImageMetric::executeFourierNewHalf (FIM.h:1078-1323)
   for n_evals (particle, transform) pairs.  transforms[e*16] is the candidate transform; on return
   out_transforms[e*16] holds transform * Translate(-shift) (FIM.h:1312-1322) and out_distance[e] the
   correlationPeak.  particle_index[e] selects the particle in d_particles (n_particles volumes). */
int ettk_ccc_eval_batch(ettk_plan* plan, const float* d_particles, int n_particles,
                        const ettk_particle* descs, int n_evals, const int* particle_index,
                        const double* transforms, float* out_distance, double* out_transforms);

/* ettk_ccc_eval_batch without any memory allocation.  No counterpart in the reference: it is for callers that
   cannot allocate device memory while they run (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  Same arguments, results and errors as ettk_ccc_eval_batch; the workspace is a fixed
   128 MiB device region per GPU, reserved when the library is loaded on that GPU and held until the process
   ends; evaluations run in chunks that fit it (at box 64 with an 8-voxel translation window, 16 per chunk).
   A size whose single evaluation does not fit is rejected with ETTK_ERR_UNSUPPORTED.  All calls on one GPU
   share the region: calls that overlap on one GPU race, so the caller serialises them.  The plan and its
   reference must have been set up beforehand (ettk_plan_create, ettk_reference_set allocate). */
int ettk_ccc_eval_batch_noalloc(ettk_plan* plan, const float* d_particles, int n_particles,
                                const ettk_particle* descs, int n_evals, const int* particle_index,
                                const double* transforms, float* out_distance, double* out_transforms);

/* S^2 projection of |FFT| (nbfProjectionRotationMetric3D::cartesianToSpherical, ProjRot.h:265-413) of each
   particle at its current alignment: out_pp is n x (2B x 2B) float, row-major [polar][azimuth]. */
int ettk_project_batch(ettk_plan* plan, const float* d_particles, int n, const ettk_particle* descs,
                       float* out_pp);

/* SO(3) correlation of S^2 functions on the Driscoll-Healy grid (ProjRot.h:1137-1295, soft-1.0
   FST_semi_memo + so3CombineCoef + Inverse_SO3_Naive): for each of n pairs, signal sig[i] and pattern
   pat[i] (2B x 2B doubles, host) give out_c[i] ((2B)^3 doubles, layout [alpha][beta][gamma]). */
int ettk_so3_correlate_batch(ettk_plan* plan, int n, const double* sig, const double* pat, double* out_c);

/* softScaled (nbfProjectionRotationMetric3D::softScaled, ProjRot.h:1457-1589): the SO(3) correlation above, then
   rescaled per pair so that its minimum and maximum equal the S^2 inner products <sig, R pat> (sin(theta)-weighted)
   of the pattern rotated to the rotations of the first minimum and first maximum in [alpha][beta][gamma] order
   (rotateFct at alpha = a pi / B, beta = (2b + 1) pi / 4B, gamma = g pi / B, rotate_so3.c:695-751). */
int ettk_so3_correlate_scaled_batch(ettk_plan* plan, int n, const double* sig, const double* pat, double* out_c);

/* nbfProjectionRotationMetric3D::harmonics(A, t, R) (ProjRot.h:1348-1454): each (2B x 2B) spherical image
   images[i] (host floats, row-major [polar][azimuth]) rotated on the sphere by the rotation part of
   transforms[i*16]: Euler angles (psi, theta, phi) taken from the matrix as the reference does (including its
   choice psi = phi = 0 when the rotation is about z alone), then rotateFct(B, B, B-1) by (-psi, -theta, -phi)
   (soft-1.0 rotate_so3.c:695-751); out[i] receives the real part.  Used by the reference for stored spherical
   wedges (nbfWedge3D::getSphericalImage, Wedge.h:364-383) and averages' spherical wedges (WAvg.h:680-683). */
int ettk_sph_rotate_batch(ettk_plan* plan, int n, const float* images, const double* transforms, float* out);

/* Full pairwise alignment of n particles against the reference: nbfProjectionRotationMetric3D::execute
   (ProjRot.h:416-1052) as called from getDistances type 0/3 (CorrMetric.h:1874-2008).
   mode 0 = plain: particle by particle, Powell refinement driven from the host (the reference method);
   mode 1 = batched: all particles' stages in shared launches, Powell refinement in lock-step batches;
   mode 2 = batched + adjoint: as 1, refinement by projected gradient ascent on the continuous CCC using
            the adjoint kernel (no reference counterpart; see ettk_ccc_grad_batch). */
int ettk_align_batch(ettk_plan* plan, const float* d_particles, int n, const ettk_particle* descs,
                     int mode, ettk_result* out);

/* nbfCorrelationImageMetric::getDistances (CorrMetric.h:659-704, worker CorrMetric.h:1911-2008) for the
   reference against n particles, by metric type:
   type 0 = S^2/SO(3) search + refinement (refinementOn);  type 3 = search without refinement;
   type 1 = refinement only, from each particle's stored transform (nbfFourierImageMetric::execute,
            FIM.h:1812-1840: the stored transform seeds refine() and is not applied while it runs);
   type 2 = translation search only at the stored alignment (nbfFourierImageMetricCore::execute,
            Core.h:35-66: one constrained CCC at the identity, its shift kept when the result is valid).
   form as `mode` of ettk_align_batch (0 plain, 1 batched, 2 batched + adjoint refinement). */
int ettk_distances_batch(ettk_plan* plan, const float* d_particles, int n, const ettk_particle* descs,
                         int type, int form, ettk_result* out);

/* Continuous constrained CCC and its gradient with respect to (rotX, rotY, rotZ, tx, ty, tz) around the
   seed transforms (degrees, voxels).  No reference counterpart: it replaces the Powell line searches of
   nbfFourierImageMetric::refine (FIM.h:1702-1809).  The score is taken in the particle's own frame: the
   reference (minus its mean) is resliced trilinearly through fin(x)^-1, fin = Euler M t(x), and correlated
   with the particle over band-pass x particle wedge x reference wedge (binary, the particle's wedge fixed in
   its frame).  params6[e*6] = (rx, ry, rz, tx, ty, tz) with t(x) = [R0 Rx Ry Rz | tau0 + (tx, ty, tz)];
   seeds[e*16]; out_value[e] = normalised score (higher is better); out_grad[e*6] its exact gradient. */
int ettk_ccc_grad_batch(ettk_plan* plan, const float* d_particles, int n_particles,
                        const ettk_particle* descs, int n_evals, const int* particle_index,
                        const double* seeds, const double* params6, double* out_value, double* out_grad);

/* Local refinement of n seed poses by n_iter evaluations of the continuous score of ettk_ccc_grad_batch
   (accept/reject gradient ascent, all seeds in lock-step batches).  out_poses[e*16], out_scores[e]. */
int ettk_refine_batch(ettk_plan* plan, const float* d_particles, int n_particles, const ettk_particle* descs,
                      int n, const int* particle_index, const double* seeds, int n_iter, double* out_poses,
                      double* out_scores);

/* Amplitude-weighted Fourier average (WAvg.h:261-356 and CorrMetric.h:927-980).
   ettk_avg_reset zeroes the (re, im, acc) half-spectrum accumulators; ettk_avg_accumulate adds n particles
   resliced with matrices[i*16] and weighted by weights[i]; ettk_avg_finalize writes the RMS-normalised
   average (N^3 float, x fastest) to d_out. */
int ettk_avg_reset(ettk_plan* plan);
int ettk_avg_accumulate(ettk_plan* plan, const float* d_particles, int n, const ettk_particle* descs,
                        const float* weights, const double* matrices);
int ettk_avg_finalize(ettk_plan* plan, float* d_out);

/* Variance map of an average: nbfCorrelationImageMetric::getImageVariance (CorrMetric.h:987-1003) through
   nbfWedgedAverageImage3D::getImageVarianceMPI (WAvg.h:440-478).  Each entry with weights[i] != 0 is
   resliced at matrices[i*16] (its own matrix reset, as for the average), T = valid voxels minus their mean, and
   var += |w_i| (T - average)^2 / sum_i |w_i| in FP64 in entry order.  d_average is the average the variance is
   taken around (N^3 floats, x fastest, e.g. the output of ettk_avg_finalize); d_out receives the map, x fastest. */
int ettk_avg_variance(ettk_plan* plan, const float* d_particles, int n, const ettk_particle* descs,
                      const float* weights, const double* matrices, const float* d_average, float* d_out);

/* Accumulated missing wedge of an average (nbfWedgedAverageImage3D::updateAccumulatedWedgeImage and
   setAccumulatedWedgeImage, WAvg.h:496-538 and :49-70; spherical counterpart updateSphericalWedgeImage,
   :614-667).  Entry e of n (the reference's (particle, candidate) order) adds the wedge of descs[e] at the
   alignment matrices[e*16]; the particle's own matrix is not used (setTransform(NULL) precedes getWedgeImage);
   entries with weights[e] == 0 are skipped, and at least one must remain.  Outputs, each optional (NULL):
   d_accum   N^3 floats, the sum of the smoothed full-grid wedges (getAccumulatedWedgeImage), array order
             A(i,j,k) with i the x frequency and k fastest, frequency (i,j,k) - N/2;
   d_wedge   N^3 floats, getWedgeImage: the sum binarised at max/2 and smoothed when the wedge is effective,
             all ones otherwise;
   out_spherical (2B)^2 host floats, getSphericalWedgeImage: the binarised spherical sum, all ones when the
             wedge is not effective;
   out_effective 1 when sum(binarised) < 0.95 N^3 (wedgeEffective). */
int ettk_avg_wedge(ettk_plan* plan, int n, const ettk_particle* descs, const float* weights,
                   const double* matrices, float* d_accum, float* d_wedge, float* out_spherical,
                   int* out_effective);

/* Reference that is an average (setInput1 with an nbfWedgedAverageImage3D): as ettk_reference_set, with the
   reference's wedge taken from its particles as ettk_avg_wedge computes it, in the CCC
   (getWedgeImageHalf, WAvg.h:602-611), in the S^2 search (getSphericalWedgeImage, ProjRot.h:481) and, binarised,
   in the continuous refinement. */
int ettk_reference_set_average(ettk_plan* plan, const float* d_ref, int n, const ettk_particle* descs,
                               const float* weights, const double* matrices);

/* Per-kernel timing of the last ettk_align_batch / ettk_ccc_eval_batch call, measured with HIP events on
   the plan's stream: total milliseconds spent in the constrained-CCC pipeline and number of CCC
   evaluations the device ran (fewer than the sum of ettk_result.n_ccc in the batched + adjoint form, which
   reuses the SO(3) peak's evaluation for the refinement's starting point). */
int ettk_last_timing(ettk_plan* plan, double* ccc_ms, long long* n_ccc, double* total_ms);

/* The path's 3-D transforms on device buffers, FFTW conventions: the r2c/c2r plans of
   nbfFourierFilter::initializeFFTWhalf (nbfFourierFilter.h:270-358) on [batch][i][j][k] volumes with the half
   spectrum along k ([batch][n][n][n/2+1] complex float, interleaved), r2c forward (e^-i), c2r backward (e^+i)
   and unnormalised.  Kinds: ETTK_FFT_R2C (in real, out complex), ETTK_FFT_C2R (in complex and preserved, out
   real), and the two transforms of the constrained CCC that never materialise the real volume (in and out
   complex, n <= 64): ETTK_FFT_MASK (r2c(c2r(in) * aux), aux an n^3 real mask in the same [i][j][k] order,
   FIM.h:1123-1153) and ETTK_FFT_SQUARE (r2c(c2r(in)^2 (-1)^(i+j+k)), FIM.h:1176-1187).  n in {16, 32, 48, 64,
   96, 128, 192, 256}.  Runs on the context's stream and returns when the result is complete. */
#define ETTK_FFT_R2C     0
#define ETTK_FFT_C2R     1
#define ETTK_FFT_MASK    2
#define ETTK_FFT_SQUARE  3
int ettk_fft3(ettk_ctx* ctx, int n, int batch, int kind, const void* in, void* out, const float* aux);

/* One iteration of multi-reference alignment and classification: nbfLoopClustering::doMra
   (inc/em/nbfLoopClustering.h:1218-1647) as MPI_Classification repeats it (src/MPI_Classification.cpp:880-1166).
   d_refs holds K references (N^3 floats each, x fastest) and receives the K new class averages.  When ref_members
   (n ints) and member_matrices (n x 16) are given, reference k is the average of the particles with
   ref_members[i] == k at member_matrices[i] (setInput1 of an nbfWedgedAverageImage3D: its wedge is their
   accumulated wedge); otherwise the references carry no wedge.
     1. every particle is aligned to every reference with getDistances of metric `type` (0 search + refinement,
        1 refinement from the stored alignment, 2 shift only, 3 search only) in execution `form`
        (as ettk_distances_batch);
     2. best-reference assignment (nbfLoopClustering.h:1523-1592): each particle under each of its K alignments
        against every reference at zero shift; the alignment with the lowest sum of distances (first minimum)
        gives out_class[i];
     3. selection (MPI_Classification.cpp:924-935, 1137-1150): in each class the particles whose distance is at most
        the value at index floor(top_percentage (m - 1)) of the class's m sorted distances (all when
        top_percentage >= 1), out_selected[i] = 1;
     4. each class's new reference: the amplitude-weighted average of its selected particles at their alignments
        to it (getImage, CorrMetric.h:803-984), RMS-normalised; a class without selected particles keeps its
        reference.
   Outputs (out_selected, out_distance, out_matrices may be NULL): out_distance[i] and out_matrices[i*16] are the
   distance and the alignment of particle i to its class; passing out_class masked by out_selected (-1 where not
   selected) and out_matrices as the next call's ref_members and member_matrices continues the iterations. */
int ettk_mra_iteration(ettk_plan* plan, int K, float* d_refs, const int* ref_members, const double* member_matrices,
                       const float* d_particles, int n, const ettk_particle* descs, int type, int form,
                       float top_percentage, int* out_class, int* out_selected, float* out_distance,
                       double* out_matrices);

/* Execution choices of a plan (no reference counterpart): evaluations per constrained-CCC chunk and whether the
   chunks are replayed from HIP graphs captured once per chunk size and particle array.  Results do not depend on
   them.  ettk_plan_tune times the plan's CCC batch for every candidate (chunks of 16 to 256 evaluations, eager or
   graph launches) on evaluations of the given particles at fixed pseudo-random rotations, keeps the fastest and
   appends it to the tuning cache (the file named by ETTK_TUNE_CACHE, else ~/.cache/ettk_amd_tune.txt) under the
   key (device architecture, box size, bandwidth, translation window); a plan created later for the same key
   starts from the recorded choice.  The plan's reference must be set. */
int ettk_plan_tune(ettk_plan* plan, const float* d_particles, int n, const ettk_particle* descs, int* out_chunk,
                   int* out_graphs, double* out_us_per_eval);
int ettk_plan_set_execution(ettk_plan* plan, int ccc_chunk, int use_graphs);
int ettk_plan_get_execution(ettk_plan* plan, int* ccc_chunk, int* use_graphs);

/* Particles of uneven box sizes: several buckets, each of one box size and bandwidth with its own reference and
   particles, aligned in one call (as MPI_Classification aligns volume lists cut at several sizes,
   src/MPI_Classification.cpp:1087, each with nbfProjectionRotationMetric3D::execute, ProjRot.h:416-1052).  The
   context keeps one plan per distinct parameter set across calls; each plan runs its buckets on its own streams
   from its own host thread, so the buckets' kernels overlap on the GPU.  mode as ettk_align_batch.  Flags:
   ETTK_BUCKETS_GRAPHS replays CCC chunks from captured graphs (unless the tuning cache says otherwise),
   ETTK_BUCKETS_TUNE runs ettk_plan_tune on each bucket's particles first. */
typedef struct ettk_bucket {
    ettk_params params;
    const float* d_reference;   /* N^3, x fastest */
    ettk_particle reference;
    const float* d_particles;   /* n x N^3 */
    int n;
    const ettk_particle* descs;
    ettk_result* out;           /* n results */
} ettk_bucket;
#define ETTK_BUCKETS_GRAPHS 1
#define ETTK_BUCKETS_TUNE   2
int ettk_align_buckets(ettk_ctx* ctx, int n_buckets, const ettk_bucket* buckets, int mode, int flags);

/* Sub-volumes cut from a tomogram: nbfWedgedSubImage3D::getImage with its tomogram loaded
   (inc/em/nbfWedgedSubImage3D.h:603-802, normalize off), as src/CutVolumes3DFromPositions.cpp and the first
   getImage of each particle do.  d_tomogram is nx x ny x nz floats, x fastest (the MRC order); for particle i,
   output voxel q of the N^3 box reads the tomogram at positions[3i..3i+2] + fin (q - N/2 - cut_offset z) with
   fin = Euler(descs[i].normal) descs[i].matrix (WSub.h:712-749), by VTK's tricubic interpolation (taps clamped to
   the tomogram); samples outside [-0.5, n - 0.5] on any axis are invalid.  The mean of the valid samples is
   subtracted and invalid voxels are set to 0 (WSub.h:794-802).  d_out receives n volumes of N^3 floats, x fastest,
   ready to be passed as particles.  Runs on the context's stream and returns when the volumes are complete. */
int ettk_extract_batch(ettk_ctx* ctx, const float* d_tomogram, int nx, int ny, int nz, int n, const double* positions,
                       const ettk_particle* descs, int N, float* d_out);

/* Measurement hooks (no reference counterpart).  ettk_profile(ctx, 1) resets and enables per-kernel HIP-event
   timing on the context's stream; ettk_profile_query returns, for one kernel name (e.g. "k_fft_strided<64>",
   "k_reslice", "k_combine_dftA"), the summed launch time in ms, the launch count and the algorithmic bytes
   those launches move (one read of every input and one write of every output). */
int ettk_profile(ettk_ctx* ctx, int enable);
int ettk_profile_query(ettk_ctx* ctx, const char* kernel, double* total_ms, long long* launches, double* bytes);
/* Restricts the HIP-event timing to the launches of one kernel (NULL or "": every instrumented kernel), so that a
   timed run carries two events per launch of that kernel only; host stages are recorded either way. */
int ettk_profile_only(ettk_ctx* ctx, const char* kernel);

#ifdef __cplusplus
}
#endif
#endif /* ETTK_ALIGN_H */
