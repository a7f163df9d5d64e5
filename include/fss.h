/*
 * fss.h -- C ABI of libfss_hip.so, the MI355X (gfx950) kernels for ShEPhERD Gaussian-overlap
 * scoring and multi-restart SE(3) alignment.
 *
 * Conventions (every entry point):
 *   - all array pointers are caller-owned DEVICE memory, densely packed, naturally aligned
 *     (float4 rows are 16-byte aligned);
 *   - nothing is allocated inside any call; every call is asynchronous on `stream`
 *     (a hipStream_t passed as void*; NULL = the legacy default stream);
 *   - the return value is 0 on success or a nonzero FSS_E* / hipError_t code; the text of the
 *     last error on the calling thread is available from fss_last_error();
 *   - calls on distinct streams are thread-safe.
 *
 * Point layout.  A "cloud" is a run of float4 rows (x, y, z, w) inside one packed array.
 *   - shape channel (atoms or surface points): w = the point's ESP value (0 when unused);
 *   - pharmacophore channel: anchors (x, y, z, 0) and direction vectors (x, y, z, 0) in two
 *     parallel float4 arrays plus one int32 type index (order of P_TYPES) per row.
 * A pair names its ref and fit clouds by (start, count) rows, so several pairs may share one
 * ref cloud (a screening query) without copies.
 *
 * Pose layout.  Pose k of pair p is row p*B + k (B = restarts per pair) of
 *   q (n_pairs*B, 4) float32 -- raw quaternion (w, x, y, z), NOT required to be unit;
 *   t (n_pairs*B, 3) float32 -- translation.  The fit cloud moves: b' = R(q/|q|) b + t.
 */
#ifndef FSS_H
#define FSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Objective modes. */
#define FSS_MODE_SHAPE        0  /* Gaussian shape Tanimoto (vol / surf)                    */
#define FSS_MODE_SHAPE_ESP    1  /* ESP-weighted shape Tanimoto (surf_esp / vol_esp)        */
#define FSS_MODE_PHARM        2  /* typed, directional pharmacophore Tanimoto / Tversky     */
#define FSS_MODE_SHAPE_PHARM  3  /* (1-w)*shape Tanimoto + w*pharm similarity (vol_color)   */
#define FSS_MODE_ESP_COMBO    4  /* ShaEP: (1-w)*shape Tanimoto + w*surface ESP comparison  *
                                  * (vol_and_surf_esp); needs the ShaEP channel below         */

/* Pharmacophore similarity. */
#define FSS_SIM_TANIMOTO 0
#define FSS_SIM_TVERSKY  1

/* Error codes (beside hipError_t values). */
#define FSS_OK           0
#define FSS_E_INVALID    9001  /* bad argument (null pointer, size, mode)                   */
#define FSS_E_TOO_LARGE  9002  /* clouds do not fit the kernel's LDS staging               */
#define FSS_E_NO_DEVICE  9003
#define FSS_E_CLOSED     9004  /* the screening service was destroyed before the query finished */
#define FSS_PENDING      1     /* fss_screen_poll: the query has not finished yet            */

#define FSS_N_PTYPES 9         /* Acceptor Donor Aromatic Hydrophobe Halogen Cation Anion ZnBinder Dummy */

typedef struct FssPairBatch {
  int32_t n_pairs;
  /* shape channel */
  const float*   ref_pts;    /* (rows, 4) */
  const float*   fit_pts;    /* (rows, 4) */
  const int32_t* ref_rng;    /* (n_pairs, 2): first row, row count */
  const int32_t* fit_rng;    /* (n_pairs, 2) */
  /* pharmacophore channel (may be NULL unless the mode uses it) */
  const float*   ref_anc;    /* (rows, 4) */
  const float*   ref_vec;    /* (rows, 4) */
  const int32_t* ref_typ;    /* (rows,)   */
  const int32_t* ref_prng;   /* (n_pairs, 2) */
  const float*   fit_anc;
  const float*   fit_vec;
  const int32_t* fit_typ;
  const int32_t* fit_prng;
  /* host-side maxima over the pairs (size the LDS staging); must be >= every count */
  int32_t max_ref, max_fit, max_ref_ph, max_fit_ph;
  /* ShaEP channel (FSS_MODE_ESP_COMBO only; may be NULL otherwise): observer surface points
   * (x, y, z, precomputed ESP) and source atoms including hydrogens (x, y, z, partial charge)
   * with their vdW radii. */
  const float*   ref_srf;    /* (rows, 4) */
  const float*   fit_srf;    /* (rows, 4) */
  const int32_t* ref_srng;   /* (n_pairs, 2) */
  const int32_t* fit_srng;   /* (n_pairs, 2) */
  const float*   ref_atm;    /* (rows, 4) */
  const float*   fit_atm;    /* (rows, 4) */
  const float*   ref_rad;    /* (rows) */
  const float*   fit_rad;    /* (rows) */
  const int32_t* ref_arng;   /* (n_pairs, 2) */
  const int32_t* fit_arng;   /* (n_pairs, 2) */
  int32_t max_ref_srf, max_fit_srf, max_ref_atm, max_fit_atm;
} FssPairBatch;

/* Scoring constants shared by every entry point. */
typedef struct FssScoreParams {
  int32_t mode;              /* FSS_MODE_*                                                   */
  int32_t similarity;        /* FSS_SIM_* for the pharmacophore channel                     */
  double  alpha;             /* shape Gaussian width (0.81 volumetric)                       */
  double  lam;               /* ESP lambda, already scaled (surface: 0.3*LAM_SCALING); for
                                FSS_MODE_ESP_COMBO the comparison's lambda (LAM_SCALING*lam) */
  double  color_weight;      /* w of FSS_MODE_SHAPE_PHARM                                    */
  double  sigma;             /* Tversky sigma                                                */
  double  ph_alpha[FSS_N_PTYPES];   /* per-type pharmacophore alpha                          */
  double  ph_k[FSS_N_PTYPES];       /* per-type (pi/(2 alpha))^1.5                           */
  int32_t ph_cat[FSS_N_PTYPES];     /* 0 isotropic, 1 directional, 2 aromatic, 3 ignored;
                                       extended points (pharmacophore_scoring.py:284-370):
                                       4 anchor (weight 1) + extended-point overlap,
                                       5 extended-point overlap only; the extended point is
                                       anchor + unit vector                                  */
  double  esp_weight;        /* w of FSS_MODE_ESP_COMBO                                      */
  double  probe_radius;      /* observer points within radius + probe of an atom are masked  */
  double  coulomb;           /* ESP = coulomb * sum q / d (1e4 / (4 * 55.263 * pi))          */
} FssScoreParams;

typedef struct FssAlignParams {
  int32_t n_restarts;        /* B, poses per pair                                            */
  int32_t max_steps;
  float   lr, beta1, beta2, eps;    /* torch.optim.Adam semantics (bias-corrected)           */
  int32_t early_stop;        /* 1: stop a pair once |dL| <= es_tol held > es_patience steps  */
  float   es_tol;
  int32_t es_patience;
  int32_t lanes_per_pose;    /* 1..64: lanes sharing one pose's N x M sum; a power of two     *
                              * unless the workgroup's lanes fit one wavefront (64)            */
  int32_t restarts_per_wg;   /* poses per workgroup; >= B when early_stop is set             */
  const int32_t* pair_order; /* device (n_pairs,) launch order of the pairs (costliest first),
                                or NULL for 0..n_pairs-1; results are always indexed by pair  */
  int32_t skip_finish;       /* 1: run the optimiser only; call fss_align_finish afterwards    */
  int32_t variant;           /* shape inner loop: 0 default (expanded form), 1 direct form,
                                2 expanded form (exp2 of 2a.b' + W_a, factor exp2(-|b'|^2)),
                                3 split-bf16 MFMA (lanes_per_pose 64), 4 expanded form over
                                clouds streamed through LDS tiles.  Clouds that do not fit LDS
                                in the requested layout run the expanded form (3 -> 2), and in
                                FSS_MODE_SHAPE, FSS_MODE_SHAPE_ESP and FSS_MODE_SHAPE_PHARM
                                (whose pharmacophore rows stay staged whole) the streamed form
                                (0, 2, 3 -> 4; pairs_per_wg 1); the other modes then fail with
                                FSS_E_TOO_LARGE                                                */
  int32_t groups_per_wg;     /* lane groups per workgroup (0: restarts_per_wg); each group runs
                                ceil(restarts_per_wg / groups_per_wg) restarts in turn every
                                step, so per-pair early stop can keep all of a pair's restarts in
                                one workgroup with fewer threads (the MFMA loop at 50 restarts) */
  int32_t pairs_per_wg;      /* pairs per workgroup (0 or 1: one); > 1 needs restarts_per_wg >=
                                n_restarts: the workgroup's lane groups work on consecutive
                                pairs of the launch order, so small clouds share one wavefront's
                                per-step pose work (e.g. 6 pairs x 10 restarts x 1 lane) */
  int32_t tile_rows;         /* streamed form: rows of each cloud per LDS tile (0: 1024, fewer
                                when the workgroup's LDS needs it)                             */
} FssAlignParams;

typedef struct FssAlignOut {
  float*   q;                /* (n_pairs*B, 4) in: seeds, out: final raw quaternions         */
  float*   t;                /* (n_pairs*B, 3) in: seeds, out: final translations            */
  double*  pose_score;       /* (n_pairs*B) similarity of every final pose: FP64 sums for the
                                restarts within 1e-3 of the pair's best (FP32 pre-pass; every
                                mode but FSS_MODE_ESP_COMBO, which scores all in FP64); FP32 sums
                                (error ~1e-6) for the rest.  The best restart is the FP64 argmax
                                either way.                                                     */
  double*  best_score;       /* (n_pairs)                                                    */
  int32_t* best_idx;         /* (n_pairs) restart index of the best final pose               */
  float*   best_tf;          /* (n_pairs, 4, 4) SE(3) matrix of the best pose, row-major      */
  int32_t* n_steps;          /* (n_pairs) optimiser steps executed                           */
} FssAlignOut;

/* Library identity and errors. */
const char* fss_version(void);
const char* fss_last_error(void);
int fss_device_count(int* count);

/*
 * Self-overlaps VAA and VBB of every pair at the identity pose (value only).
 * out: (n_pairs, 4) float64 = [VAA_shape, VBB_shape, VAA_pharm, VBB_pharm], evaluated with FP64
 * arithmetic; channels the mode does not use are written as 0.  FSS_MODE_ESP_COMBO writes the
 * two surface point counts (the ESP comparison's normaliser) in place of the pharm pair.
 * Replaces: accel/kernels/shape_triton.py:436-448 (_batch_self_overlap),
 *           accel/kernels/esp_triton.py:393-425 (_batch_self_overlap_esp),
 *           score/analytical_gradients/_torch.py:364-389, 519-541, 569-599.
 */
int fss_self_overlap(const FssPairBatch* batch, const FssScoreParams* sp, double* out,
                     void* stream);

/*
 * Overlap value and analytic gradient of every pose (one workgroup per pose).
 * pose_pair: (n_poses,) int32 pair index of each pose; q (n_poses,4), t (n_poses,3).
 * Per pose and per channel c (0 shape, 1 pharm):
 *   V  (n_poses, 2)      overlap O_AB
 *   gR (n_poses, 2, 9)   dO/dR, row-major (NULL when need_grad == 0)
 *   gT (n_poses, 2, 3)   dO/dt
 *   gQ (n_poses, 2, 4)   dO/dq_hat through the unit-quaternion Jacobians
 * Replaces: accel/kernels/dispatch.py:96-112 table entries overlap_score_grad_se3_batch
 *   (shape_triton.py:191-249), overlap_score_grad_esp_se3_batch (esp_triton.py:175-246),
 *   pharm_grad_dq_se3_batch (pharm_triton.py:403-430), pharm_score_grad_se3_batch
 *   (pharm_triton.py:433-467); arithmetic of ag/_torch.py:223-361 and :441-516.
 */
/* For FSS_MODE_ESP_COMBO the second channel is the ShaEP ESP comparison esp_1 + esp_2
 *   (score/electrostatic_scoring.py:128-227) and its analytic gradient; the mask of observer
 *   points inside the other molecule's radius + probe volume has zero gradient. */
int fss_overlap_grad(const FssPairBatch* batch, const FssScoreParams* sp, int32_t n_poses,
                     const int32_t* pose_pair, const float* q, const float* t, int32_t need_grad,
                     float* V, float* gR, float* gT, float* gQ, void* stream);

/*
 * Loss and SE(3) gradient of every pose for one optimiser step, with the Tanimoto / Tversky
 * chain rule, the unit-quaternion projection and the 1/B scaling of the reference's batched
 * objective.  loss (n_poses,) = 1 - S;  grad (n_poses, 7) = [dL/dq_raw, dL/dt] / B.
 * self_ov: (n_pairs, 4) from fss_self_overlap.
 * Replaces: ag/_torch.py:602-693 (compute_analytical_grad_se3_esp), :696-772
 *   (compute_analytical_grad_se3_shape), :939-1017 (compute_analytical_grad_se3).
 */
int fss_pose_loss_grad(const FssPairBatch* batch, const FssScoreParams* sp, int32_t n_poses,
                       const int32_t* pose_pair, int32_t n_restarts, const double* self_ov,
                       const float* q, const float* t, float* loss, float* grad, void* stream);

/*
 * One torch.optim.Adam step (bias-corrected, eps outside the square root) over n rows of
 * 7 parameters [q_raw, t]; m, v are the moment buffers and `step` the 1-based step count.
 * pair_active (n / n_restarts,) int32 or NULL: rows of pairs whose flag is 0 are left untouched
 * (a pair that stopped early keeps its pose).
 * Replaces: the torch.optim.Adam call of alignment/_torch_analytical.py:108,281,418
 *   (the reference's own accel Adam is accel/kernels/shape_triton.py:253-406).
 */
int fss_adam_step(int32_t n, float* q, float* t, const float* grad, float* m, float* v,
                  int32_t step, float lr, float beta1, float beta2, float eps,
                  const int32_t* pair_active, int32_t n_restarts, void* stream);

/*
 * The accelerated drivers' Adam, in place over n pose rows: q (n,4) unit quaternions, t (n,3),
 * their gradients dQ, dT and first / second moments.  Moments without bias correction, eps
 * inside the square root, q renormalised after the step; tangent_proj != 0 first replaces dQ
 * by dQ - q (dQ . q).
 * Replaces: fused_adam_qt / fused_adam_qt_with_tangent_proj (accel/kernels/dispatch.py:96-112;
 *   shape_triton.py:253-406, the CPU twin cpu.py:112-138).
 */
int fss_adam_qt(int32_t n, float* q, float* t, const float* dQ, const float* dT, float* m_q, float* v_q,
                float* m_t, float* v_t, double lr, double beta1, double beta2, double eps,
                int32_t tangent_proj, void* stream);

/*
 * Early-stop bookkeeping of a stepped optimiser, on device (no host synchronisation):
 * L = mean of the pair's n_restarts losses; counter = |L - last| > tol ? 0 : counter + 1;
 * last = L; n_steps = step; once counter > patience the pair's active flag drops to 0.
 * last (n_pairs) float64 starts at 1.0 (+inf for FSS_MODE_ESP_COMBO, whose reference optimiser
 * starts there: alignment/_torch.py:922); counter, n_steps start at 0; active starts at 1.
 * fss_align applies the same two starting values itself.
 * Replaces: the per-step loss.item() test of alignment/_torch_analytical.py:132-139.
 */
int fss_early_stop_update(int32_t n_pairs, int32_t n_restarts, const float* loss, double* last,
                          int32_t* counter, int32_t* active, int32_t* n_steps, int32_t step,
                          float tol, int32_t patience, void* stream);

/*
 * FP64 similarity of every pose (one wave per pose): score (n_poses,) float64.
 * pose_pair (n_poses,) or NULL for pair = pose / n_restarts.
 * Replaces: the final re-scoring of alignment/_torch_analytical.py:141-158, 445-462, 308-322
 *   (get_SE3_transform + apply_SE3_transform + get_overlap / get_overlap_esp / get_overlap_pharm).
 */
int fss_pose_score(const FssPairBatch* batch, const FssScoreParams* sp, int32_t n_poses,
                   const int32_t* pose_pair, int32_t n_restarts, const double* self_ov,
                   const float* q, const float* t, double* score, void* stream);

/*
 * Persistent multi-restart alignment: seeds in q/t, every optimiser step, per-pair early stop,
 * final re-scoring of every pose and the best-restart selection, in one launch (plus a small
 * finishing launch), one workgroup per (pair, restart-block).
 * Replaces: alignment/_torch_analytical.py:23-478 (the oracle optimisers) and the accel
 *   drivers accel/drivers/shape.py:155-313, esp.py:139-367, pharm.py:123-519.
 */
int fss_align(const FssPairBatch* batch, const FssScoreParams* sp, const FssAlignParams* ap,
              const double* self_ov, FssAlignOut* out, void* stream);

/*
 * The launch fss_align would make for these arguments, without launching: the inner loop it
 * runs (variant 1-4, after the fallbacks for clouds that do not fit LDS), the streamed form's
 * tile rows per side (0 when the clouds are staged whole), the workgroup size, the grid and the
 * LDS bytes per workgroup.  Fails as fss_align would (FSS_E_INVALID, FSS_E_TOO_LARGE).
 * No counterpart in the reference (its Triton kernels tile every pair): lets a caller see when a
 * batch takes the streamed path.
 */
typedef struct FssAlignPlanInfo {
  int32_t variant;
  int32_t tile_ref, tile_fit;
  int32_t block;
  int64_t grid;
  int64_t lds_bytes;
} FssAlignPlanInfo;
int fss_align_plan(const FssPairBatch* batch, const FssScoreParams* sp, const FssAlignParams* ap,
                   FssAlignPlanInfo* info);

/*
 * fss_align without any device allocation by the caller or the library: the self-overlaps are
 * computed inside (as fss_self_overlap does) and, with the per-restart re-scoring scores, kept
 * in a static device array of FSS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.
 * The caller passes the batch, the parameters, the seeds in out->q / out->t and the outputs
 * best_score, best_idx, best_tf and n_steps; out->pose_score is not used; skip_finish must be 0.
 * Batches needing more than the scratch (256-byte-rounded 32 B per pair + 8 B per pose) are
 * rejected with FSS_E_TOO_LARGE.  Results equal fss_align's.  Every call on one GPU shares that
 * GPU's scratch, so calls that overlap in time on one GPU race; nothing in it needs clearing.
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 */
#define FSS_NOALLOC_SCRATCH_BYTES (8u << 20)
int fss_align_noalloc(const FssPairBatch* batch, const FssScoreParams* sp, const FssAlignParams* ap,
                      FssAlignOut* out, void* stream);

/*
 * The finishing half of fss_align (FP64 re-scoring of every final pose, best restart, SE(3)
 * matrix) for a call made with skip_finish = 1; reads out->q, out->t.
 */
int fss_align_finish(const FssPairBatch* batch, const FssScoreParams* sp, int32_t n_restarts,
                     const double* self_ov, FssAlignOut* out, void* stream);

/*
 * Restart seeds on device (one wave per pair): pose r gets quat_table[r] (n_restarts x 4, w x y z)
 * and t = centroid(ref) - R(q) centroid(fit).  flags:
 *   FSS_SEED_PHARM  centroids (and axes) of the pharmacophore anchors instead of the shape clouds;
 *   FSS_SEED_SURF   ... of the ShaEP channel's surface points (vol_and_surf_esp seeds);
 *   FSS_SEED_REF    the reference's seed rule: row 0 is the identity, rows 1-4 the principal-axis
 *                   alignments of the fit onto the ref (FP64 Jacobi eigensolver; n_restarts >= 5),
 *                   and n_restarts == 1 means identity with zero translation.
 * q (n_pairs*B, 4), t (n_pairs*B, 3).
 * Replaces: _initialize_se3_params alignment/_torch.py:253-326 with
 *   quaternions_for_principal_component_alignment_np alignment/utils/pca_np.py:111-166.
 */
#define FSS_SEED_PHARM 1
#define FSS_SEED_REF   2
#define FSS_SEED_SURF  4      /* centroids and axes of the ShaEP channel's surface points */
int fss_seed_poses(const FssPairBatch* batch, int32_t flags, int32_t n_restarts,
                   const float* quat_table, float* q, float* t, void* stream);

/*
 * Differentiable similarity of fixed clouds (score-as-loss).  fit rows are the moved points.
 *   fwd: S (n_pairs,) and, for reuse by bwd, ov (n_pairs, 4) = [VAB, VAA, VBB, 0].
 *   bwd: dS/dfit (fit rows, 4 floats: x, y, z, 0) scaled by gS (n_pairs,), including the
 *        dVBB/dfit term; dS/dref likewise when d_ref != NULL.
 * Modes FSS_MODE_SHAPE and FSS_MODE_SHAPE_ESP.  The *_f64 variants take double clouds.
 * Both clouds of a pair are staged in LDS while max_ref + max_fit is at most the per-workgroup
 * LDS size / 20 (float: the point and -|p|^2) or / 32 (double) less the reduction scratch, i.e.
 * 8187 / 5114 points on gfx950; larger batches run
 * streamed kernels (each wavefront stages the 64-point block it works on), with the same sums in
 * the same order.  bwd is deterministic (fixed reduction order).
 * No reference counterpart for bwd w.r.t. coordinates; forward follows
 *   score/gaussian_overlap.py:16-106 and score/electrostatic_scoring.py:17-125.
 */
int fss_score_fwd(const FssPairBatch* batch, const FssScoreParams* sp, float* S, float* ov,
                  void* stream);
int fss_score_bwd(const FssPairBatch* batch, const FssScoreParams* sp, const float* ov,
                  const float* gS, float* d_fit, float* d_ref, void* stream);
int fss_score_fwd_f64(const FssPairBatch* batch, const FssScoreParams* sp, const double* ref_pts,
                      const double* fit_pts, double* S, double* ov, void* stream);
int fss_score_bwd_f64(const FssPairBatch* batch, const FssScoreParams* sp, const double* ref_pts,
                      const double* fit_pts, const double* ov, const double* gS, double* d_fit,
                      double* d_ref, void* stream);

/* ----------------------------------------------------------------------------------------------
 * Screening service: a conformer library resident in HBM, and a service that aligns queued
 * queries against all of it with continuous batching and keeps each query's top-k hits.
 *
 * The service owns one worker thread and one HIP stream (default priority) on the library's
 * device.  Every launch takes up to launch_pairs (query, library entry) pairs from the admitted
 * queries, oldest first (the tail of one query and the head of the next share a launch; a query
 * that arrives while others run joins as soon as capacity frees), runs fss_seed_poses,
 * fss_self_overlap and fss_align on them, and merges each query's new pair
 * scores into its device-resident top-k (ties to the lower library index).  A query's hits do not
 * depend on the queries it shares launches with.  Device memory is allocated by fss_lib_create /
 * fss_screen_create only; submit and poll allocate none.  Calls on one service may come from any
 * thread.
 * Replaces: screen / screen_many shepherd_score/screen.py:1038-1192 (shard streaming, per-query
 *   aligner calls, _TopK :898-933, Hit :74); the library replaces the streamed ProfileStore
 *   shards (screen.py:317-605) with one HBM-resident copy.
 * -------------------------------------------------------------------------------------------- */
typedef struct FssLibrary FssLibrary;
typedef struct FssScreen FssScreen;

#define FSS_LIB_SHAPE 1        /* rows (x, y, z, esp) per conformer */
#define FSS_LIB_PHARM 2        /* anchors, unit vectors, types per conformer */
#define FSS_LIB_SHAEP 4        /* ShaEP: surface rows (x, y, z, ESP), atoms with hydrogens
                                  (x, y, z, partial charge) and their radii per conformer */
#define FSS_LIB_SURFACE 8      /* surface rows (x, y, z, ESP) alone, for the surface shape modes
                                  (FSS_LIB_SHAEP holds them too); with FSS_LIB_SHAPE holding the
                                  heavy atoms, one library serves the atom and surface modes */

typedef struct FssLibSpec {
  int32_t device;              /* HIP device ordinal */
  int32_t channels;            /* FSS_LIB_SHAPE | FSS_LIB_PHARM | FSS_LIB_SHAEP | FSS_LIB_SURFACE */
  int64_t max_conformers;
  int64_t max_points;          /* total shape rows */
  int64_t max_pharm;           /* total pharmacophore rows */
  int64_t max_surface;         /* total surface rows (FSS_LIB_SHAEP or FSS_LIB_SURFACE) */
  int64_t max_atoms;           /* total ShaEP atom rows */
} FssLibSpec;

/* Rows of n conformers for fss_lib_append (host or device pointers; the channels the library
 * was created with must be present, the others are ignored). */
typedef struct FssLibRows {
  const float* pts;            /* shape: (sum(counts), 4) */
  const int32_t* counts;       /* (n) */
  const float* anc;            /* pharm: (sum(pcounts), 4) */
  const float* vec;            /* (sum(pcounts), 4) */
  const int32_t* typ;          /* (sum(pcounts)) */
  const int32_t* pcounts;      /* (n) */
  const float* srf;            /* ShaEP / surface: (sum(scounts), 4) surface points and ESP */
  const int32_t* scounts;      /* (n) */
  const float* atm;            /* ShaEP: (sum(acounts), 4) atoms with hydrogens and partial charges */
  const float* rad;            /* (sum(acounts)) vdW radii */
  const int32_t* acounts;      /* (n) */
} FssLibRows;

/* Allocates the library's device arrays for the given capacities. */
int fss_lib_create(const FssLibSpec* spec, FssLibrary** out);
/* Appends n conformers (host or device pointers; copied before return).  Shape channel: pts
 * (sum(counts), 4) rows, counts (n).  Pharm channel: anc, vec (sum(pcounts), 4), typ (sum(pcounts)),
 * pcounts (n).  *first receives the library index of the first appended conformer. */
int fss_lib_upload(FssLibrary* lib, int32_t n, const float* pts, const int32_t* counts,
                   const float* anc, const float* vec, const int32_t* typ, const int32_t* pcounts,
                   int64_t* first);
/* fss_lib_upload for every channel, the ShaEP and surface ones included.  Replaces: the with-H
 * store reader's per-molecule views (screen.py:830-846). */
int fss_lib_append(FssLibrary* lib, int32_t n, const FssLibRows* rows, int64_t* first);
int fss_lib_size(const FssLibrary* lib, int64_t* n);
/* Fails while a service on this library exists. */
int fss_lib_destroy(FssLibrary* lib);

#define FSS_SCREEN_MAX_TOPK 1024

typedef struct FssScreenParams {
  int32_t n_restarts;          /* restarts per pair (seeds: identity, then quat_table rows) */
  int32_t max_steps;
  float lr, beta1, beta2, eps;
  int32_t early_stop;          /* per-pair early stop of fss_align */
  float es_tol;
  int32_t es_patience;
  int32_t top_k;               /* 1 .. FSS_SCREEN_MAX_TOPK */
  int32_t launch_pairs;        /* pairs per launch */
  int32_t max_active;          /* queries admitted at once (the rest wait in FIFO order) */
  int32_t max_query_points;    /* shape rows of one query */
  int32_t max_query_pharm;     /* pharmacophore rows of one query */
  int32_t lanes_per_pose, restarts_per_wg, variant;   /* fss_align tiling; 0 = defaults (every
                                                         restart in one wavefront, 64 / n_restarts
                                                         lanes each, for up to 64 restarts) */
  int32_t seed_flags;          /* FSS_SEED_PHARM / FSS_SEED_SURF: seed from those rows */
  const float* quat_table;     /* (n_restarts, 4) host rotations of the restarts */
  int32_t max_query_surface;   /* surface rows of one query (FSS_MODE_ESP_COMBO, or
                                  shape_from_surface) */
  int32_t max_query_atoms;     /* ShaEP atom rows of one query (FSS_MODE_ESP_COMBO) */
  int32_t shape_from_surface;  /* 1: the shape term uses the library's surface rows and the
                                  query's srf rows: FSS_MODE_ESP_COMBO at alpha != 0.81, and
                                  the surface modes (FSS_MODE_SHAPE / FSS_MODE_SHAPE_ESP, ESP in
                                  the rows' w) on a library with surface rows.  0: the shape
                                  channel's rows. */
  int64_t query_budget_bytes;  /* device bytes for the rows of admitted queries (see
                                  fss_plan_query_pools); 0: max_active queries of the maximum
                                  size.  A queued query is admitted, oldest first, when a slot
                                  and room for its own rows are free. */
  int32_t pairs_per_wg;        /* fss_align pairs per workgroup (0 or 1: one); > 1 also orders
                                  each launch's pairs by their entry's row count, so the pairs
                                  sharing a workgroup have alike sizes */
  int32_t groups_per_wg;       /* fss_align lane groups per workgroup (0: restarts_per_wg); each
                                  runs its share of the pair's restarts in turn (the MFMA loop,
                                  variant 3, for surface modes) */
} FssScreenParams;

typedef struct FssQuery {      /* host arrays, copied by fss_screen_submit */
  int32_t n_points;
  const float* pts;            /* (n_points, 4): x, y, z, esp */
  int32_t n_pharm;
  const float* anc;            /* (n_pharm, 4) */
  const float* vec;            /* (n_pharm, 4) */
  const int32_t* typ;          /* (n_pharm) */
  int32_t n_surface;           /* FSS_MODE_ESP_COMBO: */
  const float* srf;            /* (n_surface, 4): x, y, z, ESP */
  int32_t n_atoms;
  const float* atm;            /* (n_atoms, 4): x, y, z, partial charge (with hydrogens) */
  const float* rad;            /* (n_atoms) */
  int32_t all_scores;          /* 1: also keep every library entry's best score, returned by
                                  fss_screen_poll in FssHits.all_scores (screen.py:1174-1176,
                                  scores_out) */
} FssQuery;

typedef struct FssHits {       /* host arrays of capacity entries, filled by fss_screen_poll */
  int32_t capacity;
  int32_t n;                   /* out: hits written, best first */
  double* score;
  int64_t* index;              /* library index */
  float* transform;            /* (capacity, 16): row-major 4x4 moving the library entry */
  double queue_ms;             /* out: submit -> admission */
  double latency_ms;           /* out: submit -> result */
  double* all_scores;          /* for a query submitted with all_scores: every library entry's
                                  best score in library order                                 */
  int64_t n_all;               /* in: capacity of all_scores; out: entries written (the library
                                  size at admission; 0 when the query did not ask)            */
} FssHits;

int fss_screen_create(FssLibrary* lib, const FssScoreParams* sp, const FssScreenParams* prm,
                      FssScreen** out);
/* Queues a query against the whole library (its size at admission); *ticket names it. */
int fss_screen_submit(FssScreen* svc, const FssQuery* q, int64_t* ticket);
/* Waits up to wait_ms (< 0: no limit) for the ticket's hits.  FSS_OK: hits written and the
 * ticket released; FSS_PENDING: not finished yet; an error code: the query failed.  A query
 * submitted with all_scores needs out->all_scores with n_all >= the library size: otherwise
 * FSS_E_INVALID, and the result stays held for a retry with a large enough buffer. */
int fss_screen_poll(FssScreen* svc, int64_t ticket, int32_t wait_ms, FssHits* out);
/* Launches run, (query, library entry) pairs aligned, and the optimiser steps those pairs ran
 * (summed over pairs; each step of a pair covers all its restarts).  Any pointer may be NULL. */
int fss_screen_stats(FssScreen* svc, int64_t* launches, int64_t* pairs, int64_t* pair_steps);
/* The work those steps held: summed over aligned pairs, executed steps x restarts x row pairs of
 * the shape term (query rows x entry rows), of the pharmacophore term (all row pairs, every
 * type), and of the ShaEP potential (each side's surface rows x the other side's atoms).  The
 * algorithmic FLOP of a screen is these counts times the per-term costs.  Any pointer may be
 * NULL. */
int fss_screen_terms(FssScreen* svc, double* shape_terms, double* pharm_terms, double* esp_terms);
/* Stops the worker; queries still queued or running fail with FSS_E_CLOSED.  No other call on
 * the service may be in progress or follow. */
int fss_screen_destroy(FssScreen* svc);
/* The service's launch planner (host only): fills capacity from n requests, given oldest first,
 * with remaining[i] entries left each: take[i] = min(remaining[i], what the earlier ones left).
 * For whole-library queries this FIFO filling gives lower median and tail latency than equal
 * shares (processor sharing) at the same throughput. */
int fss_plan_launch(int32_t n, const int64_t* remaining, int64_t capacity, int64_t* take);

/* Admission by memory.  An admitted query's rows live in four device pools (FSS_QPOOL_*: shape
 * rows, 16 bytes a row; pharmacophore anchor + vector + type, 36; ShaEP surface rows, 16; ShaEP
 * atom + radius, 20); each pool hands out contiguous row ranges first-fit and merges them back
 * when the query finishes.  This replaces the reference's admission by free device memory
 * (accel/batch/_bucket.py:76-98, _pad.py:24-97: split a bucket until it fits, retry on OOM) with
 * a budget fixed when the service is created, so a query never fails for memory: it waits. */
#define FSS_QPOOL_POINTS 0
#define FSS_QPOOL_PHARM 1
#define FSS_QPOOL_SURFACE 2
#define FSS_QPOOL_ATOMS 3
#define FSS_QPOOLS 4
/* Rows of each pool that fss_screen_create reserves for these parameters (host only).  Pools the
 * mode does not read get 0 rows.  With query_budget_bytes = 0, or at least max_active queries of
 * the maximum size, every pool holds max_active maximum queries; otherwise the budget is split
 * over the pools in proportion to one maximum query's bytes in each, and each pool holds at
 * least one maximum query.  FSS_E_INVALID if the budget does not hold one maximum query. */
int fss_plan_query_pools(const FssScoreParams* sp, const FssScreenParams* prm, int64_t rows[FSS_QPOOLS]);
/* The first-fit row pool of the service (host only; no device memory). */
typedef struct FssRowPool FssRowPool;
int fss_rowpool_create(int64_t rows, FssRowPool** out);
/* First free range of n rows: FSS_OK and *first, or FSS_PENDING when no range is long enough.
 * n = 0 always succeeds at row 0. */
int fss_rowpool_alloc(FssRowPool* pool, int64_t n, int64_t* first);
/* Returns a range from fss_rowpool_alloc; FSS_E_INVALID if it is not wholly allocated. */
int fss_rowpool_free(FssRowPool* pool, int64_t first, int64_t n);
int fss_rowpool_stats(const FssRowPool* pool, int64_t* free_rows, int64_t* largest_free);
int fss_rowpool_destroy(FssRowPool* pool);
/* The service's admission step (host only): queries k = 0 .. n-1, oldest first, needing
 * rows[k * n_pools + c] rows of pools[c], are admitted in order while fewer than free_slots have
 * been and every pool has a range for the next one; the first that does not fit stops the step
 * (no later query overtakes it).  first[k * n_pools + c] receives the ranges of the admitted
 * ones, *admitted their number. */
int fss_admit_fifo(FssRowPool* const* pools, int32_t n_pools, int32_t n, const int64_t* rows,
                   int32_t free_slots, int64_t* first, int32_t* admitted);
/* The service's pool sizes (rows[FSS_QPOOLS], as fss_plan_query_pools), the most queries it has
 * held admitted at once, and how many queries waited for pool room while a slot was free (each
 * counted once).  Any pointer may be NULL. */
int fss_screen_admission(FssScreen* svc, int64_t* pool_rows, int64_t* max_admitted,
                         int64_t* memory_waits);

#ifdef __cplusplus
}
#endif

#endif /* FSS_H */
