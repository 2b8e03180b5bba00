/*
 * apirl_hip.h -- C-ABI of the MI355X (gfx950) Siddon projector / backprojector and
 * MLEM / OSEM update.
 *
 * This is the drop-in boundary for APIRL's Projector API and EM drivers.  Every
 * entry point takes plain pointers and sizes; device pointers are hipMalloc'd
 * float arrays in the reference's layouts:
 *   image     float[nz][ny][nx]            (data/src/Images.cpp:34-56)
 *   sinogram  float[nsino][nang_sub][nr]   sinograms segment-major, angles of the
 *                                          selected subset only, r fastest
 *                                          (data/src/Sinogram3D.cpp:396-464,
 *                                          Sinogram3DCylindricalPet.cpp:164-197)
 * Errors: every function returns an apirl_status; 0 is success.  The message of
 * the last failure on a context is returned by apirl_last_error().  No function
 * calls exit() (the reference's CUDA path does, reconGPU/inc/helper_cuda.h:32-40).
 * Threading: contexts and plans carry all state; no globals, so distinct contexts
 * may be used from distinct threads.  Work is enqueued on the context's stream.
 */
#ifndef APIRL_HIP_H
#define APIRL_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum apirl_status {
  APIRL_OK = 0,
  APIRL_ERR_INVALID = 1,     /* bad argument or size */
  APIRL_ERR_HIP = 2,         /* HIP runtime error (message has the HIP error string) */
  APIRL_ERR_NOMEM = 3,       /* device allocation failed */
  APIRL_ERR_UNSUPPORTED = 4  /* combination not implemented (the reference's "return false") */
} apirl_status;

/* Backprojection forms. */
enum {
  APIRL_BP_SCATTER = 0, /* ray-driven, fp32 atomics into image copies: order not fixed */
  APIRL_BP_GATHER = 1,  /* voxel-driven, fixed order: deterministic; in 2D equal, bit for
                           bit, to the reference's single-thread result; in 3D the same
                           float terms summed in (angle, bin, segment, sinogram) order */
  APIRL_BP_TILE = 2     /* 3D, APIRL_PREC_REFERENCE: ray-driven walks clipped to tiles of
                           8 x 8 columns by a slab of slices, summed in LDS by one
                           wavefront per tile (no global atomics): deterministic, the
                           same float terms in the wavefront's order */
};

/* Arithmetic of the projector and of the gather backprojector. */
enum {
  APIRL_PREC_REFERENCE = 0, /* fp32, the reference's operation order (bit-faithful) */
  APIRL_PREC_KAHAN = 1,     /* fp32 with Kahan-compensated accumulation along the ray */
  APIRL_PREC_F64 = 2,       /* all double (3D): the Siddon setup, the alpha stepping, the
                               weights and the sum */
  APIRL_PREC_F64_SETUP = 3  /* the ray / plane alpha intersection setup in double (3D: alpha
                               range, entry and exit voxels, list length, first crossings and
                               increments, ray length), rounded to float; the stepping by
                               repeated float additions, float weights and a Kahan-compensated
                               float sum.  2D: as APIRL_PREC_KAHAN */
};
#define APIRL_PREC_F64_ALPHA APIRL_PREC_F64 /* earlier name of the all-double mode */

typedef struct apirl_ctx apirl_ctx;
typedef struct apirl_plan apirl_plan;

typedef struct apirl_image_desc {
  int nx, ny, nz;   /* nz = 1 for 2D images */
  float dx, dy, dz; /* voxel size in mm (Images.h SizeImage) */
} apirl_image_desc;

/* Cylindrical-scanner sinogram geometry.
 *   dims = 2: one 2D sinogram (Sinogram2DinCylindrical3Dpet, 2D LOR form,
 *             Sinogram2DinCylindrical3Dpet.cpp:249-266); z arrays ignored.
 *   dims = 3: nsino 2D sinograms (Sinogram3DCylindricalPet); LOR z taken from the
 *             per-sinogram mean ring coordinates (SiddonProjector.cpp:349-361,394-395).
 * ang_deg / r_mm may be NULL for the reference's uniform grids
 * (Sinogram2DinCylindrical3Dpet.cpp:46-60). */
typedef struct apirl_sino_desc {
  int dims;
  int nang;            /* projection angles of the full sinogram */
  int nr;              /* radial bins */
  int nsino;           /* 2D sinograms (1 when dims = 2) */
  float rfov_mm;       /* radius of the radial grid */
  float rscanner_mm;   /* scanner radius */
  float ring_width_mm; /* axial FOV / rings (3D only) */
  const float* ang_deg;   /* [nang] or NULL */
  const float* r_mm;      /* [nr] or NULL */
  const float* z1_mean_mm; /* [nsino] mean axial coordinate of ring 1 (3D) */
  const float* z2_mean_mm; /* [nsino] mean axial coordinate of ring 2 (3D) */
} apirl_sino_desc;

/* ---- context ---------------------------------------------------------------- */
/* Replaces initCuda()/checkCudaErrors of reconGPU (CuMlemSinogram3d.cu:60-137). */
apirl_status apirl_ctx_create(int device, apirl_ctx** out);
apirl_status apirl_ctx_destroy(apirl_ctx* ctx);
const char* apirl_last_error(const apirl_ctx* ctx);
/* Work is enqueued on this stream (hipStream_t; NULL = the null stream). */
apirl_status apirl_ctx_set_stream(apirl_ctx* ctx, void* stream);
apirl_status apirl_ctx_synchronize(apirl_ctx* ctx);
/* Device memory helpers for callers without a HIP runtime of their own. */
apirl_status apirl_malloc(apirl_ctx* ctx, void** dptr, size_t bytes);
apirl_status apirl_free(apirl_ctx* ctx, void* dptr);
apirl_status apirl_memcpy_h2d(apirl_ctx* ctx, void* dst, const void* src, size_t bytes);
apirl_status apirl_memcpy_d2h(apirl_ctx* ctx, void* dst, const void* src, size_t bytes);
apirl_status apirl_memset(apirl_ctx* ctx, void* dptr, int value, size_t bytes);

/* ---- geometry plan ------------------------------------------------------------ */
/* Host-side LOR tables (endpoints with the same libm calls as the reference,
 * Sinogram2DinCylindrical3Dpet.cpp:216-266, and the xy part of the Siddon setup,
 * recon/src/Siddon.cpp:83-145,177-181,253-286) computed once and kept in HBM.
 * Replaces CuMlemSinogram3d::InitGpuMemory (CuMlemSinogram3d.cu:216-368). */
apirl_status apirl_plan_create(apirl_ctx* ctx, const apirl_image_desc* image,
                               const apirl_sino_desc* sino, apirl_plan** out);
apirl_status apirl_plan_destroy(apirl_plan* plan);
/* Number of angles in subset `subset` of `nsubsets` (Sinogram3DCylindricalPet.cpp:168-171). */
int apirl_subset_num_angles(int nang, int nsubsets, int subset);
/* Bins of a (subset) sinogram: nsino * subset angles * nr. */
size_t apirl_plan_subset_bins(const apirl_plan* plan, int subset, int nsubsets);
size_t apirl_plan_image_voxels(const apirl_plan* plan);
/* Sum over the subset's rays of the reference's visited-voxel count
 * |di|+|dj|+|dk| (recon/src/Siddon.cpp:220,521); computed on the device. */
apirl_status apirl_plan_intersections(apirl_plan* plan, int subset, int nsubsets,
                                      unsigned long long* out);

/* Launch-shape tuning of a plan.  No counterpart in the reference (its CPU projector
 * has no tiles; reconGPU fixes them at compile time, CuMlemSinogram3d.cu:20-58):
 *   gather_zlanes: voxel mapping of the 3D gather backprojector.  1 = each workgroup
 *                  a 16 x 16 tile of one slice; Z in 2..64 = each 64-lane wavefront
 *                  64/Z neighbouring voxel columns times Z consecutive slices.
 *   gather_zvox:   with gather_zlanes = 1, voxels per lane along z (1..4): a lane
 *                  gathers V consecutive voxels of a column, sharing each candidate
 *                  ray's transaxial tests and sinogram search among them (float
 *                  precision modes; APIRL_PREC_F64_ALPHA always uses 1).
 *   tile_zs:       slices per item of the tile backprojector (APIRL_BP_TILE; default
 *                  about 5).
 *   tile_cache_mb: device memory (MiB, default 16384) in which the plan keeps the
 *                  tile backprojector's per-subset geometry from its first call on:
 *                  the candidate rays of every 8 x 8-column tile with their paths, and
 *                  for each (tile, slab, ray) the sinograms that reach the slab inside
 *                  the tile (2 bytes each), which later calls read instead of listing
 *                  them.  -1 keeps none and frees what is kept; a subset beyond the
 *                  budget is rebuilt on every call.  The first call on a subset waits
 *                  for the stream once (to size its lists).
 * Gather results do not depend on these (the per-voxel summation order is fixed);
 * tile results are deterministic for a given tile_zs, kept geometry or not.
 * Zero in a field leaves it unchanged. */
typedef struct apirl_tuning {
  int gather_zlanes;
  int gather_zvox;
  int gather_xcd;     /* workgroup order of the column kernel (below) */
  int tile_zs;
  int tile_cache_mb;
} apirl_tuning;
/* gather_xcd: APIRL_GATHER_XCD_ROUND_ROBIN takes the 16 x 16 x zvox tiles in grid
 * order, which the hardware deals to the 8 XCDs in turn; APIRL_GATHER_XCD_WEDGES gives
 * each XCD the tiles of one 45-degree sector of the image (every z block), so the
 * candidate rays a tile shares with its neighbours are fetched into one XCD's L2
 * (the default: 2.5 ms against 2.9 ms for one 12-angle subset of the 256x256x81 /
 * 344x252x837 workload). */
#define APIRL_GATHER_XCD_ROUND_ROBIN 1
#define APIRL_GATHER_XCD_WEDGES 2
/* Oversampled LORs (SiddonProjector(numSamples, numAxialSamples),
 * recon/src/SiddonProjector.cpp:28-32, 366-425, 491-548): every bin of a 3D plan is
 * traced as ns transaxial x nas axial LORs (Sinogram2DinCylindrical3Dpet.cpp:283-299),
 * all of whose terms go to the bin.  ns = nas = 1 restores one LOR per bin.  The
 * projector, both backprojectors (plain, sensitivity, divide), the EM driver and
 * batches follow it, in APIRL_PREC_REFERENCE.  The gather takes the ns x nas LORs of
 * a bin as ns x nas plans of one LOR per bin (built on its first call, about the
 * plan's own size each), gathered in turn into the image; batches run such jobs with
 * the single-job kernels after each grouped round. */
apirl_status apirl_plan_set_oversampling(apirl_plan* plan, int ns, int nas);

/* Scanner geometry of a 3D plan's LOR endpoints.  APIRL_SCANNER_CYLINDRICAL: the
 * plan's angle and radial grids on the cylinder of rscanner_mm
 * (Sinogram2DinCylindrical3Dpet.cpp:216-229).  APIRL_SCANNER_MMR: the Siemens
 * Biograph mMR (Sinogram2DinSiemensMmr.cpp:27-50, 175-192; constants
 * Sinogram3DSiemensMmr.cpp:9-27): arc-corrected radial positions of one detector
 * element per bin, odd bins half an angular step earlier, LORs ending on the radius
 * 328 + 9.6 mm (mean depth of interaction).  The plan's nang x nr sinogram grid
 * keeps its shape; the ring table (z1, z2 and the ring width 260 mm / rings) stays
 * the caller's.  Both backprojectors and batches take either form; oversampling
 * (apirl_plan_set_oversampling) needs the cylindrical one. */
#define APIRL_SCANNER_CYLINDRICAL 0
#define APIRL_SCANNER_MMR 1
/* APIRL_SCANNER_MMR_SUBSET: the LORs the reference's OSEM traces on mMR data.
 * Sinogram3DSiemensMmr::getSubset (Sinogram3DSiemensMmr.cpp:150-195) constructs its
 * subset with the parametric constructor, whose segments are built inside the
 * cylindrical base constructor, where the virtual inicializarSegmentos is the base
 * class's (Sinogram3DCylindricalPet): the subset's 2D sinograms are cylindrical ones
 * on the radius 328 mm holding the mMR's arc-corrected radial positions (copied,
 * :185-193), with no half-step for odd bins and no depth of interaction
 * (Sinogram2DinCylindrical3Dpet.cpp:216-229).  OsemSinogram3d projects and
 * backprojects every subset, sensitivities included, on these LORs, and only its
 * final full projection (a Copy(), which keeps the mMR class) on the mMR's.
 * apirl_em_reconstruct with osem_rules on an APIRL_SCANNER_MMR plan, and batch jobs
 * likewise, switch to this form for the sub-iterations by themselves. */
#define APIRL_SCANNER_MMR_SUBSET 2
apirl_status apirl_plan_set_scanner(apirl_plan* plan, int scanner);
apirl_status apirl_plan_get_tuning(const apirl_plan* plan, apirl_tuning* out);
apirl_status apirl_plan_set_tuning(apirl_plan* plan, const apirl_tuning* tuning);

/* ---- Projector API (recon/inc/Projector.h:48-85) ------------------------------ */
/* Project: d_sino[subset bins] := A d_img.  Replaces SiddonProjector::Project
 * (2D SiddonProjector.cpp:146-193, 3D :322-441).  2D bins whose LOR the reference
 * rejects (r > rfov) are left untouched, as there. */
apirl_status apirl_project(apirl_plan* plan, const float* d_img, float* d_sino, int subset,
                           int nsubsets, int precision);
/* Backproject: d_img += A^T d_sino (accumulates; the caller zeroes, Mlem2d.cpp:162).
 * Replaces SiddonProjector::Backproject (2D :34-87, 3D :444-568).  precision applies
 * to the gather form: APIRL_PREC_REFERENCE sums the reference's float weights in
 * float, APIRL_PREC_KAHAN with a compensation term, APIRL_PREC_F64 computes the
 * crossings and sums in double; the scatter form takes APIRL_PREC_REFERENCE only and
 * allocates its 16 per-XCD image copies on its first call on a plan. */
apirl_status apirl_backproject(apirl_plan* plan, const float* d_sino, float* d_img, int subset,
                               int nsubsets, int mode, int precision);
/* The backprojection fused with the EM update that consumes it (Backproject, then
 * MlemSinogram3d.cpp:228-242, OsemSinogram3d.cpp:217-237 or Mlem2d.cpp:171-185):
 * d_img[v] := cond ? d_img[v] * b[v] / d_sens[v] : 0, b = A^T d_sino, cond = d_sens[v]
 * > thr (strict) or >= thr.  Bit for bit apirl_backproject (mode APIRL_BP_GATHER or
 * APIRL_BP_TILE) into a zeroed image followed by apirl_em_update, without storing b.
 * One LOR per bin. */
apirl_status apirl_backproject_update(apirl_plan* plan, const float* d_sino, float* d_img,
                                      const float* d_sens, float thr, int strict, int subset,
                                      int nsubsets, int mode, int precision);
/* DivideAndBackproject: d_img += A^T r with r the reference's ratio rule per segment
 * (3D SiddonProjector.cpp:571-704; 2D :90-144).  APIRL_BP_GATHER evaluates the rule
 * per bin, then gathers (deterministic; w (m / e) instead of the reference's
 * (w m) / e, within float rounding). */
apirl_status apirl_divide_backproject(apirl_plan* plan, const float* d_meas, const float* d_est,
                                      float* d_img, int subset, int nsubsets, int mode);

/* ---- fused EM steps ------------------------------------------------------------ */
/* Flags of apirl_project_ratio. */
#define APIRL_RATIO_LIK_POSITIVE 1  /* likelihood over est > 0 (Sinogram2D.cpp:565-581);
                                       without it over est != 0 (Sinogram3D.cpp:584-607) */
#define APIRL_RATIO_ADDITIVE_LOCAL 2 /* bin (s, k, r) of the subset takes the additive
                                       sinogram's angle k (the subset-local index into
                                       the full sinogram), as OsemSinogram3d.cpp:172-173
                                       adds it through Sinogram3D::addBinToBin
                                       (Sinogram3D.cpp:857-873); without it the bin's own
                                       angle subset + nsubsets k */
/* One forward step of an EM sub-iteration, fused: est = A x (+ additive, no clamp:
 * Sinogram3D.cpp:857-873), the likelihood partial sum y log(est) - est and the ratio
 * ratio = (y != 0 && est != 0) ? y / est : 0 (Sinogram3D.cpp:830-856).
 * d_meas and d_additive are FULL sinograms (all angles); d_ratio is the subset
 * sinogram.  d_additive and d_lik may be NULL; *d_lik receives a double.
 * flags: APIRL_RATIO_* (0 and 1 keep their meaning of lik_positive_only). */
apirl_status apirl_project_ratio(apirl_plan* plan, const float* d_img, const float* d_meas,
                                 const float* d_additive, float* d_ratio, double* d_lik,
                                 int subset, int nsubsets, int flags, int precision);
/* apirl_project_ratio that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The likelihood's block partials go
 * to a file-scope device array of 2^19 doubles (4 MiB per GPU, held until the process
 * ends) instead of the context's buffer, so a call with d_lik != NULL takes at most
 * APIRL_NOALLOC_MAX_BINS subset bins (larger ones fail with APIRL_ERR_INVALID); with
 * d_lik == NULL any size works.  The plan's own buffers (allocated by
 * apirl_plan_create) are used as by apirl_project_ratio; nothing is allocated or
 * freed.  Calls that overlap on one GPU share the array and race: serialise them
 * (one stream, or events). */
#define APIRL_NOALLOC_MAX_BINS ((size_t)1 << 27)
apirl_status apirl_project_ratio_noalloc(apirl_plan* plan, const float* d_img,
                                         const float* d_meas, const float* d_additive,
                                         float* d_ratio, double* d_lik, int subset,
                                         int nsubsets, int flags, int precision);
/* x := cond ? x * b / s : 0, cond = s > thr (strict) or s >= thr
 * (Mlem2d.cpp:171-185, MlemSinogram3d.cpp:228-242, OsemSinogram3d.cpp:217-237). */
apirl_status apirl_em_update(apirl_ctx* ctx, float* d_x, const float* d_b, const float* d_s,
                             size_t n, float thr, int strict);
/* Sensitivity of a subset: d_sens := forcePositive(A^T m) with m = 1 or the subset of
 * a multiplicative sinogram d_mult (FULL layout, may be NULL)
 * (OsemSinogram3d.cpp:327-346, Mlem2d.cpp:262-278, Images.cpp:383-395).
 * force_positive = 0 skips the clamp (the MLEM drivers do not clamp). */
apirl_status apirl_sensitivity(apirl_plan* plan, const float* d_mult, float* d_sens, int subset,
                               int nsubsets, int mode, int force_positive);
/* MLEM update threshold min_nonzero + 1e-4 (max - min_nonzero) (Mlem.cpp:39-42,
 * Images.cpp:397-431); result written to *h_thr. */
apirl_status apirl_mlem_threshold(apirl_ctx* ctx, const float* d_sens, size_t n, float* h_thr);

/* ---- EM drivers (recon/src/Mlem2d.cpp, MlemSinogram3d.cpp, OsemSinogram3d.cpp) --- */
typedef struct apirl_em_options {
  int iterations;
  int subsets;         /* 1 = MLEM */
  int bp_mode;         /* APIRL_BP_SCATTER or APIRL_BP_GATHER */
  int precision;       /* APIRL_PREC_* for the forward step */
  int osem_rules;      /* 1: per-subset sensitivity with forcePositive, threshold 0 with >=
                          (OsemSinogram3d); 0: MLEM rules (threshold min+1e-4(max-min),
                          strict > for 2D, >= for 3D) */
  int additive_own_angles; /* 0: the reference's OSEM adds the additive sinogram's
                          angle k to subset bin k (APIRL_RATIO_ADDITIVE_LOCAL); 1: each
                          bin's own angle.  Equal for one subset. */
  int sensitivity_given; /* 1: d_sens_out already holds the sensitivity image(s) (the
                          reference's "sensitivity filename": OsemSinogram3d.cpp:83-95,
                          MlemSinogram3d.cpp:132-138): they are used instead of computed,
                          after forcePositive in place under OSEM rules, with the MLEM
                          threshold taken from them otherwise. */
} apirl_em_options;

/* Runs the reconstruction with every buffer resident on the device.
 * d_meas: full measured sinogram; d_init: initial image; d_img: output image;
 * d_additive / d_mult may be NULL; h_lik (may be NULL) receives
 * iterations*subsets + 1 likelihood values in the reference's order;
 * d_sens_out (may be NULL) receives the sensitivity image(s). */
apirl_status apirl_em_reconstruct(apirl_plan* plan, const apirl_em_options* opt,
                                  const float* d_meas, const float* d_init,
                                  const float* d_additive, const float* d_mult, float* d_img,
                                  double* h_lik, float* d_sens_out);
/* The same, also timing every sub-iteration as the reference's drivers log it
 * (OsemSinogram3d.cpp:155-240, 303-313; MlemSinogram3d.cpp, Mlem2d.cpp alike):
 * h_times (may be NULL) receives 4 x iterations*subsets floats in ms, the rows
 * [whole sub-iteration | projection + ratio | backprojection | update], measured with
 * HIP events on the context's stream. */
apirl_status apirl_em_reconstruct_timed(apirl_plan* plan, const apirl_em_options* opt,
                                        const float* d_meas, const float* d_init,
                                        const float* d_additive, const float* d_mult,
                                        float* d_img, double* h_lik, float* d_sens_out,
                                        float* h_times);

/* ---- Batched jobs (continuous batching) ------------------------------------------
 * No counterpart in the reference, whose EM drivers run one job at a time
 * (OsemSinogram3d.cpp:36-325, Mlem2d.cpp:60-260).  A batch advances many OSEM / MLEM
 * jobs of different geometries together: each apirl_batch_step enqueues one round,
 * one sub-iteration of every active job, as seven grouped launches (clear, fused
 * projection + ratio, axial setup, gather backprojection, update) replayed from a
 * captured hipGraph.  Jobs may be submitted between steps while others are running.
 * A job's image has the same bits as apirl_em_reconstruct with APIRL_BP_GATHER and
 * APIRL_PREC_REFERENCE gives it alone; the likelihood is not evaluated.  The batch
 * runs on its own stream; apirl_batch_submit orders the job after the work already
 * enqueued on the context's stream, and the image is final once apirl_batch_job_done
 * reports it. */
typedef struct apirl_batch apirl_batch;
typedef struct apirl_job_desc {
  apirl_plan* plan;
  int iterations, subsets;
  int osem_rules;           /* 1: OSEM update rules; 0: MLEM thresholds per subset */
  const float* d_meas;      /* full sinogram */
  const float* d_init;      /* initial image */
  const float* d_additive;  /* additive sinogram (full) or NULL */
  const float* d_mult;      /* multiplicative sinogram (sensitivity only) or NULL */
  float* d_out;             /* image, updated in place every sub-iteration */
  int additive_own_angles;  /* as apirl_em_options.additive_own_angles */
} apirl_job_desc;
/* max_sinograms bounds the 3D jobs' sinogram count (LDS of the gather launch). */
apirl_status apirl_batch_create(apirl_ctx* ctx, int max_jobs, int max_sinograms,
                                apirl_batch** out);
apirl_status apirl_batch_destroy(apirl_batch* batch);
const char* apirl_batch_last_error(const apirl_batch* batch);
/* Device memory a job of this plan holds while admitted (sensitivities, ratio,
 * backprojection, gather records), for admission by HBM footprint. */
size_t apirl_batch_job_bytes(const apirl_plan* plan, int subsets);
/* The backprojector a batch runs a job of this plan with: APIRL_BP_TILE for 3D jobs of
 * 2^21 voxels or more with one LOR per bin that the tile form takes (they run with the
 * single-job kernels after each grouped round), else APIRL_BP_GATHER.  A batched job's
 * image equals apirl_em_reconstruct's with that bp_mode, bit for bit. */
int apirl_batch_job_backprojector(const apirl_plan* plan);
/* Workgroups per grouped launch (0 keeps the current value) and graph replay on/off. */
apirl_status apirl_batch_set_launch(apirl_batch* batch, int grid, int use_graph);
/* Scheduling of the grouped launches (no reference counterpart: the batch engine is
 * this library's own): `grab` virtual blocks (256 rays, 256 voxels or one gather
 * tile each) per draw from a phase's work counter (1..64, default 2), and the order
 * in which the jobs' blocks are laid out in a round.  Results do not depend on
 * either; serving.autotune picks them by timing live rounds. */
#define APIRL_BATCH_ORDER_SUBMISSION 0
#define APIRL_BATCH_ORDER_LARGEST_FIRST 1
apirl_status apirl_batch_set_schedule(apirl_batch* batch, int grab, int order);
apirl_status apirl_batch_submit(apirl_batch* batch, const apirl_job_desc* job, int* job_id);
/* Enqueues one round; *active = jobs still to be scheduled after it. */
apirl_status apirl_batch_step(apirl_batch* batch, int* active);
apirl_status apirl_batch_job_done(apirl_batch* batch, int job_id, int* done);
apirl_status apirl_batch_synchronize(apirl_batch* batch);
/* Frees a finished job's device state for reuse by later jobs. */
apirl_status apirl_batch_release(apirl_batch* batch, int job_id);

/* ---- host helpers (no device work) --------------------------------------------- */
/* Reference grids (Sinogram2DinCylindrical3Dpet.cpp:46-60). */
void apirl_sino_grid(int nang, int nr, float rfov_mm, float* ang_deg, float* r_mm);
/* Ring-pair enumeration and mean ring coordinates of a cylindrical 3D sinogram
 * (Sinogram3D.cpp:358-367,396-464; SiddonProjector.cpp:357-360).  Returns the
 * number of sinograms, or -1 if a segment count does not match nsino_per_seg. */
int apirl_ring_table(int nrings, float zfov_mm, int nseg, const int* nsino_per_seg,
                     const int* min_rd, const int* max_rd, float* z1_mean, float* z2_mean);

#ifdef __cplusplus
}
#endif
#endif
