/*
 * apirl_oracle.h -- CPU restatement of APIRL's Siddon projector / backprojector
 * and MLEM / OSEM update, used ONLY as test infrastructure (the checker the HIP
 * path is compared with, and the "port" CPU baseline).  The product never links
 * it.  Every function follows the reference file:line named beside it, including
 * its floating-point operation order and its quirks (dropped exit voxel,
 * zero-length entry segments, upper-/lower-only bounds checks).
 *
 * Conventions shared by all functions:
 *   image  : float[nz][ny][nx], index z*nx*ny + y*nx + x        (data/src/Images.cpp:34-56)
 *   2D sino: float[nang][nr]                                      (data/inc/Sinogram2D.h:214)
 *   3D sino: float[nsino][nang][nr], sinograms segment-major      (Sinogram3D.cpp:396-464)
 *   angles : degrees, r : mm (grids built by orc_sino_grid)
 *
 * Precision: ORC_F32 functions reproduce the reference's float arithmetic bit for
 * bit; the *_f64 variants run the same rules in double and serve as the FP64
 * oracle that errors are reported against.
 */
#ifndef APIRL_ORACLE_H
#define APIRL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_image {
  int nx, ny, nz;
  float dx, dy, dz;
} orc_image;

/* One Siddon segment as the reference's SiddonSegment (recon/inc/Siddon.h:37-43). */
typedef struct orc_seg {
  int i, j, k;
  float w;
} orc_seg;

/* Sinogram grids: ang[i] = i * (180/nang), r[j] = dr/2 + j*dr - rfov, dr = 2*rfov/nr
 * (data/src/Sinogram2DinCylindrical3Dpet.cpp:46-60, Sinogram2D.cpp:190-198). */
void orc_sino_grid(int nang, int nr, float rfov, float* ang_deg, float* r_mm);

/* LOR endpoints of bin (ang, r) on a cylinder of radius rscanner
 * (Sinogram2DinCylindrical3Dpet.cpp:249-266).  Returns 0 when r > rfov (2D rule);
 * pass rfov = +inf for the 3D form (:216-229), which never rejects. */
int orc_lor_points(float ang_deg, float r, float rscanner, float rfov, float* p1x, float* p1y,
                   float* p2x, float* p2y);

/* Ring-pair enumeration of a cylindrical 3D sinogram (Sinogram3D.cpp:358-367,
 * 396-464) and the mean axial coordinates the projector uses
 * (SiddonProjector.cpp:349-361,394-395).  Writes, per sinogram, the final P1.Z and
 * P2.Z of its LORs into za/zb (nsino_total entries) and returns nsino_total, or -1
 * if a segment's count does not match nsino[seg]. */
int orc_ring_table(int nrings, float zfov, int nseg, const int* nsino, const int* minrd,
                   const int* maxrd, float* za, float* zb);

/* Siddon traversals (recon/src/Siddon.cpp:358-643 for 2D, :13-344 for 3D).
 * Emit up to cap segments into out; return the reference's list length. */
int orc_siddon2d(float p0x, float p0y, float vx, float vy, const orc_image* im, orc_seg* out,
                 int cap);
int orc_siddon3d(float p0x, float p0y, float p0z, float vx, float vy, float vz,
                 const orc_image* im, orc_seg* out, int cap);

/* Projector, 2D (SiddonProjector.cpp:146-193) and 3D (:322-441).
 * 2D leaves bins whose LOR is rejected untouched; 3D overwrites every bin. */
void orc_project2d(const orc_image* im, const float* img, int nang, const float* ang_deg, int nr,
                   const float* r_mm, float rfov, float rscanner, float* sino);
void orc_project3d(const orc_image* im, const float* img, int nang, const float* ang_deg, int nr,
                   const float* r_mm, float rscanner, int nsino, const float* za,
                   const float* zb, float* sino);

/* Backprojectors accumulate into img in the reference's single-thread order
 * (SiddonProjector.cpp:34-87 for 2D, :444-568 for 3D). */
void orc_mmr_lor_table(int nang, int nr, float* pts);
/* The LORs of the reference's mMR SUBSET sinograms (Sinogram3DSiemensMmr::getSubset,
 * Sinogram3DSiemensMmr.cpp:150-195): cylindrical (Sinogram2DinCylindrical3Dpet.cpp:
 * 216-229) on the radius 328 mm at the mMR's angles and arc-corrected r -- the subset
 * is built with cylindrical 2D sinograms; [nang][nr][4] for every full-grid angle. */
void orc_mmr_subset_lor_table(int nang, int nr, float* pts);
void orc_project3d_pts(const orc_image* im, const float* img, int nang, int nr, const float* pts,
                       int nsino, const float* za, const float* zb, float* sino);
void orc_backproject3d_pts(const orc_image* im, const float* sino, int nang, int nr,
                           const float* pts, int nsino, const float* za, const float* zb,
                           float* img);
void orc_subsample_r(int nr, const float* r, int ns, float* r_sub);
int orc_ring_table_os(int nrings, float zfov, int nseg, const int* nsino, const int* minrd,
                      const int* maxrd, int nas, float* za, float* zb);
void orc_project3d_os(const orc_image* im, const float* img, int nang, const float* ang_deg,
                      int nr, const float* r_sub, int ns, float rscanner, int nsino,
                      const float* za, const float* zb, int nas, float* sino);
void orc_backproject3d_os(const orc_image* im, const float* sino, int nang,
                          const float* ang_deg, int nr, const float* r_sub, int ns,
                          float rscanner, int nsino, const float* za, const float* zb, int nas,
                          float* img);
void orc_divide_backproject2d(const orc_image* im, const float* meas, const float* est,
                              int nang, const float* ang_deg, int nr, const float* r_mm,
                              float rfov, float rscanner, float* img);
void orc_backproject2d(const orc_image* im, const float* sino, int nang, const float* ang_deg,
                       int nr, const float* r_mm, float rfov, float rscanner, float* img);
void orc_backproject3d(const orc_image* im, const float* sino, int nang, const float* ang_deg,
                       int nr, const float* r_mm, float rscanner, int nsino, const float* za,
                       const float* zb, float* img);
/* DivideAndBackproject 3D (SiddonProjector.cpp:571-704), single-thread order. */
void orc_divide_backproject3d(const orc_image* im, const float* meas, const float* est,
                              int nang, const float* ang_deg, int nr, const float* r_mm,
                              float rscanner, int nsino, const float* za, const float* zb,
                              float* img);

/* Bin-wise sinogram operations (Sinogram3D.cpp:830-870, Sinogram2D.cpp:234-266). */
void orc_inverse_divide(float* est, const float* meas, long n);
void orc_add(float* est, const float* add, long n);
/* Likelihood sum y*log(yhat) - yhat (Sinogram2D.cpp:565-581 uses yhat>0,
 * Sinogram3D.cpp:584-607 uses yhat!=0); positive_only selects the 2D rule. */
float orc_likelihood(const float* est, const float* meas, long n, int positive_only);

/* Image helpers (Images.cpp:383-431, recon/src/Mlem.cpp:39-42). */
int orc_force_positive(float* img, long n);
float orc_min_nonzero(const float* img, long n);
float orc_max(const float* img, long n);
float orc_mlem_threshold(const float* sens, long n);
/* EM update x := cond ? x*b/s : 0, cond = (s > thr) if strict else (s >= thr)
 * (Mlem2d.cpp:171-185, MlemSinogram3d.cpp:228-242, OsemSinogram3d.cpp:217-237). */
void orc_em_update(float* x, const float* b, const float* s, long n, float thr, int strict);

/* Angle subset s of S: angles {s + S*k}, count floor(P/S) + (P mod S > s)
 * (Sinogram3DCylindricalPet.cpp:164-197).  Writes the subset's full-angle indices. */
int orc_subset_angles(int nang, int nsubsets, int subset, int* idx);

/* Drivers.  Each follows the reference loop exactly; outputs the final image,
 * the sensitivity image(s) and the likelihood trace. */
int orc_mlem2d(const orc_image* im, int nang, int nr, float rscanner, const float* meas,
               const float* init, int iters, float* img, float* sens, float* lik);
int orc_mlem3d(const orc_image* im, int nang, int nr, float rfov, float rscanner, int nsino,
               const float* za, const float* zb, const float* meas, const float* init,
               int iters, float* img, float* sens, float* lik);
/* MlemSinogram3d over LORs given by their endpoints per (angle, r) bin (as
 * orc_project3d_pts), e.g. the Siemens mMR's. */
int orc_mlem3d_pts(const orc_image* im, int nang, int nr, const float* pts, int nsino,
                   const float* za, const float* zb, const float* meas, const float* init,
                   int iters, float* img, float* sens, float* lik);
int orc_osem3d(const orc_image* im, int nang, int nr, float rfov, float rscanner, int nsino,
               const float* za, const float* zb, const float* meas, const float* init,
               int nsubsets, int iters, float* img, float* sens_all, float* lik);
/* orc_osem3d with the corrections of OsemSinogram3d.cpp:101-117, 172-173, 327-346:
 * add (full additive sinogram, may be NULL) added to each subset projection without a
 * clamp, subset bin (q, k, r) taking additive bin (q, k, r) when add_local (the
 * reference's Sinogram3D::addBinToBin on a subset, Sinogram3D.cpp:857-873) or its own
 * full angle otherwise; mult (full multiplicative sinogram, may be NULL) backprojected
 * subset by subset as the sensitivity; the final likelihood without the additive term. */
int orc_osem3d_corr(const orc_image* im, int nang, int nr, float rfov, float rscanner,
                    int nsino, const float* za, const float* zb, const float* meas,
                    const float* init, int nsubsets, int iters, const float* add,
                    const float* mult, int add_local, float* img, float* sens_all,
                    float* lik);
/* orc_osem3d_corr over LORs given by endpoints: subset bins take pts_sub (e.g.
 * orc_mmr_lor_table_subsets), the final full projection pts_full. */
int orc_osem3d_pts_corr(const orc_image* im, int nang, int nr, const float* pts_sub,
                        const float* pts_full, int nsino, const float* za, const float* zb,
                        const float* meas, const float* init, int nsubsets, int iters,
                        const float* add, const float* mult, int add_local, float* img,
                        float* sens_all, float* lik);
/* orc_mlem3d_pts with MlemSinogram3d's corrections (additive sinogram in every
 * projection but the final one, multiplicative sinogram as the sensitivity). */
int orc_mlem3d_pts_corr(const orc_image* im, int nang, int nr, const float* pts, int nsino,
                        const float* za, const float* zb, const float* meas, const float* init,
                        int iters, const float* add, const float* mult, float* img, float* sens,
                        float* lik);

/* FP64 oracle: the reference's 3D Siddon (recon/src/Siddon.cpp:13-344) carried out
 * in double from the same float LORs, with double accumulation (used to report the
 * FP64 error of every result, not for parity). */
/* The FP64 oracle's list for one LOR (P1 = p, P2 = p + v): voxels and weights of the
 * double traversal; returns the list length. */
int orc_siddon3d_f64(const float* p, const float* v, const orc_image* im, orc_seg* segs,
                     double* w, int cap);
void orc_project3d_f64(const orc_image* im, const double* img, int nang, const float* ang_deg,
                       int nr, const float* r_mm, float rscanner, int nsino, const float* za,
                       const float* zb, double* sino);
void orc_backproject3d_f64(const orc_image* im, const double* sino, int nang,
                           const float* ang_deg, int nr, const float* r_mm, float rscanner,
                           int nsino, const float* za, const float* zb, double* img);

/* FP64 oracle projection of selected bins (s_idx, a_idx, r_idx index the full
 * sinogram's sinogram, angle and radial bin). */
void orc_project3d_bins_f64(const orc_image* im, const double* img, const float* ang_deg,
                            const float* r_mm, float rscanner, const float* za, const float* zb,
                            long nbins, const int* s_idx, const int* a_idx, const int* r_idx,
                            double* out);
/* OSEM of orc_osem3d with the FP64 oracle's system matrix, all in double. */
int orc_osem3d_f64(const orc_image* im, int nang, int nr, float rfov, float rscanner, int nsino,
                   const float* za, const float* zb, const float* meas, const double* init,
                   int nsubsets, int iters, double* img);

#ifdef __cplusplus
}
#endif
#endif
