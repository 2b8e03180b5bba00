/*
 * apirl_oracle.c -- CPU restatement of APIRL's Siddon MLEM/OSEM path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared with and the
 * "port" CPU baseline.  Nothing in the shipped library calls this file.
 *
 * Written from the reference's algorithm (citations below), single-threaded, in
 * the reference's loop order, so that with the same inputs it yields the same
 * float results as the reference built with OMP_NUM_THREADS=1.  Build with
 * -ffp-contract=off (oracle/Makefile): a fused multiply-add would change which
 * side of a voxel face a rounded coordinate lands on.
 *
 * Out-of-array reads: the reference's 3D projector checks only upper bounds and
 * its 2D projector only lower bounds (SiddonProjector.cpp:178,417), so an entry
 * voxel index of -1 or n can address memory outside the image.  Such a segment
 * always has a zero or near-zero weight.  Here, as in the HIP path, a read whose
 * linear index falls outside [0, nx*ny*nz) contributes w*0; reads that land inside
 * the array (a neighbouring row or slice) are reproduced as the reference does.
 */
#define _GNU_SOURCE
#include "apirl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DEG2RAD 0.01745329251994329576f /* utils/inc/Geometry.h:12 */

static inline float fminref(float a, float b) { return (b < a) ? b : a; } /* std::min */
static inline float fmaxref(float a, float b) { return (a < b) ? b : a; } /* std::max */

void orc_sino_grid(int nang, int nr, float rfov, float* ang_deg, float* r_mm) {
  /* Sinogram2DinCylindrical3Dpet.cpp:46-60; maxAng_deg is the int 180. */
  float rinc = (2 * rfov) / nr;
  float phinc = (float)180 / nang;
  for (int i = 0; i < nang; i++) ang_deg[i] = i * phinc;
  for (int j = 0; j < nr; j++) r_mm[j] = rinc / 2 + j * rinc - rfov;
}

int orc_lor_points(float ang_deg, float r, float rscanner, float rfov, float* p1x, float* p1y,
                   float* p2x, float* p2y) {
  /* Sinogram2DinCylindrical3Dpet.cpp:249-266 (float sincosf / sqrtf as compiled). */
  float rad = ang_deg * DEG2RAD;
  float aux = sqrtf(rscanner * rscanner - r * r);
  if (r > rfov) return 0;
  float s, c;
  sincosf(rad, &s, &c);
  *p1x = r * c + s * aux;
  *p1y = r * s - c * aux;
  *p2x = r * c - s * aux;
  *p2y = r * s + c * aux;
  return 1;
}

int orc_ring_table(int nrings, float zfov, int nseg, const int* nsino, const int* minrd,
                   const int* maxrd, float* za, float* zb) {
  /* Ring grid Sinogram3D.cpp:358-367; enumeration :396-464; mean z
   * SiddonProjector.cpp:357-360; final P.Z with the axial-sample offset :394-395. */
  float* zring = (float*)malloc(sizeof(float) * nrings);
  float zinc = (float)zfov / nrings;
  for (int i = 0; i < nrings; i++) zring[i] = zinc / 2 + i * zinc;
  float w = zfov / nrings;          /* getAxialFoV_mm()/getNumRings() */
  float dzs = w / 1;                /* deltaZ_oversampling, one axial sample */
  int total = 0;
  for (int s = 0; s < nseg; s++) {
    int count = 0;
    for (int z1 = 0; z1 < 2 * nrings; z1++) {
      int n = 0, first1 = 0, first2 = 0, last1 = 0, last2 = 0;
      int z1a = z1;
      for (int z2 = 0; z2 < nrings; z2++) {
        if ((z1a - z2) <= maxrd[s] && (z1a - z2) >= minrd[s] && z1a >= 0 && z2 >= 0 &&
            z1a < nrings && z2 < nrings) {
          if (n == 0) { first1 = z1a; first2 = z2; }
          last1 = z1a; last2 = z2;
          n++;
        }
        z1a--;
      }
      if (n > 0) {
        float z1m = (zring[first1] + zring[last1]) / 2;
        float z2m = (zring[first2] + zring[last2]) / 2;
        za[total] = z1m - w / 2 + dzs / 2 + 0 * dzs;
        zb[total] = z2m - w / 2 + dzs / 2 + 0 * dzs;
        total++;
        count++;
      }
    }
    if (count != nsino[s]) { free(zring); return -1; }
  }
  free(zring);
  return total;
}

/* ---------------------------------------------------------------- Siddon */

int orc_siddon2d(float p0x, float p0y, float vx, float vy, const orc_image* im, orc_seg* out,
                 int cap) {
  /* recon/src/Siddon.cpp:358-643 */
  const float dx = im->dx, dy = im->dy;
  const float R = (dx * im->nx) / 2; /* Images.cpp:55 */
  float amin, amax;
  if (vx == 0) {
    float a1 = (-R - p0y) / vy, a2 = (R - p0y) / vy;
    if (a1 < a2) { amin = a1; amax = a2; } else { amin = a2; amax = a1; }
  } else if (vy == 0) {
    float a1 = (-R - p0x) / vx, a2 = (R - p0x) / vx;
    if (a1 < a2) { amin = a1; amax = a2; } else { amin = a2; amax = a1; }
  } else {
    float ax1 = (-R - p0x) / vx, ay1 = (-R - p0y) / vy;
    float ax2 = (R - p0x) / vx, ay2 = (R - p0y) / vy;
    amin = fmaxref(fminref(ax1, ax2), fminref(ay1, ay2));
    amax = fminref(fmaxref(ax1, ax2), fmaxref(ay1, ay2));
  }
  if (amin < 0 || amin > 1) amin = 0;
  if (amax > 1 || amax < 0) amax = 1;
  float x1 = p0x + vx * amin, y1 = p0y + vy * amin;
  float x2 = p0x + vx * amax, y2 = p0y + vy * amax;
  int imin = (int)floorf((x1 + R) / dx), jmin = (int)floorf((y1 + R) / dy);
  int imax = (int)floorf((x2 + R) / dx), jmax = (int)floorf((y2 + R) / dy);
  if ((imin < 0 && imax < 0) || (jmin < 0 && jmax < 0) || (imin >= im->nx && imax >= im->nx) ||
      (jmin >= im->ny && jmax >= im->ny))
    return 0;
  int n = abs(imax - imin) + abs(jmax - jmin);
  int iinc = vx > 0 ? 1 : (vx < 0 ? -1 : 0);
  int jinc = vy > 0 ? 1 : (vy < 0 ? -1 : 0);
  float L = sqrtf(((p0x + vx) - p0x) * ((p0x + vx) - p0x) + ((p0y + vy) - p0y) * ((p0y + vy) - p0y));
  float axu = fabsf(dx / vx), ayu = fabsf(dy / vy);
  float ax, ay;
  if (vx > 0) ax = (-R + (imin + iinc) * dx - p0x) / vx;
  else if (vx < 0) ax = (-R + imin * dx - p0x) / vx;
  else ax = FLT_MAX;
  if (ax < 0) ax = amin;
  if (vy > 0) ay = (-R + (jmin + jinc) * dy - p0y) / vy;
  else if (vy < 0) ay = (-R + jmin * dy - p0y) / vy;
  else ay = FLT_MAX;
  if (ay < 0) ay = amin;
  float ac = amin;
  int i = imin, j = jmin;
  for (int m = 0; m < n; m++) {
    float w;
    int ci = i, cj = j;
    if (ax <= ay) {
      w = (ax - ac) * L * 1.0f;
      i += iinc; ac = ax; ax += axu;
    } else {
      w = (ay - ac) * L * 1.0f;
      j += jinc; ac = ay; ay += ayu;
    }
    if (m < cap) { out[m].i = ci; out[m].j = cj; out[m].k = 0; out[m].w = w; }
  }
  return n;
}

int orc_siddon3d(float p0x, float p0y, float p0z, float vx, float vy, float vz,
                 const orc_image* im, orc_seg* out, int cap) {
  /* recon/src/Siddon.cpp:13-344 */
  const float dx = im->dx, dy = im->dy, dz = im->dz;
  const float R = (dx * im->nx) / 2;
  float amin, amax;
  if (vx == 0) {
    float a1 = (-R - p0y) / vy, a2 = (R - p0y) / vy;
    if (a1 < a2) { amin = a1; amax = a2; } else { amin = a2; amax = a1; }
  } else if (vy == 0) {
    float a1 = (-R - p0x) / vx, a2 = (R - p0x) / vx;
    if (a1 < a2) { amin = a1; amax = a2; } else { amin = a2; amax = a1; }
  } else {
    float ax1 = (-R - p0x) / vx, ay1 = (-R - p0y) / vy;
    float ax2 = (R - p0x) / vx, ay2 = (R - p0y) / vy;
    amin = fmaxref(fminref(ax1, ax2), fminref(ay1, ay2));
    amax = fminref(fmaxref(ax1, ax2), fmaxref(ay1, ay2));
  }
  if (amin < 0 || amin > 1) amin = 0;
  if (amax > 1 || amax < 0) amax = 1;
  float x1 = p0x + vx * amin, y1 = p0y + vy * amin, z1 = p0z + vz * amin;
  float x2 = p0x + vx * amax, y2 = p0y + vy * amax, z2 = p0z + vz * amax;
  float offz = dz / 2; /* :162 */
  int imin = (int)floorf((x1 + R) / dx), jmin = (int)floorf((y1 + R) / dy);
  int kmin = (int)floorf((z1 - offz) / dz);
  int imax = (int)floorf((x2 + R) / dx), jmax = (int)floorf((y2 + R) / dy);
  int kmax = (int)floorf((z2 - offz) / dz);
  int n = abs(imax - imin) + abs(jmax - jmin) + abs(kmax - kmin); /* :220 */
  int iinc = vx > 0 ? 1 : (vx < 0 ? -1 : 0);
  int jinc = vy > 0 ? 1 : (vy < 0 ? -1 : 0);
  int kinc = vz > 0 ? 1 : (vz < 0 ? -1 : 0);
  float L = sqrtf(((p0x + vx) - p0x) * ((p0x + vx) - p0x) + ((p0y + vy) - p0y) * ((p0y + vy) - p0y) +
                  ((p0z + vz) - p0z) * ((p0z + vz) - p0z));
  float axu = fabsf(dx / vx), ayu = fabsf(dy / vy), azu = fabsf(dz / vz);
  float ax, ay, az;
  /* The reference's "alpha < 0" guards are statements without effect (:267-286). */
  if (vx > 0) ax = (-R + (imin + iinc) * dx - p0x) / vx;
  else if (vx < 0) ax = (-R + imin * dx - p0x) / vx;
  else ax = FLT_MAX;
  if (vy > 0) ay = (-R + (jmin + jinc) * dy - p0y) / vy;
  else if (vy < 0) ay = (-R + jmin * dy - p0y) / vy;
  else ay = FLT_MAX;
  if (vz > 0) az = (offz + (kmin + kinc) * dz - p0z) / vz;
  else if (vz < 0) az = (offz + kmin * dz - p0z) / vz;
  else az = FLT_MAX;
  float ac = amin;
  int i = imin, j = jmin, k = kmin;
  for (int m = 0; m < n; m++) {
    float w;
    int ci = i, cj = j, ck = k;
    if (ax <= ay && ax <= az) {
      w = (ax - ac) * L;
      i += iinc; ac = ax; ax += axu;
    } else if (ay <= ax && ay <= az) {
      w = (ay - ac) * L;
      j += jinc; ac = ay; ay += ayu;
    } else {
      w = (az - ac) * L;
      k += kinc; ac = az; az += azu;
    }
    if (m < cap) { out[m].i = ci; out[m].j = cj; out[m].k = ck; out[m].w = w; }
  }
  return n;
}

/* ------------------------------------------------------------ projectors */

#define SEG_CAP 8192

static float pix_at(const float* img, long npix, long idx) {
  return (idx >= 0 && idx < npix) ? img[idx] : 0.0f;
}

void orc_project2d(const orc_image* im, const float* img, int nang, const float* ang_deg, int nr,
                   const float* r_mm, float rfov, float rscanner, float* sino) {
  /* SiddonProjector.cpp:146-193 */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  long npix = (long)im->nx * im->ny;
  for (int a = 0; a < nang; a++)
    for (int r = 0; r < nr; r++) {
      float p1x, p1y, p2x, p2y;
      if (!orc_lor_points(ang_deg[a], r_mm[r], rscanner, rfov, &p1x, &p1y, &p2x, &p2y)) continue;
      int n = orc_siddon2d(p1x, p1y, p2x - p1x, p2y - p1y, im, segs, SEG_CAP);
      float acc = 0;
      for (int l = 0; l < n; l++)
        if (segs[l].j >= 0 && segs[l].i >= 0)
          acc += 1.0f * segs[l].w * pix_at(img, npix, (long)segs[l].j * im->nx + segs[l].i);
      sino[(long)a * nr + r] = acc;
    }
  free(segs);
}

void orc_backproject2d(const orc_image* im, const float* sino, int nang, const float* ang_deg,
                       int nr, const float* r_mm, float rfov, float rscanner, float* img) {
  /* SiddonProjector.cpp:34-87 */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  for (int a = 0; a < nang; a++)
    for (int r = 0; r < nr; r++) {
      float p1x, p1y, p2x, p2y;
      if (!orc_lor_points(ang_deg[a], r_mm[r], rscanner, rfov, &p1x, &p1y, &p2x, &p2y)) continue;
      int n = orc_siddon2d(p1x, p1y, p2x - p1x, p2y - p1y, im, segs, SEG_CAP);
      float b = sino[(long)a * nr + r];
      for (int l = 0; l < n; l++)
        if (segs[l].j >= 0 && segs[l].i >= 0 && segs[l].j < im->ny && segs[l].i < im->nx)
          img[(long)segs[l].j * im->nx + segs[l].i] += segs[l].w * b * 1.0f;
    }
  free(segs);
}

void orc_divide_backproject2d(const orc_image* im, const float* meas, const float* est,
                              int nang, const float* ang_deg, int nr, const float* r_mm,
                              float rfov, float rscanner, float* img) {
  /* SiddonProjector.cpp:90-144: bins with est == 0 or meas == 0 skipped (:109-110);
   * term = ((geom * w) * m) / e with geom = 1 (:125); full bounds check (:121-122). */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  for (int a = 0; a < nang; a++)
    for (int r = 0; r < nr; r++) {
      float m = meas[(long)a * nr + r], e = est[(long)a * nr + r];
      if (e == 0 || m == 0) continue;
      float p1x, p1y, p2x, p2y;
      if (!orc_lor_points(ang_deg[a], r_mm[r], rscanner, rfov, &p1x, &p1y, &p2x, &p2y)) continue;
      int n = orc_siddon2d(p1x, p1y, p2x - p1x, p2y - p1y, im, segs, SEG_CAP);
      for (int l = 0; l < n; l++)
        if (segs[l].j >= 0 && segs[l].i >= 0 && segs[l].j < im->ny && segs[l].i < im->nx)
          img[(long)segs[l].j * im->nx + segs[l].i] += 1.0f * segs[l].w * m / e;
    }
  free(segs);
}

/* LOR of 3D bin (sinogram s, angle a, r): xy from the 2D form, z from za/zb
 * (SiddonProjector.cpp:385-402). */
static void lor3d(const float* ang_deg, const float* r_mm, float rscanner, const float* za,
                  const float* zb, int s, int a, int r, float* p, float* v) {
  float p1x, p1y, p2x, p2y;
  orc_lor_points(ang_deg[a], r_mm[r], rscanner, INFINITY, &p1x, &p1y, &p2x, &p2y);
  p[0] = p1x; p[1] = p1y; p[2] = za[s];
  v[0] = p2x - p1x; v[1] = p2y - p1y; v[2] = zb[s] - za[s];
}

void orc_project3d(const orc_image* im, const float* img, int nang, const float* ang_deg, int nr,
                   const float* r_mm, float rscanner, int nsino, const float* za,
                   const float* zb, float* sino) {
  /* SiddonProjector.cpp:322-441: FillConstant(0), then += over the list with the
   * upper-bound-only check of :417. */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  long nxy = (long)im->nx * im->ny, npix = nxy * im->nz;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        float p[3], v[3];
        lor3d(ang_deg, r_mm, rscanner, za, zb, s, a, r, p, v);
        int n = orc_siddon3d(p[0], p[1], p[2], v[0], v[1], v[2], im, segs, SEG_CAP);
        float acc = 0;
        for (int l = 0; l < n; l++)
          if (segs[l].i < im->nx && segs[l].j < im->ny && segs[l].k < im->nz)
            acc += segs[l].w * 1.0f *
                   pix_at(img, npix, (long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i);
        sino[((long)s * nang + a) * nr + r] = acc;
      }
  free(segs);
}

void orc_backproject3d(const orc_image* im, const float* sino, int nang, const float* ang_deg,
                       int nr, const float* r_mm, float rscanner, int nsino, const float* za,
                       const float* zb, float* img) {
  /* SiddonProjector.cpp:444-568: zero bins skipped (:496), full bounds check (:541). */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  long nxy = (long)im->nx * im->ny;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        float b = sino[((long)s * nang + a) * nr + r];
        if (b == 0) continue;
        float p[3], v[3];
        lor3d(ang_deg, r_mm, rscanner, za, zb, s, a, r, p, v);
        int n = orc_siddon3d(p[0], p[1], p[2], v[0], v[1], v[2], im, segs, SEG_CAP);
        for (int l = 0; l < n; l++)
          if (segs[l].k < im->nz && segs[l].j < im->ny && segs[l].i < im->nx && segs[l].k >= 0 &&
              segs[l].j >= 0 && segs[l].i >= 0)
            img[(long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i] += segs[l].w * 1.0f * b;
      }
  free(segs);
}

/* ------------------------------------------------------- oversampled bins */

void orc_subsample_r(int nr, const float* r, int ns, float* r_sub) {
  /* Sub-LOR o of bin j: r - inc/2 + inc/ns*o + inc/(2 ns)
   * (Sinogram2DinCylindrical3Dpet.cpp:283-288), inc from the neighbouring bins
   * (Sinogram2D.cpp:172-187). */
  for (int j = 0; j < nr; j++) {
    float inc;
    if (j > 0 && j < nr - 1) inc = (r[j + 1] - r[j - 1]) / 2;
    else if (j == 0) inc = r[1] - r[0];
    else inc = r[nr - 1] - r[nr - 2];
    for (int o = 0; o < ns; o++)
      r_sub[(long)j * ns + o] = r[j] - (inc / 2) + inc / ns * o + (inc / (2 * ns));
  }
}

int orc_ring_table_os(int nrings, float zfov, int nseg, const int* nsino, const int* minrd,
                      const int* maxrd, int nas, float* za, float* zb) {
  /* As orc_ring_table, with nas axial samples per sinogram: P.Z = zmean - w/2 + dz/2
   * + p dz, dz = w / nas (SiddonProjector.cpp:335-336, 394-395). */
  float* zring = (float*)malloc(sizeof(float) * nrings);
  float zinc = (float)zfov / nrings;
  for (int i = 0; i < nrings; i++) zring[i] = zinc / 2 + i * zinc;
  float w = zfov / nrings;
  float dzs = w / nas;
  int total = 0;
  for (int s = 0; s < nseg; s++) {
    int count = 0;
    for (int z1 = 0; z1 < 2 * nrings; z1++) {
      int n = 0, first1 = 0, first2 = 0, last1 = 0, last2 = 0;
      int z1a = z1;
      for (int z2 = 0; z2 < nrings; z2++) {
        if ((z1a - z2) <= maxrd[s] && (z1a - z2) >= minrd[s] && z1a >= 0 && z2 >= 0 &&
            z1a < nrings && z2 < nrings) {
          if (n == 0) { first1 = z1a; first2 = z2; }
          last1 = z1a; last2 = z2;
          n++;
        }
        z1a--;
      }
      if (n > 0) {
        float z1m = (zring[first1] + zring[last1]) / 2;
        float z2m = (zring[first2] + zring[last2]) / 2;
        for (int q = 0; q < nas; q++) {
          za[(long)total * nas + q] = z1m - w / 2 + dzs / 2 + q * dzs;
          zb[(long)total * nas + q] = z2m - w / 2 + dzs / 2 + q * dzs;
        }
        total++;
        count++;
      }
    }
    if (count != nsino[s]) { free(zring); return -1; }
  }
  free(zring);
  return total;
}

/* LOR (o, q) of an oversampled bin: xy from the sub-LOR's r, z from the axial sample. */
static void lor3d_os(const float* ang_deg, const float* r_sub, int ns, float rscanner,
                     const float* za, const float* zb, int nas, int s, int a, int r, int o,
                     int q, float* p, float* v) {
  float p1x, p1y, p2x, p2y;
  orc_lor_points(ang_deg[a], r_sub[(long)r * ns + o], rscanner, INFINITY, &p1x, &p1y, &p2x, &p2y);
  p[0] = p1x; p[1] = p1y; p[2] = za[(long)s * nas + q];
  v[0] = p2x - p1x; v[1] = p2y - p1y; v[2] = zb[(long)s * nas + q] - za[(long)s * nas + q];
}

void orc_project3d_os(const orc_image* im, const float* img, int nang, const float* ang_deg,
                      int nr, const float* r_sub, int ns, float rscanner, int nsino,
                      const float* za, const float* zb, int nas, float* sino) {
  /* SiddonProjector.cpp:366-425: every LOR (o, q) of the bin adds its terms to the bin
   * (r_sub = the bins' own r when ns == 1, getPointsFromLor). */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  long nxy = (long)im->nx * im->ny, npix = nxy * im->nz;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        float acc = 0;
        for (int o = 0; o < ns; o++)
          for (int q = 0; q < nas; q++) {
            float p[3], v[3];
            lor3d_os(ang_deg, r_sub, ns, rscanner, za, zb, nas, s, a, r, o, q, p, v);
            int n = orc_siddon3d(p[0], p[1], p[2], v[0], v[1], v[2], im, segs, SEG_CAP);
            for (int l = 0; l < n; l++)
              if (segs[l].i < im->nx && segs[l].j < im->ny && segs[l].k < im->nz)
                acc += segs[l].w * 1.0f *
                       pix_at(img, npix,
                              (long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i);
          }
        sino[((long)s * nang + a) * nr + r] = acc;
      }
  free(segs);
}

void orc_backproject3d_os(const orc_image* im, const float* sino, int nang,
                          const float* ang_deg, int nr, const float* r_sub, int ns,
                          float rscanner, int nsino, const float* za, const float* zb, int nas,
                          float* img) {
  /* SiddonProjector.cpp:491-548: zero bins skipped, every LOR (o, q) of the bin. */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  long nxy = (long)im->nx * im->ny;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        float b = sino[((long)s * nang + a) * nr + r];
        if (b == 0) continue;
        for (int o = 0; o < ns; o++)
          for (int q = 0; q < nas; q++) {
            float p[3], v[3];
            lor3d_os(ang_deg, r_sub, ns, rscanner, za, zb, nas, s, a, r, o, q, p, v);
            int n = orc_siddon3d(p[0], p[1], p[2], v[0], v[1], v[2], im, segs, SEG_CAP);
            for (int l = 0; l < n; l++)
              if (segs[l].k < im->nz && segs[l].j < im->ny && segs[l].i < im->nx &&
                  segs[l].k >= 0 && segs[l].j >= 0 && segs[l].i >= 0)
                img[(long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i] +=
                    segs[l].w * 1.0f * b;
          }
      }
  free(segs);
}

/* ------------------------------------------------------------ mMR geometry */

/* The mMR's radial positions and angles (Sinogram2DinSiemensMmr.cpp:27-50): bin =
 * 4.0891/2 mm (scaled by 344/nr), lr = bin/2 + bin (j - (float)(nr/2)),
 * r = (Rs + DOI cos(lr/Rs)) sin(lr/Rs), phi_i = i 180/nang. */
static void mmr_grids(int nang, int nr, float* r, float* ang) {
  const float Rs = 328.0f, doi = 9.6f, bin0 = 4.0891f / 2.0f;
  float bin = nr == 344 ? bin0 : bin0 * 344 / nr;
  for (int j = 0; j < nr; j++) {
    float lr = bin / 2 + (bin * (j - (float)(nr / 2)));
    r[j] = (Rs + doi * cosf(lr / Rs)) * sinf(lr / Rs);
  }
  float inc = (float)180 / nang;
  for (int a = 0; a < nang; a++) ang[a] = a * inc;
}

void orc_mmr_lor_table(int nang, int nr, float* pts) {
  /* Sinogram2DinSiemensMmr.cpp:175-192: odd bins at phi - step/2 (step =
   * getAngValue(1) - getAngValue(0)), the LOR ending on Rs + DOI.
   * pts[(a nr + j) 4 ..] = x1, y1, x2, y2. */
  const float Re = 328.0f + 9.6f;
  float* r = (float*)malloc(sizeof(float) * nr);
  float* ang = (float*)malloc(sizeof(float) * nang);
  mmr_grids(nang, nr, r, ang);
  float step = nang > 1 ? ang[1] - ang[0] : 0.f;
  for (int a = 0; a < nang; a++)
    for (int j = 0; j < nr; j++) {
      float phi = (ang[a] - (j % 2) * step / 2) * 0.01745329251994329576f;
      float aux = sqrtf(Re * Re - r[j] * r[j]);
      float* q = pts + ((long)a * nr + j) * 4;
      q[0] = r[j] * cosf(phi) + sinf(phi) * aux;
      q[1] = r[j] * sinf(phi) - cosf(phi) * aux;
      q[2] = r[j] * cosf(phi) - sinf(phi) * aux;
      q[3] = r[j] * sinf(phi) + cosf(phi) * aux;
    }
  free(r);
  free(ang);
}

void orc_mmr_subset_lor_table(int nang, int nr, float* pts) {
  /* Sinogram3DSiemensMmr::getSubset (Sinogram3DSiemensMmr.cpp:150-195) constructs
   * the subset with Sinogram3DSiemensMmr(numProj, numR, ...), whose segments are made
   * by the Sinogram3DCylindricalPet base constructor: there the virtual
   * inicializarSegmentos is the base class's, so the subset's 2D sinograms are
   * Sinogram2DinCylindrical3Dpet on radioScannerMmr_mm = 328 mm.  The full sinogram's
   * angles and arc-corrected r are copied into them (:163-193), and their LORs are
   * the cylindrical 3D form (Sinogram2DinCylindrical3Dpet.cpp:216-229): no odd-bin
   * half step, no depth of interaction. */
  float* r = (float*)malloc(sizeof(float) * nr);
  float* ang = (float*)malloc(sizeof(float) * nang);
  mmr_grids(nang, nr, r, ang);
  for (int a = 0; a < nang; a++)
    for (int j = 0; j < nr; j++) {
      float* q = pts + ((long)a * nr + j) * 4;
      orc_lor_points(ang[a], r[j], 328.0f, INFINITY, q, q + 1, q + 2, q + 3);
    }
  free(r);
  free(ang);
}

void orc_project3d_pts(const orc_image* im, const float* img, int nang, int nr, const float* pts,
                       int nsino, const float* za, const float* zb, float* sino) {
  /* SiddonProjector.cpp:322-441 with the LOR endpoints of each bin given. */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  long nxy = (long)im->nx * im->ny, npix = nxy * im->nz;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        const float* q = pts + ((long)a * nr + r) * 4;
        int n = orc_siddon3d(q[0], q[1], za[s], q[2] - q[0], q[3] - q[1], zb[s] - za[s], im, segs,
                             SEG_CAP);
        float acc = 0;
        for (int l = 0; l < n; l++)
          if (segs[l].i < im->nx && segs[l].j < im->ny && segs[l].k < im->nz)
            acc += segs[l].w * 1.0f *
                   pix_at(img, npix, (long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i);
        sino[((long)s * nang + a) * nr + r] = acc;
      }
  free(segs);
}

void orc_backproject3d_pts(const orc_image* im, const float* sino, int nang, int nr,
                           const float* pts, int nsino, const float* za, const float* zb,
                           float* img) {
  /* SiddonProjector.cpp:444-568 with the LOR endpoints of each bin given. */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  long nxy = (long)im->nx * im->ny;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        float b = sino[((long)s * nang + a) * nr + r];
        if (b == 0) continue;
        const float* q = pts + ((long)a * nr + r) * 4;
        int n = orc_siddon3d(q[0], q[1], za[s], q[2] - q[0], q[3] - q[1], zb[s] - za[s], im, segs,
                             SEG_CAP);
        for (int l = 0; l < n; l++)
          if (segs[l].k < im->nz && segs[l].j < im->ny && segs[l].i < im->nx && segs[l].k >= 0 &&
              segs[l].j >= 0 && segs[l].i >= 0)
            img[(long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i] += segs[l].w * 1.0f * b;
      }
  free(segs);
}

void orc_divide_backproject3d(const orc_image* im, const float* meas, const float* est,
                              int nang, const float* ang_deg, int nr, const float* r_mm,
                              float rscanner, int nsino, const float* za, const float* zb,
                              float* img) {
  /* SiddonProjector.cpp:571-704: bins with meas == 0 skipped (:607); per segment
   * (w*m)/e if e != 0, w*m otherwise (:669-683); upper-bound-only check (:666),
   * writes outside the array dropped. */
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  long nxy = (long)im->nx * im->ny, npix = nxy * im->nz;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        long bin = ((long)s * nang + a) * nr + r;
        float m = meas[bin], e = est[bin];
        if (m == 0) continue;
        float p[3], v[3];
        lor3d(ang_deg, r_mm, rscanner, za, zb, s, a, r, p, v);
        int n = orc_siddon3d(p[0], p[1], p[2], v[0], v[1], v[2], im, segs, SEG_CAP);
        for (int l = 0; l < n; l++)
          if (segs[l].k < im->nz && segs[l].j < im->ny && segs[l].i < im->nx) {
            long idx = (long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i;
            float val = (e != 0) ? segs[l].w * 1.0f * m / e : segs[l].w * 1.0f * m;
            if (idx >= 0 && idx < npix) img[idx] += val;
          }
      }
  free(segs);
}

/* -------------------------------------------------------- bin/image ops */

void orc_inverse_divide(float* est, const float* meas, long n) {
  for (long i = 0; i < n; i++) {
    float num = meas[i], den = est[i];
    est[i] = (num != 0 && den != 0) ? num / den : 0;
  }
}

void orc_add(float* est, const float* add, long n) {
  for (long i = 0; i < n; i++) est[i] = est[i] + add[i];
}

float orc_likelihood(const float* est, const float* meas, long n, int positive_only) {
  float L = 0;
  for (long i = 0; i < n; i++) {
    float e = est[i];
    if (positive_only ? (e > 0) : (e != 0)) L += meas[i] * logf(e) - e;
  }
  return L;
}

int orc_force_positive(float* img, long n) {
  int c = 0;
  for (long i = 0; i < n; i++)
    if (img[i] < 0) { c++; img[i] = 0; }
  return c;
}

float orc_min_nonzero(const float* p, long n) {
  /* Images.cpp:397-419, including its second branch. */
  float m = FLT_MAX;
  for (long i = 0; i < n; i++) {
    if (p[i] != 0 && m != 0) {
      if (p[i] < m) m = p[i];
    } else if (p[i] != 0) {
      m = p[0];
    }
  }
  return m;
}

float orc_max(const float* p, long n) {
  float m = p[0];
  for (long i = 0; i < n; i++)
    if (p[i] > m) m = p[i];
  return m;
}

float orc_mlem_threshold(const float* sens, long n) {
  float mn = orc_min_nonzero(sens, n), mx = orc_max(sens, n);
  return mn + (mx - mn) * 0.0001f;
}

void orc_em_update(float* x, const float* b, const float* s, long n, float thr, int strict) {
  for (long k = 0; k < n; k++) {
    int ok = strict ? (s[k] > thr) : (s[k] >= thr);
    x[k] = ok ? x[k] * b[k] / s[k] : 0;
  }
}

int orc_subset_angles(int nang, int nsubsets, int subset, int* idx) {
  int n = (int)floorf((float)nang / (float)nsubsets);
  if ((nang % nsubsets) > subset) n++;
  for (int k = 0; k < n; k++) idx[k] = subset + nsubsets * k;
  return n;
}

/* --------------------------------------------------------------- drivers */

int orc_mlem2d(const orc_image* im, int nang, int nr, float rscanner, const float* meas,
               const float* init, int iters, float* img, float* sens, float* lik) {
  /* recon/src/Mlem2d.cpp:5-14,60-278: the sinogram's radial grid is reset to the
   * image FOV; sensitivity = backprojection of ones; threshold min+1e-4(max-min);
   * update with s > thr. */
  long npix = (long)im->nx * im->ny, nb = (long)nang * nr;
  float rfov = (im->dx * im->nx) / 2;
  float* ang = (float*)malloc(sizeof(float) * nang);
  float* r = (float*)malloc(sizeof(float) * nr);
  orc_sino_grid(nang, nr, rfov, ang, r);
  float* ones = (float*)malloc(sizeof(float) * nb);
  float* est = (float*)calloc(nb, sizeof(float));
  float* bp = (float*)malloc(sizeof(float) * npix);
  for (long i = 0; i < nb; i++) ones[i] = 1;
  memset(sens, 0, sizeof(float) * npix);
  orc_backproject2d(im, ones, nang, ang, nr, r, rfov, rscanner, sens);
  float thr = orc_mlem_threshold(sens, npix);
  memcpy(img, init, sizeof(float) * npix);
  for (int t = 0; t < iters; t++) {
    orc_project2d(im, img, nang, ang, nr, r, rfov, rscanner, est);
    lik[t] = orc_likelihood(est, meas, nb, 1);
    memset(bp, 0, sizeof(float) * npix);
    orc_inverse_divide(est, meas, nb);
    orc_backproject2d(im, est, nang, ang, nr, r, rfov, rscanner, bp);
    orc_em_update(img, bp, sens, npix, thr, 1);
  }
  /* The final likelihood is taken on the last estimated projection, which holds
   * the ratio by then (Mlem2d.cpp:227-228). */
  lik[iters] = orc_likelihood(est, meas, nb, 1);
  free(ang); free(r); free(ones); free(est); free(bp);
  return 0;
}

/* LORs of a 3D sinogram: the cylindrical scanner's (angle, r) grid at radius
 * rscanner, or explicit endpoints per (angle, r) bin (pts != NULL). */
typedef struct {
  int nang, nr;
  const float *ang, *r, *pts;
  float rscanner;
} lor_source;

static void project3d_src(const orc_image* im, const float* img, const lor_source* L, int nsino,
                          const float* za, const float* zb, float* sino) {
  if (L->pts)
    orc_project3d_pts(im, img, L->nang, L->nr, L->pts, nsino, za, zb, sino);
  else
    orc_project3d(im, img, L->nang, L->ang, L->nr, L->r, L->rscanner, nsino, za, zb, sino);
}

static void backproject3d_src(const orc_image* im, const float* sino, const lor_source* L,
                              int nsino, const float* za, const float* zb, float* img) {
  if (L->pts)
    orc_backproject3d_pts(im, sino, L->nang, L->nr, L->pts, nsino, za, zb, img);
  else
    orc_backproject3d(im, sino, L->nang, L->ang, L->nr, L->r, L->rscanner, nsino, za, zb, img);
}

static int mlem3d_core(const orc_image* im, const lor_source* L, int nsino, const float* za,
                       const float* zb, const float* meas, const float* init, int iters,
                       const float* add, const float* mult, float* img, float* sens,
                       float* lik) {
  /* recon/src/MlemSinogram3d.cpp:89-331: sensitivity = backprojection of ones, or of
   * the multiplicative sinogram when there is one (:312-331); the additive sinogram
   * added to every projection without a clamp (:190-191, Sinogram3D.cpp:857-873);
   * update with s >= thr (:209-219); one extra projection, without the additive
   * term, for the final likelihood (:277-278). */
  long npix = (long)im->nx * im->ny * im->nz, nb = (long)nsino * L->nang * L->nr;
  float* ones = (float*)malloc(sizeof(float) * nb);
  float* est = (float*)calloc(nb, sizeof(float));
  float* bp = (float*)malloc(sizeof(float) * npix);
  for (long i = 0; i < nb; i++) ones[i] = mult ? mult[i] : 1;
  memset(sens, 0, sizeof(float) * npix);
  backproject3d_src(im, ones, L, nsino, za, zb, sens);
  float thr = orc_mlem_threshold(sens, npix);
  memcpy(img, init, sizeof(float) * npix);
  for (int t = 0; t < iters; t++) {
    project3d_src(im, img, L, nsino, za, zb, est);
    if (add) orc_add(est, add, nb);
    lik[t] = orc_likelihood(est, meas, nb, 0);
    memset(bp, 0, sizeof(float) * npix);
    orc_inverse_divide(est, meas, nb);
    backproject3d_src(im, est, L, nsino, za, zb, bp);
    orc_em_update(img, bp, sens, npix, thr, 0);
  }
  project3d_src(im, img, L, nsino, za, zb, est);
  lik[iters] = orc_likelihood(est, meas, nb, 0);
  free(ones); free(est); free(bp);
  return 0;
}

int orc_mlem3d(const orc_image* im, int nang, int nr, float rfov, float rscanner, int nsino,
               const float* za, const float* zb, const float* meas, const float* init,
               int iters, float* img, float* sens, float* lik) {
  float* ang = (float*)malloc(sizeof(float) * nang);
  float* r = (float*)malloc(sizeof(float) * nr);
  orc_sino_grid(nang, nr, rfov, ang, r);
  lor_source L = {nang, nr, ang, r, NULL, rscanner};
  int rc = mlem3d_core(im, &L, nsino, za, zb, meas, init, iters, NULL, NULL, img, sens, lik);
  free(ang); free(r);
  return rc;
}

int orc_mlem3d_pts(const orc_image* im, int nang, int nr, const float* pts, int nsino,
                   const float* za, const float* zb, const float* meas, const float* init,
                   int iters, float* img, float* sens, float* lik) {
  lor_source L = {nang, nr, NULL, NULL, pts, 0.0f};
  return mlem3d_core(im, &L, nsino, za, zb, meas, init, iters, NULL, NULL, img, sens, lik);
}

int orc_mlem3d_pts_corr(const orc_image* im, int nang, int nr, const float* pts, int nsino,
                        const float* za, const float* zb, const float* meas, const float* init,
                        int iters, const float* add, const float* mult, float* img, float* sens,
                        float* lik) {
  lor_source L = {nang, nr, NULL, NULL, pts, 0.0f};
  return mlem3d_core(im, &L, nsino, za, zb, meas, init, iters, add, mult, img, sens, lik);
}

int orc_osem3d(const orc_image* im, int nang, int nr, float rfov, float rscanner, int nsino,
               const float* za, const float* zb, const float* meas, const float* init,
               int nsubsets, int iters, float* img, float* sens_all, float* lik) {
  return orc_osem3d_corr(im, nang, nr, rfov, rscanner, nsino, za, zb, meas, init, nsubsets,
                         iters, NULL, NULL, 1, img, sens_all, lik);
}

static int osem3d_core(const orc_image* im, const lor_source* Lsub, const lor_source* Lfull,
                       int nsino, const float* za, const float* zb, const float* meas,
                       const float* init, int nsubsets, int iters, const float* add,
                       const float* mult, int add_local, float* img, float* sens_all,
                       float* lik);

int orc_osem3d_corr(const orc_image* im, int nang, int nr, float rfov, float rscanner,
                    int nsino, const float* za, const float* zb, const float* meas,
                    const float* init, int nsubsets, int iters, const float* add,
                    const float* mult, int add_local, float* img, float* sens_all,
                    float* lik) {
  float* ang = (float*)malloc(sizeof(float) * nang);
  float* r = (float*)malloc(sizeof(float) * nr);
  orc_sino_grid(nang, nr, rfov, ang, r);
  lor_source L = {nang, nr, ang, r, NULL, rscanner};
  int rc = osem3d_core(im, &L, &L, nsino, za, zb, meas, init, nsubsets, iters, add, mult,
                       add_local, img, sens_all, lik);
  free(ang); free(r);
  return rc;
}

int orc_osem3d_pts_corr(const orc_image* im, int nang, int nr, const float* pts_sub,
                        const float* pts_full, int nsino, const float* za, const float* zb,
                        const float* meas, const float* init, int nsubsets, int iters,
                        const float* add, const float* mult, int add_local, float* img,
                        float* sens_all, float* lik) {
  lor_source Ls = {nang, nr, NULL, NULL, pts_sub, 0.0f};
  lor_source Lf = {nang, nr, NULL, NULL, pts_full, 0.0f};
  return osem3d_core(im, &Ls, &Lf, nsino, za, zb, meas, init, nsubsets, iters, add, mult,
                     add_local, img, sens_all, lik);
}

/* The LORs of subset s's angles idx[0..na-1] taken from L (grid or endpoint form). */
static void subset_source(const lor_source* L, const int* idx, int na, float* sang,
                          float* spts, lor_source* out) {
  *out = *L;
  out->nang = na;
  if (L->pts) {
    for (int k = 0; k < na; k++)
      memcpy(spts + (long)k * L->nr * 4, L->pts + (long)idx[k] * L->nr * 4,
             sizeof(float) * 4 * L->nr);
    out->pts = spts;
  } else {
    for (int k = 0; k < na; k++) sang[k] = L->ang[idx[k]];
    out->ang = sang;
  }
}

static int osem3d_core(const orc_image* im, const lor_source* Lsub, const lor_source* Lfull,
                       int nsino, const float* za, const float* zb, const float* meas,
                       const float* init, int nsubsets, int iters, const float* add,
                       const float* mult, int add_local, float* img, float* sens_all,
                       float* lik) {
  const int nang = Lsub->nang, nr = Lsub->nr;
  /* recon/src/OsemSinogram3d.cpp:24-346: per-subset sensitivity (backprojection of
   * ones, or of the multiplicative sinogram's subset, over the subset, then
   * forcePositive; :327-346, :101-117), threshold 0 with >=, image updated after
   * every subset; the additive sinogram added to the subset's projection without a
   * clamp (:172-173) -- Sinogram3D::addBinToBin (Sinogram3D.cpp:857-873) indexes
   * the FULL additive sinogram with the subset's own angle index k, so subset bin
   * (q, k, r) gets additive bin (q, k, r) (add_local = 1; add_local = 0 takes the
   * bin's own full angle instead); final likelihood from a full projection (over
   * Lfull's LORs) without the additive term (:291-293).  Subset bins take Lsub's
   * LORs: the grid's, or for the mMR the subset's own (orc_mmr_lor_table_subsets). */
  long npix = (long)im->nx * im->ny * im->nz, nb = (long)nsino * nang * nr;
  int* idx = (int*)malloc(sizeof(int) * nang);
  float* sang = (float*)malloc(sizeof(float) * nang);
  float* spts = (float*)malloc(sizeof(float) * 4 * (size_t)nang * nr);
  float* ysub = (float*)malloc(sizeof(float) * nb);
  float* est = (float*)malloc(sizeof(float) * nb);
  float* bp = (float*)malloc(sizeof(float) * npix);
  lor_source Ls;
  for (int s = 0; s < nsubsets; s++) {
    int na = orc_subset_angles(nang, nsubsets, s, idx);
    subset_source(Lsub, idx, na, sang, spts, &Ls);
    long nbs = (long)nsino * na * nr;
    if (mult) {
      for (int q = 0; q < nsino; q++)
        for (int k = 0; k < na; k++)
          memcpy(ysub + ((long)q * na + k) * nr, mult + ((long)q * nang + idx[k]) * nr,
                 sizeof(float) * nr);
    } else {
      for (long i = 0; i < nbs; i++) ysub[i] = 1;
    }
    float* sens = sens_all + (long)s * npix;
    memset(sens, 0, sizeof(float) * npix);
    backproject3d_src(im, ysub, &Ls, nsino, za, zb, sens);
    orc_force_positive(sens, npix);
  }
  memcpy(img, init, sizeof(float) * npix);
  for (int t = 0; t < iters; t++)
    for (int s = 0; s < nsubsets; s++) {
      int na = orc_subset_angles(nang, nsubsets, s, idx);
      subset_source(Lsub, idx, na, sang, spts, &Ls);
      long nbs = (long)nsino * na * nr;
      for (int q = 0; q < nsino; q++)
        for (int k = 0; k < na; k++)
          memcpy(ysub + ((long)q * na + k) * nr, meas + ((long)q * nang + idx[k]) * nr,
                 sizeof(float) * nr);
      project3d_src(im, img, &Ls, nsino, za, zb, est);
      if (add)
        for (int q = 0; q < nsino; q++)
          for (int k = 0; k < na; k++) {
            float* e = est + ((long)q * na + k) * nr;
            const float* ad = add + ((long)q * nang + (add_local ? k : idx[k])) * nr;
            for (int j = 0; j < nr; j++) e[j] = e[j] + ad[j];
          }
      lik[t * nsubsets + s] = orc_likelihood(est, ysub, nbs, 0);
      memset(bp, 0, sizeof(float) * npix);
      orc_inverse_divide(est, ysub, nbs);
      backproject3d_src(im, est, &Ls, nsino, za, zb, bp);
      orc_em_update(img, bp, sens_all + (long)s * npix, npix, 0.0f, 0);
    }
  project3d_src(im, img, Lfull, nsino, za, zb, est);
  lik[iters * nsubsets] = orc_likelihood(est, meas, nb, 0);
  free(idx); free(sang); free(spts); free(ysub); free(est); free(bp);
  return 0;
}

/* ------------------------------------------------------------ FP64 oracle */

/* The FP64 oracle: recon/src/Siddon.cpp:13-344 (orc_siddon3d's statements) carried
 * out in double from the same float LOR -- alpha range, entry / exit voxels, list
 * length (with the dropped exit voxel), first crossings, accumulated crossings and
 * weights.  Away from ties this visits orc_siddon3d's voxels; on a LOR that lies
 * on a voxel face it decides the side in double. */
static int siddon3d_d(const orc_image* im, const float* pf, const float* vf, orc_seg* out,
                      double* wout, int cap) {
  const double p0x = pf[0], p0y = pf[1], p0z = pf[2], vx = vf[0], vy = vf[1], vz = vf[2];
  const double dx = im->dx, dy = im->dy, dz = im->dz;
  const double R = (dx * im->nx) / 2;
  double amin, amax;
  if (vx == 0) {
    double a1 = (-R - p0y) / vy, a2 = (R - p0y) / vy;
    if (a1 < a2) { amin = a1; amax = a2; } else { amin = a2; amax = a1; }
  } else if (vy == 0) {
    double a1 = (-R - p0x) / vx, a2 = (R - p0x) / vx;
    if (a1 < a2) { amin = a1; amax = a2; } else { amin = a2; amax = a1; }
  } else {
    double ax1 = (-R - p0x) / vx, ay1 = (-R - p0y) / vy;
    double ax2 = (R - p0x) / vx, ay2 = (R - p0y) / vy;
    double lo1 = ax1 < ax2 ? ax1 : ax2, lo2 = ay1 < ay2 ? ay1 : ay2;
    double hi1 = ax1 > ax2 ? ax1 : ax2, hi2 = ay1 > ay2 ? ay1 : ay2;
    amin = lo1 > lo2 ? lo1 : lo2;
    amax = hi1 < hi2 ? hi1 : hi2;
  }
  if (amin < 0 || amin > 1) amin = 0;
  if (amax > 1 || amax < 0) amax = 1;
  double x1 = p0x + vx * amin, y1 = p0y + vy * amin, z1 = p0z + vz * amin;
  double x2 = p0x + vx * amax, y2 = p0y + vy * amax, z2 = p0z + vz * amax;
  double offz = dz / 2;
  int imin = (int)floor((x1 + R) / dx), jmin = (int)floor((y1 + R) / dy);
  int kmin = (int)floor((z1 - offz) / dz);
  int imax = (int)floor((x2 + R) / dx), jmax = (int)floor((y2 + R) / dy);
  int kmax = (int)floor((z2 - offz) / dz);
  int n = abs(imax - imin) + abs(jmax - jmin) + abs(kmax - kmin);
  int iinc = vx > 0 ? 1 : (vx < 0 ? -1 : 0);
  int jinc = vy > 0 ? 1 : (vy < 0 ? -1 : 0);
  int kinc = vz > 0 ? 1 : (vz < 0 ? -1 : 0);
  double L = sqrt(vx * vx + vy * vy + vz * vz);
  double axu = fabs(dx / vx), ayu = fabs(dy / vy), azu = fabs(dz / vz);
  double ax, ay, az;
  if (vx > 0) ax = (-R + (imin + iinc) * dx - p0x) / vx;
  else if (vx < 0) ax = (-R + imin * dx - p0x) / vx;
  else ax = DBL_MAX;
  if (vy > 0) ay = (-R + (jmin + jinc) * dy - p0y) / vy;
  else if (vy < 0) ay = (-R + jmin * dy - p0y) / vy;
  else ay = DBL_MAX;
  if (vz > 0) az = (offz + (kmin + kinc) * dz - p0z) / vz;
  else if (vz < 0) az = (offz + kmin * dz - p0z) / vz;
  else az = DBL_MAX;
  double ac = amin;
  int i = imin, j = jmin, k = kmin;
  for (int m = 0; m < n; m++) {
    double w;
    int ci = i, cj = j, ck = k;
    if (ax <= ay && ax <= az) {
      w = (ax - ac) * L;
      i += iinc; ac = ax; ax += axu;
    } else if (ay <= ax && ay <= az) {
      w = (ay - ac) * L;
      j += jinc; ac = ay; ay += ayu;
    } else {
      w = (az - ac) * L;
      k += kinc; ac = az; az += azu;
    }
    if (m < cap) { out[m].i = ci; out[m].j = cj; out[m].k = ck; out[m].w = (float)w; wout[m] = w; }
  }
  return n;
}

static int f64_weights(const orc_image* im, const float* p, const float* v, orc_seg* segs,
                       double* wexact) {
  int n = siddon3d_d(im, p, v, segs, wexact, SEG_CAP);
  return n > SEG_CAP ? SEG_CAP : n;
}

int orc_siddon3d_f64(const float* p, const float* v, const orc_image* im, orc_seg* segs,
                     double* w, int cap) {
  orc_seg* all = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  double* wa = (double*)malloc(sizeof(double) * SEG_CAP);
  int n = f64_weights(im, p, v, all, wa);
  for (int l = 0; l < n && l < cap; l++) {
    segs[l] = all[l];
    w[l] = wa[l];
  }
  free(all);
  free(wa);
  return n;
}

void orc_project3d_f64(const orc_image* im, const double* img, int nang, const float* ang_deg,
                       int nr, const float* r_mm, float rscanner, int nsino, const float* za,
                       const float* zb, double* sino) {
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  double* w = (double*)malloc(sizeof(double) * SEG_CAP);
  long nxy = (long)im->nx * im->ny;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        float p[3], v[3];
        lor3d(ang_deg, r_mm, rscanner, za, zb, s, a, r, p, v);
        int n = f64_weights(im, p, v, segs, w);
        double acc = 0;
        for (int l = 0; l < n; l++)
          if (segs[l].i >= 0 && segs[l].j >= 0 && segs[l].k >= 0 && segs[l].i < im->nx &&
              segs[l].j < im->ny && segs[l].k < im->nz)
            acc += w[l] * img[(long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i];
        sino[((long)s * nang + a) * nr + r] = acc;
      }
  free(segs); free(w);
}

void orc_backproject3d_f64(const orc_image* im, const double* sino, int nang,
                           const float* ang_deg, int nr, const float* r_mm, float rscanner,
                           int nsino, const float* za, const float* zb, double* img) {
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  double* w = (double*)malloc(sizeof(double) * SEG_CAP);
  long nxy = (long)im->nx * im->ny;
  for (int s = 0; s < nsino; s++)
    for (int a = 0; a < nang; a++)
      for (int r = 0; r < nr; r++) {
        double b = sino[((long)s * nang + a) * nr + r];
        if (b == 0) continue;
        float p[3], v[3];
        lor3d(ang_deg, r_mm, rscanner, za, zb, s, a, r, p, v);
        int n = f64_weights(im, p, v, segs, w);
        for (int l = 0; l < n; l++)
          if (segs[l].i >= 0 && segs[l].j >= 0 && segs[l].k >= 0 && segs[l].i < im->nx &&
              segs[l].j < im->ny && segs[l].k < im->nz)
            img[(long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i] += w[l] * b;
      }
  free(segs); free(w);
}

void orc_project3d_bins_f64(const orc_image* im, const double* img, const float* ang_deg,
                            const float* r_mm, float rscanner, const float* za, const float* zb,
                            long nbins, const int* s_idx, const int* a_idx, const int* r_idx,
                            double* out) {
  orc_seg* segs = (orc_seg*)malloc(sizeof(orc_seg) * SEG_CAP);
  double* w = (double*)malloc(sizeof(double) * SEG_CAP);
  long nxy = (long)im->nx * im->ny;
  for (long q = 0; q < nbins; q++) {
    float p[3], v[3];
    lor3d(ang_deg, r_mm, rscanner, za, zb, s_idx[q], a_idx[q], r_idx[q], p, v);
    int n = f64_weights(im, p, v, segs, w);
    double acc = 0;
    for (int l = 0; l < n; l++)
      if (segs[l].i >= 0 && segs[l].j >= 0 && segs[l].k >= 0 && segs[l].i < im->nx &&
          segs[l].j < im->ny && segs[l].k < im->nz)
        acc += w[l] * img[(long)segs[l].k * nxy + (long)segs[l].j * im->nx + segs[l].i];
    out[q] = acc;
  }
  free(segs);
  free(w);
}

int orc_osem3d_f64(const orc_image* im, int nang, int nr, float rfov, float rscanner, int nsino,
                   const float* za, const float* zb, const float* meas, const double* init,
                   int nsubsets, int iters, double* img) {
  /* orc_osem3d's loop (OsemSinogram3d.cpp:24-346) with the FP64 oracle's system
   * matrix and every image, sinogram and ratio held in double. */
  long npix = (long)im->nx * im->ny * im->nz, nb = (long)nsino * nang * nr;
  float* ang = (float*)malloc(sizeof(float) * nang);
  float* r = (float*)malloc(sizeof(float) * nr);
  orc_sino_grid(nang, nr, rfov, ang, r);
  int* idx = (int*)malloc(sizeof(int) * nang);
  float* sang = (float*)malloc(sizeof(float) * nang);
  double* ysub = (double*)malloc(sizeof(double) * nb);
  double* est = (double*)malloc(sizeof(double) * nb);
  double* bp = (double*)malloc(sizeof(double) * npix);
  double* sens = (double*)malloc(sizeof(double) * npix * nsubsets);
  for (int s = 0; s < nsubsets; s++) {
    int na = orc_subset_angles(nang, nsubsets, s, idx);
    for (int k = 0; k < na; k++) sang[k] = ang[idx[k]];
    long nbs = (long)nsino * na * nr;
    for (long i = 0; i < nbs; i++) ysub[i] = 1;
    double* sn = sens + (long)s * npix;
    memset(sn, 0, sizeof(double) * npix);
    orc_backproject3d_f64(im, ysub, na, sang, nr, r, rscanner, nsino, za, zb, sn);
  }
  memcpy(img, init, sizeof(double) * npix);
  for (int t = 0; t < iters; t++)
    for (int s = 0; s < nsubsets; s++) {
      int na = orc_subset_angles(nang, nsubsets, s, idx);
      for (int k = 0; k < na; k++) sang[k] = ang[idx[k]];
      long nbs = (long)nsino * na * nr;
      for (int q = 0; q < nsino; q++)
        for (int k = 0; k < na; k++)
          for (int j = 0; j < nr; j++)
            ysub[((long)q * na + k) * nr + j] = meas[((long)q * nang + idx[k]) * nr + j];
      orc_project3d_f64(im, img, na, sang, nr, r, rscanner, nsino, za, zb, est);
      for (long i = 0; i < nbs; i++)
        est[i] = (ysub[i] != 0 && est[i] != 0) ? ysub[i] / est[i] : 0;
      memset(bp, 0, sizeof(double) * npix);
      orc_backproject3d_f64(im, est, na, sang, nr, r, rscanner, nsino, za, zb, bp);
      const double* sn = sens + (long)s * npix;
      for (long i = 0; i < npix; i++) img[i] = (sn[i] >= 0) ? img[i] * bp[i] / sn[i] : 0;
    }
  free(ang); free(r); free(idx); free(sang); free(ysub); free(est); free(bp); free(sens);
  return 0;
}
